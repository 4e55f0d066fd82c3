// HDF5Writer.h -- result export in the reference's HDF5 layout
// (src/Utility/HDF5Writer.h), over the C library declared in ifem_hdf5.h.
//
//   /<level>/<sim>-1/basis/<patch>            GoTools text of each patch
//   /<level>/<sim>-1/fields/<field>/<patch>   primary solution per patch
//   /<level>/<sim>-1/knotspan/<field>/<patch> element values
//   /<level>/timeinfo/level                   time
//   /log/<name>/0                             text
#pragma once

#include <string>
#include <vector>

#include "IFEMGpu.h"

struct ifem_h5_file;

namespace ifem_gpu {

//! Writes <name>.hdf5 (HDF5Writer, src/Utility/HDF5Writer.C:37-49): a file
//! not appended to is truncated when first opened; once closed, later
//! opens append (closeFile switches to read-write, :108-118).
class HDF5Writer {
public:
  explicit HDF5Writer(const std::string& name, bool append = false);
  ~HDF5Writer();
  HDF5Writer(const HDF5Writer&) = delete;
  HDF5Writer& operator=(const HDF5Writer&) = delete;

  //! -1 for a truncating writer, else the count of levels /0, /1, ...
  //! stored (getLastTimeLevel, src/Utility/HDF5Writer.C:52-76).
  int getLastTimeLevel();
  //! Opens the file and adds group /<level> (openFile, :79-105).
  void openFile(int level);
  //! Flushes and closes (closeFile, :108-118).
  void closeFile(int level);

  //! Bases and primary solutions of a single-basis model (writeSIM,
  //! :231-384): the bases at level 0 or when the geometry changed, then
  //! solution k of patch k (1-based in the file) under
  //! <simName>-1/fields/<integrand.getField1Name(11)>.
  bool writeSIM(int level, const std::string& simName, const std::vector<const ASMs3D*>& patches,
                const std::vector<std::vector<Real>>& patchSolutions, const Integrand& integrand,
                bool geometryUpdated = false);
  //! GoTools text of every patch under <name>/basis (writeBasis, :575-683).
  bool writeBasis(int level, const std::string& name, const std::vector<const ASMs3D*>& patches);
  //! Element field of patch `patch` (writeKnotspan, :538-572).
  bool writeKnotspan(int level, const std::string& basis, const std::string& field, int patch,
                     const std::vector<Real>& values);
  //! /<level>/timeinfo/level = time (writeTimeInfo, :685-696).
  bool writeTimeInfo(int level, double time);
  //! /log/<name>/0 = data (writeLog, :736-756).
  bool writeLog(const std::string& data, const std::string& name);

  const std::string& getName() const { return m_name; }
  //! Message of the last failure.
  std::string lastError() const;

private:
  std::string m_name;
  bool m_append;
  ifem_h5_file* m_file = nullptr;
};

} // namespace ifem_gpu
