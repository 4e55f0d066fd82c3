// HDF5Writer.C -- see HDF5Writer.h.
#include "HDF5Writer.h"

#include <sstream>
#include <sys/stat.h>

#include "ifem_hdf5.h"

namespace ifem_gpu {

HDF5Writer::HDF5Writer(const std::string& name, bool append) : m_name(name + ".hdf5")
{
  struct stat st;
  m_append = append && stat(m_name.c_str(), &st) == 0;
}

HDF5Writer::~HDF5Writer()
{
  if (m_file) ifem_h5_close(m_file);
}

int HDF5Writer::getLastTimeLevel()
{
  return m_append ? ifem_h5_last_level(m_name.c_str()) : -1;
}

void HDF5Writer::openFile(int level)
{
  if (m_file) return;
  if (ifem_h5_open(m_name.c_str(), m_append ? 1 : 0, level, &m_file)) m_file = nullptr;
}

void HDF5Writer::closeFile(int)
{
  if (m_file) ifem_h5_close(m_file);
  m_file = nullptr;
  m_append = true;
}

bool HDF5Writer::writeBasis(int level, const std::string& name,
                            const std::vector<const ASMs3D*>& patches)
{
  if (!m_file) return false;
  for (size_t k = 0; k < patches.size(); k++) {
    std::ostringstream os;
    if (!patches[k] || !patches[k]->write(os)) return false;
    const std::string text = os.str();
    if (ifem_h5_write_basis(m_file, level, name.c_str(), (int)k + 1, text.data(), text.size()))
      return false;
  }
  return true;
}

bool HDF5Writer::writeSIM(int level, const std::string& simName,
                          const std::vector<const ASMs3D*>& patches,
                          const std::vector<std::vector<Real>>& patchSolutions,
                          const Integrand& integrand, bool geometryUpdated)
{
  if (!m_file || patchSolutions.size() != patches.size()) return false;
  const std::string basis = simName + "-1";
  if ((level == 0 || geometryUpdated) && !this->writeBasis(level, basis, patches)) return false;
  const std::string field = integrand.getField1Name(11);
  for (size_t k = 0; k < patches.size(); k++) {
    const std::vector<Real>& v = patchSolutions[k];
    if (v.size() != (size_t)patches[k]->getNoNodes() * integrand.getNoFields()) return false;
    if (ifem_h5_write_field(m_file, level, basis.c_str(), field.c_str(), (int)k + 1, v.data(),
                            v.size()))
      return false;
  }
  return true;
}

bool HDF5Writer::writeKnotspan(int level, const std::string& basis, const std::string& field,
                               int patch, const std::vector<Real>& values)
{
  return m_file && !ifem_h5_write_knotspan(m_file, level, basis.c_str(), field.c_str(), patch,
                                           values.data(), values.size());
}

bool HDF5Writer::writeTimeInfo(int level, double time)
{
  return m_file && !ifem_h5_write_time(m_file, level, time);
}

bool HDF5Writer::writeLog(const std::string& data, const std::string& name)
{
  return m_file && !ifem_h5_write_log(m_file, name.c_str(), data.data(), data.size());
}

std::string HDF5Writer::lastError() const { return ifem_h5_last_error(); }

} // namespace ifem_gpu
