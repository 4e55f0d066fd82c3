/* ifem_hdf5.h -- result export to HDF5 in the reference's file layout.
 *
 * Replaces the file-writing half of HDF5Writer (src/Utility/HDF5Writer.C,
 * HDF5Base.C) for the GPU path: a solved model's basis (the GoTools text of
 * every patch) and its per-patch solution fields are written under
 *
 *   /<level>/<name>-1/basis/<patch>              char (H5T_NATIVE_CHAR)
 *   /<level>/<name>-1/fields/<field>/<patch>     double
 *   /<level>/<name>-1/knotspan/<field>/<patch>   double (element values)
 *   /<level>/timeinfo/level                      double (the time)
 *   /log/<name>/0                                char
 *
 * with 1-based patch numbers, so that tools reading the reference's .hdf5
 * output read these files unchanged.  Plain C, linked against the system
 * HDF5 library.  Every call returns 0 on success and a negative code on
 * failure; ifem_h5_last_error() then describes it.
 */
#ifndef IFEM_HDF5_H
#define IFEM_HDF5_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ifem_h5_file ifem_h5_file;

/* Opens `path` for writing (HDF5Writer constructor + openFile,
 * src/Utility/HDF5Writer.C:37-105): append != 0 and an existing file keep
 * its contents, else the file is truncated.  Group /<level> is created. */
int ifem_h5_open(const char* path, int append, int level, ifem_h5_file** out);

/* Flushes and closes (HDF5Writer::closeFile, src/Utility/HDF5Writer.C:108-118). */
int ifem_h5_close(ifem_h5_file* f);

/* Time levels of an existing file: the count n of groups /0 .. /n-1
 * present in a row, which is what HDF5Writer::getLastTimeLevel returns
 * for a file opened to append (src/Utility/HDF5Writer.C:52-76); -1 for a
 * missing or unreadable file. */
int ifem_h5_last_level(const char* path);

/* Basis text of one patch: /<level>/<basis>/basis/<patch>, as
 * HDF5Writer::writeBasis (src/Utility/HDF5Writer.C:575-683). */
int ifem_h5_write_basis(ifem_h5_file* f, int level, const char* basis, int patch,
                        const char* text, size_t len);

/* Control-point field of one patch: /<level>/<basis>/fields/<field>/<patch>
 * (the primary-solution branch of HDF5Writer::writeSIM,
 * src/Utility/HDF5Writer.C:352-384), len values, node-major and
 * component-minor as SIMbase::extractPatchSolution returns them. */
int ifem_h5_write_field(ifem_h5_file* f, int level, const char* basis, const char* field,
                        int patch, const double* data, size_t len);

/* Element field of one patch: /<level>/<basis>/knotspan/<field>/<patch>
 * (HDF5Writer::writeKnotspan, src/Utility/HDF5Writer.C:538-572). */
int ifem_h5_write_knotspan(ifem_h5_file* f, int level, const char* basis, const char* field,
                           int patch, const double* data, size_t len);

/* /<level>/timeinfo/level = time (HDF5Writer::writeTimeInfo,
 * src/Utility/HDF5Writer.C:685-696). */
int ifem_h5_write_time(ifem_h5_file* f, int level, double time);

/* /log/<name>/0 = text (HDF5Writer::writeLog, src/Utility/HDF5Writer.C:736-756). */
int ifem_h5_write_log(ifem_h5_file* f, const char* name, const char* text, size_t len);

/* GoTools text of a non-rational trivariate spline volume (class 700), as
 * the reference writes a patch basis (ASMs3D::write, src/ASM/ASMs3D.C:163-175,
 * GoTools SplineVolume::write): header "700 1 0 0", "3 0", per direction
 * "n order" and the knots, then one control point per line, numbers in
 * %.15g, and an empty last line.  order = p+1, ncoef = n per direction,
 * coefs (n1 n2 n3 x 3, direction 1 fastest).  Writes at most cap bytes to
 * buf (not NUL-terminated) and the full length to *len; buf may be NULL to
 * query the length. */
int ifem_g2_format(const int order[3], const int ncoef[3], const double* const knots[3],
                   const double* coefs, char* buf, size_t cap, size_t* len);

/* Description of the last failure of the calling thread. */
const char* ifem_h5_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
