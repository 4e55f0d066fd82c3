/* hdf5_writer.c -- result export to HDF5 in the reference's layout
 * (declared in include/ifem_hdf5.h; layout: src/Utility/HDF5Writer.C).
 *
 * Groups are opened or created one path component at a time (addGroup /
 * getGroup, src/Utility/HDF5Writer.C:122-140); a dataset
 * that already exists is replaced, so a time level can be written again in
 * an appended file.  Numbers of the GoTools text use %.15g, the precision
 * of the reference's basis output. */
#define _POSIX_C_SOURCE 200809L
#include "ifem_hdf5.h"

#include <hdf5.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

struct ifem_h5_file {
  hid_t fid;
};

static __thread char g_err[512];

static int fail(int code, const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

const char* ifem_h5_last_error(void) { return g_err; }

/* H5Lexists without the library's error printout for missing links */
static int link_exists(hid_t loc, const char* name)
{
  htri_t r = -1;
  H5E_BEGIN_TRY { r = H5Lexists(loc, name, H5P_DEFAULT); } H5E_END_TRY;
  return r > 0;
}

/* Opens (creating as needed) the group at '/'-separated `path` below the
 * file root; returns its id or a negative value. */
static hid_t group_path(hid_t fid, const char* path)
{
  char buf[1024];
  if (strlen(path) >= sizeof buf) return -1;
  strcpy(buf, path);
  hid_t cur = H5Gopen2(fid, "/", H5P_DEFAULT);
  char* save = NULL;
  for (char* tok = strtok_r(buf, "/", &save); tok && cur >= 0; tok = strtok_r(NULL, "/", &save)) {
    hid_t next = link_exists(cur, tok) ? H5Gopen2(cur, tok, H5P_DEFAULT)
                                       : H5Gcreate2(cur, tok, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    H5Gclose(cur);
    cur = next;
  }
  return cur;
}

/* One-dimensional dataset `name` of len elements in `group` (the serial
 * case of HDF5Writer::writeArray, src/Utility/HDF5Writer.C:143-182). */
static int write_array(hid_t group, const char* name, size_t len, const void* data, hid_t type)
{
  if (link_exists(group, name) && H5Ldelete(group, name, H5P_DEFAULT) < 0)
    return fail(-3, "cannot replace dataset %s", name);
  hsize_t n = (hsize_t)len;
  hid_t space = H5Screate_simple(1, &n, NULL);
  if (space < 0) return fail(-3, "cannot create a dataspace of %zu elements", len);
  hid_t set = H5Dcreate2(group, name, type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  int rc = 0;
  if (set < 0)
    rc = fail(-3, "cannot create dataset %s", name);
  else if (len > 0 && H5Dwrite(set, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
    rc = fail(-3, "cannot write dataset %s", name);
  if (set >= 0) H5Dclose(set);
  H5Sclose(space);
  return rc;
}

/* dataset <group path>/<patch> */
static int write_patch_array(ifem_h5_file* f, const char* gpath, int patch, size_t len,
                             const void* data, hid_t type)
{
  if (!f || f->fid < 0) return fail(-1, "file not open");
  if (len > 0 && !data) return fail(-1, "null data");
  hid_t g = group_path(f->fid, gpath);
  if (g < 0) return fail(-3, "cannot open or create group %s", gpath);
  char name[32];
  snprintf(name, sizeof name, "%d", patch);
  int rc = write_array(g, name, len, data, type);
  H5Gclose(g);
  return rc;
}

static int check_name(const char* s, const char* what)
{
  if (!s || !*s) return fail(-1, "empty %s", what);
  if (strlen(s) > 400) return fail(-1, "%s too long", what);
  return 0;
}

int ifem_h5_open(const char* path, int append, int level, ifem_h5_file** out)
{
  if (!path || !out) return fail(-1, "null argument");
  *out = NULL;
  struct stat st;
  const int exists = stat(path, &st) == 0;
  hid_t fid;
  H5E_BEGIN_TRY {
    fid = (append && exists) ? H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT)
                             : H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  } H5E_END_TRY;
  if (fid < 0) return fail(-2, "failed to open %s", path);
  ifem_h5_file* f = (ifem_h5_file*)malloc(sizeof *f);
  if (!f) {
    H5Fclose(fid);
    return fail(-2, "out of memory");
  }
  f->fid = fid;
  char gp[32];
  snprintf(gp, sizeof gp, "%d", level);
  hid_t g = group_path(fid, gp);
  if (g < 0) {
    H5Fclose(fid);
    free(f);
    return fail(-3, "cannot create group /%d", level);
  }
  H5Gclose(g);
  *out = f;
  return 0;
}

int ifem_h5_close(ifem_h5_file* f)
{
  if (!f) return 0;
  int rc = 0;
  if (f->fid >= 0) {
    H5Fflush(f->fid, H5F_SCOPE_GLOBAL);
    if (H5Fclose(f->fid) < 0) rc = fail(-2, "close failed");
  }
  free(f);
  return rc;
}

int ifem_h5_last_level(const char* path)
{
  struct stat st;
  if (!path || stat(path, &st) != 0) return -1;
  hid_t fid;
  H5E_BEGIN_TRY { fid = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT); } H5E_END_TRY;
  if (fid < 0) return -1;
  /* as the reference counts: levels /0, /1, ... present in a row */
  int n = 0;
  char name[32];
  for (;;) {
    snprintf(name, sizeof name, "/%d", n);
    if (!link_exists(fid, name)) break;
    n++;
  }
  H5Fclose(fid);
  return n;
}

int ifem_h5_write_basis(ifem_h5_file* f, int level, const char* basis, int patch,
                        const char* text, size_t len)
{
  int rc = check_name(basis, "basis name");
  if (rc) return rc;
  char gp[512];
  snprintf(gp, sizeof gp, "%d/%s/basis", level, basis);
  return write_patch_array(f, gp, patch, len, text, H5T_NATIVE_CHAR);
}

int ifem_h5_write_field(ifem_h5_file* f, int level, const char* basis, const char* field,
                        int patch, const double* data, size_t len)
{
  int rc = check_name(basis, "basis name");
  if (!rc) rc = check_name(field, "field name");
  if (rc) return rc;
  char gp[1024];
  snprintf(gp, sizeof gp, "%d/%s/fields/%s", level, basis, field);
  return write_patch_array(f, gp, patch, len, data, H5T_NATIVE_DOUBLE);
}

int ifem_h5_write_knotspan(ifem_h5_file* f, int level, const char* basis, const char* field,
                           int patch, const double* data, size_t len)
{
  int rc = check_name(basis, "basis name");
  if (!rc) rc = check_name(field, "field name");
  if (rc) return rc;
  char gp[1024];
  snprintf(gp, sizeof gp, "%d/%s/knotspan/%s", level, basis, field);
  return write_patch_array(f, gp, patch, len, data, H5T_NATIVE_DOUBLE);
}

int ifem_h5_write_time(ifem_h5_file* f, int level, double time)
{
  if (!f || f->fid < 0) return fail(-1, "file not open");
  char gp[64];
  snprintf(gp, sizeof gp, "%d/timeinfo", level);
  hid_t g = group_path(f->fid, gp);
  if (g < 0) return fail(-3, "cannot create group %s", gp);
  int rc = write_array(g, "level", 1, &time, H5T_NATIVE_DOUBLE);
  H5Gclose(g);
  return rc;
}

int ifem_h5_write_log(ifem_h5_file* f, const char* name, const char* text, size_t len)
{
  int rc = check_name(name, "log name");
  if (rc) return rc;
  char gp[512];
  snprintf(gp, sizeof gp, "log/%s", name);
  return write_patch_array(f, gp, 0, len, text, H5T_NATIVE_CHAR);
}

/* growable text buffer */
typedef struct {
  char* p;
  size_t n, cap;
} Text;

static int put(Text* t, const char* fmt, ...)
{
  char tmp[160];
  va_list ap;
  va_start(ap, fmt);
  int k = vsnprintf(tmp, sizeof tmp, fmt, ap);
  va_end(ap);
  if (k < 0 || (size_t)k >= sizeof tmp) return -1;
  if (t->n + (size_t)k > t->cap) {
    size_t c = t->cap ? 2 * t->cap : 4096;
    while (c < t->n + (size_t)k) c *= 2;
    char* q = (char*)realloc(t->p, c);
    if (!q) return -1;
    t->p = q;
    t->cap = c;
  }
  memcpy(t->p + t->n, tmp, (size_t)k);
  t->n += (size_t)k;
  return 0;
}

int ifem_g2_format(const int order[3], const int ncoef[3], const double* const knots[3],
                   const double* coefs, char* buf, size_t cap, size_t* len)
{
  if (!order || !ncoef || !knots || !coefs || !len) return fail(-1, "null argument");
  for (int d = 0; d < 3; d++)
    if (order[d] < 1 || ncoef[d] < order[d] || !knots[d])
      return fail(-1, "invalid basis in direction %d", d);
  Text t = {NULL, 0, 0};
  int bad = put(&t, "700 1 0 0\n3 0\n");
  for (int d = 0; d < 3 && !bad; d++) {
    bad = put(&t, "%d %d\n", ncoef[d], order[d]);
    const int nk = ncoef[d] + order[d];
    for (int k = 0; k < nk && !bad; k++) bad = put(&t, k + 1 < nk ? "%.15g " : "%.15g\n", knots[d][k]);
  }
  const long ncp = (long)ncoef[0] * ncoef[1] * ncoef[2];
  for (long i = 0; i < ncp && !bad; i++)
    bad = put(&t, "%.15g %.15g %.15g\n", coefs[3 * i], coefs[3 * i + 1], coefs[3 * i + 2]);
  if (!bad) bad = put(&t, "\n");
  if (bad) {
    free(t.p);
    return fail(-2, "out of memory");
  }
  *len = t.n;
  if (buf) memcpy(buf, t.p, t.n < cap ? t.n : cap);
  free(t.p);
  return 0;
}
