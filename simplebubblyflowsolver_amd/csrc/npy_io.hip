// NPY directory format of the reference's initial conditions and save
// directories (src/fileio.c, src/array.c:190-316, src/domain.c:23-75,
// src/fluid/init.c, src/fluid/save.c, src/interface/init.c,
// src/interface/save.c), and the init / save entry points that use it.
//
// A field file holds the y-interior rows of an array including its x
// ghost columns, shape (Ny, Nx + nadds_x), C order, '<f8'; scalars have
// shape ().  Reading checks ndims, shape, dtype and order as the
// reference does (src/fileio.c:200-225).
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "domain.h"
#include "fluid_solver.h"
#include "interface_solver.h"
#include "sbfs_internal.hpp"

namespace sbfs {

int allocate_fluid(const domain_t* d, fluid_t* f);        // runtime.hip
int allocate_interface(const domain_t* d, interface_t* s); // runtime.hip
int build_metrics(domain_t* d);                           // runtime.hip

static std::string npy_path(const char* dir, const char* name) {
  std::string s(dir);
  if (!s.empty() && s.back() != '/') s += '/';
  return s + name + ".npy";
}

static int npy_read(const char* dir, const char* name, const char* descr,
                    const std::vector<size_t>& shape, size_t elem, void* out) {
  const std::string path = npy_path(dir, name);
  FILE* fp = fopen(path.c_str(), "rb");
  SBFS_REQUIRE(fp != nullptr, "%s: %s", path.c_str(), strerror(errno));
  unsigned char pre[10];
  if (fread(pre, 1, 8, fp) != 8 || memcmp(pre, "\x93NUMPY", 6) != 0) {
    fclose(fp);
    SBFS_REQUIRE(false, "%s: not an NPY file", path.c_str());
  }
  size_t hlen = 0;
  if (pre[6] == 1) {
    unsigned char b[2];
    if (fread(b, 1, 2, fp) != 2) { fclose(fp); SBFS_REQUIRE(false, "%s: truncated", path.c_str()); }
    hlen = (size_t)b[0] | ((size_t)b[1] << 8);
  } else {
    unsigned char b[4];
    if (fread(b, 1, 4, fp) != 4) { fclose(fp); SBFS_REQUIRE(false, "%s: truncated", path.c_str()); }
    hlen = (size_t)b[0] | ((size_t)b[1] << 8) | ((size_t)b[2] << 16) | ((size_t)b[3] << 24);
  }
  std::string h(hlen, '\0');
  if (fread(&h[0], 1, hlen, fp) != hlen) {
    fclose(fp);
    SBFS_REQUIRE(false, "%s: truncated header", path.c_str());
  }
  // descr
  const size_t pd = h.find("'descr'");
  const size_t pf = h.find("'fortran_order'");
  const size_t ps = h.find("'shape'");
  if (pd == std::string::npos || pf == std::string::npos || ps == std::string::npos) {
    fclose(fp);
    SBFS_REQUIRE(false, "%s: malformed header", path.c_str());
  }
  const size_t q0 = h.find('\'', pd + 7), q1 = h.find('\'', q0 + 1);
  const std::string d = h.substr(q0 + 1, q1 - q0 - 1);
  const bool fortran = h.compare(h.find(':', pf) + 1, 5, " True") == 0 ||
                       h.compare(h.find(':', pf) + 1, 4, "True") == 0;
  std::vector<size_t> sh;
  {
    size_t a = h.find('(', ps), b = h.find(')', a);
    std::string t = h.substr(a + 1, b - a - 1);
    size_t k = 0;
    while (k < t.size()) {
      while (k < t.size() && (t[k] == ' ' || t[k] == ',')) k++;
      if (k >= t.size()) break;
      sh.push_back((size_t)strtoull(t.c_str() + k, nullptr, 10));
      while (k < t.size() && t[k] != ',') k++;
    }
  }
  // the checks of src/fileio.c:200-225, in its order: ndims, shape, dtype,
  // storage order
  std::string bad;
  char msg[256];
  if (sh.size() != shape.size()) {
    snprintf(msg, sizeof msg, "ndims: %zu expected, %zu obtained", shape.size(), sh.size());
    bad = msg;
  }
  for (size_t n = 0; bad.empty() && n < shape.size(); n++)
    if (sh[n] != shape[n]) {
      snprintf(msg, sizeof msg, "shape[%zu]: %zu expected, %zu obtained", n, shape[n], sh[n]);
      bad = msg;
    }
  if (bad.empty() && d != descr) bad = std::string("dtype: '") + descr + "' expected, '" + d + "' obtained";
  if (bad.empty() && fortran) bad = "is_fortran_order: false expected, true obtained";
  if (!bad.empty()) {
    fclose(fp);
    SBFS_REQUIRE(false, "%s: %s", path.c_str(), bad.c_str());
  }
  size_t count = elem;
  for (size_t s : shape) count *= s;
  const size_t got = fread(out, 1, count, fp);
  fclose(fp);
  SBFS_REQUIRE(got == count, "%s: %zu of %zu data bytes", path.c_str(), got, count);
  return 0;
}

int npy_write(const char* dir, const char* name, const char* descr,
                     const std::vector<size_t>& shape, size_t elem, const void* data) {
  std::string sh = "(";
  for (size_t k = 0; k < shape.size(); k++) {
    sh += std::to_string(shape[k]);
    sh += (shape.size() == 1) ? "," : (k + 1 < shape.size() ? ", " : "");
  }
  sh += ")";
  std::string h = std::string("{'descr': '") + descr + "', 'fortran_order': False, 'shape': " + sh + ", }";
  const size_t total = 10 + h.size() + 1;
  const size_t pad = (64 - total % 64) % 64;
  h += std::string(pad, ' ');
  h += '\n';
  const std::string path = npy_path(dir, name);
  FILE* fp = fopen(path.c_str(), "wb");
  SBFS_REQUIRE(fp != nullptr, "%s: %s", path.c_str(), strerror(errno));
  const unsigned char pre[10] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0,
                                 (unsigned char)(h.size() & 0xff), (unsigned char)(h.size() >> 8)};
  size_t count = elem;
  for (size_t s : shape) count *= s;
  bool ok = fwrite(pre, 1, 10, fp) == 10 && fwrite(h.data(), 1, h.size(), fp) == h.size() &&
            fwrite(data, 1, count, fp) == count;
  fclose(fp);
  SBFS_REQUIRE(ok, "%s: write failed", path.c_str());
  return 0;
}

// field <-> device: the file rows are the y-interior rows, all x columns
static int load_field(const domain_t* d, const char* dir, const char* name, array_t* a) {
  const int xl = a->nadds[0][0], xh = a->nadds[0][1], yl = a->nadds[1][0];
  const size_t w = d->mysizes[0] + xl + xh, ny = d->mysizes[1];
  std::vector<double> buf(w * ny);
  if (npy_read(dir, name, "<f8", {ny, w}, sizeof(double), buf.data())) return 1;
  SBFS_HIP(hipMemcpy((char*)a->data + sizeof(double) * w * yl, buf.data(), sizeof(double) * w * ny,
                     hipMemcpyHostToDevice));
  return 0;
}

static int dump_field(const domain_t* d, const char* dir, const char* name, const array_t* a) {
  const int xl = a->nadds[0][0], xh = a->nadds[0][1], yl = a->nadds[1][0];
  const size_t w = d->mysizes[0] + xl + xh, ny = d->mysizes[1];
  std::vector<double> buf(w * ny);
  SBFS_HIP(hipStreamSynchronize(stream()));
  SBFS_HIP(hipMemcpy(buf.data(), (const char*)a->data + sizeof(double) * w * yl,
                     sizeof(double) * w * ny, hipMemcpyDeviceToHost));
  return npy_write(dir, name, "<f8", {ny, w}, sizeof(double), buf.data());
}

static int env_double(const char* name, double* v) {
  const char* s = getenv(name);
  SBFS_REQUIRE(s != nullptr, "%s not found", name);
  errno = 0;
  *v = strtod(s, nullptr);
  SBFS_REQUIRE(errno == 0, "%s: invalid value as double", name);
  return 0;
}

}  // namespace sbfs

using namespace sbfs;

extern "C" {

int domain_init(const char dirname_ic[], domain_t* d) {
  SBFS_REQUIRE(dirname_ic && d, "domain_init: null argument");
  memset(d, 0, sizeof(*d));
  size_t gl[2];
  if (npy_read(dirname_ic, "glsizes", "<u8", {2}, sizeof(size_t), gl)) return 1;
  if (npy_read(dirname_ic, "lengths", "<f8", {2}, sizeof(double), d->lengths)) return 1;
  SBFS_REQUIRE(gl[0] >= 2 && gl[1] >= 4, "domain_init: glsizes %zu x %zu too small", gl[0], gl[1]);
  for (int k = 0; k < 2; k++) {
    d->glsizes[k] = d->mysizes[k] = gl[k];
    d->offsets[k] = 0;
  }
  d->xf = (double*)calloc(gl[0] + 1, sizeof(double));
  d->xc = (double*)calloc(gl[0] + 2, sizeof(double));
  SBFS_REQUIRE(d->xf && d->xc, "host allocation failed");
  if (npy_read(dirname_ic, "xf", "<f8", {gl[0] + 1}, sizeof(double), d->xf)) return 1;
  if (npy_read(dirname_ic, "xc", "<f8", {gl[0] + 2}, sizeof(double), d->xc)) return 1;
  return build_metrics(d);
}

int domain_save(const char dirname[], const domain_t* d) {
  size_t gl[2] = {d->glsizes[0], d->glsizes[1]};
  if (npy_write(dirname, "glsizes", "<u8", {2}, sizeof(size_t), gl)) return 1;
  if (npy_write(dirname, "lengths", "<f8", {2}, sizeof(double), d->lengths)) return 1;
  if (npy_write(dirname, "xf", "<f8", {gl[0] + 1}, sizeof(double), d->xf)) return 1;
  if (npy_write(dirname, "xc", "<f8", {gl[0] + 2}, sizeof(double), d->xc)) return 1;
  return 0;
}

int fluid_init(const char dirname_ic[], const domain_t* d, fluid_t* f) {
  SBFS_REQUIRE(dirname_ic && d && f, "fluid_init: null argument");
  memset(f, 0, sizeof(*f));
  if (allocate_fluid(d, f)) return 1;
  if (load_field(d, dirname_ic, "ux", &f->ux)) return 1;
  if (load_field(d, dirname_ic, "uy", &f->uy)) return 1;
  if (load_field(d, dirname_ic, "p", &f->p)) return 1;
  if (load_field(d, dirname_ic, "psi", &f->psi[0])) return 1;
  if (fluid_update_boundaries_ux(d, &f->ux)) return 1;
  if (fluid_update_boundaries_uy(d, &f->uy)) return 1;
  if (fluid_update_boundaries_p(d, &f->p)) return 1;
  if (fluid_update_boundaries_psi(d, &f->psi[0])) return 1;
  if (env_double("Re", &f->Re)) return 1;
  if (env_double("Fr", &f->Fr)) return 1;
  if (env_double("denr", &f->denr)) return 1;
  if (env_double("visr", &f->visr)) return 1;
  f->refden = fmin(1., f->denr);
  SBFS_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int fluid_save(const char dirname[], const domain_t* d, const fluid_t* f) {
  if (npy_write(dirname, "Re", "<f8", {}, sizeof(double), &f->Re)) return 1;
  if (npy_write(dirname, "Fr", "<f8", {}, sizeof(double), &f->Fr)) return 1;
  if (dump_field(d, dirname, "ux", &f->ux)) return 1;
  if (dump_field(d, dirname, "uy", &f->uy)) return 1;
  if (dump_field(d, dirname, "p", &f->p)) return 1;
  if (dump_field(d, dirname, "psi", &f->psi[0])) return 1;
  return 0;
}

int interface_init(const char dirname_ic[], const domain_t* d, interface_t* s) {
  SBFS_REQUIRE(dirname_ic && d && s, "interface_init: null argument");
  memset(s, 0, sizeof(*s));
  if (allocate_interface(d, s)) return 1;
  if (load_field(d, dirname_ic, "vof", &s->vof)) return 1;
  if (interface_update_boundaries_vof(d, &s->vof)) return 1;
  if (env_double("We", &s->We)) return 1;
  SBFS_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int interface_save(const char dirname[], const domain_t* d, const interface_t* s) {
  if (npy_write(dirname, "We", "<f8", {}, sizeof(double), &s->We)) return 1;
  return dump_field(d, dirname, "vof", &s->vof);
}

int sbfs_read_step_time(const char dirname[], size_t* step, double* time) {
  SBFS_REQUIRE(dirname && step && time, "sbfs_read_step_time: null argument");
  if (npy_read(dirname, "step", "<u8", {}, sizeof(size_t), step)) return 1;
  return npy_read(dirname, "time", "<f8", {}, sizeof(double), time);
}

int sbfs_write_step_time(const char dirname[], size_t step, double time) {
  SBFS_REQUIRE(dirname, "sbfs_write_step_time: null argument");
  if (npy_write(dirname, "step", "<u8", {}, sizeof(size_t), &step)) return 1;
  return npy_write(dirname, "time", "<f8", {}, sizeof(double), &time);
}

}  // extern "C"
