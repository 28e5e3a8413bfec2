// lceq_conf.hip -- host-side configuration files in the reference's on-disk
// format: the direct-access unformatted files of f/confsave.f:50-253 and
// f/confget.f:318-526 (oldconf).
//
// Links: one record per (mu, parity, z), mu outer, then parity (even first),
// then z; a record holds the global xy half-plane of that z in site order
// x/2 + (Lx/2) y, each link as (Re alpha, Im alpha, Re beta, Im beta).  The
// record length is 16 nc Lx Ly / 2 bytes (f/confsave.f:57), sized for 16-byte
// complex numbers.  Higgs: one record per (parity, z) with (a1, a2, a3) per
// site, record length 8 * 3 Lx Ly / 2 bytes (f/confsave.f:58).  Fortran
// direct access adds no record markers.  A build with 8-byte reals (REAL as
// REAL(8)) fills the records; the as-shipped build with 4-byte reals writes
// its values into the first half of each record (real_bytes = 4).
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/lceq.h"
#include "lceq_internal.hpp"

namespace {

struct ConfGeom {
  int lx, ly, lz;
  int64_t h2, n;  // sites of an xy half-plane, of a parity
  bool ok;
};

ConfGeom conf_geom(const int* dims) {
  ConfGeom c{0, 0, 0, 0, 0, false};
  if (!dims) return c;
  c.lx = dims[0];
  c.ly = dims[1];
  c.lz = dims[2];
  if (c.lx < 2 || c.ly < 2 || c.lz < 2 || (c.lx | c.ly | c.lz) & 1) return c;
  c.h2 = (int64_t)c.lx * c.ly / 2;
  c.n = c.h2 * c.lz;
  c.ok = true;
  return c;
}

// One field: nclass classes (links: mu*2 + parity; Higgs: parity) of ncomp
// components per site; host layout host[((cls * n + s) * ncomp + c].
int field_io(bool write, const ConfGeom& g, double* host, int ncomp, int nclass,
             const char* path, int rb, const char* what) {
  const int64_t recl = (int64_t)(ncomp == 4 ? 16 * 2 : 8 * 3) * g.h2;
  const int64_t data = (int64_t)ncomp * rb * g.h2;
  FILE* f = fopen(path, write ? "wb" : "rb");
  if (!f) return lceq::set_error(LCEQ_ERR_IO, "cannot open %s file %s", what, path);
  int rc = LCEQ_OK;
  if (!write) {
    fseek(f, 0, SEEK_END);
    long size = ftell(f);
    fseek(f, 0, SEEK_SET);
    int64_t need = (int64_t)(nclass * g.lz - 1) * recl + data;
    if (size < need) {
      fclose(f);
      return lceq::set_error(LCEQ_ERR_IO, "%s file %s has %ld bytes, a %dx%dx%d lattice needs %lld",
                             what, path, size, g.lx, g.ly, g.lz, (long long)need);
    }
  }
  std::vector<unsigned char> rec((size_t)recl, 0);
  for (int cls = 0; cls < nclass && rc == LCEQ_OK; ++cls)
    for (int z = 0; z < g.lz; ++z) {
      double* base = host + ((size_t)cls * g.n + (size_t)z * g.h2) * ncomp;
      if (write) {
        for (int64_t i = 0; i < g.h2 * ncomp; ++i) {
          if (rb == 8) {
            memcpy(&rec[(size_t)i * 8], &base[i], 8);
          } else {
            float v = (float)base[i];
            memcpy(&rec[(size_t)i * 4], &v, 4);
          }
        }
        if (fwrite(rec.data(), 1, (size_t)recl, f) != (size_t)recl)
          rc = lceq::set_error(LCEQ_ERR_IO, "short write to %s", path);
      } else {
        fseek(f, (long)(((int64_t)cls * g.lz + z) * recl), SEEK_SET);
        if (fread(rec.data(), 1, (size_t)data, f) != (size_t)data) {
          rc = lceq::set_error(LCEQ_ERR_IO, "short read from %s", path);
          break;
        }
        for (int64_t i = 0; i < g.h2 * ncomp; ++i) {
          if (rb == 8) {
            memcpy(&base[i], &rec[(size_t)i * 8], 8);
          } else {
            float v;
            memcpy(&v, &rec[(size_t)i * 4], 4);
            base[i] = v;
          }
        }
      }
    }
  if (fclose(f) != 0 && rc == LCEQ_OK) rc = lceq::set_error(LCEQ_ERR_IO, "closing %s failed", path);
  return rc;
}

int conf_io(bool write, const int* dims, double* u, double* a, const char* u_path,
            const char* a_path, int real_bytes) {
  ConfGeom g = conf_geom(dims);
  if (!g.ok) return lceq::set_error(LCEQ_ERR_ARG, "extents must be even and >= 2");
  if (real_bytes != 4 && real_bytes != 8) return lceq::set_error(LCEQ_ERR_ARG, "real_bytes is 4 or 8");
  int rc = LCEQ_OK;
  if (u_path) {
    if (!u) return lceq::set_error(LCEQ_ERR_ARG, "null link array");
    rc = field_io(write, g, u, 4, 6, u_path, real_bytes, "link");
  }
  if (rc == LCEQ_OK && a_path) {
    if (!a) return lceq::set_error(LCEQ_ERR_ARG, "null Higgs array");
    rc = field_io(write, g, a, 3, 2, a_path, real_bytes, "Higgs");
  }
  return rc;
}

// Fortran Iw.m edit of v into out (w characters, at least m digits, w
// asterisks when the value does not fit).
void fortran_iw_m(std::string& out, long v, int w, int m) {
  char digits[32];
  snprintf(digits, sizeof digits, "%0*ld", m, v < 0 ? -v : v);
  std::string s = (v < 0 ? "-" : "") + std::string(digits);
  if ((int)s.size() > w) s = std::string((size_t)w, '*');
  out += std::string((size_t)(w - (int)s.size()), ' ') + s;
}

// Fortran NINT (round half away from zero) of the product, evaluated in
// double precision as the build with 8-byte reals does
long fortran_nint(double v) { return (long)std::lround(v); }

}  // namespace

extern "C" int lceq_conf_write(const int* dims, const double* u, const double* a,
                               const char* u_path, const char* a_path, int real_bytes) {
  return conf_io(true, dims, const_cast<double*>(u), const_cast<double*>(a), u_path, a_path,
                 real_bytes);
}

extern "C" int lceq_conf_read(const int* dims, double* u, double* a, const char* u_path,
                              const char* a_path, int real_bytes) {
  return conf_io(false, dims, u, a, u_path, a_path, real_bytes);
}

extern "C" int lceq_confsave(lceq_lattice* L, const char* u_path, const char* a_path,
                             int real_bytes) {
  if (!L) return lceq::set_error(LCEQ_ERR_ARG, "null lattice");
  int dims[3];
  int rc = lceq_dims(L, dims);
  if (rc) return rc;
  int64_t n = (int64_t)dims[0] * dims[1] * dims[2] / 2;
  std::vector<double> u(u_path ? (size_t)24 * n : 0), a(a_path ? (size_t)6 * n : 0);
  if (u_path && (rc = lceq_get_links(L, u.data(), n))) return rc;
  if (a_path && (rc = lceq_get_higgs(L, a.data(), n))) return rc;
  return conf_io(true, dims, u.data(), a.data(), u_path, a_path, real_bytes);
}

extern "C" int lceq_oldconf(lceq_lattice* L, const char* u_path, const char* a_path,
                            int real_bytes) {
  if (!L) return lceq::set_error(LCEQ_ERR_ARG, "null lattice");
  int dims[3];
  int rc = lceq_dims(L, dims);
  if (rc) return rc;
  int64_t n = (int64_t)dims[0] * dims[1] * dims[2] / 2;
  std::vector<double> u(u_path ? (size_t)24 * n : 0), a(a_path ? (size_t)6 * n : 0);
  if ((rc = conf_io(false, dims, u.data(), a.data(), u_path, a_path, real_bytes))) return rc;
  if (u_path && (rc = lceq_put_links(L, u.data(), n))) return rc;
  if (a_path && (rc = lceq_put_higgs(L, a.data(), n))) return rc;
  return LCEQ_OK;
}

// File names of f/confsave.f:249-253 (upper = 1: '_BE', '_U') and of
// f/measinit.f:246-250 (upper = 0: '_be', '_u'), which oldconf opens.
extern "C" int lceq_conf_filename(int higgs, int upper, int lz, double beta, double x, double y,
                                  int sweep, char* out, int cap) {
  if (!out || cap < 1) return lceq::set_error(LCEQ_ERR_ARG, "null output buffer");
  std::string s = higgs ? "a_conf_s" : "u_conf_s";
  fortran_iw_m(s, lz, 2, 2);
  s += upper ? "_BE" : "_be";
  fortran_iw_m(s, fortran_nint(1000.0 * beta), 5, 5);
  s += "_x";
  fortran_iw_m(s, fortran_nint(10000.0 * x), 5, 5);
  s += "_y";
  fortran_iw_m(s, fortran_nint(10000.0 * y), 5, 5);
  s += upper ? "_U" : "_u";
  fortran_iw_m(s, sweep, 6, 6);
  if ((int)s.size() + 1 > cap) return lceq::set_error(LCEQ_ERR_ARG, "buffer of %d too small", cap);
  memcpy(out, s.c_str(), s.size() + 1);
  return LCEQ_OK;
}
