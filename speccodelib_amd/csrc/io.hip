// Files of the reference's runs, host side: Fortran unformatted sequential
// dumps (output.f90:103-166, input.f90:106-170: one record per field,
// ./out/ux<n>.dat, uy, uz [, phi]) and the ./post text diagnostics
// (spectral_dns_dealiased.f90:120-171: ener_time.data lines written with
// '(I6,20E12.4)', spectrum.data blocks of the time then "ll E(ll)" lines).
//
// A record is its length in bytes as a 4-byte little-endian integer, the
// data, and the length again.  Records longer than 2^31 - 9 bytes (a
// 1024^3 FP32 field) are split into subrecords as gfortran does: the
// leading marker of every subrecord but the last is negated (more follow),
// the trailing marker of every subrecord but the first is negated (one
// precedes); the reader accepts any split whose markers agree.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "plan.hpp"

namespace {

constexpr int64_t MAX_SUBRECORD = 2147483639;  // 2^31 - 9

int32_t io_error(const char* what, const std::string& path) {
  scl::set_error("%s: %s%s%s", path.c_str(), what, errno ? ": " : "", errno ? std::strerror(errno) : "");
  return SCL_EINVAL;
}

struct File {
  FILE* f = nullptr;
  File(const std::string& p, const char* mode) { f = std::fopen(p.c_str(), mode); }
  ~File() {
    if (f) std::fclose(f);
  }
};

int32_t write_record(const std::string& path, const void* data, int64_t nbytes, const char* mode) {
  errno = 0;
  File fp(path, mode);
  if (!fp.f) return io_error("cannot open for writing", path);
  const char* p = static_cast<const char*>(data);
  int64_t left = nbytes;
  bool first = true;
  do {
    const int64_t len = left < MAX_SUBRECORD ? left : MAX_SUBRECORD;
    const bool last = len == left;
    const int32_t lead = last ? (int32_t)len : -(int32_t)len;
    const int32_t trail = first ? (int32_t)len : -(int32_t)len;
    if (std::fwrite(&lead, 4, 1, fp.f) != 1 || (len && std::fwrite(p, 1, (size_t)len, fp.f) != (size_t)len) ||
        std::fwrite(&trail, 4, 1, fp.f) != 1)
      return io_error("write failed", path);
    p += len;
    left -= len;
    first = false;
  } while (left > 0);
  if (std::fflush(fp.f) != 0) return io_error("write failed", path);
  return SCL_OK;
}

// One record of exactly nbytes, the whole file: anything else (a short or
// truncated file, disagreeing markers, another length, trailing bytes) is
// SCL_EINVAL with the reason.
int32_t read_record(const std::string& path, void* data, int64_t nbytes) {
  errno = 0;
  File fp(path, "rb");
  if (!fp.f) return io_error("cannot open for reading", path);
  errno = 0;
  char* p = static_cast<char*>(data);
  int64_t got = 0;
  bool more = true, first = true;
  while (more) {
    int32_t lead = 0, trail = 0;
    if (std::fread(&lead, 4, 1, fp.f) != 1) return io_error("truncated (no record marker)", path);
    const int64_t len = lead < 0 ? -(int64_t)lead : lead;
    more = lead < 0;
    if (got + len > nbytes) {
      scl::set_error("%s: the record holds more than the %lld bytes of the field", path.c_str(), (long long)nbytes);
      return SCL_EINVAL;
    }
    if (len && std::fread(p + got, 1, (size_t)len, fp.f) != (size_t)len) return io_error("truncated record", path);
    if (std::fread(&trail, 4, 1, fp.f) != 1) return io_error("truncated (no trailing record marker)", path);
    const int64_t tlen = trail < 0 ? -(int64_t)trail : trail;
    if (tlen != len || (trail < 0) == first) {
      scl::set_error("%s: record markers disagree (%d, %d): not an unformatted sequential record", path.c_str(),
                     (int)lead, (int)trail);
      return SCL_EINVAL;
    }
    got += len;
    first = false;
  }
  if (got != nbytes) {
    scl::set_error("%s: the record holds %lld bytes, the field %lld", path.c_str(), (long long)got,
                   (long long)nbytes);
    return SCL_EINVAL;
  }
  char extra;
  if (std::fread(&extra, 1, 1, fp.f) == 1) {
    scl::set_error("%s: bytes after the record (one record per file expected)", path.c_str());
    return SCL_EINVAL;
  }
  return SCL_OK;
}

std::string dump_path(const char* dir, const char* name, int idump) {
  std::string d = dir && dir[0] ? dir : ".";
  if (d.back() != '/') d += '/';
  return d + name + std::to_string(idump) + ".dat";
}

// Fortran Ew.d editing (w = 12, d = 4): [-]0.dddd E[+-]ee, right-justified.
void fortran_e12_4(double v, char* out) {
  char buf[32];
  if (v == 0.0 || !std::isfinite(v)) {
    if (v == 0.0) std::snprintf(buf, sizeof(buf), "%s0.0000E+00", std::signbit(v) ? "-" : "");
    else std::snprintf(buf, sizeof(buf), "%s", std::isnan(v) ? "NaN" : (v > 0 ? "Infinity" : "-Infinity"));
  } else {
    // d.dddde+xx from %e with d digits after the point gives the mantissa
    // rounded to 4 significant digits; Fortran's form is 0.dddd x 10^(e+1)
    char e[32];
    std::snprintf(e, sizeof(e), "%.3e", std::fabs(v));
    const int ex = std::atoi(std::strchr(e, 'e') + 1) + 1;
    const char digits[5] = {e[0], e[2], e[3], e[4], 0};
    std::snprintf(buf, sizeof(buf), "%s0.%sE%c%02d", v < 0 ? "-" : "", digits, ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
  }
  std::snprintf(out, 13, "%12s", buf);
}

}  // namespace

extern "C" int32_t scl_record_write(const char* path, const void* data, int64_t nbytes) {
  if (!path || (!data && nbytes) || nbytes < 0) return scl::set_error("bad record arguments"), SCL_EINVAL;
  return write_record(path, data, nbytes, "wb");
}

extern "C" int32_t scl_record_read(const char* path, void* data, int64_t nbytes) {
  if (!path || (!data && nbytes) || nbytes < 0) return scl::set_error("bad record arguments"), SCL_EINVAL;
  return read_record(path, data, nbytes);
}

extern "C" int32_t scl_dns_output(scl_dns* d, const char* dir, int32_t idump) {
  if (!d) return scl::set_error("null solver"), SCL_EINVAL;
  scl_plan3d* p = d->plan;
  const size_t nb = p->ref_bytes();
  std::vector<char> a(nb), b(nb), c(nb);
  SCL_CHECK(scl_dns_get_velocity(d, a.data(), b.data(), c.data()));
  SCL_CHECK(write_record(dump_path(dir, "ux", idump), a.data(), (int64_t)nb, "wb"));
  SCL_CHECK(write_record(dump_path(dir, "uy", idump), b.data(), (int64_t)nb, "wb"));
  SCL_CHECK(write_record(dump_path(dir, "uz", idump), c.data(), (int64_t)nb, "wb"));
  if (d->nc == 4) {
    SCL_CHECK(scl_dns_get_scalar(d, a.data()));
    SCL_CHECK(write_record(dump_path(dir, "phi", idump), a.data(), (int64_t)nb, "wb"));
  }
  return SCL_OK;
}

extern "C" int32_t scl_dns_input(scl_dns* d, const char* dir, int32_t idump) {
  if (!d) return scl::set_error("null solver"), SCL_EINVAL;
  scl_plan3d* p = d->plan;
  const size_t nb = p->ref_bytes();
  std::vector<char> a(nb), b(nb), c(nb);
  // every file is read and checked before the solver's state changes
  SCL_CHECK(read_record(dump_path(dir, "ux", idump), a.data(), (int64_t)nb));
  SCL_CHECK(read_record(dump_path(dir, "uy", idump), b.data(), (int64_t)nb));
  SCL_CHECK(read_record(dump_path(dir, "uz", idump), c.data(), (int64_t)nb));
  std::vector<char> f;
  if (d->nc == 4) {
    f.resize(nb);
    SCL_CHECK(read_record(dump_path(dir, "phi", idump), f.data(), (int64_t)nb));
  }
  SCL_CHECK(scl_dns_set_velocity(d, a.data(), b.data(), c.data()));
  if (d->nc == 4) SCL_CHECK(scl_dns_set_scalar(d, f.data()));
  return SCL_OK;
}

extern "C" int32_t scl_write_ener_time(const char* path, int32_t istep, const double* vals, int32_t nvals) {
  if (!path || nvals < 0 || nvals > 20 || (nvals && !vals)) return scl::set_error("bad ener_time arguments"), SCL_EINVAL;
  errno = 0;
  File fp(path, "ab");
  if (!fp.f) return io_error("cannot open for appending", path);
  char line[8 + 20 * 12 + 2];
  int n = std::snprintf(line, sizeof(line), "%6d", (int)istep);
  if (n != 6) std::snprintf(line, sizeof(line), "******");  // I6 overflow, as Fortran prints it
  n = 6;
  for (int i = 0; i < nvals; ++i, n += 12) fortran_e12_4(vals[i], line + n);
  line[n++] = '\n';
  if (std::fwrite(line, 1, (size_t)n, fp.f) != (size_t)n) return io_error("write failed", path);
  return SCL_OK;
}

extern "C" int32_t scl_write_spectrum(const char* path, double time, const double* ek, int32_t n) {
  if (!path || n < 0 || (n && !ek)) return scl::set_error("bad spectrum arguments"), SCL_EINVAL;
  errno = 0;
  File fp(path, "ab");
  if (!fp.f) return io_error("cannot open for appending", path);
  // list-directed output: a REAL(4) value is written with 9 significant
  // digits, an integer right-justified in 12 columns
  if (std::fprintf(fp.f, "  %15.8E\n", (double)(float)time) < 0) return io_error("write failed", path);
  for (int i = 0; i < n; ++i)
    if (std::fprintf(fp.f, " %11d  %15.8E\n", i + 1, (double)(float)ek[i]) < 0) return io_error("write failed", path);
  return SCL_OK;
}
