// Core types shared by the HIP kernels and the C-ABI layer.
//
// Device layout of one complex half-spectrum field (the reference's
// complex(sp), dimension(lx1, ly, lz) with lx1 = nx/2 + 1):
//   main[z][y][x], x in [0, nx/2)   row pitch nx/2 complex (rows 128-byte
//                                   aligned for nx >= 32)
//   nyq [z][y]                      the kx = nx/2 plane, stored apart so the
//                                   main rows stay aligned
// Both parts live in one allocation: main first, nyq right after it.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/scl.h"

namespace scl {

template <class R>
struct alignas(2 * sizeof(R)) cpx {
  R x, y;
};

template <class R>
__host__ __device__ inline cpx<R> mk(R a, R b) {
  cpx<R> c;
  c.x = a;
  c.y = b;
  return c;
}
template <class R>
__host__ __device__ inline cpx<R> operator+(cpx<R> a, cpx<R> b) {
  return mk<R>(a.x + b.x, a.y + b.y);
}
template <class R>
__host__ __device__ inline cpx<R> operator-(cpx<R> a, cpx<R> b) {
  return mk<R>(a.x - b.x, a.y - b.y);
}
template <class R>
__host__ __device__ inline cpx<R> operator*(cpx<R> a, cpx<R> b) {
  return mk<R>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <class R>
__host__ __device__ inline cpx<R> operator*(R s, cpx<R> a) {
  return mk<R>(s * a.x, s * a.y);
}
template <class R>
__host__ __device__ inline cpx<R> conj(cpx<R> a) {
  return mk<R>(a.x, -a.y);
}
// i * a
template <class R>
__host__ __device__ inline cpx<R> mul_i(cpx<R> a) {
  return mk<R>(-a.y, a.x);
}
// a * conj(b) when CONJ, else a * b
template <bool CONJ, class R>
__host__ __device__ inline cpx<R> cmul(cpx<R> a, cpx<R> b) {
  if (CONJ) return mk<R>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
  return a * b;
}

// A load of data the kernel reads once, marked nontemporal so that the
// caches do not keep it.  On MI355X a streaming read of 2 GiB reaches
// 7.1 TB/s this way against 5.5 TB/s with plain loads (measured).  Only for
// accesses that cover whole 128-byte lines (the 16-column z-passes): where
// two workgroups read the halves of one line (8-column y tiles), the second
// relies on L2 keeping it, and these loads made the y-passes 20-30% slower.
#ifndef SCL_NT_LOADS
#define SCL_NT_LOADS 1
#endif
template <class R, bool NT = true>
__device__ inline cpx<R> ld_once(const cpx<R>* p) {
  if constexpr (SCL_NT_LOADS && NT) {
    typedef R V2 __attribute__((ext_vector_type(2)));
    const V2 t = __builtin_nontemporal_load(reinterpret_cast<const V2*>(p));
    return mk<R>(t.x, t.y);
  } else {
    return *p;
  }
}

// An int the compiler cannot see through: the LDS addresses a line
// transform derives from it are recomputed for that transform instead of
// being kept live (and spilled) from the previous one.
__device__ inline int opaque(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

// Geometry of a 3D plan, shared by all kernels.  Row (y, z) holds the mx
// main entries of that (y, z) and, at nmain + row, its kx = nx/2 entry.
// Rows are ordered [z >> zs][y][z & (2^zs - 1)]: with zs = 3, the 8
// consecutive z-rows of a column lie in one mx * 8-entry span, so a z-column
// of 1024 rows touches 128 such spans instead of 1024 rows 4 MB apart (each
// a separate address translation); zs = 0 is plain [z][y].
struct Geom {
  int nx, ny, nz;  // physical grid
  int mx;          // nx/2: complex entries per main row
  int zs;          // log2 of the z-block of the row order
  int64_t nrow;    // ny*nz rows
  int64_t nmain;   // mx*ny*nz complex entries in main
  int64_t ntot;    // nmain + ny*nz (main + nyq)
  __host__ __device__ int64_t row(int y, int z) const {
    return ((((int64_t)(z >> zs)) * ny + y) << zs) + (z & ((1 << zs) - 1));
  }
  __host__ __device__ void yz(int64_t r, int& y, int& z) const {
    const int64_t blk = (int64_t)ny << zs;
    const int64_t zb = r / blk, rem = r - zb * blk;
    y = (int)(rem >> zs);
    z = (int)((zb << zs) + (rem & ((1 << zs) - 1)));
  }
};

inline Geom make_geom(int nx, int ny, int nz, int zs = 0) {
  Geom g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.mx = nx / 2;
  g.zs = zs;
  g.nrow = (int64_t)ny * nz;
  g.nmain = (int64_t)g.mx * g.nrow;
  g.ntot = g.nmain + g.nrow;
  return g;
}

// Signed wavenumber of index i on an axis of length n (wavenumber.f90:109).
__host__ __device__ inline int kwave(int i, int n) { return i < n / 2 ? i : i - n; }

// Error state (thread-local message, read by scl_last_error()).
void set_error(const char* fmt, ...);
const char* last_error();

#define SCL_HIP(call)                                                             \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) {                                                       \
      scl::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #call,                 \
                     hipGetErrorString(e_));                                      \
      return SCL_EHIP;                                                            \
    }                                                                             \
  } while (0)

#define SCL_CHECK(expr)          \
  do {                           \
    int32_t s_ = (expr);         \
    if (s_ != SCL_OK) return s_; \
  } while (0)

}  // namespace scl
