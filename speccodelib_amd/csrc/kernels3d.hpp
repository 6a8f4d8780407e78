// HIP kernels of the 3D pseudo-spectral RHS and IF-RK4 step.
//
// One RHS evaluation (ratesdudt, dns-deform.f90:259-288) is five passes over
// HBM, each a batch of 1D FFTs along one axis with the pointwise work fused
// into the pass that already holds the data:
//   ZINV  z-inverse of u and of w = i k x u (vorticity built from the loaded
//         u, never read from HBM): reads 3 fields, writes 6
//   YINV  y-inverse of the 6 fields, in place
//   XPASS per x-row: c2r of the 6 fields, u x w (the packed-real cross
//         product of convec_dns.f90:30-42, times 1/(nx ny nz)), r2c of the 3
//         products: reads 6 fields, writes 3 in place
//   YFWD  y-forward of the 3 fields, in place
//   ZFWD  z-forward + forcing + symmetrize's zeroing + projection + the
//         IF-RK4 stage update (dns-deform.f90:164-237), reading v, du and the
//         stage state
// plus two O(n^2) kernels: the forcing coefficients (a reduction over the
// few forced modes) and the kx=0-plane Hermitian copy of symmetrize.
#pragma once

#include "scl_fft.hpp"

namespace scl {

template <class R>
struct FieldPtrs {
  cpx<R>* p[8];
  int n;
};

__device__ inline float exp_r(float x) { return expf(x); }
__device__ inline double exp_r(double x) { return exp(x); }
__device__ inline float expm1_r(float x) { return expm1f(x); }
__device__ inline double expm1_r(double x) { return expm1(x); }

// ----------------------------------------------------------------------------
// Column tiles.  A workgroup transforms CW columns along AXIS (1 = y, 2 = z).
// Main tiles hold CW consecutive x (one 128-byte segment per row for float);
// tail tiles hold the kx = nx/2 column for CW values of the other index.
// ----------------------------------------------------------------------------
struct Tile {
  int64_t base;  // offset of this thread's column, index 0 along the axis
  int64_t rs;    // y-pass: row stride; z-pass: stride of the rows of a z-block
  int64_t bs;    // z-pass: stride of the z-blocks (y-pass: = rs)
  int sh;        // z-pass: g.zs (y-pass: 0)
  int kx;        // x index of the column (mx for the kx = nx/2 plane)
  int oth;       // the other batch index: z for the y-pass, y for the z-pass
  bool ok;       // false for the unused lanes of a partial tail tile
  // offset of index j along the axis
  __device__ int64_t off(int j) const {
    return base + (int64_t)(j >> sh) * bs + (int64_t)(j & ((1 << sh) - 1)) * rs;
  }
};

// The z-pass addressing with 32-bit in-field offsets (< 2^30 at 1024^3) and
// the z-block shift ZS (= g.zs) known at compile time: every field gets its
// own base pointer, and a thread keeps one 32-bit offset per slot.
template <int ZS>
struct ZOff {
  uint32_t ob, rs, bs;
  __device__ uint32_t at(int j) const {
    return ob + (uint32_t)(j >> ZS) * bs + (uint32_t)(j & ((1 << ZS) - 1)) * rs;
  }
};
template <int ZS>
__device__ inline ZOff<ZS> zoff(const Tile& t) {
  return ZOff<ZS>{(uint32_t)t.base, (uint32_t)t.rs, (uint32_t)t.bs};
}
// The same with the shift known at run time (kernels off the step's path).
struct ZOffR {
  uint32_t ob, rs, bs;
  int sh;
  __device__ uint32_t at(int j) const {
    return ob + (uint32_t)(j >> sh) * bs + (uint32_t)(j & ((1 << sh) - 1)) * rs;
  }
};
__device__ inline ZOffR zoffr(const Tile& t) { return ZOffR{(uint32_t)t.base, (uint32_t)t.rs, (uint32_t)t.bs, t.sh}; }

template <int CW, int AXIS>
__device__ inline Tile col_tile(const Geom& g, int wg, int c) {
  const int xtn = g.mx / CW;
  Tile t;
  if constexpr (AXIS == 1) {
    const int nmain = xtn * g.nz;
    t.sh = 0;
    if (wg < nmain) {
      const int xt = wg % xtn, z = wg / xtn;
      t.base = g.row(0, z) * g.mx + xt * CW + c;
      t.rs = (int64_t)g.mx << g.zs;
      t.kx = xt * CW + c;
      t.oth = z;
      t.ok = true;
    } else {
      const int z = (wg - nmain) * CW + c;
      t.ok = z < g.nz;
      t.base = g.nmain + g.row(0, t.ok ? z : 0);
      t.rs = (int64_t)1 << g.zs;
      t.kx = g.mx;
      t.oth = z;
    }
    t.bs = t.rs;
  } else {
    const int nmain = xtn * g.ny;
    t.sh = g.zs;
    if (wg < nmain) {
      const int xt = wg % xtn, y = wg / xtn;
      t.base = g.row(y, 0) * g.mx + xt * CW + c;
      t.rs = g.mx;
      t.bs = ((int64_t)g.ny << g.zs) * g.mx;
      t.kx = xt * CW + c;
      t.oth = y;
      t.ok = true;
    } else {
      const int y = (wg - nmain) * CW + c;
      t.ok = y < g.ny;
      t.base = g.nmain + g.row(t.ok ? y : 0, 0);
      t.rs = 1;
      t.bs = (int64_t)g.ny << g.zs;
      t.kx = g.mx;
      t.oth = y;
    }
  }
  return t;
}

// Main tiles plus tail tiles; a partial last tail tile (batch count not a
// multiple of CW, as for one 2D member) has lanes with ok = false.
template <int CW, int AXIS>
inline int col_grid(const Geom& g) {
  return AXIS == 1 ? (g.mx / CW) * g.nz + (g.nz + CW - 1) / CW : (g.mx / CW) * g.ny + (g.ny + CW - 1) / CW;
}

// LDS addressing of the Stockham exchanges.  One padding slot after every
// 8 entries (pos + pos/8) keeps the stride-8 writes of a radix-8 pass and
// the unit-stride reads bank-conflict-free within each 16-lane group.
__host__ __device__ constexpr int lds_pad(int n) { return n + n / 8; }

template <int CW>
struct ColIdx {
  int c;
  __device__ inline int operator()(int pos) const { return (pos + (pos >> 3)) * CW + c; }
};

struct LinIdx {
  __device__ inline int operator()(int pos) const { return pos + (pos >> 3); }
};

// Column tile shape: CW columns of N points, E points per thread.  CW = 8
// gives 64-byte row segments; the two halves of a 128-byte line belong to
// logical tiles t and t+1, which xcd_swizzle() places on the same XCD (one
// L2) at nearly the same time.
#ifndef SCL_COL_CW
#define SCL_COL_CW 8
#endif
// FP64 1024-point lines take 4 columns: 8 would need 163 KB of LDS with the
// line's twiddles.
template <class R, int N, int CWv = (sizeof(R) == 8 && N >= 1024 ? 4 : SCL_COL_CW), int EV = 0>
struct ColCfg {
  static constexpr int CW = CWv;
  static constexpr int E = EV ? EV : (N <= 512 ? 8 : 16);
  static constexpr int T = N / E;
  static constexpr int BS = CW * T;
};
// The fused z-pass kernels (vorticity in, RK4 epilogue out) take CW as a
// template argument: 16 columns (whole 128-byte lines) where the grid
// allows, since z-rows are nx*ny/2 entries apart and half-line segments
// cost more there than along y.

// Workgroups are dispatched round-robin over the 8 XCDs (block b -> XCD
// b mod 8).  Map block b to logical tile (b mod 8) * (G/8) + b/8 so that each
// XCD walks a contiguous range of tiles; the last G mod 8 blocks keep their
// index.
__device__ inline int xcd_swizzle(int b, int G) {
  const int G8 = G & ~7;
  if (b >= G8) return b;
  return (b & 7) * (G8 >> 3) + (b >> 3);
}

// Pure batched column FFT, in place, one field per blockIdx.y.
template <class R, int N, int AXIS, int DIR>
__global__ __launch_bounds__((ColCfg<R, N>::BS)) void k_colfft(cpx<R>* __restrict__ f0, int64_t fstride, Geom g,
                                                              const cpx<R>* __restrict__ tw, int tws) {
  using CC = ColCfg<R, N>;
  constexpr int CW = CC::CW, E = CC::E, T = CC::T;
  __shared__ cpx<R> lds[lds_pad(N) * CW];
  __shared__ cpx<R> twl[N];  // this line length's twiddles; first read after a barrier
  for (int i = threadIdx.x; i < N; i += CC::BS) twl[i] = tw[i * tws];
  const int c = threadIdx.x % CW, rt = threadIdx.x / CW;
  cpx<R>* f = f0 + (int64_t)blockIdx.y * fstride;
  const Tile t = col_tile<CW, AXIS>(g, xcd_swizzle(blockIdx.x, gridDim.x), c);
  cpx<R> v[E];
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = t.ok ? f[t.off(rt + T * m)] : mk<R>(0, 0);
  Line<R, N, E, DIR>::run(v, lds, rt, twl, 1, ColIdx<CW>{c});
  if (!t.ok) return;
#pragma unroll
  for (int m = 0; m < E; ++m) f[t.off(Line<R, N, E, DIR>::out_pos(m, rt))] = v[m];
}

// z-inverse of a velocity column held in registers (slot m = index
// rt + T m along z, as every transform here leaves it) and of w_x, w_y of
// w = i k x u (dns-deform.f90:268-270), by linearity of the z-transform Z:
//   Z(wx) = i (ky Uz - Z(kz uy)),  Z(wy) = i (Z(kz ux) - kx Uz),  Uq = Z(uq):
// five transforms, stored to w[0..4] (u, then w_x, w_y) at stride ws.
// w_z = i (kx u_y - ky u_x) needs no kz, so the y-inverse builds it from
// the transformed u_x, u_y (k_yinv_vort) and one field less crosses HBM.
// Offsets are 32-bit (ZOff).  PARK (1024-thread workgroups, one per CU at the
// 128-VGPR cap, and 1024-point lines, one 512-thread workgroup per CU at
// 2 waves per SIMD: LDS to spare in both): Uz waits in park (E BS entries,
// each thread reading back only its own) between its store and its two
// later uses.
template <class R, int N, int E, int CW, bool PARK, int ZS>
__device__ inline void zinv_store(cpx<R> (&ux)[E], cpx<R> (&uy)[E], cpx<R> (&uz)[E], cpx<R>* lds, cpx<R>* park,
                                  const cpx<R>* twl, int rt, int c, R kx, R ky, int nz, ZOff<ZS> zo,
                                  cpx<R>* __restrict__ w, int64_t ws) {
  constexpr int T = N / E, BS = CW * T;
  using L = Line<R, N, E, +1>;
  // An empty asm that "changes" the base offset: the store addresses below
  // are then recomputed here rather than kept live (and spilled) from the
  // caller's loads of the same columns.
  asm volatile("" : "+v"(zo.ob));
  auto off = [&](int m) { return zo.at(rt + T * m); };
  auto kz = [&](int m) { return (R)kwave(rt + T * m, nz); };
  const int slot = rt * CW + c;
  cpx<R> d[E];
  L::run(uz, lds, opaque(rt), twl, 1, ColIdx<CW>{c});
#pragma unroll
  for (int m = 0; m < E; ++m) {
    (w + 2 * ws)[off(m)] = uz[m];
    if constexpr (PARK) park[m * BS + slot] = uz[m];
  }
  auto Uz = [&](int m) { return PARK ? park[m * BS + slot] : uz[m]; };
#pragma unroll
  for (int m = 0; m < E; ++m) d[m] = kz(m) * uy[m];
  L::run(d, lds, opaque(rt), twl, 1, ColIdx<CW>{c});
#pragma unroll
  for (int m = 0; m < E; ++m) (w + 3 * ws)[off(m)] = mul_i(ky * Uz(m) - d[m]);
  L::run(uy, lds, opaque(rt), twl, 1, ColIdx<CW>{c});
#pragma unroll
  for (int m = 0; m < E; ++m) (w + ws)[off(m)] = uy[m];
#pragma unroll
  for (int m = 0; m < E; ++m) d[m] = kz(m) * ux[m];
  L::run(d, lds, opaque(rt), twl, 1, ColIdx<CW>{c});
#pragma unroll
  for (int m = 0; m < E; ++m) (w + 4 * ws)[off(m)] = mul_i(d[m] - kx * Uz(m));
  L::run(ux, lds, opaque(rt), twl, 1, ColIdx<CW>{c});
#pragma unroll
  for (int m = 0; m < E; ++m) w[off(m)] = ux[m];
}

// ZINV: z-inverse of the stage state s (3 fields at stride sstride) into
// w[0..4] (stride wstride), see zinv_store.
template <class R, int N, int CWZ, int EZ = 0, int ZS = 0>
__global__ __launch_bounds__((ColCfg<R, N, CWZ, EZ>::BS)) void k_zinv_vort(const cpx<R>* __restrict__ s,
                                                                     int64_t sstride, cpx<R>* __restrict__ w,
                                                                     int64_t wstride, Geom g,
                                                                     const cpx<R>* __restrict__ tw, int tws) {
  using CC = ColCfg<R, N, CWZ, EZ>;
  constexpr int CW = CC::CW, E = CC::E, T = CC::T;
  constexpr bool PARK = CC::BS >= 1024 || (N >= 1024 && sizeof(R) == 4);
  __shared__ cpx<R> lds[lds_pad(N) * CW];
  __shared__ cpx<R> twl[N];  // this line length's twiddles; first read after a barrier
  __shared__ cpx<R> park[PARK ? E * CC::BS : 1];
  for (int i = threadIdx.x; i < N; i += CC::BS) twl[i] = tw[i * tws];
  const int c = threadIdx.x % CW, rt = threadIdx.x / CW;
  const Tile t = col_tile<CW, 2>(g, xcd_swizzle(blockIdx.x, gridDim.x), c);
  const ZOff<ZS> zo = zoff<ZS>(t);
  cpx<R> ux[E], uy[E], uz[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const uint32_t o = zo.at(rt + T * m);
    uz[m] = ld_once<R, (CW >= 16)>(s + o + 2 * sstride);
    uy[m] = ld_once<R, (CW >= 16)>(s + o + sstride);
    ux[m] = ld_once<R, (CW >= 16)>(s + o);
  }
  zinv_store<R, N, E, CW, PARK, ZS>(ux, uy, uz, lds, park, twl, rt, c, (R)t.kx, (R)kwave(t.oth, g.ny), g.nz, zo, w,
                                    wstride);
}

// YINV (u_x, u_y and w_z): y-inverse of fields 0 and 1 (u_x, u_y) and of
// w_z = i (kx Uy - ky Ux), formed before the transform while ky is still a
// spectral index, into field 5; in place at stride fs.  Fields 2..4 (u_z,
// w_x, w_y) take the plain column kernel.
template <class R, int N>
__global__ __launch_bounds__((ColCfg<R, N>::BS)) void k_yinv_vort(cpx<R>* __restrict__ W, int64_t fs, Geom g,
                                                                const cpx<R>* __restrict__ tw, int tws) {
  using CC = ColCfg<R, N>;
  constexpr int CW = CC::CW, E = CC::E, T = CC::T;
  __shared__ cpx<R> lds[lds_pad(N) * CW];
  __shared__ cpx<R> twl[N];  // first read after a barrier inside the transform
  for (int i = threadIdx.x; i < N; i += CC::BS) twl[i] = tw[i * tws];
  const int c = threadIdx.x % CW, rt = threadIdx.x / CW;
  const Tile t = col_tile<CW, 1>(g, xcd_swizzle(blockIdx.x, gridDim.x), c);
  using L = Line<R, N, E, +1>;
  auto load = [&](cpx<R>(&v)[E], int f) {
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = t.ok ? W[t.off(rt + T * m) + f * fs] : mk<R>(0, 0);
  };
  auto store = [&](cpx<R>(&v)[E], int f) {
    if (!t.ok) return;
#pragma unroll
    for (int m = 0; m < E; ++m) W[t.off(L::out_pos(m, rt)) + f * fs] = v[m];
  };
  const R kx = (R)t.kx;
  cpx<R> a[E], b[E], d[E];
  load(a, 0);
  load(b, 1);
#pragma unroll
  for (int m = 0; m < E; ++m) d[m] = mul_i(kx * b[m] - (R)kwave(rt + T * m, g.ny) * a[m]);
  L::run(d, lds, rt, twl, 1, ColIdx<CW>{c});
  store(d, 5);
  L::run(a, lds, rt, twl, 1, ColIdx<CW>{c});
  store(a, 0);
  L::run(b, lds, rt, twl, 1, ColIdx<CW>{c});
  store(b, 1);
}

// The values one stage update of one component reads: v and the
// accumulator du (IF-RK4, MODE 1..4) or v and the AB2 history (MODE 5..7:
// v0 from DU, the last N from S).  Loaded apart from the update so that a
// caller can issue them ahead of the stores of the slots before.
template <class R>
struct StageIn {
  cpx<R> v, h;
};

// MODE 4 reads du only inside the cut.  NT: nontemporal loads (whole-line
// accesses only, ld_once).
template <int MODE, class R, class I, bool NT = false>
__device__ inline StageIn<R> stage_load(I o, const cpx<R>* V, const cpx<R>* DU, const cpx<R>* S, bool cut) {
  StageIn<R> in{mk<R>(0, 0), mk<R>(0, 0)};
  if constexpr (MODE != 4) in.v = ld_once<R, NT>(V + o);
  if constexpr (MODE == 2 || MODE == 3 || MODE == 6) in.h = ld_once<R, NT>(DU + o);
  if constexpr (MODE == 4) {
    if (!cut) in.h = ld_once<R, NT>(DU + o);
  }
  if constexpr (MODE == 7) in.h = ld_once<R, NT>(S + o);
  return in;
}

// One IF-RK4 stage of one component in the division-free form (see
// k_zfwd_rhs): a = dt N, em1 = expm1(x), x = -rate dt/2, so E y = y + em1 y
// and E^2 y = y + expm1(2x) y, each exact to the rounding of the result.
// Returns the new stage state (MODE 1..3, stored to S only where store_s)
// or the new v (MODE 4, always stored); in holds what stage_load read.
// MODE 4 with nf set: a non-finite new v raises the sticky flag *nf (read
// first, so that a field full of NaN does not make every lane an atomic).
template <int MODE, class R, class I>
__device__ inline cpx<R> stage_apply(const StageIn<R>& in, cpx<R> a, I o, cpx<R>* V, cpx<R>* DU, cpx<R>* S, R em1,
                                     R x, bool cut, bool store_s, unsigned* nf = nullptr) {
  cpx<R> out;
  if constexpr (MODE == 1) {
    const cpx<R> v = in.v;
    const cpx<R> u = v + ((R)1 / (R)6) * a, q = v + (R)0.5 * a;
    DU[o] = u + expm1_r((R)2 * x) * u;
    out = q + em1 * q;
  } else if constexpr (MODE == 2 || MODE == 3) {
    const cpx<R> v = in.v;
    const cpx<R> q = ((R)1 / (R)3) * a;
    DU[o] = in.h + (q + em1 * q);
    const cpx<R> ev = v + em1 * v;
    if constexpr (MODE == 2) {
      out = ev + (R)0.5 * a;
    } else {
      const cpx<R> u = ev + a;
      out = u + em1 * u;
    }
  } else {
    out = cut ? mk<R>(0, 0) : in.h + ((R)1 / (R)6) * a;
    V[o] = out;
    if (nf && !(isfinite(out.x) && isfinite(out.y)) && *(volatile unsigned*)nf == 0u) atomicOr(nf, 1u);
    return out;
  }
  if (store_s) S[o] = out;
  return out;
}

// Parameters of the fused ZFWD epilogue.
template <class R>
struct RhsParams {
  R dt;
  R nu;
  R fk2;        // fkmax^2 (forcing where k2 < fk2)
  R kc2;        // 2/3 cut: zero where k2 >= kc2
  int forced;   // 0 none, 1 force(), 2 force_heli()
  int dealias;  // 1: zero k2 >= kc2 in the RHS as well
  unsigned* nonfinite = nullptr;  // sticky flag set by stage 4 on a non-finite v (null: no check)
  const double* ab = nullptr;     // IF-AB2 (MODE 5..7): the step's coefficients on the device (AB_*)
};

// Slots of the IF-AB2 device state (scl_dns::ab): time, this step's dt, dto,
// umax, twrite, the step count, and the weights of the update.
enum { AB_TIME = 0, AB_DT = 1, AB_DTO = 2, AB_UMAX = 3, AB_TWRITE = 4, AB_ISTEP = 5, AB_S = 6, AB_EK = 7,
       AB_DELTA = 8, AB_NSLOT = 16 };

// One IF-AB2 update of one component (spectral_dns_dealiased.f90:237-285),
// a = N (symmetrized, projected), with e1 = expm1(-nu k2 dt) and
// e2 = expm1(-nu k2 dto), so E w = w + e w exactly to the rounding of the
// result:
//   MODE 5 (step 0):  v1 = (v0 + dt a) E(dt);  DU = v0, S = a (N0)
//   MODE 6 (step 1):  v2 = [cS E(dto) v0 + cek v1 + cdel a] E(dt)  (v0 from DU)
//   MODE 7 (>= 2):    v = [v + cS a + cek E(dto) O] E(dt);  S = a (O = S)
// The new v is 0 where cut (k2 >= kcut^2, and the mean).  in and nf: as stage_apply.
template <int MODE, class R, class I>
__device__ inline cpx<R> ab_apply(const StageIn<R>& in, cpx<R> a, I o, cpx<R>* V, cpx<R>* DU, cpx<R>* S, R dt, R cS,
                                  R cek, R cdel, R e1, R e2, bool cut, unsigned* nf) {
  const cpx<R> v = in.v;
  cpx<R> y;
  if constexpr (MODE == 5) {
    DU[o] = v;
    S[o] = a;
    y = v + dt * a;
  } else if constexpr (MODE == 6) {
    const cpx<R> v0 = in.h;
    y = cS * (v0 + e2 * v0) + cek * v + cdel * a;
  } else {
    const cpx<R> ob = in.h;
    y = v + cS * a + cek * (ob + e2 * ob);
    S[o] = a;
  }
  const cpx<R> out = cut ? mk<R>(0, 0) : y + e1 * y;
  V[o] = out;
  if (nf && !(isfinite(out.x) && isfinite(out.y)) && *(volatile unsigned*)nf == 0u) atomicOr(nf, 1u);
  return out;
}


// ZFWD: z-forward of the three transformed products + forcing + the
// pointwise part of symmetrize + projection (+ optional RHS truncation),
// then, for MODE 1..4, the IF-RK4 stage update; MODE 0 writes N(u) back
// into W.  st is the stage state the forcing uses (v for stage 1).
// The scheme of dns-deform.f90:164-237 (k1 = dt N(v), s2 = E (v + k1/2),
// k2 = dt N(s2)/E, s3 = E (v + k2/2), k3 = dt N(s3)/E, s4 = E^2 (v + k3),
// k4 = dt N(s4)/E^2, v' = E^2 (v + (k1 + 2 k2 + 2 k3 + k4)/6), cut at
// kc2), E = exp(-nu k2 dt/2) (:90), is kept with the accumulator du holding
// E^2 v + ... instead, a = dt N:
//   MODE 1: du = E^2 (v + a/6);  S = E (v + a/2)
//   MODE 2: du += E a/3;         S = E v + a/2
//   MODE 3: du += E a/3;         S = E (E v + a)
//   MODE 4: v = du + a/6, 0 where k2 >= kc2
// so the last stage reads no v, and no stage divides by E.
//
// FUSED (MODE 1..4): the new stage state (MODE 4: the new v) stays in
// registers and its z-inverse goes to W[0..4] for the next RHS
// (zinv_store), so that RHS needs no z-inverse pass of its own; S is then
// stored only where the forcing reads it (k2 < fk2).  Each workgroup reads
// all of its columns of W before it writes them.
template <class R, int N, int MODE, int CWZ, int EZ = 0, bool FUSED = false, int ZS = 0>
__global__ __launch_bounds__((ColCfg<R, N, CWZ, EZ>::BS)) void k_zfwd_rhs(cpx<R>* W, int64_t fstride, cpx<R>* V,
                                                                    cpx<R>* DU, cpx<R>* S, const cpx<R>* st,
                                                                    const double* __restrict__ coef, RhsParams<R> p,
                                                                    Geom g, const cpx<R>* __restrict__ tw, int tws) {
  using CC = ColCfg<R, N, CWZ, EZ>;
  constexpr int CW = CC::CW, E = CC::E, T = CC::T;
  __shared__ cpx<R> lds[lds_pad(N) * CW];
  __shared__ cpx<R> twl[N];  // this line length's twiddles; first read after a barrier
  for (int i = threadIdx.x; i < N; i += CC::BS) twl[i] = tw[i * tws];
  const int c = threadIdx.x % CW, rt = threadIdx.x / CW;
  const Tile t = col_tile<CW, 2>(g, xcd_swizzle(blockIdx.x, gridDim.x), c);
  // 32-bit in-field offsets: the three components go through their own
  // base pointers, so a thread keeps one 32-bit offset per slot instead of
  // 64-bit ones (which spilled).
  const ZOff<ZS> zo = zoff<ZS>(t);
  cpx<R>* const W1 = W + fstride;
  cpx<R>* const W2 = W + 2 * fstride;
  // 1024-thread workgroups (and 1024-point lines) run one per CU, which
  // leaves LDS to spare: the transformed x component waits there (each
  // thread reads back only its own entries) instead of holding 2E registers
  // through the other two transforms, and z is loaded after x is parked.
  constexpr bool PARK = CC::BS >= 1024 || (N >= 1024 && sizeof(R) == 4);
  __shared__ cpx<R> park[PARK ? E * CC::BS : 1];
  cpx<R> nx_[E], ny_[E], nz_[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const uint32_t o = zo.at(rt + T * m);
    nx_[m] = ld_once<R, (CW >= 16)>(W + o);
    ny_[m] = ld_once<R, (CW >= 16)>(W1 + o);
    if constexpr (!PARK) nz_[m] = ld_once<R, (CW >= 16)>(W2 + o);
  }
  using L = Line<R, N, E, -1>;
  L::run(nx_, lds, opaque(rt), twl, 1, ColIdx<CW>{c});
  if constexpr (PARK) {
    // offsets recomputed, not kept live through the transform above
    ZOff<ZS> zn = zo;
    asm volatile("" : "+v"(zn.ob));
#pragma unroll
    for (int m = 0; m < E; ++m) {
      park[m * CC::BS + threadIdx.x] = nx_[m];
      nz_[m] = ld_once<R, (CW >= 16)>(W2 + zn.at(rt + T * m));
    }
  }
  L::run(ny_, lds, opaque(rt), twl, 1, ColIdx<CW>{c});
  // The stage update's loads (v, du or the AB2 history) of all three
  // components of slot m are issued together, before any of the slot's
  // stores: V, DU and S may alias as far as the compiler knows, so it would
  // otherwise wait out one load round trip per component.
  constexpr bool NT = CW >= 16;
  auto stage_in = [&](int m, const ZOff<ZS>& zz, StageIn<R> (&dst)[3]) {
    const int z = L::out_pos(m, rt);
    const int iy = kwave(t.oth, g.ny), iz = kwave(z, g.nz);
    const int ik2 = t.kx * t.kx + iy * iy + iz * iz;
    // MODE 4 reads du only inside the cut
    const bool cut = (ik2 == 0 ? std::numeric_limits<R>::min() : (R)ik2) >= p.kc2;
    const uint32_t o = zz.at(z);
    dst[0] = stage_load<MODE, R, uint32_t, NT>(o, V, DU, S, cut);
    dst[1] = stage_load<MODE, R, uint32_t, NT>(o, V + fstride, DU + fstride, S + fstride, cut);
    dst[2] = stage_load<MODE, R, uint32_t, NT>(o, V + 2 * fstride, DU + 2 * fstride, S + 2 * fstride, cut);
  };
  L::run(nz_, lds, opaque(rt), twl, 1, ColIdx<CW>{c});
  // the epilogue recomputes its offsets (not kept live from the loads)
  ZOff<ZS> ze = zo;
  asm volatile("" : "+v"(ze.ob));

  const int ikx = t.kx, iky = kwave(t.oth, g.ny);
  const R kx = (R)ikx, ky = (R)iky;
  const R tiny = std::numeric_limits<R>::min();
  const bool plane_zero = (t.oth == g.ny / 2) || (t.kx == g.mx);
  R fa = 0, fb = 0;
  if (p.forced) {
    fa = (R)coef[0];
    fb = (R)coef[1];
  }
  R ab_dt = 0, ab_dto = 0, ab_s = 0, ab_ek = 0, ab_del = 0;
  if constexpr (MODE >= 5) {
    ab_dt = (R)p.ab[AB_DT];
    ab_dto = (R)p.ab[AB_DTO];
    ab_s = (R)p.ab[AB_S];
    ab_ek = (R)p.ab[AB_EK];
    ab_del = (R)p.ab[AB_DELTA];
  }
  // With the park, the new x stage state goes back to its park slot (just
  // read) until the z-inverse, which keeps 2E registers free here.
  auto keep_x = [&](int m, cpx<R> v) {
    if constexpr (PARK && FUSED) park[m * CC::BS + threadIdx.x] = v;
    else nx_[m] = v;
  };
#pragma unroll
  for (int m = 0; m < E; ++m) {
    StageIn<R> in[3];
    if constexpr (MODE > 0) stage_in(m, ze, in);
    const int z = L::out_pos(m, rt);
    const int ikz = kwave(z, g.nz);
    const R kz = (R)ikz;
    const int ik2 = ikx * ikx + iky * iky + ikz * ikz;
    const R k2 = ik2 == 0 ? tiny : (R)ik2;
    const uint32_t o = ze.at(z);
    cpx<R> a = PARK ? park[m * CC::BS + threadIdx.x] : nx_[m], b = ny_[m], d = nz_[m];
    // forcing (force.f90:25-29 / force_heli.f90:44-52)
    if (p.forced && k2 < p.fk2) {
      const cpx<R> sx = st[o], sy = st[fstride + o], sz = st[2 * fstride + o];
      a = a + fa * sx;
      b = b + fa * sy;
      d = d + fa * sz;
      if (p.forced == 2) {
        a = a + fb * mul_i(ky * sz - kz * sy);
        b = b + fb * mul_i(kz * sx - kx * sz);
        d = d + fb * mul_i(kx * sy - ky * sx);
      }
    }
    // symmetrize.f90:41-48: ky=-ny/2, kz=-nz/2, kx=nx/2 planes and the mean
    if (plane_zero || z == g.nz / 2 || ik2 == 0) {
      a = mk<R>(0, 0);
      b = a;
      d = a;
    }
    // projection.f90:16-30
    {
      const R den = kx * kx + ky * ky + kz * kz + tiny;
      const cpx<R> q = mk<R>((kx * a.x + ky * b.x + kz * d.x) / den, (kx * a.y + ky * b.y + kz * d.y) / den);
      a = a - kx * q;
      b = b - ky * q;
      d = d - kz * q;
    }
    if (p.dealias && k2 >= p.kc2) {
      a = mk<R>(0, 0);
      b = a;
      d = a;
    }
    if constexpr (MODE == 0) {
      W[o] = a;
      W1[o] = b;
      W2[o] = d;
    } else if constexpr (MODE >= 5) {
      const R xk = -p.nu * k2;
      const R e1 = expm1_r(xk * ab_dt);
      const R e2 = MODE == 5 ? (R)0 : expm1_r(xk * ab_dto);
      const bool cut = k2 >= p.kc2 || ik2 == 0;
      keep_x(m, ab_apply<MODE, R>(in[0], a, o, V, DU, S, ab_dt, ab_s, ab_ek, ab_del, e1, e2, cut, p.nonfinite));
      ny_[m] = ab_apply<MODE, R>(in[1], b, o, V + fstride, DU + fstride, S + fstride, ab_dt, ab_s, ab_ek, ab_del, e1,
                                 e2, cut, p.nonfinite);
      nz_[m] = ab_apply<MODE, R>(in[2], d, o, V + 2 * fstride, DU + 2 * fstride, S + 2 * fstride, ab_dt, ab_s, ab_ek,
                                 ab_del, e1, e2, cut, p.nonfinite);
    } else {
      // E = exp(x), x = -nu k2 dt/2.  E is within 1e-3 of 1 for most modes,
      // where FP32 spacing (6e-8) makes a rounded E or E*E a biased factor
      // that compounds every step; E*w is applied as w + expm1(x)*w, exact
      // to the rounding of the result.
      const R x = -p.nu * k2 * p.dt / (R)2;
      const R em1 = expm1_r(x);
      const bool cut = k2 >= p.kc2;
      const bool store_s = !FUSED || (p.forced && k2 < p.fk2);
      keep_x(m, stage_apply<MODE, R>(in[0], p.dt * a, o, V, DU, S, em1, x, cut, store_s, p.nonfinite));
      ny_[m] = stage_apply<MODE, R>(in[1], p.dt * b, o, V + fstride, DU + fstride, S + fstride, em1, x, cut, store_s,
                                    p.nonfinite);
      nz_[m] = stage_apply<MODE, R>(in[2], p.dt * d, o, V + 2 * fstride, DU + 2 * fstride, S + 2 * fstride, em1, x,
                                    cut, store_s, p.nonfinite);
    }
  }
  if constexpr (FUSED && MODE > 0) {
    if constexpr (PARK) {
#pragma unroll
      for (int m = 0; m < E; ++m) nx_[m] = park[m * CC::BS + threadIdx.x];
    }
    zinv_store<R, N, E, CW, PARK, ZS>(nx_, ny_, nz_, lds, park, twl, rt, c, kx, ky, g.nz, zo, W, fstride);
  }
}

// Forcing coefficients from the stage state: a reduction over the box of
// modes that can satisfy k2 < fk2 (force.f90:16-24, force_heli.f90:15-42),
// accumulated in FP64 in a fixed order.  coef = {a, b, E_f, Omega_f, uw}.
template <class R>
__global__ __launch_bounds__(256) void k_force_coef(const cpx<R>* __restrict__ st, int64_t fstride, Geom g,
                                                    R fk2, int fbox, double eps, double eta, int type,
                                                    double* __restrict__ coef) {
  __shared__ double red[3][256];
  const int nxb = min(fbox, g.mx) + 1;
  const int nyb = min(2 * fbox + 1, g.ny), nzb = min(2 * fbox + 1, g.nz);
  const int total = nxb * nyb * nzb;
  double e = 0, en = 0, uw = 0;
  const R tiny = std::numeric_limits<R>::min();
  for (int i = threadIdx.x; i < total; i += 256) {
    const int ix = i % nxb, jy = (i / nxb) % nyb, jz = i / (nxb * nyb);
    const int iy = nyb == g.ny ? jy : (jy - fbox + g.ny) % g.ny;
    const int iz = nzb == g.nz ? jz : (jz - fbox + g.nz) % g.nz;
    const int kx = ix, ky = kwave(iy, g.ny), kz = kwave(iz, g.nz);
    const int ik2 = kx * kx + ky * ky + kz * kz;
    const R k2 = ik2 == 0 ? tiny : (R)ik2;
    if (!(k2 < fk2)) continue;
    const int64_t o = ix < g.mx ? g.row(iy, iz) * g.mx + ix : g.nmain + g.row(iy, iz);
    const cpx<R> sx = st[o], sy = st[o + fstride], sz = st[o + 2 * fstride];
    const double wgt = ix == 0 ? 0.5 : 1.0;
    const double X[3][2] = {{sx.x, sx.y}, {sy.x, sy.y}, {sz.x, sz.y}};
    e += wgt * (X[0][0] * X[0][0] + X[0][1] * X[0][1] + X[1][0] * X[1][0] + X[1][1] * X[1][1] + X[2][0] * X[2][0] +
                X[2][1] * X[2][1]);
    if (type == 2) {
      // w = i k x s in FP64
      const double k[3] = {(double)kx, (double)ky, (double)kz};
      double wv[3][2];
      for (int q = 0; q < 3; ++q) {
        const int a1 = (q + 1) % 3, a2 = (q + 2) % 3;
        const double re = k[a1] * X[a2][0] - k[a2] * X[a1][0];
        const double im = k[a1] * X[a2][1] - k[a2] * X[a1][1];
        wv[q][0] = -im;
        wv[q][1] = re;
      }
      for (int q = 0; q < 3; ++q) {
        en += wgt * (wv[q][0] * wv[q][0] + wv[q][1] * wv[q][1]);
        uw += wgt * 2.0 * (X[q][0] * wv[q][0] + X[q][1] * wv[q][1]);
      }
    }
  }
  red[0][threadIdx.x] = e;
  red[1][threadIdx.x] = en;
  red[2][threadIdx.x] = uw;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
      red[2][threadIdx.x] += red[2][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double E = red[0][0], O = red[1][0], U = red[2][0];
    if (type == 2) {
      const double den = 4.0 * E * O - U * U;
      coef[0] = (4.0 * eps * O - eta * U) / (2.0 * den);
      coef[1] = (eta * E - eps * U) / den;
    } else {
      coef[0] = eps / (2.0 * E);
      coef[1] = 0.0;
    }
    coef[2] = E;
    coef[3] = O;
    coef[4] = U;
  }
}

// IF-AB2 energy of v for the umax refresh (spectral_dns_dealiased.f90:
// 129-142: sum' |v|^2, half weight on kx = 0), per-workgroup FP64 partials;
// only on steps that refresh umax (the step count is read on the device,
// so a captured step decides by itself).
template <class R>
__global__ __launch_bounds__(256) void k_ab_energy(const cpx<R>* __restrict__ v, int64_t fstride, Geom g,
                                                   const double* __restrict__ ab, int ieout,
                                                   double* __restrict__ part) {
  const int64_t istep = (int64_t)ab[AB_ISTEP];
  if (ieout <= 0 || istep % ieout != 0) return;
  double e = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < g.ntot; i += (int64_t)gridDim.x * 256) {
    const bool x0 = i < g.nmain && i % g.mx == 0;
    const cpx<R> a = v[i], b = v[i + fstride], c = v[i + 2 * fstride];
    const double t = (double)a.x * a.x + (double)a.y * a.y + (double)b.x * b.x + (double)b.y * b.y +
                     (double)c.x * c.x + (double)c.y * c.y;
    e += x0 ? 0.5 * t : t;
  }
  __shared__ double red[256];
  red[threadIdx.x] = e;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Constants of the IF-AB2 stepper.
struct AbConsts {
  int stepper;      // 1 CFL steps, 2 fixed dt
  int ieout;
  double betadx;    // beta pi / lx
  double dtw;
  double myeps;     // the output-time tolerance (constant_sp.f90: 2 epsilon(1._sp))
  double dt_fixed;
};

// Start of an IF-AB2 step (one thread): the umax refresh, the output time,
// dt (spectral_dns_dealiased.f90:129-149, 226-232; fixed dt with the half
// steps of specdns-forced-omp-ft.f90:224-246) and the weights of the update.
static __global__ void k_ab_begin(double* __restrict__ ab, const double* __restrict__ part, int nparts, AbConsts c) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int64_t istep = (int64_t)ab[AB_ISTEP];
  const double time = ab[AB_TIME], dto = ab[AB_DTO];
  double dt;
  if (c.stepper == 2) {
    dt = istep < 2 ? 0.5 * c.dt_fixed : c.dt_fixed;
  } else {
    if (c.ieout > 0 && istep % c.ieout == 0) {
      double e = 0;
      for (int b = 0; b < nparts; ++b) e += part[b];
      ab[AB_UMAX] = 3.5 * sqrt(2.0 * e);
    }
    if (c.dtw > 0 && fabs(time - ab[AB_TWRITE]) <= c.myeps) ab[AB_TWRITE] += c.dtw;
    dt = fmin(c.betadx / ab[AB_UMAX], 3.0 * dto);
    if (c.dtw > 0) {
      const double rem = ab[AB_TWRITE] - time;
      if (dt >= rem) dt = rem;
      else if (rem - dt < 0.2 * dt) dt = 0.5 * rem;
    }
  }
  double S = 0, ek = 0, del = 0;
  if (istep == 1) {
    const double dtt = dt / dto;
    S = 0.5 * (1.0 + dtt * dtt);
    ek = 0.5 * (1.0 - dtt * dtt);
    del = 0.5 * dto * (1.0 + dtt) * (1.0 + dtt);
  } else if (istep >= 2) {
    const double dtt = dt / dto;
    S = (1.0 + 0.5 * dtt) * dt;
    ek = -0.5 * dtt * dt;
  }
  ab[AB_DT] = dt;
  ab[AB_S] = S;
  ab[AB_EK] = ek;
  ab[AB_DELTA] = del;
}

// End of an IF-AB2 step: time += dt, dto (step 0: dt; step 1: dt + dto;
// later: dt), the step count.
static __global__ void k_ab_end(double* __restrict__ ab) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int64_t istep = (int64_t)ab[AB_ISTEP];
  const double dt = ab[AB_DT];
  ab[AB_TIME] += dt;
  ab[AB_DTO] = istep == 1 ? dt + ab[AB_DTO] : dt;
  ab[AB_ISTEP] = (double)(istep + 1);
}

// Hermitian copy on the kx = 0 plane (symmetrize.f90:51-61): for kz < 0
// (z in [nz/2+1, nz-1]) and, on kz = 0, for ky < 0, c(0,y,z) = conj(c(0,-y,-z)).
template <class R>
__global__ __launch_bounds__(256) void k_herm_fix(FieldPtrs<R> fp, Geom g) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t n1 = (int64_t)(g.nz / 2 - 1) * g.ny;
  const int64_t n2 = g.ny / 2 - 1;
  int y, z, yp, zp;
  if (i < n1) {
    z = g.nz / 2 + 1 + (int)(i / g.ny);
    y = (int)(i % g.ny);
    zp = g.nz - z;
    yp = (g.ny - y) % g.ny;
  } else if (i < n1 + n2) {
    z = 0;
    y = g.ny / 2 + 1 + (int)(i - n1);
    zp = 0;
    yp = g.ny - y;
  } else {
    return;
  }
  const int64_t od = g.row(y, z) * g.mx, os = g.row(yp, zp) * g.mx;
  for (int f = 0; f < fp.n; ++f) fp.p[f][od] = conj(fp.p[f][os]);
}

inline int64_t herm_count(const Geom& g) { return (int64_t)(g.nz / 2 - 1) * g.ny + g.ny / 2 - 1; }

// ----------------------------------------------------------------------------
// x-rows.  A row of nx reals is the main row of M = nx/2 complex (packed
// x[2n] + i x[2n+1]) and, in spectral form, X[0..M-1] plus X[M] in nyq.
// ----------------------------------------------------------------------------
// T = M/8 threads own a row (T <= 64, so a row never spans two wavefronts);
// thread rt holds points rt + T*m.  The last Stockham pass leaves its
// outputs in that same layout (slot m = index rt + T*m), so an inverse
// transform feeds the pointwise product and the forward transform with no
// data exchange, and the index pairs (k, M-k) of the real-FFT packing are
// one lane shuffle apart.
template <class R, int M>
struct RowCfg {
  static constexpr int E = 8;
  static constexpr int TPR = M / 8;
#ifndef SCL_ROW_BS
#define SCL_ROW_BS 64
#endif
  // one wavefront per workgroup: a row's exchanges never wait on other waves
  static constexpr int BS = SCL_ROW_BS;
  static constexpr int RPW = BS / TPR;
};

template <class R>
__device__ inline cpx<R> shfl_c(cpx<R> v, int src, int width) {
  return mk<R>(__shfl(v.x, src, width), __shfl(v.y, src, width));
}

// Entry at index (M - k) mod M for k = rt + T*m of a row in the entry
// layout: lane T - rt, slot 7 - m when rt > 0; own slot (8 - m) mod 8 when
// rt = 0.  Every lane of the row must call it (it shuffles).
template <class R, int M>
__device__ inline cpx<R> row_partner(const cpx<R> (&v)[8], int m, int rt) {
  constexpr int T = M / 8;
  if constexpr (T == 1) {
    return v[(8 - m) & 7];
  } else {
    const cpx<R> s = shfl_c(v[7 - m], (T - rt) & (T - 1), T);
    return rt == 0 ? v[(8 - m) & 7] : s;
  }
}

// From X[0..M] (x: X[rt + T*m]; xm: X[M], read by rt = 0), builds
// Z[k] = E_k + i O_k, E_k = X[k] + conj X[M-k],
// O_k = (X[k] - conj X[M-k]) exp(+2 pi i k/nx), whose length-M inverse is
// x[2n] + i x[2n+1].  Im X[0] and Im X[M] are ignored, as by FFTW's
// halfcomplex c2r.
template <class R, int M>
__device__ inline void row_c2r_prep(cpx<R> (&v)[8], const cpx<R> (&x)[8], cpx<R> xm, int rt, const cpx<R>* tw,
                                    int twsN) {
  constexpr int T = M / 8;
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int k = rt + T * m;
    cpx<R> a = x[m];
    cpx<R> b = row_partner<R, M>(x, m, rt);
    if (m == 0 && rt == 0) {
      b = xm;
      a.y = 0;
      b.y = 0;
    }
    const cpx<R> bc = conj(b);
    const cpx<R> e = a + bc;
    const cpx<R> o = cmul<true>(a - bc, tw[k * twsN]);
    v[m] = e + mul_i(o);
  }
}

// After the length-M forward FFT of packed reals (z: Z[rt + T*m]):
// X[k] = (Zk + conj Z_{M-k})/2 - (i/2) exp(-2 pi i k/nx) (Zk - conj Z_{M-k}),
// k = 0..M-1 to the main row, X[M] = Re Z0 - Im Z0 to nyq.
template <class R, int M>
__device__ inline void row_r2c_post(const cpx<R> (&z)[8], cpx<R>* mainrow, cpx<R>* nyqp, int rt, const cpx<R>* tw,
                                    int twsN) {
  constexpr int T = M / 8;
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int k = rt + T * m;
    const cpx<R> zk = z[m], zc = conj(row_partner<R, M>(z, m, rt));
    const cpx<R> s = zk + zc, d = zk - zc;
    mainrow[k] = (R)0.5 * (s - mul_i(tw[k * twsN] * d));
  }
  if (rt == 0) *nyqp = mk<R>(z[0].x - z[0].y, (R)0);
}

// XPASS: per row, c2r of 6 fields, u x w * scale, r2c of 3, in place into
// fields 0..2.  Fields at W + f*fstride.  TAPE: u, w in physical space also
// to tape (stride fstride, the layout k_row_c2r leaves: the adjoint's record
// of the forward).
template <class R, int M, bool TAPE = false>
__global__ __launch_bounds__((RowCfg<R, M>::BS)) void k_row_convec(cpx<R>* W, int64_t fstride, Geom g, R scale,
                                                                 const cpx<R>* __restrict__ tw, int twsM, int twsN,
                                                                 cpx<R>* __restrict__ tape) {
  using C = RowCfg<R, M>;
  constexpr int T = C::TPR;
  __shared__ cpx<R> lds[C::RPW * lds_pad(M)];
  // twiddles of the length-M FFT and of the real-FFT packing, once per
  // workgroup (the grid is persistent: each workgroup walks row groups
  // blockIdx.x, blockIdx.x + gridDim.x, ...)
  __shared__ cpx<R> twM[M], twN[M];
  for (int i = threadIdx.x; i < M; i += C::BS) {
    twM[i] = tw[i * twsM];
    twN[i] = tw[i * twsN];
  }
  __syncthreads();
  const int slot = threadIdx.x / T, rt = threadIdx.x % T;
  cpx<R>* lrow = lds + slot * lds_pad(M);
  using L = Line<R, M, 8, +1>;
  using LF = Line<R, M, 8, -1>;
  const int64_t ngroups = g.nrow / C::RPW;
  for (int64_t grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
    const int64_t row = grp * C::RPW + slot;
    // all six rows are requested before any transform starts
    cpx<R> raw[6][8], xm[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      const cpx<R>* base = W + f * fstride;
#pragma unroll
      for (int m = 0; m < 8; ++m) raw[f][m] = base[row * M + rt + T * m];
      xm[f] = rt == 0 ? base[g.nmain + row] : mk<R>(0, 0);
    }
    cpx<R> a[6][8];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      row_c2r_prep<R, M>(a[f], raw[f], xm[f], rt, twN, 1);
      L::run(a[f], lrow, rt, twM, 1, LinIdx{});
      if constexpr (TAPE) {
        cpx<R>* tb = tape + f * fstride + row * M;
#pragma unroll
        for (int m = 0; m < 8; ++m) tb[L::out_pos(m, rt)] = a[f][m];
      }
    }
    // u x w on the packed reals: real and imaginary parts are the even and
    // odd x points (convec_dns.f90:30-42)
    cpx<R> p[3][8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const cpx<R> ux = a[0][m], uy = a[1][m], uz = a[2][m], wx = a[3][m], wy = a[4][m], wz = a[5][m];
      p[0][m] = mk<R>((uy.x * wz.x - uz.x * wy.x) * scale, (uy.y * wz.y - uz.y * wy.y) * scale);
      p[1][m] = mk<R>((uz.x * wx.x - ux.x * wz.x) * scale, (uz.y * wx.y - ux.y * wz.y) * scale);
      p[2][m] = mk<R>((ux.x * wy.x - uy.x * wx.x) * scale, (ux.y * wy.y - uy.y * wx.y) * scale);
    }
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      cpx<R>* base = W + f * fstride;
      LF::run(p[f], lrow, rt, twM, 1, LinIdx{});
      row_r2c_post<R, M>(p[f], base + row * M, base + g.nmain + row, rt, twN, 1);
    }
  }
}

// Standalone x-direction c2r (one field per blockIdx.y): spectral rows ->
// packed real rows in main.
template <class R, int M>
__global__ __launch_bounds__((RowCfg<R, M>::BS)) void k_row_c2r(cpx<R>* W, int64_t fstride, Geom g,
                                                              const cpx<R>* __restrict__ tw, int twsM, int twsN) {
  using C = RowCfg<R, M>;
  constexpr int T = C::TPR;
  __shared__ cpx<R> lds[C::RPW * lds_pad(M)];
  const int slot = threadIdx.x / T, rt = threadIdx.x % T;
  const int64_t row = (int64_t)blockIdx.x * C::RPW + slot;
  cpx<R>* lrow = lds + slot * lds_pad(M);
  cpx<R>* base = W + (int64_t)blockIdx.y * fstride;
  using L = Line<R, M, 8, +1>;
  cpx<R> raw[8], a[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) raw[m] = base[row * M + rt + T * m];
  const cpx<R> xm = rt == 0 ? base[g.nmain + row] : mk<R>(0, 0);
  row_c2r_prep<R, M>(a, raw, xm, rt, tw, twsN);
  L::run(a, lrow, rt, tw, twsM, LinIdx{});
#pragma unroll
  for (int m = 0; m < 8; ++m) base[row * M + L::out_pos(m, rt)] = a[m];
}

// Standalone x-direction r2c: packed real rows in main -> spectral rows.
template <class R, int M>
__global__ __launch_bounds__((RowCfg<R, M>::BS)) void k_row_r2c(cpx<R>* W, int64_t fstride, Geom g,
                                                              const cpx<R>* __restrict__ tw, int twsM, int twsN) {
  using C = RowCfg<R, M>;
  constexpr int T = C::TPR;
  __shared__ cpx<R> lds[C::RPW * lds_pad(M)];
  const int slot = threadIdx.x / T, rt = threadIdx.x % T;
  const int64_t row = (int64_t)blockIdx.x * C::RPW + slot;
  cpx<R>* lrow = lds + slot * lds_pad(M);
  cpx<R>* base = W + (int64_t)blockIdx.y * fstride;
  cpx<R> a[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) a[m] = base[row * M + rt + T * m];
  Line<R, M, 8, -1>::run(a, lrow, rt, tw, twsM, LinIdx{});
  row_r2c_post<R, M>(a, base + row * M, base + g.nmain + row, rt, tw, twsN);
}

// ----------------------------------------------------------------------------
// Layout conversion and small pointwise kernels (host-array entry points).
// ----------------------------------------------------------------------------
// ref (lx1, ly, lz) <-> main + nyq
template <class R>
__global__ __launch_bounds__(256) void k_ref_to_split(const cpx<R>* __restrict__ ref, cpx<R>* __restrict__ f, Geom g) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t n = g.nrow * (g.mx + 1);
  if (i >= n) return;
  const int64_t rr = i / (g.mx + 1);  // reference row z*ny + y
  const int x = (int)(i % (g.mx + 1));
  const int64_t row = g.row((int)(rr % g.ny), (int)(rr / g.ny));
  if (x < g.mx) f[row * g.mx + x] = ref[i];
  else f[g.nmain + row] = ref[i];
}

template <class R>
__global__ __launch_bounds__(256) void k_split_to_ref(const cpx<R>* __restrict__ f, cpx<R>* __restrict__ ref, Geom g) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t n = g.nrow * (g.mx + 1);
  if (i >= n) return;
  const int64_t rr = i / (g.mx + 1);  // reference row z*ny + y
  const int x = (int)(i % (g.mx + 1));
  const int64_t row = g.row((int)(rr % g.ny), (int)(rr / g.ny));
  ref[i] = x < g.mx ? f[row * g.mx + x] : f[g.nmain + row];
}

// Decode entry i of the split layout into (x, y, z).
__device__ inline void split_xyz(const Geom& g, int64_t i, int& x, int& y, int& z) {
  int64_t row;
  if (i < g.nmain) {
    row = i / g.mx;
    x = (int)(i % g.mx);
  } else {
    row = i - g.nmain;
    x = g.mx;
  }
  g.yz(row, y, z);
}

// padd (padd.f90:9-13) on the device, split layouts: every entry of the
// padded field gb is written, with the entry of g at the same signed
// wavenumber (kx <= nx/2, -ny/2 <= ky < ny/2, -nz/2 <= kz < nz/2) or 0.
template <class R>
__global__ __launch_bounds__(256) void k_pad(const cpx<R>* __restrict__ u, Geom g, cpx<R>* __restrict__ ub, Geom gb) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= gb.ntot) return;
  int x, y, z;
  split_xyz(gb, i, x, y, z);
  const int ky = kwave(y, gb.ny), kz = kwave(z, gb.nz);
  cpx<R> v = mk<R>(0, 0);
  if (x <= g.mx && ky >= -g.ny / 2 && ky < g.ny / 2 && kz >= -g.nz / 2 && kz < g.nz / 2) {
    const int64_t row = g.row((ky + g.ny) % g.ny, (kz + g.nz) % g.nz);
    v = u[x < g.mx ? row * g.mx + x : g.nmain + row];
  }
  ub[i] = v;
}

// unpadd (unpadd.f90:9-12): every entry of u from the padded entry at its
// signed wavenumber.
template <class R>
__global__ __launch_bounds__(256) void k_unpad(const cpx<R>* __restrict__ ub, Geom gb, cpx<R>* __restrict__ u, Geom g) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.ntot) return;
  int x, y, z;
  split_xyz(g, i, x, y, z);
  const int64_t row = gb.row((kwave(y, g.ny) + gb.ny) % gb.ny, (kwave(z, g.nz) + gb.nz) % gb.nz);
  u[i] = x < gb.mx ? ub[row * gb.mx + x] : ub[gb.nmain + row];
}

// The strain of smag_dealias.f90:21-26 as written there (off-diagonal
// terms without the 1/2 of the symmetric strain): 11, 12, 13, 22, 23, 33
// into S[q * fs].
template <class R>
__global__ __launch_bounds__(256) void k_strain_smagd(const cpx<R>* __restrict__ v, cpx<R>* __restrict__ S, int64_t fs,
                                                      Geom g) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.ntot) return;
  int x, y, z;
  split_xyz(g, i, x, y, z);
  const R kx = (R)x, ky = (R)kwave(y, g.ny), kz = (R)kwave(z, g.nz);
  const cpx<R> a = v[i], b = v[i + fs], c = v[i + 2 * fs];
  S[i] = mul_i(kx * a);
  S[i + fs] = mul_i(kx * b + ky * a);
  S[i + 2 * fs] = mul_i(kx * c + kz * a);
  S[i + 3 * fs] = mul_i(ky * b);
  S[i + 4 * fs] = mul_i(ky * c + kz * b);
  S[i + 5 * fs] = mul_i(kz * c);
}

// smag_dealias.f90:28-50 on the padded physical fields (packed reals of the
// main rows): |S| = sqrt(2 (S11^2 + S22^2 + S33^2 + 2 (S12^2 + S13^2 + S23^2)))
// and S_q <- -2 c1 delta^2 |S| S_q * scale, in place.
template <class R>
__global__ __launch_bounds__(256) void k_smagd_phys(cpx<R>* __restrict__ S, int64_t fs, int64_t n, R coef) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  cpx<R> s[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) s[q] = S[i + q * fs];
  const R ax = sqrt((R)2 * (s[0].x * s[0].x + s[3].x * s[3].x + s[5].x * s[5].x +
                            (R)2 * (s[1].x * s[1].x + s[2].x * s[2].x + s[4].x * s[4].x)));
  const R ay = sqrt((R)2 * (s[0].y * s[0].y + s[3].y * s[3].y + s[5].y * s[5].y +
                            (R)2 * (s[1].y * s[1].y + s[2].y * s[2].y + s[4].y * s[4].y)));
#pragma unroll
  for (int q = 0; q < 6; ++q) S[i + q * fs] = mk<R>(coef * ax * s[q].x, coef * ay * s[q].y);
}

// symmetrize.f90:41-48 zeroing (the Hermitian copy is k_herm_fix), and,
// when thr > 0, dealiasing.f90:35's k2 >= thr truncation.
template <class R>
__global__ __launch_bounds__(256) void k_sym_zero(cpx<R>* __restrict__ f, Geom g, R thr) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.ntot) return;
  int x, y, z;
  split_xyz(g, i, x, y, z);
  const int kx = x, ky = kwave(y, g.ny), kz = kwave(z, g.nz);
  const int ik2 = kx * kx + ky * ky + kz * kz;
  const R k2 = ik2 == 0 ? std::numeric_limits<R>::min() : (R)ik2;
  if (x == g.mx || y == g.ny / 2 || z == g.nz / 2 || ik2 == 0 || (thr > 0 && k2 >= thr)) f[i] = mk<R>(0, 0);
}

template <class R>
__global__ __launch_bounds__(256) void k_dealias_mask(cpx<R>* __restrict__ f, Geom g, R thr) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.ntot) return;
  int x, y, z;
  split_xyz(g, i, x, y, z);
  const int kx = x, ky = kwave(y, g.ny), kz = kwave(z, g.nz);
  const int ik2 = kx * kx + ky * ky + kz * kz;
  const R k2 = ik2 == 0 ? std::numeric_limits<R>::min() : (R)ik2;
  if (k2 >= thr) f[i] = mk<R>(0, 0);
}

// projection.f90:16-30 on three fields at stride fstride
template <class R>
__global__ __launch_bounds__(256) void k_projection(cpx<R>* __restrict__ v, int64_t fstride, Geom g) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.ntot) return;
  int x, y, z;
  split_xyz(g, i, x, y, z);
  const R kx = (R)x, ky = (R)kwave(y, g.ny), kz = (R)kwave(z, g.nz);
  cpx<R> a = v[i], b = v[i + fstride], d = v[i + 2 * fstride];
  const R den = kx * kx + ky * ky + kz * kz + std::numeric_limits<R>::min();
  const cpx<R> q = mk<R>((kx * a.x + ky * b.x + kz * d.x) / den, (kx * a.y + ky * b.y + kz * d.y) / den);
  v[i] = a - kx * q;
  v[i + fstride] = b - ky * q;
  v[i + 2 * fstride] = d - kz * q;
}

// force (type 1): w += a v on k2 < fk2;  force_heli (type 2): f = a v + b w
// on k2 < fk2, 0 elsewhere and at k = 0 (written into w's fields).
template <class R>
__global__ __launch_bounds__(256) void k_force_apply(cpx<R>* __restrict__ w, const cpx<R>* __restrict__ v,
                                                     int64_t fstride, Geom g, const double* __restrict__ coef, R fk2,
                                                     int type) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.ntot) return;
  int x, y, z;
  split_xyz(g, i, x, y, z);
  const int ky = kwave(y, g.ny), kz = kwave(z, g.nz);
  const int ik2 = x * x + ky * ky + kz * kz;
  const R k2 = ik2 == 0 ? std::numeric_limits<R>::min() : (R)ik2;
  const R a = (R)coef[0], b = (R)coef[1];
  const bool in = k2 < fk2;
  for (int q = 0; q < 3; ++q) {
    const int64_t o = i + q * fstride;
    if (type == 1) {
      if (in) w[o] = w[o] + a * v[o];
    } else {
      w[o] = (in && ik2 != 0) ? a * v[o] + b * w[o] : mk<R>(0, 0);
    }
  }
}

}  // namespace scl
