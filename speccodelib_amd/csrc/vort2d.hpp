// 2D vorticity ensembles (spectral_dns_2d.f90:94,161-188).
//
// B members of one N x N periodic grid are stored as one split-layout field
// of Geom(N, N, B): member b takes the place of the z index, so member b's
// row y is main row b*N + y and its kx = N/2 entry is nyq[b][y].  One RHS
// evaluation is three passes over the batch:
//   k_yinv2d     y-inverse of u = i ky w/k^2, v = -i kx w/k^2 and w
//                (spectral_dns_2d.f90:161-162,166-168, y part)
//   k_row_adv2d  per row: c2r of u, v, w (x part), the products u w and v w
//                scaled by 1/N^2, r2c of both (x part) (:166-175)
//   k_yfwd2d     y-forward of both products, N = -(i kx F + i ky G) plus
//                the band forcing eps/(conj w + tiny) (:177-185), the
//                pointwise part of symmetrize2d and the 2/3 cut (:188), the
//                kx = 0 Hermitian copy, then one Lawson IF-RK4 stage.
// The reference negates the divergence only inside the forced band (:183),
// which flips the advection sign elsewhere; here N = -div(u w) + f on every
// mode.  Its stepper is IF-AB2 (:192-219) on a field that the in-place c2r
// has already overwritten; the ensembles step with IF-RK4 instead.
#pragma once

#include <limits>

#include "kernels3d.hpp"

namespace scl {

// Column shape of the y-passes: CW columns of N points, E points per thread.
template <class R, int N, int E, int CW>
struct Col2D {
  static constexpr int T = N / E;
  static constexpr int BS = CW * T;
  static constexpr size_t lds_bytes = (size_t)(lds_pad(N) * CW + N) * sizeof(cpx<R>);
  // radix schedules exist for E = 8 and E = 16 up to N = 1024 (at 1024
  // points and E = 8, 8 columns make one 1024-thread workgroup: 4 waves
  // per SIMD where E = 16 runs 2)
  static constexpr bool valid = (E == 8 ? N <= 1024 : (E == 16 && N >= 16 && N <= 1024)) && BS <= 1024 &&
                                BS >= 1 && lds_bytes <= 160 * 1024 && CW <= N / 2;
};

// Integer power by squaring (exponent >= 0).
template <class R>
__device__ inline R ipow(R b, int e) {
  R r = (R)1;
  while (e > 0) {
    if (e & 1) r *= b;
    b *= b;
    e >>= 1;
  }
  return r;
}

// Damping rate of a mode: c_hyper k^n_hyper + c_hypo k^-n_hypo + damp, the
// reference's (k sigma)^nhyper + (nu/k)^nhypo (spectral_dns_2d.f90:94) with
// c_hyper = sigma^nhyper, c_hypo = nu^nhypo, plus linear damping.
template <class R>
struct Rate2D {
  R c_hyper, c_hypo, damp;
  int n_hyper, n_hypo;
  __device__ inline R operator()(R k2) const {
    auto kpow = [&](int n) -> R { return (n & 1) ? ipow(sqrt(k2), n) : ipow(k2, n / 2); };
    R r = damp;
    if (c_hyper != (R)0) r += c_hyper * kpow(n_hyper);
    if (c_hypo != (R)0) r += c_hypo / kpow(n_hypo);
    return r;
  }
};

template <class R>
struct Rhs2D {
  R dt;
  R fk2min, fk2max;  // forced band fkmin^2 <= k^2 <= fkmax^2
  R eps;             // forcing amplitude
  R kc2;             // 2/3 cut: zero where k^2 >= kc2
  int forced;
  Rate2D<R> rate;
};

// y-inverse of u = i ky w/k^2, v = -i kx w/k^2 and w (spectral_dns_2d.f90:
// 161-162,166-168, y part) from the vorticity column o held in registers
// (slot m = index rt + T m along y, as every transform leaves it), stored
// to the work fields w[0..2] at stride fstride.  One field is formed,
// transformed and stored at a time, so at most two columns of E points are
// live per thread.
template <class R, int N, int E, int CW>
__device__ inline void yinv2d_store(cpx<R> (&o)[E], cpx<R>* lds, const cpx<R>* twl, int rt, int c, const Tile& t,
                                    cpx<R>* __restrict__ w, int64_t fstride) {
  constexpr int T = N / E;
  const R kx = (R)t.kx;
  const R tiny = std::numeric_limits<R>::min();
  using L = Line<R, N, E, +1>;
  auto ik2r = [&](int m) {
    const int iky = kwave(rt + T * m, N);
    const int ik2 = t.kx * t.kx + iky * iky;
    return ik2 == 0 ? tiny : (R)ik2;
  };
  auto store = [&](cpx<R>(&v)[E], int f) {
    if (!t.ok) return;
#pragma unroll
    for (int m = 0; m < E; ++m) w[t.base + (int64_t)L::out_pos(m, rt) * t.rs + f * fstride] = v[m];
  };
  cpx<R> a[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const R ky = (R)kwave(rt + T * m, N);
    a[m] = mk<R>(-(ky * o[m].y) / ik2r(m), (ky * o[m].x) / ik2r(m));
  }
  L::run(a, lds, opaque(rt), twl, 1, ColIdx<CW>{c});
  store(a, 0);
#pragma unroll
  for (int m = 0; m < E; ++m) a[m] = mk<R>((kx * o[m].y) / ik2r(m), -(kx * o[m].x) / ik2r(m));
  L::run(a, lds, opaque(rt), twl, 1, ColIdx<CW>{c});
  store(a, 1);
  L::run(o, lds, opaque(rt), twl, 1, ColIdx<CW>{c});
  store(o, 2);
}

// exp(i pi s/n) for s = kx + ky: s/n is a dyadic rational, exact in R.
template <class R>
__device__ inline cpx<R> half_cell_phase(int s, int n) {
  R sn, cs;
  if constexpr (sizeof(R) == 4) sincospif((float)s / (float)n, &sn, &cs);
  else sincospi((double)s / (double)n, &sn, &cs);
  return mk<R>(cs, sn);
}

// YINV.  s: the stage vorticity; w: three work fields (u, v, w) at stride
// fstride.  shift: the fields are taken at points moved by half a cell along
// x and y, the spectrum times exp(-i pi (kx + ky)/N) (the phase shift of
// turb_2d.f90:369-371, in the normalised coefficients of the real field).
template <class R, int N, int E, int CW>
__global__ __launch_bounds__((Col2D<R, N, E, CW>::BS)) void k_yinv2d(const cpx<R>* __restrict__ s,
                                                                   cpx<R>* __restrict__ w, int64_t fstride, Geom g,
                                                                   const cpx<R>* __restrict__ tw, int tws,
                                                                   int shift) {
  using CC = Col2D<R, N, E, CW>;
  constexpr int T = CC::T;
  __shared__ cpx<R> lds[lds_pad(N) * CW];
  __shared__ cpx<R> twl[N];  // first read after a barrier inside the transform
  for (int i = threadIdx.x; i < N; i += CC::BS) twl[i] = tw[i * tws];
  const int c = threadIdx.x % CW, rt = threadIdx.x / CW;
  const Tile t = col_tile<CW, 1>(g, xcd_swizzle(blockIdx.x, gridDim.x), c);
  cpx<R> o[E];
#pragma unroll
  for (int m = 0; m < E; ++m) o[m] = t.ok ? s[t.base + (int64_t)(rt + T * m) * t.rs] : mk<R>(0, 0);
  if (shift) {
#pragma unroll
    for (int m = 0; m < E; ++m) o[m] = o[m] * conj(half_cell_phase<R>(t.kx + kwave(rt + T * m, N), N));
  }
  yinv2d_store<R, N, E, CW>(o, lds, twl, rt, c, t, w, fstride);
}

// XPASS: per row, c2r of u, v, w; u w and v w times scale; r2c of both
// into work fields 0 and 1.  Persistent grid over groups of RPW rows.
template <class R, int M>
__global__ __launch_bounds__((RowCfg<R, M>::BS)) void k_row_adv2d(cpx<R>* W, int64_t fstride, Geom g, R scale,
                                                                const cpx<R>* __restrict__ tw, int twsM, int twsN) {
  using C = RowCfg<R, M>;
  constexpr int T = C::TPR;
  __shared__ cpx<R> lds[C::RPW * lds_pad(M)];
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
    cpx<R> raw[3][8], xm[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      const cpx<R>* base = W + f * fstride;
#pragma unroll
      for (int m = 0; m < 8; ++m) raw[f][m] = base[row * M + rt + T * m];
      xm[f] = rt == 0 ? base[g.nmain + row] : mk<R>(0, 0);
    }
    cpx<R> a[3][8];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      row_c2r_prep<R, M>(a[f], raw[f], xm[f], rt, twN, 1);
      L::run(a[f], lrow, rt, twM, 1, LinIdx{});
    }
    // real and imaginary parts are the even and odd x points
    // (spectral_dns_2d.f90:170-171)
    cpx<R> p[2][8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const cpx<R> uu = a[0][m], vv = a[1][m], ww = a[2][m];
      p[0][m] = mk<R>(uu.x * ww.x * scale, uu.y * ww.y * scale);
      p[1][m] = mk<R>(vv.x * ww.x * scale, vv.y * ww.y * scale);
    }
#pragma unroll
    for (int f = 0; f < 2; ++f) {
      cpx<R>* base = W + f * fstride;
      LF::run(p[f], lrow, rt, twM, 1, LinIdx{});
      row_r2c_post<R, M>(p[f], base + row * M, base + g.nmain + row, rt, twN, 1);
    }
  }
}

// YFWD + stage.  W: the two product fields (in), WS: state w, ACC: stage
// accumulator, S: stage vorticity (written for the next stage), st: the
// vorticity of this stage (forcing), equal to WS in mode 1 and to S after.
// With E = exp(-rate dt/2) and a = dt N (division-free Lawson form):
//   mode 0: W[0] = N
//   mode 1: ACC = E^2 (w + a/6);  S = E (w + a/2)
//   mode 2: ACC += E a/3;         S = E w + a/2
//   mode 3: ACC += E a/3;         S = E (E w + a)
//   mode 4: w = ACC + a/6, zero where k^2 >= kc2
// which is dns-deform.f90:164-237's scheme multiplied through by E, E and
// E^2 (the reference form divides by E, which underflows under k^8
// hyperviscosity).  E x is applied as x + expm1(.) x.
// FUSED (modes 1..4): the new stage vorticity (S, or w in mode 4) stays in
// registers, its kx = 0 column made Hermitian as it is stored, and its
// y-inverse (yinv2d_store) goes to W[0..2] for the next RHS, which then
// needs no y-inverse pass; S is stored only in the forced band, where the
// forcing reads it.
template <class R, int N, int E, int CW, bool FUSED = false>
__global__ __launch_bounds__((Col2D<R, N, E, CW>::BS)) void k_yfwd2d(cpx<R>* W, int64_t fstride, cpx<R>* WS,
                                                                   cpx<R>* ACC, cpx<R>* S, const cpx<R>* st, int mode,
                                                                   Rhs2D<R> p, Geom g,
                                                                   const cpx<R>* __restrict__ tw, int tws) {
  using CC = Col2D<R, N, E, CW>;
  constexpr int T = CC::T;
  __shared__ cpx<R> lds[lds_pad(N) * CW];
  __shared__ cpx<R> twl[N];
  for (int i = threadIdx.x; i < N; i += CC::BS) twl[i] = tw[i * tws];
  const int c = threadIdx.x % CW, rt = threadIdx.x / CW;
  const int wg = xcd_swizzle(blockIdx.x, gridDim.x);
  const int xtn = g.mx / CW;
  const bool kx0_tile = wg < xtn * g.nz && wg % xtn == 0;  // uniform over the workgroup
  const Tile t = col_tile<CW, 1>(g, wg, c);
  using L = Line<R, N, E, -1>;
  cpx<R> f0[E], f1[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int64_t o = t.base + (int64_t)(rt + T * m) * t.rs;
    f0[m] = t.ok ? W[o] : mk<R>(0, 0);
    f1[m] = t.ok ? W[o + fstride] : mk<R>(0, 0);
  }
  L::run(f0, lds, opaque(rt), twl, 1, ColIdx<CW>{c});
  L::run(f1, lds, opaque(rt), twl, 1, ColIdx<CW>{c});

  const R tiny = std::numeric_limits<R>::min();
  const int ikx = t.kx;
  const R kx = (R)ikx;
  // Outputs go straight to memory.  In the tile holding the kx = 0 column,
  // lane c = 0 also stages its outputs in LDS; after a barrier it rewrites
  // its entries y > N/2 as c(0, -ky) = conj c(0, ky) (symmetrize.f90:25-27).
  // That equals symmetrizing N first: every output is a real-linear function
  // of N and of inputs that already have this symmetry.
  if (kx0_tile) __syncthreads();  // the transforms' LDS reads are done
  auto in_band = [&](int y) {
    const int iky = kwave(y, N);
    const int ik2 = ikx * ikx + iky * iky;
    const R k2 = ik2 == 0 ? tiny : (R)ik2;
    return p.forced && k2 >= p.fk2min && k2 <= p.fk2max;
  };
  auto put = [&](uint32_t y, cpx<R> v1, cpx<R> v2) {
    const int64_t o = t.base + (int64_t)y * t.rs;
    if (mode == 0) {
      W[o] = v1;
    } else if (mode == 4) {
      WS[o] = v1;
    } else {
      ACC[o] = v1;
      if (!FUSED || in_band(y)) S[o] = v2;
    }
  };
  cpx<R> ns[FUSED ? E : 1];  // the new stage vorticity (FUSED)
  // The stage state w and accumulator of slot m + 1 are loaded before the
  // stores of slot m (which the compiler would otherwise not move them
  // above), so each slot's loads are in flight through the slot before.
  auto stage_in = [&](int m, cpx<R>& wv, cpx<R>& acc) {
    wv = mk<R>(0, 0);
    acc = wv;
    if (mode == 0 || !t.ok) return;
    const int y = L::out_pos(m, rt);
    const int iky = kwave(y, N);
    const int ik2 = ikx * ikx + iky * iky;
    const int64_t o = t.base + (int64_t)y * t.rs;
    if (mode <= 3) wv = WS[o];
    if (mode == 2 || mode == 3 || (mode == 4 && !((ik2 == 0 ? tiny : (R)ik2) >= p.kc2))) acc = ACC[o];
  };
  cpx<R> wv_c, acc_c;
  stage_in(0, wv_c, acc_c);
#pragma unroll
  for (int m = 0; m < E; ++m) {
    cpx<R> wv_n, acc_n;
    if (m + 1 < E) stage_in(m + 1, wv_n, acc_n);
    const int y = L::out_pos(m, rt);
    const int iky = kwave(y, N);
    const R ky = (R)iky;
    const int ik2 = ikx * ikx + iky * iky;
    const R k2 = ik2 == 0 ? tiny : (R)ik2;
    const int64_t o = t.base + (int64_t)y * t.rs;
    // -(i kx F + i ky G) (spectral_dns_2d.f90:177-178)
    const cpx<R> d = kx * f0[m] + ky * f1[m];
    cpx<R> r = mk<R>(d.y, -d.x);
    if (p.forced && t.ok && k2 >= p.fk2min && k2 <= p.fk2max) {
      // eps/(conj(w) + tiny) (:182-183)
      const cpx<R> ws = st[o];
      const R a = ws.x + tiny, b = ws.y;
      const R den = a * a + b * b;
      r = r + mk<R>(p.eps * a / den, p.eps * b / den);
    }
    // symmetrize2d (symmetrize.f90:17-22) and the 2/3 cut
    if (y == N / 2 || t.kx == g.mx || ik2 == 0 || k2 >= p.kc2) r = mk<R>(0, 0);
    cpx<R> o1 = r, o2 = mk<R>(0, 0);
    if (mode != 0 && t.ok) {
      const R x = -p.rate(k2) * p.dt / (R)2;
      const R em1 = expm1_r(x);
      const cpx<R> wv = wv_c;
      const cpx<R> a = p.dt * r;
      if (mode == 1) {
        const cpx<R> u = wv + ((R)1 / (R)6) * a, q = wv + (R)0.5 * a;
        o1 = u + expm1_r((R)2 * x) * u;
        o2 = q + em1 * q;
      } else if (mode == 2 || mode == 3) {
        const cpx<R> q = ((R)1 / (R)3) * a;
        o1 = acc_c + (q + em1 * q);
        const cpx<R> ew = wv + em1 * wv;
        if (mode == 2) {
          o2 = ew + (R)0.5 * a;
        } else {
          const cpx<R> u = ew + a;
          o2 = u + em1 * u;
        }
      } else {
        o1 = k2 >= p.kc2 ? mk<R>(0, 0) : acc_c + ((R)1 / (R)6) * a;
      }
    }
    if (t.ok) put(y, o1, o2);
    if constexpr (FUSED) ns[m] = mode == 4 ? o1 : o2;
    if (kx0_tile && c == 0) {
      lds[y] = o1;
      lds[N + y] = o2;
    }
    wv_c = wv_n;
    acc_c = acc_n;
  }
  if (kx0_tile) {
    __syncthreads();
    if (c == 0) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int y = L::out_pos(m, rt);
        if (y > N / 2) {
          const cpx<R> v1 = conj(lds[N - y]), v2 = conj(lds[2 * N - y]);
          put(y, v1, v2);
          if constexpr (FUSED) ns[m] = mode == 4 ? v1 : v2;
        }
      }
    }
  }
  if constexpr (FUSED) {
    if (kx0_tile) __syncthreads();  // the staged kx = 0 column is read
    yinv2d_store<R, N, E, CW>(ns, lds, twl, rt, c, t, W, fstride);
  }
}

// turb_2d.f90's step (one rank), in the normalised coefficients D of the
// real fields; its own coefficients are C = 2 conj(D) (Ooura transforms,
// exp(+i k.x) forward, kx = n/2 dropped), so i k C <-> -i k D, the phase
// exp(i phi) C <-> exp(-i phi) D and etak/conj(C) <-> (etak/4)/conj(D).
template <class R>
struct Turb2DP {
  R dt;
  R sigma, nu;  // L = (k sigma)^16 + (k/nu)^-4
  R etak4;      // etak/4, etak = eta/104
  R kf;         // forced band kf <= |k| <= kf + 3
  R kc2;        // dealiasing: zero where k2 >= (2/9) n^2
  int forced;
};

enum { TURB_HALF = 0, TURB_STEP0 = 1, TURB_STEP1 = 2, TURB_AB2 = 3, TURB_RHS = 4 };

// YFWD of one half of the phase-shift dealiased nonlinear term
// (turb_2d.f90:287-421) and the stepper (:430-483).  W: the two product
// fields; NA: the unshifted half; WS: state w; OWZ: the previous nonlinear
// term owz; W0: the state before the first step.
//   TURB_HALF: NA = -(i/2)(kx F + ky G) of the unshifted products;
//   otherwise N = NA - (i/2) exp(i phi)(kx F + ky G) of the shifted ones,
//   zeroed on k2 >= kc2, ky = -n/2, kx = n/2 and k = 0 (dealiasing and
//   symmetrize, :428-429), then with f(a) = etak4/conj(a) in the band:
//   TURB_STEP0: W0 = w; OWZ = N; w = w (1 - L dt/2) + dt/2 (N + f(w))
//   TURB_STEP1: w = W0 + dt (N + f(w)) - L dt w
//   TURB_AB2:   w = [w + dt/2 (3 (N + f(w)) - E (owz + f(owz)))] E,
//               E = exp(-L dt) (1 at k = 0); OWZ = N
//   TURB_RHS:   NA = N.
// The kx = 0 column of N and of the new w is made Hermitian as in k_yfwd2d.
template <class R, int N, int E, int CW>
__global__ __launch_bounds__((Col2D<R, N, E, CW>::BS)) void k_yfwd2d_turb(
    const cpx<R>* W, int64_t fstride, cpx<R>* NA, cpx<R>* WS, cpx<R>* OWZ, cpx<R>* W0, int mode, Turb2DP<R> p, Geom g,
    const cpx<R>* __restrict__ tw, int tws) {
  using CC = Col2D<R, N, E, CW>;
  constexpr int T = CC::T;
  __shared__ cpx<R> lds[lds_pad(N) * CW];
  __shared__ cpx<R> twl[N];
  for (int i = threadIdx.x; i < N; i += CC::BS) twl[i] = tw[i * tws];
  const int c = threadIdx.x % CW, rt = threadIdx.x / CW;
  const int wg = xcd_swizzle(blockIdx.x, gridDim.x);
  const int xtn = g.mx / CW;
  const bool kx0_tile = wg < xtn * g.nz && wg % xtn == 0;
  const Tile t = col_tile<CW, 1>(g, wg, c);
  using L = Line<R, N, E, -1>;
  cpx<R> f0[E], f1[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int64_t o = t.base + (int64_t)(rt + T * m) * t.rs;
    f0[m] = t.ok ? W[o] : mk<R>(0, 0);
    f1[m] = t.ok ? W[o + fstride] : mk<R>(0, 0);
  }
  L::run(f0, lds, rt, twl, 1, ColIdx<CW>{c});
  L::run(f1, lds, rt, twl, 1, ColIdx<CW>{c});
  if (kx0_tile) __syncthreads();  // the transforms' LDS reads are done
  const int ikx = t.kx;
  const R kx = (R)ikx;
  auto force = [&](cpx<R> a) {
    const R den = a.x * a.x + a.y * a.y;
    return mk<R>(p.etak4 * a.x / den, p.etak4 * a.y / den);
  };
  auto put = [&](int y, cpx<R> v1, cpx<R> v2) {
    const int64_t o = t.base + (int64_t)y * t.rs;
    if (mode == TURB_HALF || mode == TURB_RHS) {
      NA[o] = v1;
    } else {
      WS[o] = v1;
      if (mode != TURB_STEP1) OWZ[o] = v2;
    }
  };
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int y = L::out_pos(m, rt);
    const int iky = kwave(y, N);
    const int ik2 = ikx * ikx + iky * iky;
    const R k2 = ik2 == 0 ? (R)0.5 : (R)ik2;  // wavenumber(): k2(0, 0) = 0.5
    const int64_t o = t.base + (int64_t)y * t.rs;
    const cpx<R> d = kx * f0[m] + (R)iky * f1[m];
    cpx<R> r = mk<R>((R)0.5 * d.y, (R)-0.5 * d.x);  // -(i/2) d
    cpx<R> o1 = r, o2 = mk<R>(0, 0);
    if (mode != TURB_HALF) {
      r = (t.ok ? NA[o] : mk<R>(0, 0)) + r * half_cell_phase<R>(ikx + iky, N);
      if (y == N / 2 || ikx == g.mx || ik2 == 0 || k2 >= p.kc2) r = mk<R>(0, 0);
      o1 = r;
      if (mode != TURB_RHS && t.ok) {
        const cpx<R> w = WS[o];
        const R k = sqrt(k2);
        const R ks = k * p.sigma, kn = k / p.nu;
        const R lam = ipow(ks, 16) + (R)1 / ipow(kn, 4);
        const bool band = p.forced && k >= p.kf && k <= p.kf + (R)3;
        const R dth = (R)0.5 * p.dt;
        const cpx<R> fw = band ? force(w) : mk<R>(0, 0);
        if (mode == TURB_STEP0) {
          o1 = ((R)1 - lam * dth) * w + dth * (r + fw);
          o2 = r;
          W0[o] = w;
        } else if (mode == TURB_STEP1) {
          o1 = W0[o] + p.dt * (r + fw) - (lam * p.dt) * w;
        } else {
          const R ef = ik2 == 0 ? (R)1 : exp(-lam * p.dt);
          const cpx<R> ow = OWZ[o];
          const cpx<R> fo = band ? force(ow) : mk<R>(0, 0);
          o1 = ef * (w + dth * ((R)3 * (r + fw) - ef * (ow + fo)));
          o2 = r;
        }
      }
    }
    if (t.ok) put(y, o1, o2);
    if (kx0_tile && c == 0) {
      lds[y] = o1;
      lds[N + y] = o2;
    }
  }
  if (kx0_tile) {
    __syncthreads();
    if (c == 0) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int y = L::out_pos(m, rt);
        if (y > N / 2) put(y, conj(lds[N - y]), conj(lds[2 * N - y]));
      }
    }
  }
}

// Per-member FP64 sums in a fixed order: out[2b] = sum' |w|^2/k^2 (twice the
// kinetic energy's half-plane sum, as spectral_dns_2d.f90:125-128 with
// kx = 0 halved), out[2b+1] = sum' |w|^2.  One workgroup per member.
template <class R>
__global__ __launch_bounds__(256) void k_ens_diag(const cpx<R>* __restrict__ w, Geom g, double* __restrict__ out) {
  const int b = blockIdx.x;
  const int64_t nm = (int64_t)g.ny * g.mx;
  double e = 0.0, z = 0.0;
  for (int64_t i = threadIdx.x; i < nm + g.ny; i += 256) {
    int kx, y;
    cpx<R> v;
    if (i < nm) {
      kx = (int)(i % g.mx);
      y = (int)(i / g.mx);
      v = w[(int64_t)b * nm + i];
    } else {
      kx = g.mx;
      y = (int)(i - nm);
      v = w[g.nmain + (int64_t)b * g.ny + y];
    }
    const int ky = kwave(y, g.ny);
    const int ik2 = kx * kx + ky * ky;
    const double k2 = ik2 == 0 ? std::numeric_limits<double>::min() : (double)ik2;
    const double a = (double)v.x * v.x + (double)v.y * v.y;
    const double wt = kx == 0 ? 0.5 : 1.0;
    e += wt * a / k2;
    z += wt * a;
  }
  __shared__ double se[256], sz[256];
  se[threadIdx.x] = e;
  sz[threadIdx.x] = z;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      se[threadIdx.x] += se[threadIdx.x + s];
      sz[threadIdx.x] += sz[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * b] = se[0];
    out[2 * b + 1] = sz[0];
  }
}

}  // namespace scl
