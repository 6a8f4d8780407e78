// Register-resident Stockham FFT for one line (a column along y/z or an
// x-row), executed by a group of T = N/E threads that each hold E points.
//
// Thread rt holds, on entry, points rt + T*m (m = 0..E-1) of the line: a
// coalesced load pattern.  Each pass of radix RAD does E/RAD butterflies per
// thread in registers (Govindaraju et al., "High performance discrete Fourier
// transforms on graphics processors", SC'08, Stockham autosort form):
//   j = rt + T*b;  inputs j + r*N/RAD;  twiddle W_{Ns*RAD}^{r*(j mod Ns)};
//   outputs (j/Ns)*Ns*RAD + (j mod Ns) + r*Ns.
// Between passes the points go through LDS and come back in the entry
// pattern, so every pass reads its inputs from registers.  After the last
// pass slot m holds output index out_pos(m, rt); the caller stores from there.
//
// Twiddles come from one table tw[i] = exp(-2 pi i i/L) (L = largest line
// length of the plan, built in FP64 on the host); a line of length N reads it
// with stride L/N.  DIR = -1 is the forward (exp(-i..)) transform, +1 the
// unnormalised inverse.
#pragma once

#include "scl_core.hpp"

namespace scl {

template <int DIR, class R>
__device__ inline cpx<R> dir_i(cpx<R> a) {  // (DIR * i) * a
  return DIR > 0 ? mk<R>(-a.y, a.x) : mk<R>(a.y, -a.x);
}

template <int DIR, class R>
__device__ inline void bfly2(cpx<R>* a) {
  cpx<R> t = a[0] - a[1];
  a[0] = a[0] + a[1];
  a[1] = t;
}

template <int DIR, class R>
__device__ inline void bfly4(cpx<R>& a0, cpx<R>& a1, cpx<R>& a2, cpx<R>& a3) {
  cpx<R> t0 = a0 + a2, t1 = a0 - a2, t2 = a1 + a3, t3 = dir_i<DIR>(a1 - a3);
  a0 = t0 + t2;
  a2 = t0 - t2;
  a1 = t1 + t3;
  a3 = t1 - t3;
}

template <int DIR, class R>
__device__ inline void bfly4(cpx<R>* a) {
  bfly4<DIR>(a[0], a[1], a[2], a[3]);
}

// (x + iy) * exp(DIR * i * theta) with (c, s) = (cos theta, sin theta)
template <int DIR, class R>
__device__ inline cpx<R> rot(cpx<R> a, R c, R s) {
  return DIR > 0 ? mk<R>(a.x * c - a.y * s, a.y * c + a.x * s) : mk<R>(a.x * c + a.y * s, a.y * c - a.x * s);
}

template <int DIR, class R>
__device__ inline void bfly8(cpx<R>* a) {
  const R h = (R)0.70710678118654752440;
  cpx<R> e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
  cpx<R> o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
  bfly4<DIR>(e0, e1, e2, e3);
  bfly4<DIR>(o0, o1, o2, o3);
  o1 = rot<DIR>(o1, h, h);
  o2 = dir_i<DIR>(o2);
  o3 = rot<DIR>(o3, -h, h);
  a[0] = e0 + o0;
  a[4] = e0 - o0;
  a[1] = e1 + o1;
  a[5] = e1 - o1;
  a[2] = e2 + o2;
  a[6] = e2 - o2;
  a[3] = e3 + o3;
  a[7] = e3 - o3;
}

template <int DIR, class R>
__device__ inline void bfly16(cpx<R>* a) {
  const R c1 = (R)0.92387953251128675613, s1 = (R)0.38268343236508977173;
  const R h = (R)0.70710678118654752440;
  cpx<R> e[8], o[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    e[i] = a[2 * i];
    o[i] = a[2 * i + 1];
  }
  bfly8<DIR>(e);
  bfly8<DIR>(o);
  o[1] = rot<DIR>(o[1], c1, s1);
  o[2] = rot<DIR>(o[2], h, h);
  o[3] = rot<DIR>(o[3], s1, c1);
  o[4] = dir_i<DIR>(o[4]);
  o[5] = rot<DIR>(o[5], -s1, c1);
  o[6] = rot<DIR>(o[6], -h, h);
  o[7] = rot<DIR>(o[7], -c1, s1);
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    a[i] = e[i] + o[i];
    a[i + 8] = e[i] - o[i];
  }
}

template <int RAD, int DIR, class R>
__device__ inline void bfly(cpx<R>* a) {
  if constexpr (RAD == 2) bfly2<DIR>(a);
  else if constexpr (RAD == 4) bfly4<DIR>(a);
  else if constexpr (RAD == 8) bfly8<DIR>(a);
  else if constexpr (RAD == 16) bfly16<DIR>(a);
  else static_assert(RAD == 2, "unsupported radix");
}

template <int... RS>
struct RList {};

// Radix schedule per (line length N, points per thread E).
template <int N, int E>
struct PlanOf;
// columns (E = 16)
template <> struct PlanOf<16, 16> { using L = RList<16>; };
template <> struct PlanOf<32, 16> { using L = RList<8, 4>; };
template <> struct PlanOf<64, 16> { using L = RList<8, 8>; };
template <> struct PlanOf<128, 16> { using L = RList<16, 8>; };
template <> struct PlanOf<256, 16> { using L = RList<16, 16>; };
template <> struct PlanOf<512, 16> { using L = RList<8, 8, 8>; };
template <> struct PlanOf<1024, 16> { using L = RList<16, 8, 8>; };
// x-rows (E = 8, N = nx/2)
template <> struct PlanOf<8, 8> { using L = RList<8>; };
template <> struct PlanOf<16, 8> { using L = RList<8, 2>; };
template <> struct PlanOf<32, 8> { using L = RList<8, 4>; };
template <> struct PlanOf<64, 8> { using L = RList<8, 8>; };
template <> struct PlanOf<128, 8> { using L = RList<8, 4, 4>; };
template <> struct PlanOf<256, 8> { using L = RList<8, 8, 4>; };
template <> struct PlanOf<512, 8> { using L = RList<8, 8, 8>; };
// 1024-point z-columns at 8 points per thread (1024-thread workgroups)
template <> struct PlanOf<1024, 8> { using L = RList<8, 8, 4, 4>; };

// One Stockham pass in registers (no data exchange).
template <class R, int N, int E, int DIR, int NS, int RAD>
__device__ inline void pass_compute(cpx<R> (&v)[E], int rt, const cpx<R>* tw, int tws) {
  constexpr int T = N / E;
  constexpr int NB = E / RAD;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    cpx<R> a[RAD];
#pragma unroll
    for (int r = 0; r < RAD; ++r) a[r] = v[b + r * NB];
    if constexpr (NS > 1) {
      const int k = (rt + T * b) & (NS - 1);
#pragma unroll
      for (int r = 1; r < RAD; ++r) {
        const cpx<R> w = tw[(r * k * (N / (NS * RAD))) * tws];
        a[r] = cmul<(DIR > 0)>(a[r], w);
      }
    }
    bfly<RAD, DIR>(a);
#pragma unroll
    for (int r = 0; r < RAD; ++r) v[b + r * NB] = a[r];
  }
}

// Output index of register slot m after a pass (NS, RAD).
template <int N, int E, int NS, int RAD>
__device__ inline int pass_out_pos(int m, int rt) {
  constexpr int T = N / E;
  constexpr int NB = E / RAD;
  const int b = m % NB, r = m / NB;
  const int j = rt + T * b;
  return (j / NS) * NS * RAD + (j & (NS - 1)) + r * NS;
}

template <class R, int N, int E, int DIR, int NS, class L>
struct Passes;

template <class R, int N, int E, int DIR, int NS, int RAD, int... REST>
struct Passes<R, N, E, DIR, NS, RList<RAD, REST...>> {
  using Next = Passes<R, N, E, DIR, NS * RAD, RList<REST...>>;
  template <class IDX>
  __device__ static inline void run(cpx<R> (&v)[E], cpx<R>* lds, int rt, const cpx<R>* tw, int tws,
                                    const IDX& idx) {
    pass_compute<R, N, E, DIR, NS, RAD>(v, rt, tw, tws);
    if constexpr (sizeof...(REST) > 0) {
      constexpr int T = N / E;
#pragma unroll
      for (int m = 0; m < E; ++m) lds[idx(pass_out_pos<N, E, NS, RAD>(m, rt))] = v[m];
      __syncthreads();
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = lds[idx(rt + T * m)];
      __syncthreads();
      Next::run(v, lds, rt, tw, tws, idx);
    }
  }
  __device__ static inline int out_pos(int m, int rt) {
    if constexpr (sizeof...(REST) > 0) return Next::out_pos(m, rt);
    else return pass_out_pos<N, E, NS, RAD>(m, rt);
  }
};

// Full line transform.  LDS must hold N points addressed through idx; all
// threads of the block must call it (it synchronises the block).  On return
// the LDS region is free for reuse.
template <class R, int N, int E, int DIR>
struct Line {
  using P = Passes<R, N, E, DIR, 1, typename PlanOf<N, E>::L>;
  template <class IDX>
  __device__ static inline void run(cpx<R> (&v)[E], cpx<R>* lds, int rt, const cpx<R>* tw, int tws,
                                    const IDX& idx) {
    P::run(v, lds, rt, tw, tws, idx);
  }
  __device__ static inline int out_pos(int m, int rt) { return P::out_pos(m, rt); }
};

}  // namespace scl
