// Discrete adjoint of the IF-RK4 step (dns-deform.f90:160-237 stepping) of
// the momentum RHS N(v) = P Sym [ FFT(u x w)/N^3 - i k_j tau_ij ] with the
// static Smagorinsky stress tau_ij = FFT(-2 cs2 D^2 |S| S_ij)/N^3
// (spectral_les.f90:205-238, dynsmag.f90:196-233), or without it.
//
// Gradients are taken in the real inner product of the stored half
// spectrum, <a, b> = sum Re(conj(a) b) over every stored entry, so the
// gradient of an entry is dL/dRe + i dL/dIm.  With that product:
//   adjoint of c2r (FFTW halfcomplex: Im of the kx = 0 and kx = nx/2 lines
//     ignored)  = r2c followed by a weight 2 on 0 < kx < nx/2, 1 on kx = 0
//     and kx = nx/2;
//   adjoint of r2c = c2r after the inverse weights (1/2 interior, 1 on the
//     two planes);
//   multiplication by a complex number a -> by conj(a): the curl i k x and
//     the projection are self-adjoint;
//   symmetrize (zero planes, then c(0,-k) = conj c(0,k)) -> for every copy
//     target t with source s: s += conj(t), t = 0; then zero the planes.
// One vector-Jacobian product J(s)^T y runs as prep (spectral) -> c2r ->
// pointwise (physical) -> r2c -> combine (spectral), like the forward RHS.
#pragma once

#include "les.hpp"

namespace scl {

// Field counts of one vector-Jacobian product.  Inputs to the inverse
// transforms: u(3) w(3) [S_ij(6)] ybar(3) [taubar_ij(6)]; outputs of the
// pointwise pass: ubar(3) wbar(3) [Sbar_ij(6)].
struct AdjLayout {
  int les;
  __host__ __device__ constexpr int nin() const { return 9 + 12 * les; }
  __host__ __device__ constexpr int nout() const { return 6 + 6 * les; }
  __host__ __device__ constexpr int y_in() const { return 6 + 6 * les; }
  __host__ __device__ constexpr int tau_in() const { return 9 + 6 * les; }
};

__device__ inline bool sym_zero_mode(const Geom& g, int x, int y, int z, int ik2) {
  return x == g.mx || y == g.ny / 2 || z == g.nz / 2 || ik2 == 0;
}

// v <- P v = v - k (k.v)/(k^2 + tiny)  (projection.f90:16-30)
template <class R>
__device__ inline void project(cpx<R> (&v)[3], R kx, R ky, R kz) {
  const R den = kx * kx + ky * ky + kz * kz + std::numeric_limits<R>::min();
  const cpx<R> q = mk<R>((kx * v[0].x + ky * v[1].x + kz * v[2].x) / den, (kx * v[0].y + ky * v[1].y + kz * v[2].y) / den);
  v[0] = v[0] - kx * q;
  v[1] = v[1] - ky * q;
  v[2] = v[2] - kz * q;
}

// Spectral inputs of J(s)^T: the stage state's u, w = i k x u, S_ij, and
// ybar = Z(Y) (Y = P nbar after the Hermitian transpose; Z zeroes the
// symmetrize planes), then taubar_ij from ybar (adjoint of -i k_j tau_ij),
// both times the inverse weights of the r2c adjoint.
template <class R>
__global__ __launch_bounds__(256) void k_adj_prep(const cpx<R>* __restrict__ s, const cpx<R>* __restrict__ Y,
                                                  cpx<R>* __restrict__ W, int64_t fs, Geom g, AdjLayout L) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.ntot) return;
  int x, y, z, ikx, iky, ikz;
  ks_of(g, i, x, y, z, ikx, iky, ikz);
  const R kx = (R)ikx, ky = (R)iky, kz = (R)ikz;
  const cpx<R> v[3] = {s[i], s[i + fs], s[i + 2 * fs]};
  W[i] = v[0];
  W[i + fs] = v[1];
  W[i + 2 * fs] = v[2];
  W[i + 3 * fs] = mul_i(ky * v[2] - kz * v[1]);
  W[i + 4 * fs] = mul_i(kz * v[0] - kx * v[2]);
  W[i + 5 * fs] = mul_i(kx * v[1] - ky * v[0]);
  if (L.les) {
    cpx<R> st[6];
    strain_of<R>(v, kx, ky, kz, st);
#pragma unroll
    for (int q = 0; q < 6; ++q) W[i + (6 + q) * fs] = st[q];
  }
  const bool zero = sym_zero_mode(g, x, y, z, ikx * ikx + iky * iky + ikz * ikz);
  const R wi = (x == 0 || x == g.mx) ? (R)1 : (R)0.5;
  cpx<R> yb[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) yb[q] = zero ? mk<R>(0, 0) : wi * Y[i + q * fs];
  const int yo = L.y_in();
  W[i + yo * fs] = yb[0];
  W[i + (yo + 1) * fs] = yb[1];
  W[i + (yo + 2) * fs] = yb[2];
  if (L.les) {
    // forward: N_x -= i (kx t11 + ky t12 + kz t13), N_y -= i (kx t12 + ky t22 + kz t23),
    //          N_z -= i (kx t13 + ky t23 + kz t33)
    const int t = L.tau_in();
    W[i + t * fs] = mul_i(kx * yb[0]);
    W[i + (t + 1) * fs] = mul_i(ky * yb[0] + kx * yb[1]);
    W[i + (t + 2) * fs] = mul_i(kz * yb[0] + kx * yb[2]);
    W[i + (t + 3) * fs] = mul_i(ky * yb[1]);
    W[i + (t + 4) * fs] = mul_i(kz * yb[1] + ky * yb[2]);
    W[i + (t + 5) * fs] = mul_i(kz * yb[2]);
  }
}

// Physical pointwise part, in place on the packed main rows.  With the
// forward's q = scale u x w and sigma_ij = c |S| S_ij (c = -2 cs2 D^2 scale):
//   qbar = scale * c2r(ybar);  ubar = w x qbar;  wbar = qbar x u;
//   Sbar_q = c (|S| sigmabar_q + 2 m_q S_q (sum_p sigmabar_p S_p)/|S|),
//   m = 1 on the diagonal, 2 off it;
// and the cs2 derivative sum over points of sigmabar_p dsigma_p/dcs2 =
// cc |S| sum_p sigmabar_p S_p (cc = -2 D^2 scale), per block added into
// part[] (each block owns its slot; one sum over the slots at the end).
// The state inputs u, w, S_ij come from P (the physical-space tape of the
// recomputed forward, stride fs) when it is given, else from W.  The
// layout is a template parameter so that the per-thread arrays stay in
// registers.
template <class R, int LES>
__global__ __launch_bounds__(256) void k_adj_phys(cpx<R>* __restrict__ W, int64_t fs, Geom g, R scale, R c, R cc,
                                                  double* __restrict__ part, const cpx<R>* __restrict__ P) {
  constexpr AdjLayout L{LES};
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double acc = 0.0;
  if (i < g.nmain) {
    cpx<R> in[21];
    const cpx<R>* const st = P ? P : W;
#pragma unroll
    for (int q = 0; q < 21; ++q)
      if (q < L.nin()) in[q] = (q < L.y_in() ? st : W)[i + q * fs];
    cpx<R> out[12];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      auto re = [&](int q) -> R { return h ? in[q].y : in[q].x; };
      R o[12];
      const R u0 = re(0), u1 = re(1), u2 = re(2), w0 = re(3), w1 = re(4), w2 = re(5);
      const int yo = L.y_in();
      const R q0 = scale * re(yo), q1 = scale * re(yo + 1), q2 = scale * re(yo + 2);
      o[0] = w1 * q2 - w2 * q1;
      o[1] = w2 * q0 - w0 * q2;
      o[2] = w0 * q1 - w1 * q0;
      o[3] = q1 * u2 - q2 * u1;
      o[4] = q2 * u0 - q0 * u2;
      o[5] = q0 * u1 - q1 * u0;
      if (L.les) {
        R sv[6], sb[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) {
          sv[q] = re(6 + q);
          sb[q] = re(L.tau_in() + q);
        }
        const R a = abs_strain<R>(sv);
        R t = 0;
#pragma unroll
        for (int q = 0; q < 6; ++q) t += sb[q] * sv[q];
        const R f = a > (R)0 ? (R)2 * t / a : (R)0;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
          const R m = (q == 0 || q == 3 || q == 5) ? (R)1 : (R)2;
          o[6 + q] = c * (a * sb[q] + m * f * sv[q]);
        }
        acc += (double)cc * (double)a * (double)t;
      }
#pragma unroll
      for (int q = 0; q < 12; ++q) {
        if (h) out[q].y = o[q];
        else out[q].x = o[q];
      }
    }
#pragma unroll
    for (int q = 0; q < 12; ++q)
      if (q < L.nout()) W[i + q * fs] = out[q];
  }
  if (part) {
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
      if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] += red[0];
  }
}

// XPASS of a vector-Jacobian product with the physical tape: per x-row,
// the c2r of ybar and taubar (W[12..20]; ybar alone at W[6..8] without the
// Smagorinsky term), the products of k_adj_phys with u, w, S_ij read from P
// (stride fs, in the layout k_row_c2r leaves), the r2c of ubar, wbar,
// Sbar_ij into W[0..11], and the cs2 derivative's partial of the workgroup
// added into part[blockIdx.x] (LES only).  S_ij is read from P twice, one component at a time
// (|S|^2 and the contraction with sigmabar accumulate together; then the
// outputs), rather than kept in registers; the second read is of rows the
// workgroup has just read.  The grid is persistent, as for k_row_convec.
template <class R, int M, int LES>
__global__ __launch_bounds__((RowCfg<R, M>::BS)) void k_row_adj(cpx<R>* W, int64_t fs, Geom g, R scale, R c, R cc,
                                                              double* __restrict__ part,
                                                              const cpx<R>* __restrict__ P,
                                                              const cpx<R>* __restrict__ tw, int twsM, int twsN) {
  using C = RowCfg<R, M>;
  constexpr AdjLayout L{LES};
  constexpr int T = C::TPR;
  __shared__ cpx<R> lds[C::RPW * lds_pad(M)];
  __shared__ cpx<R> twM[M], twN[M];
  __shared__ double red[C::BS / 64];
  for (int i = threadIdx.x; i < M; i += C::BS) {
    twM[i] = tw[i * twsM];
    twN[i] = tw[i * twsN];
  }
  __syncthreads();
  const int slot = threadIdx.x / T, rt = threadIdx.x % T;
  cpx<R>* lrow = lds + slot * lds_pad(M);
  using Li = Line<R, M, 8, +1>;
  using Lf = Line<R, M, 8, -1>;
  double acc = 0.0;
  const int64_t ngroups = g.nrow / C::RPW;
  for (int64_t grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
    const int64_t row = grp * C::RPW + slot;
    auto load = [&](int f, cpx<R>(&raw)[8], cpx<R>& xm) {
      const cpx<R>* base = W + f * fs;
#pragma unroll
      for (int m = 0; m < 8; ++m) raw[m] = base[row * M + rt + T * m];
      xm = rt == 0 ? base[g.nmain + row] : mk<R>(0, 0);
    };
    auto c2r = [&](cpx<R>(&raw)[8], cpx<R> xm, cpx<R>(&a)[8]) {
      row_c2r_prep<R, M>(a, raw, xm, rt, twN, 1);
      Li::run(a, lrow, rt, twM, 1, LinIdx{});
    };
    // dep: a value the read must wait for (it orders the reads of S_ij
    // one component at a time; the reads are never merged or hoisted)
    auto tape = [&](int q, cpx<R>(&a)[8], R dep) {
      const cpx<R>* tb = P + q * fs + row * M;
      asm volatile("" : "+v"(tb) : "v"(dep) : "memory");
#pragma unroll
      for (int m = 0; m < 8; ++m) a[m] = tb[Li::out_pos(m, rt)];
    };
    auto r2c = [&](cpx<R>(&p)[8], int f) {
      cpx<R>* base = W + f * fs;
      Lf::run(p, lrow, rt, twM, 1, LinIdx{});
      row_r2c_post<R, M>(p, base + row * M, base + g.nmain + row, rt, twN, 1);
    };
    // qbar = scale * c2r(ybar); ubar = w x qbar; wbar = qbar x u
    cpx<R> qv[3][8], v[3][8], p[8];
    {
      cpx<R> raw[3][8], xm[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) load(L.y_in() + j, raw[j], xm[j]);
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        c2r(raw[j], xm[j], qv[j]);
#pragma unroll
        for (int m = 0; m < 8; ++m) qv[j][m] = scale * qv[j][m];
      }
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) tape(3 + j, v[j], (R)0);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
#pragma unroll
      for (int m = 0; m < 8; ++m)
        p[m] = mk<R>(v[j1][m].x * qv[j2][m].x - v[j2][m].x * qv[j1][m].x,
                     v[j1][m].y * qv[j2][m].y - v[j2][m].y * qv[j1][m].y);
      r2c(p, j);
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) tape(j, v[j], (R)0);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
#pragma unroll
      for (int m = 0; m < 8; ++m)
        p[m] = mk<R>(qv[j1][m].x * v[j2][m].x - qv[j2][m].x * v[j1][m].x,
                     qv[j1][m].y * v[j2][m].y - qv[j2][m].y * v[j1][m].y);
      r2c(p, 3 + j);
    }
    if constexpr (LES != 0) {
      // Sbar_q = c (|S| sigmabar_q + 2 m_q S_q (sum_p sigmabar_p S_p)/|S|)
      cpx<R> av[8], tv[8], sb[6][8];
      {
        cpx<R> xm[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) load(L.tau_in() + q, sb[q], xm[q]);
#pragma unroll
        for (int q = 0; q < 6; ++q) {
          cpx<R> raw[8];
#pragma unroll
          for (int m = 0; m < 8; ++m) raw[m] = sb[q][m];
          c2r(raw, xm[q], sb[q]);
        }
      }
      // |S|^2 = 2 (S11^2 + S22^2 + S33^2 + 2 (S12^2 + S13^2 + S23^2)) and t
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        av[m] = mk<R>(0, 0);
        tv[m] = mk<R>(0, 0);
      }
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        tape(6 + q, v[0], tv[7].y);
        const R w2 = (q == 0 || q == 3 || q == 5) ? (R)1 : (R)2;
#pragma unroll
        for (int m = 0; m < 8; ++m) {
          av[m] = mk<R>(av[m].x + w2 * v[0][m].x * v[0][m].x, av[m].y + w2 * v[0][m].y * v[0][m].y);
          tv[m] = mk<R>(tv[m].x + sb[q][m].x * v[0][m].x, tv[m].y + sb[q][m].y * v[0][m].y);
        }
      }
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        av[m] = mk<R>(sqrt((R)2 * av[m].x), sqrt((R)2 * av[m].y));
        acc += (double)cc * (double)av[m].x * (double)tv[m].x;
        acc += (double)cc * (double)av[m].y * (double)tv[m].y;
        // 2 t / |S|: FP32 through the hardware reciprocal (about 1 ulp): the
        // expansion of 16 correctly rounded divisions side by side does not
        // fit in the register budget of two waves per SIMD
        if constexpr (sizeof(R) == 4)
          tv[m] = mk<R>(av[m].x > (R)0 ? (R)2 * tv[m].x * __builtin_amdgcn_rcpf(av[m].x) : (R)0,
                        av[m].y > (R)0 ? (R)2 * tv[m].y * __builtin_amdgcn_rcpf(av[m].y) : (R)0);
        else
          tv[m] = mk<R>(av[m].x > (R)0 ? (R)2 * tv[m].x / av[m].x : (R)0,
                        av[m].y > (R)0 ? (R)2 * tv[m].y / av[m].y : (R)0);
      }
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        const R mq = (q == 0 || q == 3 || q == 5) ? (R)1 : (R)2;
        tape(6 + q, v[0], p[7].y);
#pragma unroll
        for (int m = 0; m < 8; ++m)
          p[m] = mk<R>(c * (av[m].x * sb[q][m].x + mq * tv[m].x * v[0][m].x),
                       c * (av[m].y * sb[q][m].y + mq * tv[m].y * v[0][m].y));
        r2c(p, 6 + q);
      }
    }
  }
  if constexpr (LES == 0) return;
  // fixed-order reduction of the workgroup's partial
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < C::BS / 64; ++w) s += red[w];
    part[blockIdx.x] += s;
  }
}

// Parameters of the adjoint stage bookkeeping.
template <class R>
struct AdjParams {
  R dt, nu, kc2;
};

// Reverse of the stage updates (E = exp(-nu k^2 dt/2), M = [k^2 < kc2]):
// forward  k1 = dt N(v),        s2 = E (v + k1/2)
//          k2 = dt N(s2)/E,     s3 = E (v + k2/2)
//          k3 = dt N(s3)/E,     s4 = E^2 (v + k3)
//          k4 = dt N(s4)/E^2,   v' = M E^2 (v + (k1 + 2 k2 + 2 k3 + k4)/6)
// reverse  g = M E^2 lam; vbar = g; nbar4 = dt g/(6 E^2)                (init)
//          sbar4 = J(s4)^T nbar4: vbar += E^2 sbar4; nbar3 = dt (g/3 + E^2 sbar4)/E     (MODE 4)
//          sbar3:                 vbar += E sbar3;   nbar2 = dt (g/3 + E sbar3/2)/E     (MODE 3)
//          sbar2:                 vbar += E sbar2;   nbar1 = dt (g/6 + E sbar2/2)       (MODE 2)
//          sbar1 = J(v)^T nbar1:  vbar += sbar1                                          (MODE 1)
// Each nbar is written as Y = P nbar (the projection part of the RHS adjoint).
template <class R>
__device__ inline void adj_factors(const Geom& g, int64_t i, const AdjParams<R>& p, R& kx, R& ky, R& kz, R& E, R& E2,
                                   bool& keep) {
  int x, y, z, ikx, iky, ikz;
  ks_of(g, i, x, y, z, ikx, iky, ikz);
  kx = (R)ikx;
  ky = (R)iky;
  kz = (R)ikz;
  const int ik2 = ikx * ikx + iky * iky + ikz * ikz;
  const R k2 = ik2 == 0 ? std::numeric_limits<R>::min() : (R)ik2;
  E = exp_r(-p.nu * k2 * p.dt / (R)2);
  E2 = exp_r(-p.nu * k2 * p.dt);
  keep = k2 < p.kc2;
}

template <class R>
__global__ __launch_bounds__(256) void k_adj_init(const cpx<R>* __restrict__ lam, cpx<R>* __restrict__ vb,
                                                  cpx<R>* __restrict__ Y, int64_t fs, Geom g, AdjParams<R> p) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.ntot) return;
  R kx, ky, kz, E, E2;
  bool keep;
  adj_factors(g, i, p, kx, ky, kz, E, E2, keep);
  cpx<R> n[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const cpx<R> gq = keep ? E2 * lam[i + q * fs] : mk<R>(0, 0);
    vb[i + q * fs] = gq;
    n[q] = (p.dt / ((R)6 * E2)) * gq;
  }
  project(n, kx, ky, kz);
#pragma unroll
  for (int q = 0; q < 3; ++q) Y[i + q * fs] = n[q];
}

// The stage bookkeeping of MODE on the entry at offset o, given sbar there
// (0: vbar = sbar only).
template <class R, int MODE>
__device__ inline void adj_update(const cpx<R> (&sb)[3], int64_t o, int64_t fs, const cpx<R>* __restrict__ lam,
                                  cpx<R>* __restrict__ vb, cpx<R>* __restrict__ Y, R kx, R ky, R kz, R E, R E2,
                                  bool keep, const AdjParams<R>& p) {
  if constexpr (MODE == 0) {
#pragma unroll
    for (int q = 0; q < 3; ++q) vb[o + q * fs] = sb[q];
  } else if constexpr (MODE == 1) {
#pragma unroll
    for (int q = 0; q < 3; ++q) vb[o + q * fs] = vb[o + q * fs] + sb[q];
  } else {
    cpx<R> n[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const cpx<R> gq = keep ? E2 * lam[o + q * fs] : mk<R>(0, 0);
      if constexpr (MODE == 4) {
        vb[o + q * fs] = vb[o + q * fs] + E2 * sb[q];
        n[q] = (p.dt / E) * ((R)(1.0 / 3.0) * gq + E2 * sb[q]);
      } else if constexpr (MODE == 3) {
        vb[o + q * fs] = vb[o + q * fs] + E * sb[q];
        n[q] = (p.dt / E) * ((R)(1.0 / 3.0) * gq + ((R)0.5 * E) * sb[q]);
      } else {
        vb[o + q * fs] = vb[o + q * fs] + E * sb[q];
        n[q] = p.dt * ((R)(1.0 / 6.0) * gq + ((R)0.5 * E) * sb[q]);
      }
    }
    project(n, kx, ky, kz);
#pragma unroll
    for (int q = 0; q < 3; ++q) Y[o + q * fs] = n[q];
  }
}

// Spectral end of J^T: sbar = w (ubar^ + i k x wbar^ + S^T Sbar^) with the
// c2r-adjoint weight w (2 interior, 1 on kx = 0 and kx = nx/2), then the
// stage bookkeeping of MODE.
template <class R, int MODE>
__global__ __launch_bounds__(256) void k_adj_combine(const cpx<R>* __restrict__ W, int64_t fs,
                                                     const cpx<R>* __restrict__ lam, cpx<R>* __restrict__ vb,
                                                     cpx<R>* __restrict__ Y, Geom g, AdjLayout L, AdjParams<R> p) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.ntot) return;
  R kx, ky, kz, E, E2;
  bool keep;
  adj_factors(g, i, p, kx, ky, kz, E, E2, keep);
  const int x = (int)(i < g.nmain ? i % g.mx : g.mx);
  const R wt = (x == 0 || x == g.mx) ? (R)1 : (R)2;
  const cpx<R> o0 = W[i + 3 * fs], o1 = W[i + 4 * fs], o2 = W[i + 5 * fs];
  cpx<R> sb[3];
  sb[0] = W[i] + mul_i(ky * o2 - kz * o1);
  sb[1] = W[i + fs] + mul_i(kz * o0 - kx * o2);
  sb[2] = W[i + 2 * fs] + mul_i(kx * o1 - ky * o0);
  if (L.les) {
    // adjoint of S_ij = (i/2)(k_i u_j + k_j u_i): multiplication by -i/2 k
    const cpx<R> s11 = W[i + 6 * fs], s12 = W[i + 7 * fs], s13 = W[i + 8 * fs];
    const cpx<R> s22 = W[i + 9 * fs], s23 = W[i + 10 * fs], s33 = W[i + 11 * fs];
    const R h = (R)0.5;
    sb[0] = sb[0] - mul_i(kx * s11 + h * (ky * s12 + kz * s13));
    sb[1] = sb[1] - mul_i(ky * s22 + h * (kx * s12 + kz * s23));
    sb[2] = sb[2] - mul_i(kz * s33 + h * (kx * s13 + ky * s23));
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) sb[q] = wt * sb[q];
  adj_update<R, MODE>(sb, i, fs, lam, vb, Y, kx, ky, kz, E, E2, keep, p);
}

// ZINV of the VJP inputs: k_adj_prep fused with the z-pass of their inverse
// transform.  Per column tile the stage state s is loaded and u, w, S_ij
// are formed in registers, z-transformed and stored (STATE; without it
// they come from the physical-space tape of the recomputed forward and s
// is not read); then Y is loaded and ybar (zeroed on the symmetrize
// planes, times the inverse weight) and taubar_ij are formed and
// transformed the same way.  The y- and x-passes follow as for any field.
template <class R, int N, bool STATE>
__global__ __launch_bounds__((ColCfg<R, N>::BS)) void k_adj_zinv(const cpx<R>* __restrict__ s,
                                                               const cpx<R>* __restrict__ Y, cpx<R>* __restrict__ W,
                                                               int64_t fs, Geom g, AdjLayout L,
                                                               const cpx<R>* __restrict__ tw, int tws) {
  using CC = ColCfg<R, N>;
  constexpr int CW = CC::CW, E = CC::E, T = CC::T;
  __shared__ cpx<R> lds[lds_pad(N) * CW];
  __shared__ cpx<R> twl[N];  // first read after a barrier inside the transform
  for (int i = threadIdx.x; i < N; i += CC::BS) twl[i] = tw[i * tws];
  const int c = threadIdx.x % CW, rt = threadIdx.x / CW;
  const Tile t = col_tile<CW, 2>(g, xcd_swizzle(blockIdx.x, gridDim.x), c);
  using Ln = Line<R, N, E, +1>;
  const int ikx = t.kx, iky = kwave(t.oth, g.ny);
  const R kx = (R)ikx, ky = (R)iky;
  auto kzi = [&](int m) { return kwave(rt + T * m, g.nz); };
  auto kz = [&](int m) { return (R)kzi(m); };
  const ZOffR zo = zoffr(t);
  auto store = [&](cpx<R>(&d)[E], int q) {
    if (!t.ok) return;
    ZOffR zs = zo;
    asm volatile("" : "+v"(zs.ob));  // recomputed here, not kept live from the loads
    cpx<R>* const wq = W + q * fs;
#pragma unroll
    for (int m = 0; m < E; ++m) wq[zs.at(Ln::out_pos(m, rt))] = d[m];
  };
  cpx<R> v[3][E], d[E];
  if constexpr (STATE) {
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const uint32_t o = zo.at(rt + T * m);
#pragma unroll
      for (int q = 0; q < 3; ++q) v[q][m] = t.ok ? (s + q * fs)[o] : mk<R>(0, 0);
    }
    // w = i k x u
#pragma unroll
    for (int m = 0; m < E; ++m) d[m] = mul_i(ky * v[2][m] - kz(m) * v[1][m]);
    Ln::run(d, lds, rt, twl, 1, ColIdx<CW>{c});
    store(d, 3);
#pragma unroll
    for (int m = 0; m < E; ++m) d[m] = mul_i(kz(m) * v[0][m] - kx * v[2][m]);
    Ln::run(d, lds, rt, twl, 1, ColIdx<CW>{c});
    store(d, 4);
#pragma unroll
    for (int m = 0; m < E; ++m) d[m] = mul_i(kx * v[1][m] - ky * v[0][m]);
    Ln::run(d, lds, rt, twl, 1, ColIdx<CW>{c});
    store(d, 5);
    if (L.les) {
      constexpr int P[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        const int i = P[q][0], j = P[q][1];
#pragma unroll
        for (int m = 0; m < E; ++m) {
          const R k[3] = {kx, ky, kz(m)};
          d[m] = i == j ? mul_i(k[i] * v[i][m]) : mul_i((R)0.5 * (k[i] * v[j][m] + k[j] * v[i][m]));
        }
        Ln::run(d, lds, rt, twl, 1, ColIdx<CW>{c});
        store(d, 6 + q);
      }
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      Ln::run(v[q], lds, rt, twl, 1, ColIdx<CW>{c});
      store(v[q], q);
    }
  }
  // ybar = Z(Y) times the inverse weights; v's registers are free again
  const R wi = (ikx == 0 || ikx == g.mx) ? (R)1 : (R)0.5;
  cpx<R>(&yb)[3][E] = v;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int ikz = kzi(m);
    const bool zero = sym_zero_mode(g, ikx, t.oth, rt + T * m, ikx * ikx + iky * iky + ikz * ikz);
    const uint32_t o = zo.at(rt + T * m);
#pragma unroll
    for (int q = 0; q < 3; ++q) yb[q][m] = (zero || !t.ok) ? mk<R>(0, 0) : wi * (Y + q * fs)[o];
  }
  if (L.les) {
    // taubar from ybar (adjoint of N_i -= i k_j tau_ij), order 11 12 13 22 23 33
    const int tq = L.tau_in();
#pragma unroll
    for (int q = 0; q < 6; ++q) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const R kzm = kz(m);
        switch (q) {
          case 0: d[m] = mul_i(kx * yb[0][m]); break;
          case 1: d[m] = mul_i(ky * yb[0][m] + kx * yb[1][m]); break;
          case 2: d[m] = mul_i(kzm * yb[0][m] + kx * yb[2][m]); break;
          case 3: d[m] = mul_i(ky * yb[1][m]); break;
          case 4: d[m] = mul_i(kzm * yb[1][m] + ky * yb[2][m]); break;
          default: d[m] = mul_i(kzm * yb[2][m]); break;
        }
      }
      Ln::run(d, lds, rt, twl, 1, ColIdx<CW>{c});
      store(d, tq + q);
    }
  }
  const int yo = L.y_in();
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    Ln::run(yb[q], lds, rt, twl, 1, ColIdx<CW>{c});
    store(yb[q], yo + q);
  }
}

// Transpose of the kx = 0 Hermitian copy (symmetrize.f90:51-61) on Y:
// for each target t = (0, y, z) (kz < 0, or kz = 0 and ky < 0) with source
// s = (0, -y, -z): Y[s] += conj(Y[t]), Y[t] = 0.  Sources are never targets.
template <class R>
__global__ __launch_bounds__(256) void k_adj_hermT(cpx<R>* __restrict__ Y, int64_t fs, Geom g) {
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
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const cpx<R> t = Y[od + q * fs];
    Y[os + q * fs] = Y[os + q * fs] + conj(t);
    Y[od + q * fs] = mk<R>(0, 0);
  }
}

}  // namespace scl
