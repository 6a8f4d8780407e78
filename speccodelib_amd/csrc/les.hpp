// Kernels of the LES + passive-scalar RHS and of the dynamic Smagorinsky
// coefficient.  These run as separate passes around multi-field 3D
// transforms (prep -> c2r -> physical pointwise -> r2c -> combine), unlike
// the fused five-pass DNS RHS.
//
// Physical-space fields are the main rows in packed form: entry i holds the
// reals x[2n] (.x) and x[2n+1] (.y); pointwise products act on .x and .y
// separately, as the reference's packed-real statements do
// (dynsmag.f90:43-48, smag.f90:44-57).
#pragma once

#include "kernels3d.hpp"

namespace scl {

// Field counts of the extended RHS: inputs u(3) w(3) [S_ij(6)] [phi(1)],
// outputs u x w(3) [tau_ij(6)] [u_j phi(3)].
// Inputs u(0..2) w(3..5) [S_ij(6..11)] [phi]; outputs N(0..2) [u_j phi]
// [tau_ij]: the fused x-pass writes each output field after the row's input
// in that field has been read (the flux over w, tau over S_ij).
struct ExtLayout {
  int les, scalar;
  __host__ __device__ constexpr int nin() const { return 6 + 6 * les + scalar; }
  __host__ __device__ constexpr int nout() const { return 3 + 6 * les + 3 * scalar; }
  __host__ __device__ constexpr int s_in() const { return 6; }                 // S_ij inputs
  __host__ __device__ constexpr int phi_in() const { return 6 + 6 * les; }     // phi input
  __host__ __device__ constexpr int flux_out() const { return 3; }             // u_j phi outputs
  __host__ __device__ constexpr int tau_out() const { return 3 + 3 * scalar; } // tau_ij outputs
};

__device__ inline void ks_of(const Geom& g, int64_t i, int& x, int& y, int& z, int& kx, int& ky, int& kz) {
  split_xyz(g, i, x, y, z);
  kx = x;
  ky = kwave(y, g.ny);
  kz = kwave(z, g.nz);
}

// S_ij = (i/2)(k_i v_j + k_j v_i), order 11 12 13 22 23 33 (dynsmag.f90:198-203)
template <class R>
__device__ inline void strain_of(const cpx<R> (&v)[3], R kx, R ky, R kz, cpx<R> (&s)[6]) {
  const R k[3] = {kx, ky, kz};
  const int P[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const int i = P[q][0], j = P[q][1];
    if (i == j) s[q] = mul_i(k[i] * v[i]);
    else s[q] = mul_i((R)0.5 * (k[i] * v[j] + k[j] * v[i]));
  }
}

// Spectral inputs of the extended RHS from the stage state st (velocity at
// st, st+fs, st+2fs; scalar at st+3fs).
template <class R>
__global__ __launch_bounds__(256) void k_ext_prep(const cpx<R>* __restrict__ st, cpx<R>* __restrict__ W, int64_t fs,
                                                  Geom g, ExtLayout L) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.ntot) return;
  int x, y, z, ikx, iky, ikz;
  ks_of(g, i, x, y, z, ikx, iky, ikz);
  const R kx = (R)ikx, ky = (R)iky, kz = (R)ikz;
  const cpx<R> v[3] = {st[i], st[i + fs], st[i + 2 * fs]};
  W[i] = v[0];
  W[i + fs] = v[1];
  W[i + 2 * fs] = v[2];
  W[i + 3 * fs] = mul_i(ky * v[2] - kz * v[1]);
  W[i + 4 * fs] = mul_i(kz * v[0] - kx * v[2]);
  W[i + 5 * fs] = mul_i(kx * v[1] - ky * v[0]);
  if (L.les) {
    cpx<R> s[6];
    strain_of<R>(v, kx, ky, kz, s);
#pragma unroll
    for (int q = 0; q < 6; ++q) W[i + (L.s_in() + q) * fs] = s[q];
  }
  if (L.scalar) W[i + L.phi_in() * fs] = st[i + 3 * fs];
}

// |S| = sqrt(2 S_ij S_ij) of one packed component
template <class R>
__device__ inline R abs_strain(const R (&s)[6]) {
  return sqrt((R)2 * (s[0] * s[0] + s[3] * s[3] + s[5] * s[5] + (R)2 * (s[1] * s[1] + s[2] * s[2] + s[4] * s[4])));
}

// Physical pointwise work, in place on the main rows: u x w, tau_ij =
// -2 cs2 D^2 |S| S_ij (dynsmag.f90:206-225), u_j phi (convecphi_dns.f90:23,
// 33, 42), all times scale = 1/(nx ny nz).  cs2 comes from device memory
// (the dynamic update writes it there).
// With tape given, the velocity inputs u, w and S_ij are also stored there
// (stride fs): the adjoint's record of the recomputed forward.  The layout
// is a template parameter so that the per-thread arrays stay in registers.
template <class R, int LES, int SC>
__global__ __launch_bounds__(256) void k_ext_phys(cpx<R>* __restrict__ W, int64_t fs, Geom g, R scale, R d2,
                                                  const double* __restrict__ cs2p, cpx<R>* __restrict__ tape) {
  constexpr ExtLayout L{LES, SC};
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.nmain) return;
  const R c = L.les ? (R)(-2.0 * cs2p[0]) * d2 * scale : (R)0;
  cpx<R> in[13];
#pragma unroll
  for (int q = 0; q < 13; ++q)
    if (q < L.nin()) in[q] = W[i + q * fs];
  if (tape) {
#pragma unroll
    for (int q = 0; q < 12; ++q)
      if (q < L.s_in() + 6 * L.les) tape[i + q * fs] = in[q];
  }
  cpx<R> out[12];
#pragma unroll
  for (int h = 0; h < 2; ++h) {  // the two packed reals
    auto re = [&](int q) -> R { return h ? in[q].y : in[q].x; };
    R o[12];
    const R u0 = re(0), u1 = re(1), u2 = re(2), w0 = re(3), w1 = re(4), w2 = re(5);
    o[0] = (u1 * w2 - u2 * w1) * scale;
    o[1] = (u2 * w0 - u0 * w2) * scale;
    o[2] = (u0 * w1 - u1 * w0) * scale;
    if (L.les) {
      R s[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) s[q] = re(L.s_in() + q);
      const R a = abs_strain<R>(s);
#pragma unroll
      for (int q = 0; q < 6; ++q) o[L.tau_out() + q] = c * a * s[q];
    }
    if (L.scalar) {
      const R p = re(L.phi_in());
      o[L.flux_out() + 0] = u0 * p * scale;
      o[L.flux_out() + 1] = u1 * p * scale;
      o[L.flux_out() + 2] = u2 * p * scale;
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

// XPASS of the multi-pass RHS: per x-row, the c2r of the inputs, the
// products of k_ext_phys and the r2c of the outputs, in place (the layout
// lets every output overwrite an input the row has already read).  Phase 1
// transforms u, w [phi] and writes N [and u_j phi]; phase 2 transforms
// S_ij and writes tau_ij.  TAPE: u, w, S_ij in physical space also to tape,
// in the layout k_row_c2r leaves (the adjoint's record of the forward).
// The grid is persistent, as for k_row_convec.
template <class R, int M, int LES, int SC, bool TAPE>
__global__ __launch_bounds__((RowCfg<R, M>::BS)) void k_row_ext(cpx<R>* W, int64_t fs, Geom g, R scale, R d2,
                                                              const double* __restrict__ cs2p,
                                                              cpx<R>* __restrict__ tape, const cpx<R>* __restrict__ tw,
                                                              int twsM, int twsN) {
  using C = RowCfg<R, M>;
  constexpr ExtLayout L{LES, SC};
  constexpr int T = C::TPR;
  __shared__ cpx<R> lds[C::RPW * lds_pad(M)];
  __shared__ cpx<R> twM[M], twN[M];
  for (int i = threadIdx.x; i < M; i += C::BS) {
    twM[i] = tw[i * twsM];
    twN[i] = tw[i * twsN];
  }
  __syncthreads();
  const R c = LES ? (R)(-2.0 * cs2p[0]) * d2 * scale : (R)0;
  const int slot = threadIdx.x / T, rt = threadIdx.x % T;
  cpx<R>* lrow = lds + slot * lds_pad(M);
  using Li = Line<R, M, 8, +1>;
  using Lf = Line<R, M, 8, -1>;
  const int64_t ngroups = g.nrow / C::RPW;
  for (int64_t grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
    const int64_t row = grp * C::RPW + slot;
    auto load = [&](int f, cpx<R>(&raw)[8], cpx<R>& xm) {
      const cpx<R>* base = W + f * fs;
#pragma unroll
      for (int m = 0; m < 8; ++m) raw[m] = base[row * M + rt + T * m];
      xm = rt == 0 ? base[g.nmain + row] : mk<R>(0, 0);
    };
    auto c2r = [&](cpx<R>(&raw)[8], cpx<R> xm, cpx<R>(&a)[8], int q) {
      row_c2r_prep<R, M>(a, raw, xm, rt, twN, 1);
      Li::run(a, lrow, rt, twM, 1, LinIdx{});
      if constexpr (TAPE) {
        if (q < 0) return;
        cpx<R>* tb = tape + q * fs + row * M;
#pragma unroll
        for (int m = 0; m < 8; ++m) tb[Li::out_pos(m, rt)] = a[m];
      }
    };
    auto r2c = [&](cpx<R>(&p)[8], int f) {
      cpx<R>* base = W + f * fs;
      Lf::run(p, lrow, rt, twM, 1, LinIdx{});
      row_r2c_post<R, M>(p, base + row * M, base + g.nmain + row, rt, twN, 1);
    };
    // phase 1: u, w -> N, then phi -> u_j phi
    cpx<R> raw[6][8], xm[6], a[7][8], p[8];
#pragma unroll
    for (int f = 0; f < 6; ++f) load(f, raw[f], xm[f]);
#pragma unroll
    for (int f = 0; f < 6; ++f) c2r(raw[f], xm[f], a[f], f);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
#pragma unroll
      for (int m = 0; m < 8; ++m)
        p[m] = mk<R>((a[j1][m].x * a[3 + j2][m].x - a[j2][m].x * a[3 + j1][m].x) * scale,
                     (a[j1][m].y * a[3 + j2][m].y - a[j2][m].y * a[3 + j1][m].y) * scale);
      r2c(p, j);
    }
    if constexpr (SC != 0) {
      load(L.phi_in(), raw[0], xm[0]);
      c2r(raw[0], xm[0], a[6], -1);
#pragma unroll
      for (int j = 0; j < 3; ++j) {
#pragma unroll
        for (int m = 0; m < 8; ++m)
          p[m] = mk<R>(a[j][m].x * a[6][m].x * scale, a[j][m].y * a[6][m].y * scale);
        r2c(p, L.flux_out() + j);
      }
    }
    if constexpr (LES != 0) {
      // phase 2: S_ij -> tau_ij = c |S| S_ij
      cpx<R> sraw[6][8], sxm[6], s[6][8];
#pragma unroll
      for (int q = 0; q < 6; ++q) load(L.s_in() + q, sraw[q], sxm[q]);
#pragma unroll
      for (int q = 0; q < 6; ++q) c2r(sraw[q], sxm[q], s[q], 6 + q);
      cpx<R> ca[8];
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const R sx[6] = {s[0][m].x, s[1][m].x, s[2][m].x, s[3][m].x, s[4][m].x, s[5][m].x};
        const R sy[6] = {s[0][m].y, s[1][m].y, s[2][m].y, s[3][m].y, s[4][m].y, s[5][m].y};
        ca[m] = mk<R>(c * abs_strain<R>(sx), c * abs_strain<R>(sy));
      }
#pragma unroll
      for (int q = 0; q < 6; ++q) {
#pragma unroll
        for (int m = 0; m < 8; ++m) p[m] = mk<R>(ca[m].x * s[q][m].x, ca[m].y * s[q][m].y);
        r2c(p, L.tau_out() + q);
      }
    }
  }
}

// Parameters of the extended combine kernel.
template <class R>
struct ExtParams {
  RhsParams<R> r;
  R nu_phi;   // scalar diffusivity nu/Sc
  R grad;     // mean scalar gradient G (source -G u_z)
  int sign_phi;  // +1: R = -i k.F (physical advection)
};

// The update of one component in the LES / scalar combine: the IF-RK4 stage
// MODE 1..4 (stage_apply), or the IF-AB2 step of MODE 5..7 (ab_apply, the
// weights this step's k_ab_begin left in r.ab) as spectral_les.f90:241-304
// and spectral_dns_scalar.f90:257-338 write it, with the component's
// viscosity nu (nu for u, nu/Sc for the scalar); AB2 also zeroes the mean.
// in holds what stage_load read for the component (ext_update loads it).
template <int MODE, class R, class I>
__device__ inline void ext_apply(const StageIn<R>& in, cpx<R> a, I o, cpx<R>* V, cpx<R>* DU, cpx<R>* S, R nu, R k2,
                                 bool cut, bool mean, const RhsParams<R>& r) {
  if constexpr (MODE >= 5) {
    const R dt = (R)r.ab[AB_DT];
    const R xk = -nu * k2;
    const R e1 = expm1_r(xk * dt);
    const R e2 = MODE == 5 ? (R)0 : expm1_r(xk * (R)r.ab[AB_DTO]);
    (void)ab_apply<MODE, R>(in, a, o, V, DU, S, dt, (R)r.ab[AB_S], (R)r.ab[AB_EK], (R)r.ab[AB_DELTA], e1, e2,
                            cut || mean, r.nonfinite);
  } else {
    const R x = -nu * k2 * r.dt / (R)2;
    (void)stage_apply<MODE, R>(in, r.dt * a, o, V, DU, S, expm1_r(x), x, cut, true, r.nonfinite);
  }
}

template <int MODE, class R, class I>
__device__ inline void ext_update(cpx<R> a, I o, cpx<R>* V, cpx<R>* DU, cpx<R>* S, R nu, R k2, bool cut, bool mean,
                                  const RhsParams<R>& r) {
  ext_apply<MODE, R>(stage_load<MODE, R, I>(o, V, DU, S, cut), a, o, V, DU, S, nu, k2, cut, mean, r);
}

// Combine (spectral): N = u x w + f - i k_j tau_ij (spectral_les.f90:232-234),
// symmetrize's zeroing, projection; R_phi = -i k . F - G u_z
// (spectral_dns_scalar.f90:236,245), zeroing; then the IF-RK4 stage (MODE
// 1..4) for velocity (factor exp(-nu k2 dt/2)) and scalar (exp(-nu/Sc k2
// dt/2)), or MODE 0: N and R_phi into W[0..3].
template <class R, int MODE>
__global__ __launch_bounds__(256) void k_ext_combine(cpx<R>* W, int64_t fs, cpx<R>* V, cpx<R>* DU, cpx<R>* S,
                                                     const cpx<R>* st, const double* __restrict__ coef, ExtParams<R> p,
                                                     Geom g, ExtLayout L) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.ntot) return;
  int x, y, z, ikx, iky, ikz;
  ks_of(g, i, x, y, z, ikx, iky, ikz);
  const R kx = (R)ikx, ky = (R)iky, kz = (R)ikz;
  const int ik2 = ikx * ikx + iky * iky + ikz * ikz;
  const R tiny = std::numeric_limits<R>::min();
  const R k2 = ik2 == 0 ? tiny : (R)ik2;
  cpx<R> a = W[i], b = W[i + fs], d = W[i + 2 * fs];
  if (p.r.forced && k2 < p.r.fk2) {
    const R fa = (R)coef[0], fb = (R)coef[1];
    const cpx<R> sx = st[i], sy = st[i + fs], sz = st[i + 2 * fs];
    a = a + fa * sx;
    b = b + fa * sy;
    d = d + fa * sz;
    if (p.r.forced == 2) {
      a = a + fb * mul_i(ky * sz - kz * sy);
      b = b + fb * mul_i(kz * sx - kx * sz);
      d = d + fb * mul_i(kx * sy - ky * sx);
    }
  }
  if (L.les) {
    const int t = L.tau_out();
    const cpx<R> t11 = W[i + t * fs], t12 = W[i + (t + 1) * fs], t13 = W[i + (t + 2) * fs];
    const cpx<R> t22 = W[i + (t + 3) * fs], t23 = W[i + (t + 4) * fs], t33 = W[i + (t + 5) * fs];
    a = a - mul_i(kx * t11 + ky * t12 + kz * t13);
    b = b - mul_i(kx * t12 + ky * t22 + kz * t23);
    d = d - mul_i(kx * t13 + ky * t23 + kz * t33);
  }
  const bool zero = x == g.mx || y == g.ny / 2 || z == g.nz / 2 || ik2 == 0;
  if (zero) {
    a = mk<R>(0, 0);
    b = a;
    d = a;
  }
  {
    const R den = kx * kx + ky * ky + kz * kz + tiny;
    const cpx<R> q = mk<R>((kx * a.x + ky * b.x + kz * d.x) / den, (kx * a.y + ky * b.y + kz * d.y) / den);
    a = a - kx * q;
    b = b - ky * q;
    d = d - kz * q;
  }
  cpx<R> rp = mk<R>(0, 0);
  if (L.scalar) {
    const int f = L.flux_out();
    const cpx<R> fx = W[i + f * fs], fy = W[i + (f + 1) * fs], fz = W[i + (f + 2) * fs];
    rp = (R)(-p.sign_phi) * mul_i(kx * fx + ky * fy + kz * fz) - p.grad * st[i + 2 * fs];
    if (zero) rp = mk<R>(0, 0);
  }
  if (p.r.dealias && k2 >= p.r.kc2) {
    a = mk<R>(0, 0);
    b = a;
    d = a;
    rp = a;
  }
  if constexpr (MODE == 0) {
    W[i] = a;
    W[i + fs] = b;
    W[i + 2 * fs] = d;
    if (L.scalar) W[i + 3 * fs] = rp;
  } else {
    const bool cut = k2 >= p.r.kc2, mean = ik2 == 0;
    ext_update<MODE, R>(a, i, V, DU, S, p.r.nu, k2, cut, mean, p.r);
    ext_update<MODE, R>(b, i + fs, V, DU, S, p.r.nu, k2, cut, mean, p.r);
    ext_update<MODE, R>(d, i + 2 * fs, V, DU, S, p.r.nu, k2, cut, mean, p.r);
    if (L.scalar) ext_update<MODE, R>(rp, i + 3 * fs, V, DU, S, p.nu_phi, k2, cut, mean, p.r);
  }
}

// ZINV of the extended inputs: k_ext_prep fused with the z-pass of their
// inverse transform.  Per column tile the stage state u (st, st+fs, st+2fs)
// is loaded once; with the Smagorinsky term the fields that need kz are
// formed in registers, z-transformed and stored: i kz u_x, i kz u_y (to
// W[3], W[4]) and i kz u_z (= S_33, to W[11]), then u itself and phi at
// st+3fs; k_yinv_les forms w and the other S_ij in the y-pass.  Without it
// (scalar only) w = i k x u is formed here.
template <class R, int N>
__global__ __launch_bounds__((ColCfg<R, N>::BS)) void k_ext_zinv(const cpx<R>* __restrict__ st,
                                                               cpx<R>* __restrict__ W, int64_t fs, Geom g,
                                                               ExtLayout L, const cpx<R>* __restrict__ tw, int tws) {
  using CC = ColCfg<R, N>;
  constexpr int CW = CC::CW, E = CC::E, T = CC::T;
  __shared__ cpx<R> lds[lds_pad(N) * CW];
  __shared__ cpx<R> twl[N];  // first read after a barrier inside the transform
  for (int i = threadIdx.x; i < N; i += CC::BS) twl[i] = tw[i * tws];
  const int c = threadIdx.x % CW, rt = threadIdx.x / CW;
  const Tile t = col_tile<CW, 2>(g, xcd_swizzle(blockIdx.x, gridDim.x), c);
  using Ln = Line<R, N, E, +1>;
  const R kx = (R)t.kx, ky = (R)kwave(t.oth, g.ny);
  auto kz = [&](int m) { return (R)kwave(rt + T * m, g.nz); };
  // 32-bit in-field offsets (< 2^30 at 1024^3), one base pointer per field
  const ZOffR zo = zoffr(t);
  cpx<R> v[3][E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const uint32_t o = zo.at(rt + T * m);
#pragma unroll
    for (int q = 0; q < 3; ++q) v[q][m] = t.ok ? (st + q * fs)[o] : mk<R>(0, 0);
  }
  auto store = [&](cpx<R>(&d)[E], int q) {
    if (!t.ok) return;
    ZOffR zs = zo;
    asm volatile("" : "+v"(zs.ob));  // recomputed here, not kept live from the loads
    cpx<R>* const wq = W + q * fs;
#pragma unroll
    for (int m = 0; m < E; ++m) wq[zs.at(Ln::out_pos(m, rt))] = d[m];
  };
  cpx<R> d[E];
  if (L.les) {
    // i kz u_j to W[3], W[4], W[L.s_in() + 5]
#pragma unroll
    for (int j = 0; j < 3; ++j) {
#pragma unroll
      for (int m = 0; m < E; ++m) d[m] = mul_i(kz(m) * v[j][m]);
      Ln::run(d, lds, rt, twl, 1, ColIdx<CW>{c});
      store(d, j < 2 ? 3 + j : L.s_in() + 5);
    }
  } else {
    // w = i k x u (as k_ext_prep)
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
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    Ln::run(v[q], lds, rt, twl, 1, ColIdx<CW>{c});
    store(v[q], q);
  }
  if (L.scalar) {
#pragma unroll
    for (int m = 0; m < E; ++m) d[m] = t.ok ? (st + 3 * fs)[zo.at(rt + T * m)] : mk<R>(0, 0);
    Ln::run(d, lds, rt, twl, 1, ColIdx<CW>{c});
    store(d, L.phi_in());
  }
}

// YINV of the LES inputs after k_ext_zinv: the products with kx and ky are
// formed in the y-pass, before the y-inverse.  PART 0 loads u_x, u_y and
// writes w_z, S_11, S_12, S_22 and u_x, u_y; PART 1 loads u_z, i kz u_x,
// i kz u_y and writes w_x, w_y, S_13, S_23 and u_z (S_33 and phi take a
// plain column pass).  The parts touch disjoint fields, and each reads its
// fields before writing them.  Same S_ij and w as k_ext_prep (strain_of),
// with the kz terms transformed along z first.
template <class R, int N, int PART>
__global__ __launch_bounds__((ColCfg<R, N>::BS)) void k_yinv_les(cpx<R>* __restrict__ W, int64_t fs, Geom g,
                                                               const cpx<R>* __restrict__ tw, int tws) {
  using CC = ColCfg<R, N>;
  constexpr int CW = CC::CW, E = CC::E, T = CC::T;
  constexpr ExtLayout L{1, 0};
  __shared__ cpx<R> lds[lds_pad(N) * CW];
  __shared__ cpx<R> twl[N];  // first read after a barrier inside the transform
  for (int i = threadIdx.x; i < N; i += CC::BS) twl[i] = tw[i * tws];
  const int c = threadIdx.x % CW, rt = threadIdx.x / CW;
  const Tile t = col_tile<CW, 1>(g, xcd_swizzle(blockIdx.x, gridDim.x), c);
  using Ln = Line<R, N, E, +1>;
  auto load = [&](cpx<R>(&v)[E], int f) {
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = t.ok ? W[t.off(rt + T * m) + f * fs] : mk<R>(0, 0);
  };
  auto put = [&](cpx<R>(&v)[E], int f) {
    Ln::run(v, lds, rt, twl, 1, ColIdx<CW>{c});
    if (!t.ok) return;
#pragma unroll
    for (int m = 0; m < E; ++m) W[t.off(Ln::out_pos(m, rt)) + f * fs] = v[m];
  };
  const R kx = (R)t.kx;
  auto ky = [&](int m) { return (R)kwave(rt + T * m, g.ny); };
  const int s0 = L.s_in();
  cpx<R> d[E];
  if constexpr (PART == 0) {
    cpx<R> a[E], b[E];  // u_x, u_y
    load(a, 0);
    load(b, 1);
#pragma unroll
    for (int m = 0; m < E; ++m) d[m] = mul_i(kx * b[m] - ky(m) * a[m]);
    put(d, 5);  // w_z
#pragma unroll
    for (int m = 0; m < E; ++m) d[m] = mul_i(kx * a[m]);
    put(d, s0);  // S_11
#pragma unroll
    for (int m = 0; m < E; ++m) d[m] = mul_i((R)0.5 * (kx * b[m] + ky(m) * a[m]));
    put(d, s0 + 1);  // S_12
#pragma unroll
    for (int m = 0; m < E; ++m) d[m] = mul_i(ky(m) * b[m]);
    put(d, s0 + 3);  // S_22
    put(a, 0);
    put(b, 1);
  } else {
    cpx<R> a[E], p[E], q[E];  // u_z, i kz u_x, i kz u_y
    load(a, 2);
    load(p, 3);
    load(q, 4);
#pragma unroll
    for (int m = 0; m < E; ++m) d[m] = mul_i(ky(m) * a[m]) - q[m];
    put(d, 3);  // w_x
#pragma unroll
    for (int m = 0; m < E; ++m) d[m] = p[m] - mul_i(kx * a[m]);
    put(d, 4);  // w_y
#pragma unroll
    for (int m = 0; m < E; ++m) d[m] = (R)0.5 * (mul_i(kx * a[m]) + p[m]);
    put(d, s0 + 2);  // S_13
#pragma unroll
    for (int m = 0; m < E; ++m) d[m] = (R)0.5 * (mul_i(ky(m) * a[m]) + q[m]);
    put(d, s0 + 4);  // S_23
    put(a, 2);
  }
}

// YFWD of the extended outputs (after the x-pass), folding the kx and ky
// terms of the combine: per y-column tile each output is y-transformed in
// turn; PART 0 (Smagorinsky term) accumulates N_i -= i (kx tau_i1 +
// ky tau_i2) in registers and writes N' to W[0..2]; PART 1 (scalar)
// accumulates R' = -sign i (kx F_x + ky F_y) and writes it to the first flux
// slot.  The fields the z-pass still needs (tau_13, tau_23, tau_33, F_z) are
// stored transformed in place.  The parts touch disjoint fields; each reads
// a field of the tile before writing its slot.  PART 0 is held to two waves
// per SIMD (a few registers spill at some line lengths).
template <class R, int N, int PART, int SC>
__global__ __launch_bounds__((ColCfg<R, N>::BS)) __attribute__((amdgpu_waves_per_eu(2))) void k_yfwd_ext(
    cpx<R>* __restrict__ W, int64_t fs, Geom g, R sg, const cpx<R>* __restrict__ tw, int tws) {
  using CC = ColCfg<R, N>;
  constexpr int CW = CC::CW, E = CC::E, T = CC::T;
  constexpr ExtLayout L{1, SC};
  __shared__ cpx<R> lds[lds_pad(N) * CW];
  __shared__ cpx<R> twl[N];  // first read after a barrier inside the transform
  for (int i = threadIdx.x; i < N; i += CC::BS) twl[i] = tw[i * tws];
  const int c = threadIdx.x % CW, rt = threadIdx.x / CW;
  const Tile t = col_tile<CW, 1>(g, xcd_swizzle(blockIdx.x, gridDim.x), c);
  using Ln = Line<R, N, E, -1>;
  auto load = [&](cpx<R>(&v)[E], int f) {
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = t.ok ? W[t.off(rt + T * m) + f * fs] : mk<R>(0, 0);
    Ln::run(v, lds, rt, twl, 1, ColIdx<CW>{c});
  };
  auto store = [&](cpx<R>(&v)[E], int f) {
    if (!t.ok) return;
#pragma unroll
    for (int m = 0; m < E; ++m) W[t.off(Ln::out_pos(m, rt)) + f * fs] = v[m];
  };
  const R kx = (R)t.kx;
  auto ky = [&](int m) { return (R)kwave(Ln::out_pos(m, rt), g.ny); };
  cpx<R> f[E];
  if constexpr (PART == 0) {
    cpx<R> a[E], b[E], d[E];
    load(a, 0);
    load(b, 1);
    load(d, 2);
    const int t0 = L.tau_out();
    load(f, t0);  // tau_11
#pragma unroll
    for (int m = 0; m < E; ++m) a[m] = a[m] - mul_i(kx * f[m]);
    load(f, t0 + 1);  // tau_12
#pragma unroll
    for (int m = 0; m < E; ++m) {
      a[m] = a[m] - mul_i(ky(m) * f[m]);
      b[m] = b[m] - mul_i(kx * f[m]);
    }
    load(f, t0 + 2);  // tau_13
#pragma unroll
    for (int m = 0; m < E; ++m) d[m] = d[m] - mul_i(kx * f[m]);
    store(f, t0 + 2);
    load(f, t0 + 3);  // tau_22
#pragma unroll
    for (int m = 0; m < E; ++m) b[m] = b[m] - mul_i(ky(m) * f[m]);
    load(f, t0 + 4);  // tau_23
#pragma unroll
    for (int m = 0; m < E; ++m) d[m] = d[m] - mul_i(ky(m) * f[m]);
    store(f, t0 + 4);
    load(f, t0 + 5);  // tau_33
    store(f, t0 + 5);
    store(a, 0);
    store(b, 1);
    store(d, 2);
  } else {
    const int f0 = L.flux_out();
    cpx<R> r[E];
    load(f, f0);
#pragma unroll
    for (int m = 0; m < E; ++m) r[m] = sg * mul_i(kx * f[m]);
    load(f, f0 + 1);
#pragma unroll
    for (int m = 0; m < E; ++m) r[m] = r[m] + sg * mul_i(ky(m) * f[m]);
    load(f, f0 + 2);
    store(f, f0 + 2);
    store(r, f0);
  }
}

// ZFWD of the extended outputs: the z-pass of their forward transform fused
// with k_ext_combine.  After k_yfwd_ext (x- and y-transformed, the kx and ky
// terms folded) each remaining field is z-transformed in turn and folded
// into the three velocity components N and the scalar rate R_phi held in
// registers (N_i -= i kz tau_i3, R_phi -= sign i kz F_z), then the combine's
// epilogue runs per entry: forcing, symmetrize's zeroing, projection, the
// scalar source -G u_z, the optional cut, and the IF-RK4 stage (MODE 1..4)
// or N, R_phi into W[0..3] (MODE 0).  Every field's columns are read before
// W is written.
template <class R, int N, int MODE>
__global__ __launch_bounds__((ColCfg<R, N>::BS)) void k_ext_zfwd(cpx<R>* W, int64_t fs, cpx<R>* V, cpx<R>* DU,
                                                               cpx<R>* S, const cpx<R>* st,
                                                               const double* __restrict__ coef, ExtParams<R> p,
                                                               Geom g, ExtLayout L, const cpx<R>* __restrict__ tw,
                                                               int tws) {
  using CC = ColCfg<R, N>;
  constexpr int CW = CC::CW, E = CC::E, T = CC::T;
  __shared__ cpx<R> lds[lds_pad(N) * CW];
  __shared__ cpx<R> twl[N];  // first read after a barrier inside the transform
  for (int i = threadIdx.x; i < N; i += CC::BS) twl[i] = tw[i * tws];
  const int c = threadIdx.x % CW, rt = threadIdx.x / CW;
  const Tile t = col_tile<CW, 2>(g, xcd_swizzle(blockIdx.x, gridDim.x), c);
  using Ln = Line<R, N, E, -1>;
  const int ikx = t.kx, iky = kwave(t.oth, g.ny);
  const R kx = (R)ikx, ky = (R)iky;
  auto kz = [&](int m) { return (R)kwave(Ln::out_pos(m, rt), g.nz); };
  const ZOffR zo = zoffr(t);
  auto load = [&](cpx<R>(&d)[E], int q) {
    ZOffR zl = zo;
    asm volatile("" : "+v"(zl.ob));  // offsets recomputed per field, not kept live
    const cpx<R>* const wq = W + q * fs;
#pragma unroll
    for (int m = 0; m < E; ++m) d[m] = t.ok ? wq[zl.at(rt + T * m)] : mk<R>(0, 0);
    Ln::run(d, lds, rt, twl, 1, ColIdx<CW>{c});
  };
  // k_yfwd_ext has folded the kx and ky terms: N' at 0..2, R' at the
  // first flux slot; the kz terms remain (tau_13, tau_23, tau_33, F_z)
  cpx<R> a[E], b[E], dz[E], rp[E], f[E];
  load(a, 0);
  load(b, 1);
  load(dz, 2);
  if (L.les) {
    // N_i -= i kz tau_i3
    const int t0 = L.tau_out();
    load(f, t0 + 2);  // tau_13
#pragma unroll
    for (int m = 0; m < E; ++m) a[m] = a[m] - mul_i(kz(m) * f[m]);
    load(f, t0 + 4);  // tau_23
#pragma unroll
    for (int m = 0; m < E; ++m) b[m] = b[m] - mul_i(kz(m) * f[m]);
    load(f, t0 + 5);  // tau_33
#pragma unroll
    for (int m = 0; m < E; ++m) dz[m] = dz[m] - mul_i(kz(m) * f[m]);
  }
  if (L.scalar) {
    // R_phi = -sign i k.F (convecphi_dns.f90:26-47's form, physical sign
    // for the solver): the kz term
    const int f0 = L.flux_out();
    const R sg = (R)(-p.sign_phi);
    load(rp, f0);
    load(f, f0 + 2);
#pragma unroll
    for (int m = 0; m < E; ++m) rp[m] = rp[m] + sg * mul_i(kz(m) * f[m]);
  }
  if (!t.ok) return;
  const R tiny = std::numeric_limits<R>::min();
  R fa = 0, fb = 0;
  if (p.r.forced) {
    fa = (R)coef[0];
    fb = (R)coef[1];
  }
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int z = Ln::out_pos(m, rt);
    const int ikz = kwave(z, g.nz);
    const R kzm = (R)ikz;
    const int ik2 = ikx * ikx + iky * iky + ikz * ikz;
    const R k2 = ik2 == 0 ? tiny : (R)ik2;
    const int64_t o = zo.at(z);
    // the stage update's loads of all components of this slot are issued
    // together, before any of its stores (V, DU, S may alias as far as the
    // compiler knows, which would otherwise cost a round trip per component)
    StageIn<R> in[4];
    if constexpr (MODE > 0) {
      const bool cut = k2 >= p.r.kc2;
#pragma unroll
      for (int q = 0; q < 3; ++q) in[q] = stage_load<MODE, R, int64_t>(o + q * fs, V, DU, S, cut);
      if (L.scalar) in[3] = stage_load<MODE, R, int64_t>(o + 3 * fs, V, DU, S, cut);
    }
    cpx<R> ax = a[m], by = b[m], cz = dz[m];
    if (p.r.forced && k2 < p.r.fk2) {
      const cpx<R> sx = st[o], sy = st[o + fs], sz = st[o + 2 * fs];
      ax = ax + fa * sx;
      by = by + fa * sy;
      cz = cz + fa * sz;
      if (p.r.forced == 2) {
        ax = ax + fb * mul_i(ky * sz - kzm * sy);
        by = by + fb * mul_i(kzm * sx - kx * sz);
        cz = cz + fb * mul_i(kx * sy - ky * sx);
      }
    }
    const bool zero = t.kx == g.mx || t.oth == g.ny / 2 || z == g.nz / 2 || ik2 == 0;
    if (zero) {
      ax = mk<R>(0, 0);
      by = ax;
      cz = ax;
    }
    {
      const R den = kx * kx + ky * ky + kzm * kzm + tiny;
      const cpx<R> q = mk<R>((kx * ax.x + ky * by.x + kzm * cz.x) / den, (kx * ax.y + ky * by.y + kzm * cz.y) / den);
      ax = ax - kx * q;
      by = by - ky * q;
      cz = cz - kzm * q;
    }
    cpx<R> r = mk<R>(0, 0);
    if (L.scalar) {
      r = rp[m] - p.grad * st[o + 2 * fs];
      if (zero) r = mk<R>(0, 0);
    }
    if (p.r.dealias && k2 >= p.r.kc2) {
      ax = mk<R>(0, 0);
      by = ax;
      cz = ax;
      r = ax;
    }
    if constexpr (MODE == 0) {
      W[o] = ax;
      W[o + fs] = by;
      W[o + 2 * fs] = cz;
      if (L.scalar) W[o + 3 * fs] = r;
    } else {
      const bool cut = k2 >= p.r.kc2, mean = ik2 == 0;
      ext_apply<MODE, R>(in[0], ax, o, V, DU, S, p.r.nu, k2, cut, mean, p.r);
      ext_apply<MODE, R>(in[1], by, o + fs, V, DU, S, p.r.nu, k2, cut, mean, p.r);
      ext_apply<MODE, R>(in[2], cz, o + 2 * fs, V, DU, S, p.r.nu, k2, cut, mean, p.r);
      if (L.scalar) ext_apply<MODE, R>(in[3], r, o + 3 * fs, V, DU, S, p.nu_phi, k2, cut, mean, p.r);
    }
  }
}

// --------------------------------------------------------------------------
// Dynamic coefficient (dynsmag.f90:27-194) building blocks.
// --------------------------------------------------------------------------
// D[0..2] = v, D[3..5] = g v  (dynsmag.f90:28-33)
template <class R>
__global__ __launch_bounds__(256) void k_dyn_vel(const cpx<R>* __restrict__ st, const R* __restrict__ gt,
                                                 cpx<R>* __restrict__ D, int64_t fs, Geom g) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.ntot) return;
  const R f = gt[i];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const cpx<R> v = st[i + q * fs];
    D[i + q * fs] = v;
    D[i + (3 + q) * fs] = f * v;
  }
}

// physical: P[q] = u_i u_j for the six pairs (dynsmag.f90:43-48)
template <class R>
__global__ __launch_bounds__(256) void k_dyn_products(const cpx<R>* __restrict__ U, cpx<R>* __restrict__ P, int64_t fs,
                                                      Geom g) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.nmain) return;
  const cpx<R> u[3] = {U[i], U[i + fs], U[i + 2 * fs]};
  const int PR[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
#pragma unroll
  for (int q = 0; q < 6; ++q) P[i + q * fs] = mk<R>(u[PR[q][0]].x * u[PR[q][1]].x, u[PR[q][0]].y * u[PR[q][1]].y);
}

// spectral: A[q] = g * B[q] * scale for q < n (in place when A == B)
template <class R>
__global__ __launch_bounds__(256) void k_dyn_filter(const cpx<R>* B, cpx<R>* A, int64_t fs, int n,
                                                    const R* __restrict__ gt, R scale, Geom g) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.ntot) return;
  const R f = gt[i] * scale;
  for (int q = 0; q < n; ++q) A[i + q * fs] = f * B[i + q * fs];
}

// physical: L[q] -= ut_i ut_j (dynsmag.f90:71-76)
template <class R>
__global__ __launch_bounds__(256) void k_dyn_leonard(cpx<R>* __restrict__ Lf, const cpx<R>* __restrict__ Ut,
                                                     int64_t fs, Geom g) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.nmain) return;
  const cpx<R> u[3] = {Ut[i], Ut[i + fs], Ut[i + 2 * fs]};
  const int PR[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const cpx<R> l = Lf[i + q * fs];
    Lf[i + q * fs] = mk<R>(l.x - u[PR[q][0]].x * u[PR[q][1]].x, l.y - u[PR[q][0]].y * u[PR[q][1]].y);
  }
}

// spectral: A[0..5] = S_ij(v), B[0..5] = g S_ij(v) (dynsmag.f90:79-94)
template <class R>
__global__ __launch_bounds__(256) void k_dyn_strain(const cpx<R>* __restrict__ st, const R* __restrict__ gt,
                                                    cpx<R>* __restrict__ A, cpx<R>* __restrict__ B, int64_t fs,
                                                    Geom g) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.ntot) return;
  int x, y, z, ikx, iky, ikz;
  ks_of(g, i, x, y, z, ikx, iky, ikz);
  const cpx<R> v[3] = {st[i], st[i + fs], st[i + 2 * fs]};
  cpx<R> s[6];
  strain_of<R>(v, (R)ikx, (R)iky, (R)ikz, s);
  const R f = gt[i];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    A[i + q * fs] = s[q];
    B[i + q * fs] = f * s[q];
  }
}

// physical, in place: S_ij <- c |S| S_ij (dynsmag.f90:104-117, 133-146)
template <class R>
__global__ __launch_bounds__(256) void k_dyn_smag(cpx<R>* __restrict__ A, int64_t fs, R c, Geom g) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.nmain) return;
  cpx<R> s[6];
  R sx[6], sy[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    s[q] = A[i + q * fs];
    sx[q] = s[q].x;
    sy[q] = s[q].y;
  }
  const R ax = abs_strain<R>(sx), ay = abs_strain<R>(sy);
#pragma unroll
  for (int q = 0; q < 6; ++q) A[i + q * fs] = mk<R>(c * ax * s[q].x, c * ay * s[q].y);
}

// physical: A[q] += B[q], q < 6 (dynsmag.f90:164-169)
template <class R>
__global__ __launch_bounds__(256) void k_dyn_add(cpx<R>* __restrict__ A, const cpx<R>* __restrict__ B, int64_t fs,
                                                 Geom g) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.nmain) return;
#pragma unroll
  for (int q = 0; q < 6; ++q) A[i + q * fs] = A[i + q * fs] + B[i + q * fs];
}

// Grid sums mm = sum M:M and lm = sum L:M (off-diagonals twice) over all
// physical points (dynsmag.f90:171-184), FP64: per-block partials in a
// fixed order, then k_diag_final.
template <class R>
__global__ __launch_bounds__(256) void k_dyn_dots(const cpx<R>* __restrict__ Lf, const cpx<R>* __restrict__ M,
                                                  int64_t fs, Geom g, double* __restrict__ part) {
  double mm = 0, lm = 0;
  const double w[6] = {1, 2, 2, 1, 2, 1};
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < g.nmain; i += (int64_t)gridDim.x * 256) {
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const cpx<R> m = M[i + q * fs], l = Lf[i + q * fs];
      mm += w[q] * ((double)m.x * m.x + (double)m.y * m.y);
      lm += w[q] * ((double)l.x * m.x + (double)l.y * m.y);
    }
  }
  __shared__ double red[2][256];
  red[0][threadIdx.x] = mm;
  red[1][threadIdx.x] = lm;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = red[0][0];
    part[2 * blockIdx.x + 1] = red[1][0];
  }
}

// out[j] = sum over blocks b of part[2b + j], j = 0, 1, in block order
static __global__ void k_sum_pairs(const double* __restrict__ part, int nblocks, double* __restrict__ out) {
  const int j = threadIdx.x;
  if (j >= 2) return;
  double s = 0;
  for (int b = 0; b < nblocks; ++b) s += part[2 * b + j];
  out[j] = s;
}

// cs2 = lm / mm from the summed partials, into coef[5..7] = {cs2, mm, lm}
static __global__ void k_dyn_cs2(const double* __restrict__ sums, double* __restrict__ coef) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    coef[6] = sums[0];
    coef[7] = sums[1];
    coef[5] = sums[1] / sums[0];
  }
}

// tau_ij = -cs2 * D[q] (spectral), cs2 = coef[5]
template <class R>
__global__ __launch_bounds__(256) void k_tau_from_s(cpx<R>* __restrict__ D, int64_t fs, Geom g,
                                                    const double* __restrict__ coef) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.ntot) return;
  const R c = (R)(-coef[5]);
#pragma unroll
  for (int q = 0; q < 6; ++q) D[i + q * fs] = c * D[i + q * fs];
}

// g = 1 where k2 <= kt2 (spectral_les.f90:105-109), in the split layout
template <class R>
__global__ __launch_bounds__(256) void k_gtest(R* __restrict__ gt, Geom g, R kt2) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.ntot) return;
  int x, y, z, ikx, iky, ikz;
  ks_of(g, i, x, y, z, ikx, iky, ikz);
  const int ik2 = ikx * ikx + iky * iky + ikz * ikz;
  const R k2 = ik2 == 0 ? std::numeric_limits<R>::min() : (R)ik2;
  gt[i] = k2 <= kt2 ? (R)1 : (R)0;
}

// a real array in the reference layout (lx1, ly, lz) -> split layout
template <class R>
__global__ __launch_bounds__(256) void k_ref_to_split_real(const R* __restrict__ ref, R* __restrict__ f, Geom g) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t n = g.nrow * (g.mx + 1);
  if (i >= n) return;
  const int64_t rr = i / (g.mx + 1);  // reference row z*ny + y
  const int x = (int)(i % (g.mx + 1));
  const int64_t row = g.row((int)(rr % g.ny), (int)(rr / g.ny));
  if (x < g.mx) f[row * g.mx + x] = ref[i];
  else f[g.nmain + row] = ref[i];
}

// convecphi_dns.f90:26-47 in the reference's sign: W[0] = i k . F where
// F = W[0..2] (spectral fluxes, already scaled)
template <class R>
__global__ __launch_bounds__(256) void k_divflux(cpx<R>* __restrict__ W, int64_t fs, Geom g) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.ntot) return;
  int x, y, z, ikx, iky, ikz;
  ks_of(g, i, x, y, z, ikx, iky, ikz);
  W[i] = mul_i((R)ikx * W[i] + (R)iky * W[i + fs] + (R)ikz * W[i + 2 * fs]);
}

// physical: F_j = u_j phi * scale into W[0..2] from u = W[0..2], phi = W[3]
template <class R>
__global__ __launch_bounds__(256) void k_flux(cpx<R>* __restrict__ W, int64_t fs, R scale, Geom g) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.nmain) return;
  const cpx<R> p = W[i + 3 * fs];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const cpx<R> u = W[i + q * fs];
    W[i + q * fs] = mk<R>(u.x * p.x * scale, u.y * p.y * scale);
  }
}

}  // namespace scl

namespace scl {

// Scalar variance sum' |phi|^2 and sum' k^2 |phi|^2 (half weight on kx = 0),
// FP64 per-block partials (2 per block) summed by k_sum_pairs.
template <class R>
__global__ __launch_bounds__(256) void k_scalar_diag(const cpx<R>* __restrict__ phi, Geom g, double* __restrict__ part) {
  double v = 0, d = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < g.ntot; i += (int64_t)gridDim.x * 256) {
    int x, y, z, kx, ky, kz;
    ks_of(g, i, x, y, z, kx, ky, kz);
    const cpx<R> a = phi[i];
    const double w = (x == 0 ? 0.5 : 1.0) * ((double)a.x * a.x + (double)a.y * a.y);
    v += w;
    d += ((double)kx * kx + (double)ky * ky + (double)kz * kz) * w;
  }
  __shared__ double red[2][256];
  red[0][threadIdx.x] = v;
  red[1][threadIdx.x] = d;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = red[0][0];
    part[2 * blockIdx.x + 1] = red[1][0];
  }
}

}  // namespace scl
