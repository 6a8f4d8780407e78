// Launchers of the LES + scalar RHS and the dynamic Smagorinsky update.
#pragma once

#include "les.hpp"
#include "ops.hpp"

namespace scl {

template <class R>
struct ExtOps {
  using C = cpx<R>;
  static dim3 grid_all(const Geom& g) { return dim3((unsigned)((g.ntot + 255) / 256)); }
  static dim3 grid_main(const Geom& g) { return dim3((unsigned)((g.nmain + 255) / 256)); }

  // The physical pointwise pass (k_ext_phys) of the layout L.
  static int32_t phys(const scl_plan3d& pl, C* W, int64_t fs, ExtLayout L, R scale, R d2, const double* cs2p,
                      C* tape) {
    const Geom& g = pl.g;
    if (L.les && L.scalar)
      hipLaunchKernelGGL((k_ext_phys<R, 1, 1>), grid_main(g), dim3(256), 0, pl.stream, W, fs, g, scale, d2, cs2p, tape);
    else if (L.les)
      hipLaunchKernelGGL((k_ext_phys<R, 1, 0>), grid_main(g), dim3(256), 0, pl.stream, W, fs, g, scale, d2, cs2p, tape);
    else if (L.scalar)
      hipLaunchKernelGGL((k_ext_phys<R, 0, 1>), grid_main(g), dim3(256), 0, pl.stream, W, fs, g, scale, d2, cs2p, tape);
    else
      return set_error("ext_phys: neither a Smagorinsky term nor a scalar"), SCL_EINVAL;
    return check_launch("ext_phys");
  }

  // N(st) of the extended RHS into W (MODE 0) or one RK4 stage (MODE 1..4).
  // W holds max(nin, nout) fields.  phys_tape (optional): where the
  // physical-space u, w, S_ij of this evaluation are kept (k_ext_phys).
  static int32_t rhs(const scl_plan3d& pl, int mode, ExtLayout L, C* W, int64_t fs, C* V, C* DU, C* S, const C* st,
                     const double* coef, const ExtParams<R>& ep, R delta2, C* phys_tape = nullptr) {
    const Geom& g = pl.g;
    // 1024-point z- or y-lines keep the separate prep and combine: the
    // fused passes spill there
    const bool fused = pl.zfused;
    if (fused) {
      // prep fused into the z-inverse
      SCL_CHECK(Ops<R>::ext_zinv(pl, st, W, fs, L));
    } else {
      hipLaunchKernelGGL((k_ext_prep<R>), grid_all(g), dim3(256), 0, pl.stream, st, W, fs, g, L);
      SCL_CHECK(check_launch("ext_prep"));
      SCL_CHECK(Ops<R>::colfft(pl, W, fs, L.nin(), 2, +1));
    }
    // y-inverse (fused: w and S_ij formed there); the x-pass with the
    // physical products; y-forward
    const R scale = (R)1 / (R)((double)g.nx * g.ny * g.nz);
    if (fused) SCL_CHECK(Ops<R>::ext_yinv(pl, W, fs, L));
    else SCL_CHECK(Ops<R>::colfft(pl, W, fs, L.nin(), 1, +1));
    SCL_CHECK(Ops<R>::row_ext(pl, W, fs, L, scale, delta2, coef + 5, phys_tape));
    // fused: the y-forward folds the kx, ky terms and the z-forward the kz
    // terms and the combine; else plain passes and the combine kernel
    if (fused) {
      SCL_CHECK(Ops<R>::ext_yfwd(pl, W, fs, L, (R)(-ep.sign_phi)));
      return Ops<R>::ext_zfwd(pl, mode, W, fs, V, DU, S, st, coef, ep, L);
    }
    SCL_CHECK(Ops<R>::colfft(pl, W, fs, L.nout(), 1, -1));
    SCL_CHECK(Ops<R>::colfft(pl, W, fs, L.nout(), 2, -1));
    switch (mode) {
      case 0: hipLaunchKernelGGL((k_ext_combine<R, 0>), grid_all(g), dim3(256), 0, pl.stream, W, fs, V, DU, S, st, coef, ep, g, L); break;
      case 1: hipLaunchKernelGGL((k_ext_combine<R, 1>), grid_all(g), dim3(256), 0, pl.stream, W, fs, V, DU, S, st, coef, ep, g, L); break;
      case 2: hipLaunchKernelGGL((k_ext_combine<R, 2>), grid_all(g), dim3(256), 0, pl.stream, W, fs, V, DU, S, st, coef, ep, g, L); break;
      case 3: hipLaunchKernelGGL((k_ext_combine<R, 3>), grid_all(g), dim3(256), 0, pl.stream, W, fs, V, DU, S, st, coef, ep, g, L); break;
      case 4: hipLaunchKernelGGL((k_ext_combine<R, 4>), grid_all(g), dim3(256), 0, pl.stream, W, fs, V, DU, S, st, coef, ep, g, L); break;
      case 5: hipLaunchKernelGGL((k_ext_combine<R, 5>), grid_all(g), dim3(256), 0, pl.stream, W, fs, V, DU, S, st, coef, ep, g, L); break;
      case 6: hipLaunchKernelGGL((k_ext_combine<R, 6>), grid_all(g), dim3(256), 0, pl.stream, W, fs, V, DU, S, st, coef, ep, g, L); break;
      case 7: hipLaunchKernelGGL((k_ext_combine<R, 7>), grid_all(g), dim3(256), 0, pl.stream, W, fs, V, DU, S, st, coef, ep, g, L); break;
      default: set_error("bad mode %d", mode); return SCL_EINVAL;
    }
    return check_launch("ext_combine");
  }

  // Dynamic coefficient from the velocity st (dynsmag.f90:27-194) into
  // coef[5..7] = {cs2, mm, lm}; tau (for the host-array routine) is
  // -cs2 * D[0..5] afterwards (spectral 2 D^2 |S| S_ij / N^3).
  // D: 24 scratch fields; gt: test filter (real, split layout); part: FP64
  // partial-sum scratch of 2 * DIAG_BLOCKS + 2 doubles.
  static int32_t dyn_update(const scl_plan3d& pl, const C* st, C* D, int64_t fs, const R* gt, R delta, R delta_test,
                            double* coef, double* part) {
    const Geom& g = pl.g;
    const R scale = (R)1 / (R)((double)g.nx * g.ny * g.nz);
    C* U = D;             // 0..2 u, 3..5 filtered u
    C* Lf = D + 6 * fs;   // 6..11 L_ij
    C* Sg = D + 12 * fs;  // 12..17 filtered S_ij, then -2 Dt^2 |St| St_ij
    C* M = D + 18 * fs;   // 18..23 M_ij
    hipLaunchKernelGGL((k_dyn_vel<R>), grid_all(g), dim3(256), 0, pl.stream, st, gt, U, fs, g);
    SCL_CHECK(check_launch("dyn_vel"));
    SCL_CHECK(Ops<R>::c2r3d(pl, U, fs, 6));
    hipLaunchKernelGGL((k_dyn_products<R>), grid_main(g), dim3(256), 0, pl.stream, U, Lf, fs, g);
    SCL_CHECK(check_launch("dyn_products"));
    SCL_CHECK(Ops<R>::r2c3d(pl, Lf, fs, 6));
    hipLaunchKernelGGL((k_dyn_filter<R>), grid_all(g), dim3(256), 0, pl.stream, Lf, Lf, fs, 6, gt, scale, g);
    SCL_CHECK(check_launch("dyn_filter"));
    SCL_CHECK(Ops<R>::c2r3d(pl, Lf, fs, 6));
    hipLaunchKernelGGL((k_dyn_leonard<R>), grid_main(g), dim3(256), 0, pl.stream, Lf, U + 3 * fs, fs, g);
    SCL_CHECK(check_launch("dyn_leonard"));
    // 2 D^2 |S| S_ij / N^3 (spectral) in D[0..5]; filtered S in Sg
    hipLaunchKernelGGL((k_dyn_strain<R>), grid_all(g), dim3(256), 0, pl.stream, st, gt, U, Sg, fs, g);
    SCL_CHECK(check_launch("dyn_strain"));
    SCL_CHECK(Ops<R>::c2r3d(pl, U, fs, 6));
    hipLaunchKernelGGL((k_dyn_smag<R>), grid_main(g), dim3(256), 0, pl.stream, U, fs, (R)2 * delta * delta * scale, g);
    SCL_CHECK(check_launch("dyn_smag"));
    SCL_CHECK(Ops<R>::r2c3d(pl, U, fs, 6));
    SCL_CHECK(Ops<R>::c2r3d(pl, Sg, fs, 6));
    hipLaunchKernelGGL((k_dyn_smag<R>), grid_main(g), dim3(256), 0, pl.stream, Sg, fs,
                       (R)-2 * delta_test * delta_test, g);
    SCL_CHECK(check_launch("dyn_smag_t"));
    hipLaunchKernelGGL((k_dyn_filter<R>), grid_all(g), dim3(256), 0, pl.stream, U, M, fs, 6, gt, (R)1, g);
    SCL_CHECK(check_launch("dyn_filter_s"));
    SCL_CHECK(Ops<R>::c2r3d(pl, M, fs, 6));
    hipLaunchKernelGGL((k_dyn_add<R>), grid_main(g), dim3(256), 0, pl.stream, M, Sg, fs, g);
    SCL_CHECK(check_launch("dyn_add"));
    hipLaunchKernelGGL((k_dyn_dots<R>), dim3(DIAG_BLOCKS_DYN), dim3(256), 0, pl.stream, Lf, M, fs, g, part);
    SCL_CHECK(check_launch("dyn_dots"));
    double* sums = part + 2 * DIAG_BLOCKS_DYN;
    hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(2), 0, pl.stream, part, DIAG_BLOCKS_DYN, sums);
    SCL_CHECK(check_launch("dyn_sum"));
    hipLaunchKernelGGL(k_dyn_cs2, dim3(1), dim3(1), 0, pl.stream, sums, coef);
    return check_launch("dyn_cs2");
  }

  static constexpr int DIAG_BLOCKS_DYN = 512;

  static int32_t gtest(const scl_plan3d& pl, R* gt, R k_test) {
    hipLaunchKernelGGL((k_gtest<R>), grid_all(pl.g), dim3(256), 0, pl.stream, gt, pl.g, k_test * k_test);
    return check_launch("gtest");
  }
};

}  // namespace scl
