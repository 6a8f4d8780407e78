// Definitions of Ops<R>; instantiated once per precision (ops_sp.hip,
// ops_dp.hip) so the two precisions compile in parallel.
#pragma once

#include "adjoint.hpp"
#include "les.hpp"
#include "ops.hpp"

namespace scl {

template <int V>
using IC = std::integral_constant<int, V>;

// Column/line lengths supported per precision (FP64 1024-point lines run
// on 4-column tiles: ColCfg).
template <class R>
constexpr int max_line() {
  return 1024;
}

template <class R, class F>
int32_t with_line(int n, F&& f) {
  switch (n) {
    case 16: return f(IC<16>{});
    case 32: return f(IC<32>{});
    case 64: return f(IC<64>{});
    case 128: return f(IC<128>{});
    case 256: return f(IC<256>{});
    case 512:
      if constexpr (max_line<R>() >= 512) return f(IC<512>{});
      break;
    case 1024:
      if constexpr (max_line<R>() >= 1024) return f(IC<1024>{});
      break;
    default: break;
  }
  set_error("unsupported line length %d for this precision", n);
  return SCL_EINVAL;
}

template <class R, class F>
int32_t with_half(int m, F&& f) {  // m = nx/2
  switch (m) {
    case 8: return f(IC<8>{});
    case 16: return f(IC<16>{});
    case 32: return f(IC<32>{});
    case 64: return f(IC<64>{});
    case 128: return f(IC<128>{});
    case 256:
      if constexpr (max_line<R>() >= 512) return f(IC<256>{});
      break;
    case 512:
      if constexpr (max_line<R>() >= 1024) return f(IC<512>{});
      break;
    default: break;
  }
  set_error("unsupported nx/2 = %d for this precision", m);
  return SCL_EINVAL;
}

template <class R>
static inline const cpx<R>* twp(const scl_plan3d& pl) {
  return reinterpret_cast<const cpx<R>*>(pl.tw);
}

static inline dim3 grid1(int64_t n, int bs = 256) { return dim3((unsigned)((n + bs - 1) / bs)); }

template <class R, int N, int CW = ColCfg<R, N>::CW>
static int32_t check_cols(const Geom& g) {
  if (g.mx < CW || g.mx % CW || g.ny % CW || g.nz % CW) {
    set_error("grid (%d,%d,%d) too small for %d-column tiles", g.nx, g.ny, g.nz, CW);
    return SCL_EINVAL;
  }
  return SCL_OK;
}

template <class R>
int32_t Ops<R>::colfft(const scl_plan3d& pl, C* f0, int64_t fstride, int nf, int axis, int dir) {
  const Geom& g = pl.g;
  const int n = axis == 1 ? g.ny : g.nz;
  return with_line<R>(n, [&](auto NC) -> int32_t {
    constexpr int N = decltype(NC)::value;
    using CC = ColCfg<R, N>;
    SCL_CHECK((check_cols<R, N>(g)));
    const dim3 blk(CC::BS);
    const cpx<R>* tw = twp<R>(pl);
    const int tws = pl.L / N;
    if (axis == 1) {
      const dim3 grd(col_grid<CC::CW, 1>(g), nf);
      if (dir < 0) hipLaunchKernelGGL((k_colfft<R, N, 1, -1>), grd, blk, 0, pl.stream, f0, fstride, g, tw, tws);
      else hipLaunchKernelGGL((k_colfft<R, N, 1, 1>), grd, blk, 0, pl.stream, f0, fstride, g, tw, tws);
    } else {
      const dim3 grd(col_grid<CC::CW, 2>(g), nf);
      if (dir < 0) hipLaunchKernelGGL((k_colfft<R, N, 2, -1>), grd, blk, 0, pl.stream, f0, fstride, g, tw, tws);
      else hipLaunchKernelGGL((k_colfft<R, N, 2, 1>), grd, blk, 0, pl.stream, f0, fstride, g, tw, tws);
    }
    return check_launch("colfft");
  });
}

// y-inverse of the five z-pass outputs plus w_z: u_x, u_y, w_z in
// k_yinv_vort (three fields live), then u_z, w_x, w_y in the plain column
// kernel (one field live, twice the waves per CU; measured 1193 vs 1237 us
// per RHS at 512^3 for one kernel doing all six).
template <class R>
int32_t Ops<R>::yinv_vort(const scl_plan3d& pl, C* W, int64_t fstride) {
  const Geom& g = pl.g;
  return with_line<R>(g.ny, [&](auto NC) -> int32_t {
    constexpr int N = decltype(NC)::value;
    using CC = ColCfg<R, N>;
    SCL_CHECK((check_cols<R, N>(g)));
    hipLaunchKernelGGL((k_yinv_vort<R, N>), dim3(col_grid<CC::CW, 1>(g)), dim3(CC::BS), 0, pl.stream, W, fstride, g,
                       twp<R>(pl), pl.L / N);
    SCL_CHECK(check_launch("yinv_vort"));
    return colfft(pl, W + 2 * fstride, fstride, 3, 1, 1);
  });
}

// Column count of the fused z-pass kernels for a grid.  1024-point lines
// stay at 8 columns: 16 would make 1024-thread workgroups, whose 128-VGPR
// cap spills these kernels.
#ifndef SCL_ZCOL16
#define SCL_ZCOL16 1
#endif
template <class R>
static bool zcol16(const Geom& g) {
  return SCL_ZCOL16 && sizeof(R) == 4 && g.mx % 16 == 0 && g.ny % 16 == 0 && g.nz <= 512;
}

// Points per thread of the 16-column fused z-pass kernels (0: the line
// length's default).
#ifndef SCL_Z16_E
#define SCL_Z16_E 0
#endif
// Points per thread of the 1024-point FP32 z-lines (0: 16).  8 points make
// 1024-thread workgroups at 4 waves per SIMD, but measured slower at 1024^3
// than 16 points at 2 waves per SIMD (13.1 vs 12.6 ms z-inverse, 19.0 vs
// 17.6 ms z-forward): one more LDS exchange per transform.
#ifndef SCL_Z1024_E
#define SCL_Z1024_E 0
#endif
template <class R, int N, int CW>
constexpr int zpass_e() {
  if (sizeof(R) == 4 && N == 1024 && CW == 8 && SCL_Z1024_E) return SCL_Z1024_E;
  return (CW == 16 && SCL_Z16_E && N >= SCL_Z16_E) ? SCL_Z16_E : 0;
}

// The fused z-pass kernels are compiled for the row orders zs = 0 and 3.
inline int32_t check_zs(const Geom& g) {
  if (g.zs != 0 && g.zs != 3) return set_error("row order zs = %d: the z-pass kernels take 0 or 3", g.zs), SCL_EINVAL;
  return SCL_OK;
}

template <class R>
int32_t Ops<R>::zinv_vort(const scl_plan3d& pl, const C* s, int64_t sstride, C* w, int64_t wstride) {
  const Geom& g = pl.g;
  return with_line<R>(g.nz, [&](auto NC) -> int32_t {
    constexpr int N = decltype(NC)::value;
    auto go = [&](auto CWC) -> int32_t {
      constexpr int CW = decltype(CWC)::value;
      constexpr int EZ = zpass_e<R, N, CW>();
      using CC = ColCfg<R, N, CW, EZ>;
      SCL_CHECK((check_cols<R, N, CW>(g)));
      SCL_CHECK(check_zs(g));
      if (g.zs == 3)
        hipLaunchKernelGGL((k_zinv_vort<R, N, CW, EZ, 3>), dim3(col_grid<CW, 2>(g)), dim3(CC::BS), 0, pl.stream, s,
                           sstride, w, wstride, g, twp<R>(pl), pl.L / N);
      else
        hipLaunchKernelGGL((k_zinv_vort<R, N, CW, EZ, 0>), dim3(col_grid<CW, 2>(g)), dim3(CC::BS), 0, pl.stream, s,
                           sstride, w, wstride, g, twp<R>(pl), pl.L / N);
      return check_launch("zinv_vort");
    };
    if constexpr (sizeof(R) == 4 && N <= 512) {
      if (zcol16<R>(g)) return go(IC<16>{});
    }
    return go(IC<ColCfg<R, N>::CW>{});
  });
}

template <class R>
int32_t Ops<R>::zfwd_rhs(const scl_plan3d& pl, int mode, C* W, int64_t fstride, C* V, C* DU, C* S, const C* st,
                         const double* coef, const RhsParams<R>& p, bool fused) {
  const Geom& g = pl.g;
  return with_line<R>(g.nz, [&](auto NC) -> int32_t {
    constexpr int N = decltype(NC)::value;
    auto go = [&](auto CWC) -> int32_t {
      constexpr int CW = decltype(CWC)::value;
      constexpr int EZ = zpass_e<R, N, CW>();
      using CC = ColCfg<R, N, CW, EZ>;
      SCL_CHECK((check_cols<R, N, CW>(g)));
      SCL_CHECK(check_zs(g));
      const dim3 grd(col_grid<CW, 2>(g)), blk(CC::BS);
      const cpx<R>* tw = twp<R>(pl);
      const int tws = pl.L / N;
      auto launch = [&](auto MC, auto FC, auto ZC) {
        constexpr int MODE = decltype(MC)::value, ZS = decltype(ZC)::value;
        constexpr bool FUSED = decltype(FC)::value != 0;
        hipLaunchKernelGGL((k_zfwd_rhs<R, N, MODE, CW, EZ, FUSED, ZS>), grd, blk, 0, pl.stream, W, fstride, V, DU, S, st,
                           coef, p, g, tw, tws);
      };
      auto by_zs = [&](auto MC, auto FC) {
        if (g.zs == 3) launch(MC, FC, IC<3>{});
        else launch(MC, FC, IC<0>{});
      };
      if (fused) {
        switch (mode) {
          case 1: by_zs(IC<1>{}, IC<1>{}); break;
          case 2: by_zs(IC<2>{}, IC<1>{}); break;
          case 3: by_zs(IC<3>{}, IC<1>{}); break;
          case 4: by_zs(IC<4>{}, IC<1>{}); break;
          case 5: by_zs(IC<5>{}, IC<1>{}); break;
          case 6: by_zs(IC<6>{}, IC<1>{}); break;
          case 7: by_zs(IC<7>{}, IC<1>{}); break;
          default: set_error("fused z-forward: mode %d is not a stage 1..7", mode); return SCL_EINVAL;
        }
        return check_launch("zfwd_zinv");
      }
      switch (mode) {
        case 0: by_zs(IC<0>{}, IC<0>{}); break;
        case 1: by_zs(IC<1>{}, IC<0>{}); break;
        case 2: by_zs(IC<2>{}, IC<0>{}); break;
        case 3: by_zs(IC<3>{}, IC<0>{}); break;
        case 4: by_zs(IC<4>{}, IC<0>{}); break;
        default: set_error("bad mode %d", mode); return SCL_EINVAL;
      }
      return check_launch("zfwd_rhs");
    };
    if constexpr (sizeof(R) == 4 && N <= 512) {
      if (zcol16<R>(g)) return go(IC<16>{});
    }
    return go(IC<ColCfg<R, N>::CW>{});
  });
}

template <class R>
int32_t Ops<R>::ext_zinv(const scl_plan3d& pl, const C* st, C* W, int64_t fs, const ExtLayout& L) {
  const Geom& g = pl.g;
  return with_line<R>(g.nz, [&](auto NC) -> int32_t {
    constexpr int N = decltype(NC)::value;
    using CC = ColCfg<R, N>;
    SCL_CHECK((check_cols<R, N>(g)));
    if constexpr (N >= 1024) {
      return set_error("fused LES z-pass: 1024-point lines take the unfused path"), SCL_EINVAL;
    } else {
      hipLaunchKernelGGL((k_ext_zinv<R, N>), dim3(col_grid<CC::CW, 2>(g)), dim3(CC::BS), 0, pl.stream, st, W, fs, g, L,
                         twp<R>(pl), pl.L / N);
      return check_launch("ext_zinv");
    }
  });
}

template <class R>
int32_t Ops<R>::ext_yinv(const scl_plan3d& pl, C* W, int64_t fs, const ExtLayout& L) {
  const Geom& g = pl.g;
  if (!L.les) return colfft(pl, W, fs, L.nin(), 1, +1);
  return with_line<R>(g.ny, [&](auto NC) -> int32_t {
    constexpr int N = decltype(NC)::value;
    using CC = ColCfg<R, N>;
    SCL_CHECK((check_cols<R, N>(g)));
    const dim3 grd(col_grid<CC::CW, 1>(g)), blk(CC::BS);
    hipLaunchKernelGGL((k_yinv_les<R, N, 0>), grd, blk, 0, pl.stream, W, fs, g, twp<R>(pl), pl.L / N);
    SCL_CHECK(check_launch("yinv_les"));
    hipLaunchKernelGGL((k_yinv_les<R, N, 1>), grd, blk, 0, pl.stream, W, fs, g, twp<R>(pl), pl.L / N);
    SCL_CHECK(check_launch("yinv_les"));
    // S_33 (= i kz u_z after the z-pass) and phi: a plain y-inverse
    return colfft(pl, W + (L.s_in() + 5) * fs, fs, 1 + L.scalar, 1, +1);
  });
}

template <class R>
int32_t Ops<R>::ext_yfwd(const scl_plan3d& pl, C* W, int64_t fs, const ExtLayout& L, R sg) {
  const Geom& g = pl.g;
  return with_line<R>(g.ny, [&](auto NC) -> int32_t {
    constexpr int N = decltype(NC)::value;
    using CC = ColCfg<R, N>;
    SCL_CHECK((check_cols<R, N>(g)));
    const dim3 grd(col_grid<CC::CW, 1>(g)), blk(CC::BS);
    const cpx<R>* tw = twp<R>(pl);
    const int tws = pl.L / N;
    if (!L.les && !L.scalar) return set_error("ext_yfwd: neither a Smagorinsky term nor a scalar"), SCL_EINVAL;
    if (L.les) {
      if (L.scalar)
        hipLaunchKernelGGL((k_yfwd_ext<R, N, 0, 1>), grd, blk, 0, pl.stream, W, fs, g, sg, tw, tws);
      else
        hipLaunchKernelGGL((k_yfwd_ext<R, N, 0, 0>), grd, blk, 0, pl.stream, W, fs, g, sg, tw, tws);
      SCL_CHECK(check_launch("yfwd_ext"));
    } else {
      SCL_CHECK(colfft(pl, W, fs, 3, 1, -1));  // N alone
    }
    if (!L.scalar) return SCL_OK;
    // the flux slots are W[3..5] with or without the Smagorinsky term
    hipLaunchKernelGGL((k_yfwd_ext<R, N, 1, 1>), grd, blk, 0, pl.stream, W, fs, g, sg, tw, tws);
    return check_launch("yfwd_ext");
  });
}

template <class R>
int32_t Ops<R>::ext_zfwd(const scl_plan3d& pl, int mode, C* W, int64_t fs, C* V, C* DU, C* S, const C* st,
                         const double* coef, const ExtParams<R>& ep, const ExtLayout& L) {
  const Geom& g = pl.g;
  return with_line<R>(g.nz, [&](auto NC) -> int32_t {
    constexpr int N = decltype(NC)::value;
    using CC = ColCfg<R, N>;
    SCL_CHECK((check_cols<R, N>(g)));
    if constexpr (N >= 1024) {
      return set_error("fused LES z-pass: 1024-point lines take the unfused path"), SCL_EINVAL;
    } else {
      const dim3 grd(col_grid<CC::CW, 2>(g)), blk(CC::BS);
      const cpx<R>* tw = twp<R>(pl);
      const int tws = pl.L / N;
      switch (mode) {
        case 0: hipLaunchKernelGGL((k_ext_zfwd<R, N, 0>), grd, blk, 0, pl.stream, W, fs, V, DU, S, st, coef, ep, g, L, tw, tws); break;
        case 1: hipLaunchKernelGGL((k_ext_zfwd<R, N, 1>), grd, blk, 0, pl.stream, W, fs, V, DU, S, st, coef, ep, g, L, tw, tws); break;
        case 2: hipLaunchKernelGGL((k_ext_zfwd<R, N, 2>), grd, blk, 0, pl.stream, W, fs, V, DU, S, st, coef, ep, g, L, tw, tws); break;
        case 3: hipLaunchKernelGGL((k_ext_zfwd<R, N, 3>), grd, blk, 0, pl.stream, W, fs, V, DU, S, st, coef, ep, g, L, tw, tws); break;
        case 4: hipLaunchKernelGGL((k_ext_zfwd<R, N, 4>), grd, blk, 0, pl.stream, W, fs, V, DU, S, st, coef, ep, g, L, tw, tws); break;
        case 5: hipLaunchKernelGGL((k_ext_zfwd<R, N, 5>), grd, blk, 0, pl.stream, W, fs, V, DU, S, st, coef, ep, g, L, tw, tws); break;
        case 6: hipLaunchKernelGGL((k_ext_zfwd<R, N, 6>), grd, blk, 0, pl.stream, W, fs, V, DU, S, st, coef, ep, g, L, tw, tws); break;
        case 7: hipLaunchKernelGGL((k_ext_zfwd<R, N, 7>), grd, blk, 0, pl.stream, W, fs, V, DU, S, st, coef, ep, g, L, tw, tws); break;
        default: set_error("bad mode %d", mode); return SCL_EINVAL;
      }
      return check_launch("ext_zfwd");
    }
  });
}

template <class R>
int32_t Ops<R>::adj_zinv(const scl_plan3d& pl, const C* s, const C* Y, C* W, int64_t fs, const AdjLayout& L,
                         bool state) {
  const Geom& g = pl.g;
  return with_line<R>(g.nz, [&](auto NC) -> int32_t {
    constexpr int N = decltype(NC)::value;
    using CC = ColCfg<R, N>;
    SCL_CHECK((check_cols<R, N>(g)));
    if constexpr (N >= 1024) {
      return set_error("fused adjoint z-pass: 1024-point lines take the unfused path"), SCL_EINVAL;
    } else {
      if (state)
        hipLaunchKernelGGL((k_adj_zinv<R, N, true>), dim3(col_grid<CC::CW, 2>(g)), dim3(CC::BS), 0, pl.stream, s, Y, W,
                           fs, g, L, twp<R>(pl), pl.L / N);
      else
        hipLaunchKernelGGL((k_adj_zinv<R, N, false>), dim3(col_grid<CC::CW, 2>(g)), dim3(CC::BS), 0, pl.stream, s, Y, W,
                           fs, g, L, twp<R>(pl), pl.L / N);
      return check_launch("adj_zinv");
    }
  });
}

template <class R>
int32_t Ops<R>::force_coef(const scl_plan3d& pl, const C* st, int64_t fstride, double fkmax, double eps, double eta,
                           int type, double* coef) {
  const R fk = (R)fkmax;
  const R fk2 = fk * fk;
  int fbox = (int)std::floor(fkmax);
  if (fbox < 0) fbox = 0;
  hipLaunchKernelGGL((k_force_coef<R>), dim3(1), dim3(256), 0, pl.stream, st, fstride, pl.g, fk2, fbox, eps, eta, type,
                     coef);
  return check_launch("force_coef");
}

template <class R>
int32_t Ops<R>::herm_fix(const scl_plan3d& pl, const FieldPtrs<R>& fp) {
  const int64_t n = herm_count(pl.g);
  if (n <= 0) return SCL_OK;
  hipLaunchKernelGGL((k_herm_fix<R>), grid1(n), dim3(256), 0, pl.stream, fp, pl.g);
  return check_launch("herm_fix");
}

template <class R>
int32_t Ops<R>::row_convec(const scl_plan3d& pl, C* W, int64_t fstride, C* tape) {
  const Geom& g = pl.g;
  const R scale = (R)1 / (R)((double)g.nx * g.ny * g.nz);
  return with_half<R>(g.mx, [&](auto MC) -> int32_t {
    constexpr int M = decltype(MC)::value;
    using RC = RowCfg<R, M>;
    // persistent grid: as many workgroups as can be resident at once (the
    // tag gives each variant its own cached occupancy)
    auto go = [&](auto tag, auto kern) -> int32_t {
      (void)tag;
      static int resident = 0;
      if (!resident) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, RC::BS, 0) != hipSuccess || per_cu < 1)
          per_cu = 1;
        resident = per_cu;
      }
      const int64_t groups = g.nrow / RC::RPW;
      const int64_t grid = std::min<int64_t>(groups, (int64_t)resident * pl.ncu);
      hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(RC::BS), 0, pl.stream, W, fstride, g, scale, twp<R>(pl),
                         pl.L / M, pl.L / g.nx, tape);
      return check_launch("row_convec");
    };
    if (tape) return go(IC<1>{}, k_row_convec<R, M, true>);
    return go(IC<0>{}, k_row_convec<R, M, false>);
  });
}

template <class R>
int32_t Ops<R>::row_ext(const scl_plan3d& pl, C* W, int64_t fs, const ExtLayout& L, R scale, R d2,
                        const double* cs2p, C* tape) {
  const Geom& g = pl.g;
  if (!L.les && !L.scalar) return set_error("row_ext: neither a Smagorinsky term nor a scalar"), SCL_EINVAL;
  return with_half<R>(g.mx, [&](auto MC) -> int32_t {
    constexpr int M = decltype(MC)::value;
    using RC = RowCfg<R, M>;
    // persistent grid, as for row_convec (the tag gives each variant its own
    // cached occupancy)
    auto go = [&](auto tag, auto kern) -> int32_t {
      (void)tag;
      static int per_cu = 0;
      if (!per_cu && (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, RC::BS, 0) != hipSuccess ||
                      per_cu < 1))
        per_cu = 1;
      const int64_t groups = g.nrow / RC::RPW;
      const int64_t grid = std::min<int64_t>(groups, (int64_t)per_cu * pl.ncu);
      hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(RC::BS), 0, pl.stream, W, fs, g, scale, d2, cs2p, tape,
                         twp<R>(pl), pl.L / M, pl.L / g.nx);
      return check_launch("row_ext");
    };
    if (tape) {
      if (L.les && L.scalar) return go(IC<0>{}, k_row_ext<R, M, 1, 1, true>);
      if (L.les) return go(IC<1>{}, k_row_ext<R, M, 1, 0, true>);
      return set_error("row_ext: the tape needs the Smagorinsky term"), SCL_EINVAL;
    }
    if (L.les && L.scalar) return go(IC<2>{}, k_row_ext<R, M, 1, 1, false>);
    if (L.les) return go(IC<3>{}, k_row_ext<R, M, 1, 0, false>);
    return go(IC<4>{}, k_row_ext<R, M, 0, 1, false>);
  });
}

template <class R>
int32_t Ops<R>::row_adj(const scl_plan3d& pl, C* W, int64_t fs, R scale, R c, R cc, double* part, const C* P,
                        int les) {
  const Geom& g = pl.g;
  return with_half<R>(g.mx, [&](auto MC) -> int32_t {
    constexpr int M = decltype(MC)::value;
    using RC = RowCfg<R, M>;
    // persistent grid, as for row_convec
    auto go = [&](auto tag, auto kern) -> int32_t {
      (void)tag;
      static int per_cu = 0;
      if (!per_cu && (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, RC::BS, 0) != hipSuccess ||
                      per_cu < 1))
        per_cu = 1;
      const int64_t groups = g.nrow / RC::RPW;
      const int64_t grid = std::min<int64_t>(groups, (int64_t)per_cu * pl.ncu);
      hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(RC::BS), 0, pl.stream, W, fs, g, scale, c, cc, part, P,
                         twp<R>(pl), pl.L / M, pl.L / g.nx);
      return check_launch("row_adj");
    };
    if (les) return go(IC<1>{}, k_row_adj<R, M, 1>);
    return go(IC<0>{}, k_row_adj<R, M, 0>);
  });
}

template <class R>
int32_t Ops<R>::row_c2r(const scl_plan3d& pl, C* W, int64_t fstride, int nf) {
  const Geom& g = pl.g;
  return with_half<R>(g.mx, [&](auto MC) -> int32_t {
    constexpr int M = decltype(MC)::value;
    using RC = RowCfg<R, M>;
    hipLaunchKernelGGL((k_row_c2r<R, M>), dim3((unsigned)(g.nrow / RC::RPW), nf), dim3(RC::BS), 0, pl.stream, W,
                       fstride, g, twp<R>(pl), pl.L / M, pl.L / g.nx);
    return check_launch("row_c2r");
  });
}

template <class R>
int32_t Ops<R>::row_r2c(const scl_plan3d& pl, C* W, int64_t fstride, int nf) {
  const Geom& g = pl.g;
  return with_half<R>(g.mx, [&](auto MC) -> int32_t {
    constexpr int M = decltype(MC)::value;
    using RC = RowCfg<R, M>;
    hipLaunchKernelGGL((k_row_r2c<R, M>), dim3((unsigned)(g.nrow / RC::RPW), nf), dim3(RC::BS), 0, pl.stream, W,
                       fstride, g, twp<R>(pl), pl.L / M, pl.L / g.nx);
    return check_launch("row_r2c");
  });
}

template <class R>
int32_t Ops<R>::ref_to_split(const scl_plan3d& pl, const C* ref, C* f) {
  hipLaunchKernelGGL((k_ref_to_split<R>), grid1(pl.g.nrow * (pl.g.mx + 1)), dim3(256), 0, pl.stream, ref, f, pl.g);
  return check_launch("ref_to_split");
}

template <class R>
int32_t Ops<R>::split_to_ref(const scl_plan3d& pl, const C* f, C* ref) {
  hipLaunchKernelGGL((k_split_to_ref<R>), grid1(pl.g.nrow * (pl.g.mx + 1)), dim3(256), 0, pl.stream, f, ref, pl.g);
  return check_launch("split_to_ref");
}

template <class R>
int32_t Ops<R>::sym_zero(const scl_plan3d& pl, C* f, R thr) {
  hipLaunchKernelGGL((k_sym_zero<R>), grid1(pl.g.ntot), dim3(256), 0, pl.stream, f, pl.g, thr);
  return check_launch("sym_zero");
}

template <class R>
int32_t Ops<R>::dealias_mask(const scl_plan3d& pl, C* f, R thr) {
  hipLaunchKernelGGL((k_dealias_mask<R>), grid1(pl.g.ntot), dim3(256), 0, pl.stream, f, pl.g, thr);
  return check_launch("dealias_mask");
}

template <class R>
int32_t Ops<R>::projection(const scl_plan3d& pl, C* v, int64_t fstride) {
  hipLaunchKernelGGL((k_projection<R>), grid1(pl.g.ntot), dim3(256), 0, pl.stream, v, fstride, pl.g);
  return check_launch("projection");
}

template <class R>
int32_t Ops<R>::force_apply(const scl_plan3d& pl, C* w, const C* v, int64_t fstride, const double* coef, R fk2,
                            int type) {
  hipLaunchKernelGGL((k_force_apply<R>), grid1(pl.g.ntot), dim3(256), 0, pl.stream, w, v, fstride, pl.g, coef, fk2,
                     type);
  return check_launch("force_apply");
}

template <class R>
int32_t Ops<R>::c2r3d(const scl_plan3d& pl, C* f0, int64_t fstride, int nf) {
  SCL_CHECK(colfft(pl, f0, fstride, nf, 2, +1));
  SCL_CHECK(colfft(pl, f0, fstride, nf, 1, +1));
  return row_c2r(pl, f0, fstride, nf);
}

template <class R>
int32_t Ops<R>::r2c3d(const scl_plan3d& pl, C* f0, int64_t fstride, int nf) {
  SCL_CHECK(row_r2c(pl, f0, fstride, nf));
  SCL_CHECK(colfft(pl, f0, fstride, nf, 1, -1));
  return colfft(pl, f0, fstride, nf, 2, -1);
}

}  // namespace scl
