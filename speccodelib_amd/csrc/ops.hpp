// Typed launchers: runtime sizes -> kernel template instances.
#pragma once

#include <type_traits>

#include "kernels3d.hpp"
#include "plan.hpp"

namespace scl {

struct ExtLayout;
template <class R>
struct ExtParams;
struct AdjLayout;

template <class R>
struct Ops {
  using C = cpx<R>;
  static int32_t colfft(const scl_plan3d& pl, C* f0, int64_t fstride, int nf, int axis, int dir);
  static int32_t zinv_vort(const scl_plan3d& pl, const C* s, int64_t sstride, C* w, int64_t wstride);
  static int32_t yinv_vort(const scl_plan3d& pl, C* W, int64_t fstride);
  // fused (mode 1..4): also z-inverse the new stage state (mode 4: the new
  // v) into W[0..4] for the next RHS (k_zfwd_rhs FUSED)
  static int32_t zfwd_rhs(const scl_plan3d& pl, int mode, C* W, int64_t fstride, C* V, C* DU, C* S, const C* st,
                          const double* coef, const RhsParams<R>& p, bool fused = false);
  // the LES / scalar RHS's z-passes fused with its prep and combine
  // (k_ext_zinv, k_ext_zfwd); st and W share the field stride fs
  static int32_t ext_zinv(const scl_plan3d& pl, const C* st, C* W, int64_t fs, const ExtLayout& L);
  // the y-inverse after ext_zinv (k_yinv_les with the Smagorinsky term)
  static int32_t ext_yinv(const scl_plan3d& pl, C* W, int64_t fs, const ExtLayout& L);
  // the y-forward before ext_zfwd, folding the kx, ky terms (k_yfwd_ext);
  // sg = -sign of the scalar flux divergence
  static int32_t ext_yfwd(const scl_plan3d& pl, C* W, int64_t fs, const ExtLayout& L, R sg);
  // the x-pass of the LES / scalar RHS with its physical products
  // (k_row_ext); tape (optional): u, w, S_ij in physical space
  static int32_t row_ext(const scl_plan3d& pl, C* W, int64_t fs, const ExtLayout& L, R scale, R d2,
                         const double* cs2p, C* tape);
  static int32_t ext_zfwd(const scl_plan3d& pl, int mode, C* W, int64_t fs, C* V, C* DU, C* S, const C* st,
                          const double* coef, const ExtParams<R>& ep, const ExtLayout& L);
  // the adjoint VJP's z-inverse fused with its prep (k_adj_zinv); all
  // fields share the stride fs; state false: only ybar and taubar (the
  // state's u, w, S_ij come from a physical-space tape)
  static int32_t adj_zinv(const scl_plan3d& pl, const C* s, const C* Y, C* W, int64_t fs, const AdjLayout& L,
                          bool state = true);
  // the x-pass of the Smagorinsky VJP with the physical tape P (k_row_adj)
  static int32_t row_adj(const scl_plan3d& pl, C* W, int64_t fs, R scale, R c, R cc, double* part, const C* P,
                         int les);
  static int32_t force_coef(const scl_plan3d& pl, const C* st, int64_t fstride, double fkmax, double eps, double eta,
                            int type, double* coef);
  static int32_t herm_fix(const scl_plan3d& pl, const FieldPtrs<R>& fp);
  // tape (optional): u, w in physical space (k_row_convec TAPE)
  static int32_t row_convec(const scl_plan3d& pl, C* W, int64_t fstride, C* tape = nullptr);
  static int32_t row_c2r(const scl_plan3d& pl, C* W, int64_t fstride, int nf);
  static int32_t row_r2c(const scl_plan3d& pl, C* W, int64_t fstride, int nf);
  static int32_t ref_to_split(const scl_plan3d& pl, const C* ref, C* f);
  static int32_t split_to_ref(const scl_plan3d& pl, const C* f, C* ref);
  static int32_t sym_zero(const scl_plan3d& pl, C* f, R thr);
  static int32_t dealias_mask(const scl_plan3d& pl, C* f, R thr);
  static int32_t projection(const scl_plan3d& pl, C* v, int64_t fstride);
  static int32_t force_apply(const scl_plan3d& pl, C* w, const C* v, int64_t fstride, const double* coef, R fk2,
                             int type);
  // full 3D transforms on fields in the split layout, in place
  static int32_t c2r3d(const scl_plan3d& pl, C* f0, int64_t fstride, int nf);
  static int32_t r2c3d(const scl_plan3d& pl, C* f0, int64_t fstride, int nf);
};

extern template struct Ops<float>;
extern template struct Ops<double>;

int32_t check_launch(const char* what);

}  // namespace scl
