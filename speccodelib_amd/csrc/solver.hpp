// Device fields of the IF-RK4 solver and one stage of it, shared by the
// solver (capi.hip) and the discrete adjoint (adjoint.hip).
#pragma once

#include "plan.hpp"

namespace scl {

// Solver fields: V(nc) DU(nc) S(nc) W(...), nc = 3 velocity components
// (+1 scalar), each fs entries; coef: forcing {a, b, E_f, Omega_f, uw} and
// Smagorinsky {cs2, mm, lm} (8 doubles on the device).  s_out, when set,
// replaces S as the stage output (the adjoint's recomputation writes the
// stage states straight into its tape).
struct DevFields {
  void* base;
  int64_t fs;
  int nc = 3;
  double* coef = nullptr;
  void* s_out = nullptr;
  void* phys_out = nullptr;  // the x-pass keeps u, w [, S_ij] in physical space here (adjoint tape)
  unsigned* nonfinite = nullptr;  // sticky flag: stage 4 found a non-finite v (null: not checked)
  template <class R>
  cpx<R>* f(int i) const {
    return reinterpret_cast<cpx<R>*>(base) + i * fs;
  }
  template <class R>
  cpx<R>* V() const { return f<R>(0); }
  template <class R>
  cpx<R>* DU() const { return f<R>(nc); }
  template <class R>
  cpx<R>* S() const { return s_out ? reinterpret_cast<cpx<R>*>(s_out) : f<R>(2 * nc); }
  template <class R>
  cpx<R>* W() const { return f<R>(3 * nc); }
};

// IF-RK4 stage `mode` (1..4; 0: N(st) into W) of the solver configuration
// prm on the fields d, reading the stage state st: the fused DNS RHS, or
// the multi-pass RHS when prm asks for a Smagorinsky term or a scalar.
template <class R>
int32_t solver_stage(scl_plan3d* p, const DevFields& d, int mode, const scl_dns_params& prm, const cpx<R>* st);

}  // namespace scl
