// Host-side plan and solver objects behind the opaque C handles.
#pragma once

#include <vector>

#include "scl_core.hpp"

// Extra complex entries between consecutive fields.  Packed fields sit a
// multiple of a large power of two apart at the common grid sizes
// (2^29 + 2^21 bytes at 512^3), so the same entry of V, DU, S and W falls on
// the same HBM channels; a 64 KB (FP32) offset spreads them (3% faster
// steps at 512^3, measured).
#ifndef SCL_FIELD_PAD
#define SCL_FIELD_PAD 8192
#endif

struct scl_plan3d {
  int prec = SCL_SP;
  scl::Geom g{};
  int device = 0;
  int ncu = 256;  // compute units of the device
  // the LES / scalar RHS and the adjoint products with their prep and
  // combine fused into the z- and y-passes (off for 1024-point z- or
  // y-lines, where those kernels spill)
  bool zfused = true;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  void* tw = nullptr;  // twiddle table exp(-2 pi i j/L), j < L, in the plan's precision
  int L = 0;
  void* scratch = nullptr;  // host-array entry points: 7 fields (6 work + 1 staging)
  scl_plan3d* padded = nullptr;  // the padded grid of the 3/2-rule products (made on first use)
  size_t scratch_fields = 0;
  double* coef = nullptr;     // forcing coefficients (8 doubles)
  void* diag = nullptr;       // diagnostics partial sums
  size_t diag_bytes = 0;
  size_t esize() const { return prec == SCL_SP ? 2 * sizeof(float) : 2 * sizeof(double); }
  // complex entries from one field to the next in the split device layout
  int64_t fstride() const { return g.ntot + SCL_FIELD_PAD; }
  // bytes of one field in the reference layout (lx1, ly, lz)
  size_t ref_bytes() const { return (size_t)(g.mx + 1) * g.nrow * esize(); }
};

namespace scl {
int32_t plan_scratch(scl_plan3d* p, size_t nfields, void** out);
int32_t plan_diag(scl_plan3d* p, size_t bytes, void** out);
}  // namespace scl

struct scl_dns {
  scl_plan3d* plan = nullptr;
  scl_dns_params p{};
  // nc state components (3 velocity [+ 1 scalar]); fields V(nc) DU(nc)
  // S(nc) W(nw), each plan->fstride() entries
  int nc = 3, nw = 6;
  bool ext = false;          // LES and/or scalar: the multi-pass RHS
  void* fields = nullptr;
  void* dyn = nullptr;       // dynamic coefficient: 24 scratch fields
  void* gt = nullptr;        // test filter (real, split layout)
  double* part = nullptr;    // FP64 partial sums of the dynamic update
  double* coef = nullptr;    // forcing and Smagorinsky coefficients (8 doubles), then the
                             // sticky non-finite flag (an unsigned at coef + 8)
  double time = 0.0;
  int64_t steps = 0;
  // W[0..4] hold the z-inverse of the current v (u and w_x, w_y, left there
  // by the last step's fused stage 4): the next step's stage 1 needs no
  // z-inverse pass.  Cleared whenever v or W change outside a step.
  bool chained = false;
  // IF-AB2 (p.stepper 1, 2): device state (AB_* slots), energy partials, and
  // the steps taken since the state was set (selects the bootstrap modes)
  double* ab = nullptr;
  double* ab_part = nullptr;
  int64_t ab_steps = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<hipEvent_t> events;
};
