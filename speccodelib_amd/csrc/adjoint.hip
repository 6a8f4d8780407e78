// C ABI of the discrete adjoint: checkpointed forward, per-segment
// recomputation into a stage tape, reverse sweep of vector-Jacobian products.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "adjoint.hpp"
#include "ops.hpp"
#include "solver.hpp"

namespace scl {

// Fixed-order sum of n partials into *acc (one workgroup).
__global__ __launch_bounds__(256) void k_adj_sum(const double* __restrict__ part, int n, double* __restrict__ acc) {
  __shared__ double red[256];
  double s = 0.0;
  for (int j = threadIdx.x; j < n; j += 256) s += part[j];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) *acc = red[0];
}

}  // namespace scl

using namespace scl;

struct scl_adjoint {
  scl_plan3d* plan = nullptr;
  scl_dns_params p{};
  int nsteps = 0, ckpt = 0, nseg = 0;
  int nw = 6;                 // work fields of the forward stage
  AdjLayout L{0};
  void* fields = nullptr;     // forward: V(3) DU(3) S(3) W(nw)
  void* work = nullptr;       // adjoint: WA(L.nin()) LAM(3) VB(3) Y(3)
  void* tape = nullptr;       // ckpt steps x 4 stage states x 3 fields
  void* ptape = nullptr;      // ckpt steps x 4 stages x pf physical fields (u, w [, S_ij]), or none
  int pf = 6;                 // fields per stage on the physical tape: 6, 12 with the Smagorinsky term
  void* ckpts = nullptr;      // nseg x 3 fields
  void* staging = nullptr;    // one field in the reference layout
  double* coef = nullptr;     // 8 doubles (cs2 at [5])
  double* part = nullptr;     // per-block partials of the cs2 derivative
  double* acc = nullptr;      // the cs2 derivative
  int nparts = 0;
  double cs2 = 0.0;           // C_s^2 of the kernels' coefficient slot (coef[5])
  double fwd_cs2 = 0.0;       // C_s^2 of the last forward: the backward's
  bool have_forward = false;
  bool last_taped = false;    // the tape holds the last segment of the last forward
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double fwd_ms = 0.0, bwd_ms = 0.0;
};

namespace {

template <class R>
cpx<R>* fld(void* base, int64_t fs, int i) {
  return reinterpret_cast<cpx<R>*>(base) + i * fs;
}

// reference layout (device or host) -> split field
template <class R>
int32_t load_field(scl_adjoint* a, const void* src, cpx<R>* dst, bool dev) {
  if (!src) return set_error("null array"), SCL_EINVAL;
  scl_plan3d* p = a->plan;
  const cpx<R>* ref = reinterpret_cast<const cpx<R>*>(src);
  if (!dev) {
    SCL_HIP(hipMemcpyAsync(a->staging, src, p->ref_bytes(), hipMemcpyHostToDevice, p->stream));
    ref = reinterpret_cast<const cpx<R>*>(a->staging);
  }
  return Ops<R>::ref_to_split(*p, ref, dst);
}

template <class R>
int32_t store_field(scl_adjoint* a, const cpx<R>* src, void* dst, bool dev) {
  if (!dst) return set_error("null array"), SCL_EINVAL;
  scl_plan3d* p = a->plan;
  if (dev) return Ops<R>::split_to_ref(*p, src, reinterpret_cast<cpx<R>*>(dst));
  SCL_CHECK(Ops<R>::split_to_ref(*p, src, reinterpret_cast<cpx<R>*>(a->staging)));
  SCL_HIP(hipMemcpyAsync(dst, a->staging, p->ref_bytes(), hipMemcpyDeviceToHost, p->stream));
  return SCL_OK;
}

template <class R>
int32_t set_cs2(scl_adjoint* a, double cs2) {
  a->cs2 = cs2;
  double c[8] = {0, 0, 0, 0, 0, cs2, 0, 0};
  SCL_HIP(hipMemcpyAsync(a->coef, c, sizeof(c), hipMemcpyHostToDevice, a->plan->stream));
  SCL_HIP(hipStreamSynchronize(a->plan->stream));  // c is on the host stack
  return SCL_OK;
}

DevFields dev(const scl_adjoint* a) {
  DevFields d{a->fields, a->plan->fstride(), 3, a->coef};
  return d;
}

template <class R>
cpx<R>* tape_slot(const scl_adjoint* a, int step, int stage) {
  const int64_t fs = a->plan->fstride();
  return fld<R>(a->tape, fs, 3 * (4 * step + stage));
}

// Physical-space inputs of the pointwise pass of stage `stage` (1..4) of a
// step, or null without the physical tape.
template <class R>
cpx<R>* ptape_slot(const scl_adjoint* a, int step, int stage) {
  if (!a->ptape) return nullptr;
  const int64_t fs = a->plan->fstride();
  return fld<R>(a->ptape, fs, a->pf * (4 * step + stage - 1));
}

// One forward step of the solver configuration on V; with tape_step >= 0
// the stage states v, s2, s3, s4 go to the tape, and with the physical tape
// each stage's u, w, S_ij in physical space to it.
template <class R>
int32_t fwd_step(scl_adjoint* a, int tape_step) {
  DevFields d = dev(a);
  scl_plan3d* p = a->plan;
  const int64_t fs = p->fstride();
  if (tape_step >= 0)
    SCL_HIP(hipMemcpyAsync(tape_slot<R>(a, tape_step, 0), d.V<R>(), 3 * fs * sizeof(cpx<R>),
                           hipMemcpyDeviceToDevice, p->stream));
  for (int stage = 1; stage <= 4; ++stage) {
    const cpx<R>* st = stage == 1 ? d.V<R>() : d.S<R>();
    if (tape_step >= 0) {
      st = tape_slot<R>(a, tape_step, stage - 1);
      d.s_out = stage < 4 ? (void*)tape_slot<R>(a, tape_step, stage) : nullptr;
      d.phys_out = ptape_slot<R>(a, tape_step, stage);
    }
    SCL_CHECK(solver_stage<R>(p, d, stage, a->p, st));
  }
  return SCL_OK;
}

// The cs2 derivative: the sum of the per-block partials that k_adj_phys
// adds up over the products since part[] was cleared.
int32_t sum_parts(scl_adjoint* a) {
  hipLaunchKernelGGL(k_adj_sum, dim3(1), dim3(256), 0, a->plan->stream, a->part, a->nparts, a->acc);
  return check_launch("adj_sum");
}

template <class R>
AdjParams<R> adj_params(const scl_adjoint* a) {
  AdjParams<R> ap;
  ap.dt = (R)a->p.dt;
  ap.nu = (R)a->p.nu;
  const float kcut = 2.0f * (float)a->plan->g.mx / 3.0f;  // the forward's cut (dns-deform.f90:75-76)
  ap.kc2 = (R)kcut * (R)kcut;
  return ap;
}

// J(s)^T applied to the Y field (P nbar, before the Hermitian transpose),
// then the bookkeeping of `mode`.  ph: u, w, S_ij of s in physical space
// (the physical tape), or null to transform them here.
template <class R>
int32_t vjp(scl_adjoint* a, const cpx<R>* s, int mode, const cpx<R>* ph = nullptr) {
  scl_plan3d* p = a->plan;
  const Geom& g = p->g;
  const int64_t fs = p->fstride();
  cpx<R>* WA = fld<R>(a->work, fs, 0);
  const int nin = a->L.nin();
  cpx<R>* LAM = fld<R>(a->work, fs, nin);
  cpx<R>* VB = fld<R>(a->work, fs, nin + 3);
  cpx<R>* Y = fld<R>(a->work, fs, nin + 6);
  const int64_t nh = (int64_t)(g.nz / 2 - 1) * g.ny + g.ny / 2 - 1;
  hipLaunchKernelGGL((k_adj_hermT<R>), dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, p->stream, Y, fs, g);
  SCL_CHECK(check_launch("adj_hermT"));
  // prep fused into the z-inverse (1024-point z-lines, or SCL_EXT_UNFUSED:
  // separate passes).
  // The combine stays a separate pass: fused with the z-forward of the 12
  // outputs it made the 128^3 backward slower, at one wave per SIMD (568 vs
  // 536 ms) and again at two with a few spilled registers (343 vs 337 ms).
  const double delta = a->p.delta > 0 ? a->p.delta : 2.0 * M_PI / g.nx;
  const R scale = (R)1 / (R)((double)g.nx * g.ny * g.nz);
  const R d2 = (R)(delta * delta);
  const R c = (R)(-2.0 * a->cs2) * d2 * scale;
  const R cc = (R)-2 * d2 * scale;
  const AdjParams<R> ap = adj_params<R>(a);
  if (ph) {
    // only ybar [and taubar] (from W[y_in]) are transformed; the x-pass
    // reads u, w [, S_ij] from the tape and does the physical products
    const int q0 = a->L.y_in();
    SCL_CHECK(Ops<R>::adj_zinv(*p, s, Y, WA, fs, a->L, false));
    SCL_CHECK(Ops<R>::colfft(*p, WA + q0 * fs, fs, nin - q0, 1, +1));
    SCL_CHECK(Ops<R>::row_adj(*p, WA, fs, scale, c, cc, a->part, ph, a->L.les));
    SCL_CHECK(Ops<R>::colfft(*p, WA, fs, a->L.nout(), 1, -1));
    SCL_CHECK(Ops<R>::colfft(*p, WA, fs, a->L.nout(), 2, -1));
  } else {
    if (p->zfused) {
      SCL_CHECK(Ops<R>::adj_zinv(*p, s, Y, WA, fs, a->L));
      SCL_CHECK(Ops<R>::colfft(*p, WA, fs, nin, 1, +1));
      SCL_CHECK(Ops<R>::row_c2r(*p, WA, fs, nin));
    } else {
      hipLaunchKernelGGL((k_adj_prep<R>), dim3((unsigned)((g.ntot + 255) / 256)), dim3(256), 0, p->stream, s, Y, WA,
                         fs, g, a->L);
      SCL_CHECK(check_launch("adj_prep"));
      SCL_CHECK(Ops<R>::c2r3d(*p, WA, fs, nin));
    }
    const unsigned nb = (unsigned)((g.nmain + 255) / 256);
    if (a->L.les)
      hipLaunchKernelGGL((k_adj_phys<R, 1>), dim3(nb), dim3(256), 0, p->stream, WA, fs, g, scale, c, cc, a->part,
                         nullptr);
    else
      hipLaunchKernelGGL((k_adj_phys<R, 0>), dim3(nb), dim3(256), 0, p->stream, WA, fs, g, scale, c, cc, nullptr,
                         nullptr);
    SCL_CHECK(check_launch("adj_phys"));
    SCL_CHECK(Ops<R>::r2c3d(*p, WA, fs, a->L.nout()));
  }
  const dim3 ga((unsigned)((g.ntot + 255) / 256));
  switch (mode) {
    case 0: hipLaunchKernelGGL((k_adj_combine<R, 0>), ga, dim3(256), 0, p->stream, WA, fs, LAM, VB, Y, g, a->L, ap); break;
    case 1: hipLaunchKernelGGL((k_adj_combine<R, 1>), ga, dim3(256), 0, p->stream, WA, fs, LAM, VB, Y, g, a->L, ap); break;
    case 2: hipLaunchKernelGGL((k_adj_combine<R, 2>), ga, dim3(256), 0, p->stream, WA, fs, LAM, VB, Y, g, a->L, ap); break;
    case 3: hipLaunchKernelGGL((k_adj_combine<R, 3>), ga, dim3(256), 0, p->stream, WA, fs, LAM, VB, Y, g, a->L, ap); break;
    case 4: hipLaunchKernelGGL((k_adj_combine<R, 4>), ga, dim3(256), 0, p->stream, WA, fs, LAM, VB, Y, g, a->L, ap); break;
    default: return set_error("bad adjoint mode %d", mode), SCL_EINVAL;
  }
  return check_launch("adj_combine");
}

template <class R>
int32_t forward(scl_adjoint* a, double cs2, const void* const* u0, void* const* uT, bool devp) {
  scl_plan3d* p = a->plan;
  const int64_t fs = p->fstride();
  DevFields d = dev(a);
  SCL_CHECK(set_cs2<R>(a, cs2));
  a->fwd_cs2 = cs2;
  for (int q = 0; q < 3; ++q) SCL_CHECK(load_field<R>(a, u0[q], d.V<R>() + q * fs, devp));
  // The last segment is taped on the way: the backward starts with it and
  // need not recompute it.
  a->last_taped = false;
  const int nlast = (a->nseg - 1) * a->ckpt;
  SCL_HIP(hipEventRecord(a->e0, p->stream));
  for (int n = 0; n < a->nsteps; ++n) {
    if (n % a->ckpt == 0)
      SCL_HIP(hipMemcpyAsync(fld<R>(a->ckpts, fs, 3 * (n / a->ckpt)), d.V<R>(), 3 * fs * sizeof(cpx<R>),
                             hipMemcpyDeviceToDevice, p->stream));
    SCL_CHECK(fwd_step<R>(a, n >= nlast ? n - nlast : -1));
  }
  a->last_taped = true;
  SCL_HIP(hipEventRecord(a->e1, p->stream));
  for (int q = 0; q < 3; ++q) SCL_CHECK(store_field<R>(a, d.V<R>() + q * fs, uT[q], devp));
  SCL_HIP(hipStreamSynchronize(p->stream));
  float ms = 0;
  SCL_HIP(hipEventElapsedTime(&ms, a->e0, a->e1));
  a->fwd_ms = ms;
  a->have_forward = true;
  return SCL_OK;
}

template <class R>
int32_t backward(scl_adjoint* a, const void* const* gT, void* const* g0, double* dcs2, bool devp) {
  if (!a->have_forward) return set_error("scl_adjoint_backward needs a preceding scl_adjoint_forward"), SCL_EINVAL;
  scl_plan3d* p = a->plan;
  const Geom& g = p->g;
  const int64_t fs = p->fstride();
  DevFields d = dev(a);
  const int nin = a->L.nin();
  cpx<R>* LAM = fld<R>(a->work, fs, nin);
  cpx<R>* VB = fld<R>(a->work, fs, nin + 3);
  cpx<R>* Y = fld<R>(a->work, fs, nin + 6);
  // the coefficient slot may hold another C_s^2 since (scl_adjoint_vjp sets
  // its own): recomputation and products run at the last forward's
  SCL_CHECK(set_cs2<R>(a, a->fwd_cs2));
  for (int q = 0; q < 3; ++q) SCL_CHECK(load_field<R>(a, gT[q], LAM + q * fs, devp));
  SCL_HIP(hipMemsetAsync(a->part, 0, a->nparts * sizeof(double), p->stream));
  const AdjParams<R> ap = adj_params<R>(a);
  const dim3 ga((unsigned)((g.ntot + 255) / 256));
  SCL_HIP(hipEventRecord(a->e0, p->stream));
  for (int seg = a->nseg - 1; seg >= 0; --seg) {
    const int n0 = seg * a->ckpt, n1 = std::min(a->nsteps, n0 + a->ckpt);
    if (!(seg == a->nseg - 1 && a->last_taped)) {
      a->last_taped = false;
      SCL_HIP(hipMemcpyAsync(d.V<R>(), fld<R>(a->ckpts, fs, 3 * seg), 3 * fs * sizeof(cpx<R>),
                             hipMemcpyDeviceToDevice, p->stream));
      for (int n = n0; n < n1; ++n) SCL_CHECK(fwd_step<R>(a, n - n0));
    }
    for (int n = n1 - 1; n >= n0; --n) {
      hipLaunchKernelGGL((k_adj_init<R>), ga, dim3(256), 0, p->stream, LAM, VB, Y, fs, g, ap);
      SCL_CHECK(check_launch("adj_init"));
      for (int stage = 4; stage >= 1; --stage)
        SCL_CHECK(vjp<R>(a, tape_slot<R>(a, n - n0, stage - 1), stage, ptape_slot<R>(a, n - n0, stage)));
      // vbar is the gradient with respect to u_n: the next (earlier) step's lambda
      SCL_HIP(hipMemcpyAsync(LAM, VB, 3 * fs * sizeof(cpx<R>), hipMemcpyDeviceToDevice, p->stream));
    }
  }
  SCL_CHECK(sum_parts(a));
  SCL_HIP(hipEventRecord(a->e1, p->stream));
  for (int q = 0; q < 3; ++q) SCL_CHECK(store_field<R>(a, LAM + q * fs, g0[q], devp));
  double c = 0.0;
  SCL_HIP(hipMemcpyAsync(&c, a->acc, sizeof(double), hipMemcpyDeviceToHost, p->stream));
  SCL_HIP(hipStreamSynchronize(p->stream));
  if (dcs2) *dcs2 = c;
  float ms = 0;
  SCL_HIP(hipEventElapsedTime(&ms, a->e0, a->e1));
  a->bwd_ms = ms;
  return SCL_OK;
}

template <class R>
int32_t vjp_host(scl_adjoint* a, double cs2, const void* const* s, const void* const* nb, void* const* out,
                 double* dcs2, bool devp) {
  scl_plan3d* p = a->plan;
  const int64_t fs = p->fstride();
  const int nin = a->L.nin();
  cpx<R>* VB = fld<R>(a->work, fs, nin + 3);
  cpx<R>* Y = fld<R>(a->work, fs, nin + 6);
  cpx<R>* S = fld<R>(a->tape, fs, 0);
  a->last_taped = false;  // S overwrites the first tape slot
  SCL_CHECK(set_cs2<R>(a, cs2));
  for (int q = 0; q < 3; ++q) SCL_CHECK(load_field<R>(a, s[q], S + q * fs, devp));
  for (int q = 0; q < 3; ++q) SCL_CHECK(load_field<R>(a, nb[q], Y + q * fs, devp));
  SCL_CHECK(Ops<R>::projection(*p, Y, fs));
  SCL_HIP(hipMemsetAsync(a->part, 0, a->nparts * sizeof(double), p->stream));
  SCL_CHECK(vjp<R>(a, S, 0));
  SCL_CHECK(sum_parts(a));
  for (int q = 0; q < 3; ++q) SCL_CHECK(store_field<R>(a, VB + q * fs, out[q], devp));
  double c = 0.0;
  SCL_HIP(hipMemcpyAsync(&c, a->acc, sizeof(double), hipMemcpyDeviceToHost, p->stream));
  SCL_HIP(hipStreamSynchronize(p->stream));
  if (dcs2) *dcs2 = c;
  return SCL_OK;
}

template <class F>
int32_t by_prec_a(const scl_adjoint* a, F&& f) {
  if (a->plan->prec == SCL_SP) return f(float{});
  return f(double{});
}

}  // namespace

extern "C" int32_t scl_adjoint_destroy(scl_adjoint* a) {
  if (!a) return SCL_OK;
  if (a->plan && a->plan->stream) (void)hipStreamSynchronize(a->plan->stream);
  for (void* ptr : {a->fields, a->work, a->tape, a->ptape, a->ckpts, a->staging}) {
    if (ptr) (void)hipFree(ptr);
  }
  if (a->coef) (void)hipFree(a->coef);
  if (a->part) (void)hipFree(a->part);
  if (a->acc) (void)hipFree(a->acc);
  if (a->e0) (void)hipEventDestroy(a->e0);
  if (a->e1) (void)hipEventDestroy(a->e1);
  delete a;
  return SCL_OK;
}

extern "C" int32_t scl_adjoint_create(scl_plan3d* p, const scl_dns_params* prm, int32_t nsteps, int32_t ckpt_every,
                                      scl_adjoint** out) {
  if (!out || !prm || !p) return set_error("null argument"), SCL_EINVAL;
  *out = nullptr;
  if (prm->forced || prm->scalar || prm->dealias_rhs)
    return set_error("the adjoint covers the unforced momentum equation without scalar or RHS truncation"), SCL_EINVAL;
  if (prm->sgs != 0 && prm->sgs != 1) return set_error("the adjoint needs sgs = 0 or 1 (static Smagorinsky)"), SCL_EINVAL;
  if (nsteps < 1 || ckpt_every < 1) return set_error("nsteps and ckpt_every must be >= 1"), SCL_EINVAL;
  auto* a = new scl_adjoint();
  a->plan = p;
  a->p = *prm;
  a->nsteps = nsteps;
  a->ckpt = std::min(ckpt_every, nsteps);
  a->nseg = (nsteps + a->ckpt - 1) / a->ckpt;
  a->L = AdjLayout{prm->sgs ? 1 : 0};
  a->nw = prm->sgs ? ExtLayout{1, 0}.nin() : 6;
  const int64_t fs = p->fstride();
  const size_t fb = (size_t)fs * p->esize();
  auto alloc = [&](void** ptr, size_t bytes) -> bool {
    if (hipMalloc(ptr, bytes) != hipSuccess) {
      (void)hipGetLastError();
      set_error("device allocation of %zu bytes failed", bytes);
      return false;
    }
    return true;
  };
  const Geom& g = p->g;
  a->nparts = (int)((g.nmain + 255) / 256);
  bool ok = alloc(&a->fields, fb * (9 + a->nw)) && alloc(&a->work, fb * (a->L.nin() + 9)) &&
            alloc(&a->tape, fb * 12 * a->ckpt) && alloc(&a->ckpts, fb * 3 * a->nseg) &&
            alloc(&a->staging, p->ref_bytes()) && alloc((void**)&a->coef, 8 * sizeof(double)) &&
            alloc((void**)&a->part, a->nparts * sizeof(double)) && alloc((void**)&a->acc, sizeof(double));
  if (!ok) {
    scl_adjoint_destroy(a);
    return SCL_ENOMEM;
  }
  // The physical tape (plans with the fused z-passes) spares the backward
  // the inverse transforms of the state's fields (u, w [, S_ij]: 6 or 12)
  // in every vector-Jacobian product; it is kept when it takes at most half
  // the free device memory (8.2 GB for the LES at 128^3 FP32 with 20-step
  // segments).  SCL_ADJ_PHYS_TAPE=0 turns it off; =1 requires it (creation
  // fails without it).
  a->pf = a->L.les ? 12 : 6;
  const char* pt = std::getenv("SCL_ADJ_PHYS_TAPE");
  const bool tape_off = pt && pt[0] == '0', tape_req = pt && pt[0] == '1';
  if (p->zfused && !tape_off) {
    const size_t bytes = fb * 4 * a->pf * (size_t)a->ckpt;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess && bytes <= fr / 2 && hipMalloc(&a->ptape, bytes) != hipSuccess) {
      (void)hipGetLastError();
      a->ptape = nullptr;
    }
  }
  if (tape_req && !a->ptape) {
    scl_adjoint_destroy(a);
    return set_error("SCL_ADJ_PHYS_TAPE=1: no physical tape (it needs z- and y-lines below 1024 points and memory)"),
           SCL_EINVAL;
  }
  if (hipMemset(a->fields, 0, fb * (9 + a->nw)) != hipSuccess || hipMemset(a->coef, 0, 8 * sizeof(double)) != hipSuccess ||
      hipEventCreate(&a->e0) != hipSuccess || hipEventCreate(&a->e1) != hipSuccess) {
    scl_adjoint_destroy(a);
    return set_error("adjoint set-up failed"), SCL_EHIP;
  }
  *out = a;
  return SCL_OK;
}

extern "C" int32_t scl_adjoint_forward(scl_adjoint* a, double cs2, const void* u0x, const void* u0y,
                                       const void* u0z, void* uTx, void* uTy, void* uTz, int32_t on_device) {
  if (!a) return set_error("null adjoint"), SCL_EINVAL;
  const void* u0[3] = {u0x, u0y, u0z};
  void* uT[3] = {uTx, uTy, uTz};
  return by_prec_a(a, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    return forward<R>(a, cs2, u0, uT, on_device != 0);
  });
}

extern "C" int32_t scl_adjoint_backward(scl_adjoint* a, const void* gx, const void* gy, const void* gz, void* g0x,
                                        void* g0y, void* g0z, double* dcs2, int32_t on_device) {
  if (!a) return set_error("null adjoint"), SCL_EINVAL;
  const void* gT[3] = {gx, gy, gz};
  void* g0[3] = {g0x, g0y, g0z};
  return by_prec_a(a, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    return backward<R>(a, gT, g0, dcs2, on_device != 0);
  });
}

extern "C" int32_t scl_adjoint_vjp(scl_adjoint* a, double cs2, const void* sx, const void* sy, const void* sz,
                                   const void* nx, const void* ny, const void* nz, void* ox, void* oy, void* oz,
                                   double* dcs2, int32_t on_device) {
  if (!a) return set_error("null adjoint"), SCL_EINVAL;
  const void* s[3] = {sx, sy, sz};
  const void* nb[3] = {nx, ny, nz};
  void* o[3] = {ox, oy, oz};
  return by_prec_a(a, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    return vjp_host<R>(a, cs2, s, nb, o, dcs2, on_device != 0);
  });
}

extern "C" int32_t scl_adjoint_timing(scl_adjoint* a, double* forward_ms, double* backward_ms) {
  if (!a || !forward_ms || !backward_ms) return set_error("null argument"), SCL_EINVAL;
  *forward_ms = a->fwd_ms;
  *backward_ms = a->bwd_ms;
  return SCL_OK;
}
