// The reference's RHS-callback convention (rk5.f90:1-12: derivs(x, y, dydx,
// ny), the driver owns the stepping and derivs the physics) for the
// spectral model, and a device Lawson IF-RK4 driver that takes such a
// callback.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <mutex>

#include "ops.hpp"
#include "solver.hpp"

namespace scl {

// out[q] = N[q] - nu_q k^2 y[q] over nc fields (nu_q = nu for velocity,
// nu_phi for the scalar, field 3).
template <class R>
__global__ __launch_bounds__(256) void k_rhs_linear(const cpx<R>* __restrict__ N, const cpx<R>* __restrict__ y,
                                                    cpx<R>* __restrict__ out, int64_t fs, Geom g, int nc, R nu,
                                                    R nu_phi) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.ntot) return;
  int x, yy, z;
  split_xyz(g, i, x, yy, z);
  const int ky = kwave(yy, g.ny), kz = kwave(z, g.nz);
  const R k2 = (R)(x * x + ky * ky + kz * kz);
  for (int q = 0; q < nc; ++q) {
    const R r = (q == 3 ? nu_phi : nu) * k2;
    out[i + q * fs] = N[i + q * fs] - r * y[i + q * fs];
  }
}

// nu_q k^2 per entry (real), the linear rates of the solver's state.
template <class R>
__global__ __launch_bounds__(256) void k_rates(R* __restrict__ rate, int64_t fs, Geom g, int nc, R nu, R nu_phi) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.ntot) return;
  int x, yy, z;
  split_xyz(g, i, x, yy, z);
  const int ky = kwave(yy, g.ny), kz = kwave(z, g.nz);
  const R k2 = (R)(x * x + ky * ky + kz * kz);
  for (int q = 0; q < nc; ++q) rate[i + q * fs] = (q == 3 ? nu_phi : nu) * k2;
}

// One stage of Lawson IF-RK4 for y' = -rate y + f, division-free form
// (E = exp(-rate dt/2), a = dt f):
//   1: acc = E^2 (y + a/6), s = E (y + a/2)     2: acc += E a/3, s = E y + a/2
//   3: acc += E a/3, s = E (E y + a)            4: y = acc + a/6
// rate == nullptr: E = 1 (classical RK4).
template <class R, int MODE>
__global__ __launch_bounds__(256) void k_ifrk4_stage(int64_t n, cpx<R>* __restrict__ y, const R* __restrict__ rate,
                                                     const cpx<R>* __restrict__ f, cpx<R>* __restrict__ acc,
                                                     cpx<R>* __restrict__ s, R dt) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const R x = rate ? -rate[i] * dt / (R)2 : (R)0;
  const R em1 = expm1_r(x);
  const cpx<R> a = dt * f[i];
  const cpx<R> yv = y[i];
  if constexpr (MODE == 1) {
    const cpx<R> u = yv + ((R)1 / (R)6) * a, q = yv + (R)0.5 * a;
    acc[i] = u + expm1_r((R)2 * x) * u;
    s[i] = q + em1 * q;
  } else if constexpr (MODE == 2 || MODE == 3) {
    const cpx<R> q = ((R)1 / (R)3) * a;
    acc[i] = acc[i] + (q + em1 * q);
    const cpx<R> ew = yv + em1 * yv;
    if constexpr (MODE == 2) {
      s[i] = ew + (R)0.5 * a;
    } else {
      const cpx<R> u = ew + a;
      s[i] = u + em1 * u;
    }
  } else {
    y[i] = acc[i] + ((R)1 / (R)6) * a;
  }
}

}  // namespace scl

using namespace scl;

namespace {

template <class F>
int32_t by_prec_p(int prec, F&& f) {
  if (prec == SCL_SP) return f(float{});
  return f(double{});
}

// staging field of the plan's scratch (reference layout)
template <class R>
int32_t staging_of(scl_plan3d* p, cpx<R>** st) {
  void* s;
  SCL_CHECK(plan_scratch(p, 0, &s));
  *st = reinterpret_cast<cpx<R>*>(p->scratch) + p->scratch_fields * p->fstride();
  return SCL_OK;
}

// N(st) - nu k^2 st (linear = true) or N(st) into out (split layout, nc
// fields), st = S of the solver.
template <class R>
int32_t rhs_split(scl_dns* d, const cpx<R>* st, cpx<R>* out, bool linear) {
  d->chained = false;  // W is overwritten
  scl_plan3d* p = d->plan;
  DevFields f{d->fields, p->fstride(), d->nc, d->coef};
  SCL_CHECK(solver_stage<R>(p, f, 0, d->p, st));
  const R nu = linear ? (R)d->p.nu : (R)0;
  const R nu_phi = linear ? (R)(d->p.nu / (d->p.schmidt > 0 ? d->p.schmidt : 1.0)) : (R)0;
  hipLaunchKernelGGL((k_rhs_linear<R>), dim3((unsigned)((p->g.ntot + 255) / 256)), dim3(256), 0, p->stream,
                     f.W<R>(), st, out, p->fstride(), p->g, d->nc, nu, nu_phi);
  return check_launch("rhs_linear");
}

template <class R>
int32_t dns_rhs(scl_dns* d, const void* const* y, void* const* dydt, bool dev) {
  scl_plan3d* p = d->plan;
  const int64_t fs = p->fstride();
  DevFields f{d->fields, fs, d->nc, d->coef};
  cpx<R>* S = f.S<R>();
  cpx<R>* st = nullptr;
  SCL_CHECK(staging_of<R>(p, &st));
  for (int q = 0; q < d->nc; ++q) {
    if (!y[q] || !dydt[q]) return set_error("null array"), SCL_EINVAL;
    const cpx<R>* src = reinterpret_cast<const cpx<R>*>(y[q]);
    if (!dev) {
      SCL_HIP(hipMemcpyAsync(st, y[q], p->ref_bytes(), hipMemcpyHostToDevice, p->stream));
      src = st;
    }
    SCL_CHECK(Ops<R>::ref_to_split(*p, src, S + q * fs));
  }
  // the result goes to DU (free between steps), then out in the reference layout
  cpx<R>* out = f.DU<R>();
  SCL_CHECK(rhs_split<R>(d, S, out, true));
  for (int q = 0; q < d->nc; ++q) {
    if (dev) {
      SCL_CHECK(Ops<R>::split_to_ref(*p, out + q * fs, reinterpret_cast<cpx<R>*>(dydt[q])));
    } else {
      SCL_CHECK(Ops<R>::split_to_ref(*p, out + q * fs, st));
      SCL_HIP(hipMemcpyAsync(dydt[q], st, p->ref_bytes(), hipMemcpyDeviceToHost, p->stream));
    }
  }
  SCL_HIP(hipStreamSynchronize(p->stream));
  return SCL_OK;
}

std::mutex g_derivs_mu;
scl_dns* g_derivs = nullptr;

}  // namespace

extern "C" int32_t scl_dns_rhs(scl_dns* d, double t, const void* const* y, void* const* dydt, int32_t on_device) {
  (void)t;  // the spectral system is autonomous
  if (!d || !y || !dydt) return set_error("null argument"), SCL_EINVAL;
  return by_prec_p(d->plan->prec, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    return dns_rhs<R>(d, y, dydt, on_device != 0);
  });
}

extern "C" int32_t scl_dns_state(scl_dns* d, void** state, int64_t* n) {
  if (!d || !state || !n) return set_error("null argument"), SCL_EINVAL;
  *state = d->fields;
  *n = (int64_t)d->nc * d->plan->fstride();
  d->chained = false;  // the caller may change the state through the pointer
  return SCL_OK;
}

extern "C" int32_t scl_dns_rates(scl_dns* d, void* rate) {
  if (!d || !rate) return set_error("null argument"), SCL_EINVAL;
  scl_plan3d* p = d->plan;
  return by_prec_p(p->prec, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    const R nu = (R)d->p.nu;
    const R nu_phi = (R)(d->p.nu / (d->p.schmidt > 0 ? d->p.schmidt : 1.0));
    hipLaunchKernelGGL((k_rates<R>), dim3((unsigned)((p->g.ntot + 255) / 256)), dim3(256), 0, p->stream,
                       reinterpret_cast<R*>(rate), p->fstride(), p->g, d->nc, nu, nu_phi);
    SCL_CHECK(check_launch("rates"));
    SCL_HIP(hipStreamSynchronize(p->stream));
    return SCL_OK;
  });
}

extern "C" int32_t scl_dns_derivs_cb(void* ctx, double t, const void* y, void* dydt) {
  (void)t;
  auto* d = reinterpret_cast<scl_dns*>(ctx);
  if (!d || !y || !dydt) return set_error("null argument"), SCL_EINVAL;
  scl_plan3d* p = d->plan;
  return by_prec_p(p->prec, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    const int64_t fs = p->fstride();
    DevFields f{d->fields, fs, d->nc, d->coef};
    cpx<R>* S = f.S<R>();
    SCL_HIP(hipMemcpyAsync(S, y, d->nc * fs * sizeof(cpx<R>), hipMemcpyDeviceToDevice, p->stream));
    SCL_CHECK(rhs_split<R>(d, S, reinterpret_cast<cpx<R>*>(dydt), false));
    SCL_HIP(hipStreamSynchronize(p->stream));
    return SCL_OK;
  });
}

extern "C" int32_t scl_ifrk4_integrate(int32_t prec, int64_t n, void* y, const void* rate, double t0, double dt,
                                       int32_t nsteps, scl_derivs_fn f, void* ctx) {
  if (!y || !f || n < 1 || nsteps < 0) return set_error("bad integrate arguments"), SCL_EINVAL;
  if (prec != SCL_SP && prec != SCL_DP) return set_error("bad precision %d", prec), SCL_EINVAL;
  return by_prec_p(prec, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    using C = cpx<R>;
    C* work = nullptr;
    if (hipMalloc(&work, 3 * n * sizeof(C)) != hipSuccess) {
      (void)hipGetLastError();
      return set_error("device allocation of %zu bytes failed", (size_t)(3 * n * sizeof(C))), SCL_ENOMEM;
    }
    C* fv = work;
    C* acc = work + n;
    C* s = work + 2 * n;
    C* yv = reinterpret_cast<C*>(y);
    const R* rt = reinterpret_cast<const R*>(rate);
    const dim3 grid((unsigned)((n + 255) / 256));
    int32_t st = SCL_OK;
    double t = t0;
    for (int step = 0; step < nsteps && st == SCL_OK; ++step) {
      for (int stage = 1; stage <= 4 && st == SCL_OK; ++stage) {
        const double ts = stage == 1 ? t : (stage == 4 ? t + dt : t + 0.5 * dt);
        if (hipDeviceSynchronize() != hipSuccess) {
          st = SCL_EHIP;
          set_error("device error before the callback");
          break;
        }
        st = f(ctx, ts, stage == 1 ? (const void*)yv : (const void*)s, fv);
        if (st != SCL_OK) break;
        switch (stage) {
          case 1: hipLaunchKernelGGL((k_ifrk4_stage<R, 1>), grid, dim3(256), 0, 0, n, yv, rt, fv, acc, s, (R)dt); break;
          case 2: hipLaunchKernelGGL((k_ifrk4_stage<R, 2>), grid, dim3(256), 0, 0, n, yv, rt, fv, acc, s, (R)dt); break;
          case 3: hipLaunchKernelGGL((k_ifrk4_stage<R, 3>), grid, dim3(256), 0, 0, n, yv, rt, fv, acc, s, (R)dt); break;
          default: hipLaunchKernelGGL((k_ifrk4_stage<R, 4>), grid, dim3(256), 0, 0, n, yv, rt, fv, acc, s, (R)dt); break;
        }
        st = check_launch("ifrk4_stage");
      }
      t += dt;
    }
    (void)hipDeviceSynchronize();
    (void)hipFree(work);
    return st;
  });
}

extern "C" int32_t scl_derivs_bind(scl_dns* d) {
  std::lock_guard<std::mutex> lk(g_derivs_mu);
  if (d && d->plan->prec != SCL_DP) return set_error("derivs works on real(dp) data: bind an FP64 solver"), SCL_EINVAL;
  g_derivs = d;
  return SCL_OK;
}

extern "C" void scl_derivs_(const double* x, const double* y, double* dydx, const int32_t* ny) {
  scl_dns* d;
  {
    std::lock_guard<std::mutex> lk(g_derivs_mu);
    d = g_derivs;
  }
  if (!d) {
    fprintf(stderr, "scl_derivs: no solver bound (scl_derivs_bind)\n");
    abort();
  }
  const scl_plan3d* p = d->plan;
  const int64_t per = 2 * (int64_t)(p->g.mx + 1) * p->g.nrow;  // reals per field
  if (!ny || (int64_t)*ny != per * d->nc) {
    fprintf(stderr, "scl_derivs: ny = %d, the bound solver's state has %lld reals\n", ny ? *ny : -1,
            (long long)(per * d->nc));
    abort();
  }
  const void* yi[4];
  void* yo[4];
  for (int q = 0; q < d->nc; ++q) {
    yi[q] = y + q * per;
    yo[q] = dydx + q * per;
  }
  if (scl_dns_rhs(d, x ? *x : 0.0, yi, yo, 0) != SCL_OK) {
    fprintf(stderr, "scl_derivs: %s\n", scl::last_error());
    abort();
  }
}
