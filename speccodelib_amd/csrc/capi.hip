// C ABI: plans, host-array entry points under the reference's routine names,
// and the device-resident IF-RK4 solver.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <cstdlib>
#include <limits>
#include <mutex>
#include <vector>

#include "diag.hpp"
#include "ic.hpp"
#include "ops.hpp"
#include "ops_ext.hpp"
#include "solver.hpp"

namespace scl {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

const char* last_error() { return g_err; }

int32_t check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("launch of %s failed: %s", what, hipGetErrorString(e));
    return SCL_EHIP;
  }
  return SCL_OK;
}

static bool pow2_in(int n, int lo, int hi) { return n >= lo && n <= hi && (n & (n - 1)) == 0; }

template <class F>
static int32_t by_prec(int prec, F&& f) {
  if (prec == SCL_SP) return f(float{});
  return f(double{});
}

// exp(-2 pi i j / L) for j < L, evaluated in FP64 with octant symmetry so
// that the quarter- and eighth-turn values are exact images of each other.
void twiddles(int L, std::vector<double>& re, std::vector<double>& im) {
  re.resize(L);
  im.resize(L);
  const double twopi = 6.283185307179586476925286766559;
  const int Q = L / 4;
  for (int j = 0; j < L; ++j) {
    const int q = j / Q, r = j % Q;
    double c, s;
    if (2 * r <= Q) {
      c = std::cos(twopi * r / L);
      s = std::sin(twopi * r / L);
    } else {
      c = std::sin(twopi * (Q - r) / L);
      s = std::cos(twopi * (Q - r) / L);
    }
    if (r == 0) {
      c = 1.0;
      s = 0.0;
    }
    double cr, sr;  // rotate (c, s) by q quarter turns
    switch (q) {
      case 0: cr = c; sr = s; break;
      case 1: cr = -s; sr = c; break;
      case 2: cr = -c; sr = -s; break;
      default: cr = s; sr = -c; break;
    }
    re[j] = cr;
    im[j] = -sr;  // exp(-i theta)
  }
}

int32_t plan_scratch(scl_plan3d* p, size_t nfields, void** out) {
  if (!p->scratch || p->scratch_fields < nfields) {
    if (p->scratch) {
      (void)hipStreamSynchronize(p->stream);
      (void)hipFree(p->scratch);
      p->scratch = nullptr;
      p->scratch_fields = 0;
    }
    const size_t bytes = nfields * (size_t)p->fstride() * p->esize() + p->ref_bytes();
    if (hipMalloc(&p->scratch, bytes) != hipSuccess) {
      (void)hipGetLastError();
      set_error("device allocation of %zu bytes failed", bytes);
      return SCL_ENOMEM;
    }
    p->scratch_fields = nfields;
  }
  *out = p->scratch;
  return SCL_OK;
}

int32_t plan_diag(scl_plan3d* p, size_t bytes, void** out) {
  if (p->diag_bytes < bytes) {
    if (p->diag) {
      (void)hipStreamSynchronize(p->stream);
      (void)hipFree(p->diag);
      p->diag = nullptr;
      p->diag_bytes = 0;
    }
    if (hipMalloc(&p->diag, bytes) != hipSuccess) {
      (void)hipGetLastError();
      set_error("device allocation of %zu bytes failed", bytes);
      return SCL_ENOMEM;
    }
    p->diag_bytes = bytes;
  }
  *out = p->diag;
  return SCL_OK;
}

// staging buffer of the plan's scratch (after nfields work fields)
template <class R>
static cpx<R>* staging(scl_plan3d* p) {
  return reinterpret_cast<cpx<R>*>(p->scratch) + p->scratch_fields * p->fstride();
}

template <class R>
static int32_t upload(scl_plan3d* p, const void* host, cpx<R>* dst) {
  if (!host) {
    set_error("null host array");
    return SCL_EINVAL;
  }
  cpx<R>* st = staging<R>(p);
  SCL_HIP(hipMemcpyAsync(st, host, p->ref_bytes(), hipMemcpyHostToDevice, p->stream));
  return Ops<R>::ref_to_split(*p, st, dst);
}

template <class R>
static int32_t download(scl_plan3d* p, const cpx<R>* src, void* host) {
  if (!host) {
    set_error("null host array");
    return SCL_EINVAL;
  }
  cpx<R>* st = staging<R>(p);
  SCL_CHECK(Ops<R>::split_to_ref(*p, src, st));
  SCL_HIP(hipMemcpyAsync(host, st, p->ref_bytes(), hipMemcpyDeviceToHost, p->stream));
  SCL_HIP(hipStreamSynchronize(p->stream));
  return SCL_OK;
}

static std::mutex g_plan_mu;
static scl_plan3d* g_plan = nullptr;

static scl_plan3d* resolve(scl_plan3d* p) {
  if (p) return p;
  std::lock_guard<std::mutex> lk(g_plan_mu);
  return g_plan;
}

#define SCL_PLAN(p)                                                 \
  do {                                                              \
    p = resolve(p);                                                 \
    if (!p) {                                                       \
      set_error("no plan: call scl_fftwplan3d (fftwplan3d) first"); \
      return SCL_ENOPLAN;                                           \
    }                                                               \
  } while (0)

static int32_t check_dims(const scl_plan3d* p, int lx1, int ly, int lz) {
  if (lx1 != p->g.mx + 1 || ly != p->g.ny || lz != p->g.nz) {
    set_error("array shape (%d,%d,%d) does not match the plan (%d,%d,%d)", lx1, ly, lz, p->g.mx + 1, p->g.ny,
              p->g.nz);
    return SCL_EINVAL;
  }
  return SCL_OK;
}

// The wavenumber arrays passed to a reference routine must be wavenumber()'s.
template <class R>
static int32_t check_k1d(const scl_plan3d* p, const void* kx, const void* ky, const void* kz) {
  const R* a = (const R*)kx;
  const R* b = (const R*)ky;
  const R* c = (const R*)kz;
  if (!a || !b || !c) {
    set_error("null wavenumber array");
    return SCL_EINVAL;
  }
  for (int i = 0; i <= p->g.mx; ++i)
    if (a[i] != (R)i) return set_error("kx(%d) is not wavenumber()'s", i + 1), SCL_EINVAL;
  for (int i = 0; i < p->g.ny; ++i)
    if (b[i] != (R)kwave(i, p->g.ny)) return set_error("ky(%d) is not wavenumber()'s", i + 1), SCL_EINVAL;
  for (int i = 0; i < p->g.nz; ++i)
    if (c[i] != (R)kwave(i, p->g.nz)) return set_error("kz(%d) is not wavenumber()'s", i + 1), SCL_EINVAL;
  return SCL_OK;
}

template <class R>
static int32_t check_k2(const scl_plan3d* p, const void* k2v) {
  const R* k2 = (const R*)k2v;
  if (!k2) return set_error("null k2"), SCL_EINVAL;
  const Geom& g = p->g;
  // spot checks at the corners and a few interior points
  const int pts[6][3] = {{0, 0, 0}, {1, 0, 0}, {g.mx, g.ny - 1, g.nz - 1}, {3 % (g.mx + 1), 2, 1},
                         {g.mx / 2, g.ny / 2, g.nz / 3}, {1, g.ny - 1, 1}};
  for (auto& q : pts) {
    const int x = q[0], y = q[1], z = q[2];
    const int ik2 = x * x + kwave(y, g.ny) * kwave(y, g.ny) + kwave(z, g.nz) * kwave(z, g.nz);
    const R want = ik2 == 0 ? std::numeric_limits<R>::min() : (R)ik2;
    if (k2[((int64_t)z * g.ny + y) * (g.mx + 1) + x] != want)
      return set_error("k2(%d,%d,%d) is not wavenumber()'s", x + 1, y + 1, z + 1), SCL_EINVAL;
  }
  return SCL_OK;
}

}  // namespace scl

using namespace scl;

// ---------------------------------------------------------------------------
extern "C" const char* scl_last_error(void) { return scl::last_error(); }
extern "C" int32_t scl_version(void) { return 100; }

extern "C" int32_t scl_plan3d_create(int32_t nx, int32_t ny, int32_t nz, int32_t prec, scl_plan3d** out) {
  if (!out) return set_error("null plan pointer"), SCL_EINVAL;
  *out = nullptr;
  if (prec != SCL_SP && prec != SCL_DP) return set_error("bad precision %d", prec), SCL_EINVAL;
  if (!pow2_in(nx, 16, 1024) || !pow2_in(ny, 16, 1024) || !pow2_in(nz, 16, 1024))
    return set_error("grid (%d,%d,%d): each size must be a power of two in [16,1024]", nx, ny, nz), SCL_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    (void)hipGetLastError();
    return set_error("no HIP device"), SCL_ENODEV;
  }
  // Row order (Geom::zs): z-blocks of 8 rows for 1024-point z-lines, where
  // the z-passes otherwise spend their time on address translation; plain
  // [z][y] below.  SCL_ZBLOCK (1 or 8) overrides it, for tests of both
  // orders at small sizes.
  int zs = nz >= 1024 ? 3 : 0;
  if (const char* zb = std::getenv("SCL_ZBLOCK")) {
    const int b = std::atoi(zb);
    if (b != 1 && b != 8) return set_error("SCL_ZBLOCK=%s: must be 1 or 8", zb), SCL_EINVAL;
    zs = b == 8 ? 3 : 0;
  }
  auto* p = new scl_plan3d();
  p->prec = prec;
  p->g = make_geom(nx, ny, nz, zs);
  // The LES / scalar and adjoint passes with their pointwise work fused into
  // the z- and y-passes, up to 512-point z- and y-lines (at 1024 points those
  // kernels spill).  SCL_EXT_UNFUSED=1: the separate passes at any size, for
  // tests of that path against the oracle
  const char* uf = std::getenv("SCL_EXT_UNFUSED");
  p->zfused = nz < 1024 && ny < 1024 && !(uf && uf[0] == '1');
  (void)hipGetDevice(&p->device);
  {
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, p->device) == hipSuccess && ncu > 0)
      p->ncu = ncu;
  }
  if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
    delete p;
    return set_error("stream creation failed"), SCL_EHIP;
  }
  p->own_stream = true;
  p->L = std::max(nx, std::max(ny, nz));
  std::vector<double> re, im;
  twiddles(p->L, re, im);
  int32_t st = by_prec(prec, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    std::vector<cpx<R>> h(p->L);
    for (int j = 0; j < p->L; ++j) h[j] = mk<R>((R)re[j], (R)im[j]);
    SCL_HIP(hipMalloc(&p->tw, p->L * sizeof(cpx<R>)));
    SCL_HIP(hipMemcpy(p->tw, h.data(), p->L * sizeof(cpx<R>), hipMemcpyHostToDevice));
    SCL_HIP(hipMalloc((void**)&p->coef, 8 * sizeof(double)));
    SCL_HIP(hipMemset(p->coef, 0, 8 * sizeof(double)));
    return SCL_OK;
  });
  if (st != SCL_OK) {
    scl_plan3d_destroy(p);
    return st;
  }
  *out = p;
  return SCL_OK;
}

extern "C" int32_t scl_plan3d_destroy(scl_plan3d* p) {
  if (!p) return SCL_OK;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  if (p->tw) (void)hipFree(p->tw);
  if (p->coef) (void)hipFree(p->coef);
  if (p->scratch) (void)hipFree(p->scratch);
  if (p->diag) (void)hipFree(p->diag);
  if (p->padded) scl_plan3d_destroy(p->padded);
  if (p->stream && p->own_stream) (void)hipStreamDestroy(p->stream);
  delete p;
  return SCL_OK;
}

extern "C" int32_t scl_plan3d_set_stream(scl_plan3d* p, void* stream) {
  if (!p) return set_error("null plan"), SCL_EINVAL;
  (void)hipStreamSynchronize(p->stream);
  if (p->own_stream) (void)hipStreamDestroy(p->stream);
  if (stream) {
    p->stream = (hipStream_t)stream;
    p->own_stream = false;
  } else {
    SCL_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    p->own_stream = true;
  }
  return SCL_OK;
}

extern "C" int32_t scl_fftwplan3d_prec(int32_t ix, int32_t iy, int32_t iz, int32_t prec) {
  scl_plan3d* p = nullptr;
  SCL_CHECK(scl_plan3d_create(ix, iy, iz, prec, &p));
  std::lock_guard<std::mutex> lk(g_plan_mu);
  if (g_plan) scl_plan3d_destroy(g_plan);
  g_plan = p;
  return SCL_OK;
}

extern "C" int32_t scl_fftwplan3d(int32_t ix, int32_t iy, int32_t iz) { return scl_fftwplan3d_prec(ix, iy, iz, SCL_SP); }

extern "C" int32_t scl_destroyplan3d(void) {
  std::lock_guard<std::mutex> lk(g_plan_mu);
  scl_plan3d_destroy(g_plan);
  g_plan = nullptr;
  return SCL_OK;
}

// ---------------------------------------------------------------------------
// FFTW-equivalent transforms on host arrays (real view (nx+2, ny, nz))
extern "C" int32_t scl_c2r3d(scl_plan3d* p, void* a) {
  SCL_PLAN(p);
  return by_prec(p->prec, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    void* s;
    SCL_CHECK(plan_scratch(p, 1, &s));
    cpx<R>* f = (cpx<R>*)s;
    SCL_CHECK(upload<R>(p, a, f));
    SCL_CHECK(Ops<R>::c2r3d(*p, f, p->fstride(), 1));
    return download<R>(p, f, a);
  });
}

extern "C" int32_t scl_r2c3d(scl_plan3d* p, void* a) {
  SCL_PLAN(p);
  return by_prec(p->prec, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    void* s;
    SCL_CHECK(plan_scratch(p, 1, &s));
    cpx<R>* f = (cpx<R>*)s;
    SCL_CHECK(upload<R>(p, a, f));
    SCL_CHECK(Ops<R>::r2c3d(*p, f, p->fstride(), 1));
    return download<R>(p, f, a);
  });
}

extern "C" int32_t scl_wavenumber3d(int32_t prec, void* kx, void* ky, void* kz, void* k2, int32_t lx1, int32_t ly,
                                   int32_t lz) {
  if (!kx || !ky || !kz || lx1 < 2 || ly < 2 || lz < 2) return set_error("bad wavenumber arguments"), SCL_EINVAL;
  return by_prec(prec, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    R* a = (R*)kx;
    R* b = (R*)ky;
    R* c = (R*)kz;
    for (int i = 0; i < lx1; ++i) a[i] = (R)i;
    for (int i = 0; i < ly; ++i) b[i] = (R)kwave(i, ly);
    for (int i = 0; i < lz; ++i) c[i] = (R)kwave(i, lz);
    if (k2) {
      R* d = (R*)k2;
      for (int z = 0; z < lz; ++z)
        for (int y = 0; y < ly; ++y)
          for (int x = 0; x < lx1; ++x) d[((int64_t)z * ly + y) * lx1 + x] = a[x] * a[x] + b[y] * b[y] + c[z] * c[z];
      d[0] = std::numeric_limits<R>::min();
    }
    return SCL_OK;
  });
}

extern "C" int32_t scl_convec_dns(scl_plan3d* p, const void* vx, const void* vy, const void* vz, void* wx, void* wy,
                                  void* wz, int32_t lx1, int32_t ly, int32_t lz, int32_t nx, int32_t ny, int32_t nz) {
  SCL_PLAN(p);
  SCL_CHECK(check_dims(p, lx1, ly, lz));
  if (nx != p->g.nx || ny != p->g.ny || nz != p->g.nz) return set_error("nx,ny,nz do not match the plan"), SCL_EINVAL;
  return by_prec(p->prec, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    void* s;
    SCL_CHECK(plan_scratch(p, 6, &s));
    cpx<R>* W = (cpx<R>*)s;
    const int64_t fs = p->fstride();
    const void* in[6] = {vx, vy, vz, wx, wy, wz};
    for (int f = 0; f < 6; ++f) SCL_CHECK(upload<R>(p, in[f], W + f * fs));
    SCL_CHECK(Ops<R>::colfft(*p, W, fs, 6, 2, +1));
    SCL_CHECK(Ops<R>::colfft(*p, W, fs, 6, 1, +1));
    SCL_CHECK(Ops<R>::row_convec(*p, W, fs));
    SCL_CHECK(Ops<R>::colfft(*p, W, fs, 3, 1, -1));
    SCL_CHECK(Ops<R>::colfft(*p, W, fs, 3, 2, -1));
    void* out[3] = {wx, wy, wz};
    for (int f = 0; f < 3; ++f) SCL_CHECK(download<R>(p, W + f * fs, out[f]));
    return SCL_OK;
  });
}

extern "C" int32_t scl_symmetrize3d(scl_plan3d* p, void* c, int32_t lx1, int32_t ly, int32_t lz) {
  SCL_PLAN(p);
  SCL_CHECK(check_dims(p, lx1, ly, lz));
  return by_prec(p->prec, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    void* s;
    SCL_CHECK(plan_scratch(p, 1, &s));
    cpx<R>* f = (cpx<R>*)s;
    SCL_CHECK(upload<R>(p, c, f));
    SCL_CHECK(Ops<R>::sym_zero(*p, f, (R)0));
    FieldPtrs<R> fp{};
    fp.p[0] = f;
    fp.n = 1;
    SCL_CHECK(Ops<R>::herm_fix(*p, fp));
    return download<R>(p, f, c);
  });
}

extern "C" int32_t scl_dealiasing(scl_plan3d* p, void* c, const void* k2, int32_t lx1, int32_t ly, int32_t lz) {
  SCL_PLAN(p);
  SCL_CHECK(check_dims(p, lx1, ly, lz));
  return by_prec(p->prec, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    SCL_CHECK(check_k2<R>(p, k2));
    void* s;
    SCL_CHECK(plan_scratch(p, 1, &s));
    cpx<R>* f = (cpx<R>*)s;
    SCL_CHECK(upload<R>(p, c, f));
    SCL_CHECK(Ops<R>::sym_zero(*p, f, (R)0));
    FieldPtrs<R> fp{};
    fp.p[0] = f;
    fp.n = 1;
    SCL_CHECK(Ops<R>::herm_fix(*p, fp));
    // dealiasing.f90:35: k2 >= real(ly*ly)/9
    const R thr = (R)(ly * ly) / (R)9;
    SCL_CHECK(Ops<R>::dealias_mask(*p, f, thr));
    return download<R>(p, f, c);
  });
}

extern "C" int32_t scl_projection(scl_plan3d* p, void* vx, void* vy, void* vz, const void* kx, const void* ky,
                                  const void* kz, int32_t lx1, int32_t ly, int32_t lz) {
  SCL_PLAN(p);
  SCL_CHECK(check_dims(p, lx1, ly, lz));
  return by_prec(p->prec, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    SCL_CHECK(check_k1d<R>(p, kx, ky, kz));
    void* s;
    SCL_CHECK(plan_scratch(p, 3, &s));
    cpx<R>* V = (cpx<R>*)s;
    const int64_t fs = p->fstride();
    void* a[3] = {vx, vy, vz};
    for (int f = 0; f < 3; ++f) SCL_CHECK(upload<R>(p, a[f], V + f * fs));
    SCL_CHECK(Ops<R>::projection(*p, V, fs));
    for (int f = 0; f < 3; ++f) SCL_CHECK(download<R>(p, V + f * fs, a[f]));
    return SCL_OK;
  });
}

extern "C" int32_t scl_force(scl_plan3d* p, void* wx, void* wy, void* wz, const void* vx, const void* vy,
                             const void* vz, const void* k2, int32_t lx1, int32_t ly, int32_t lz, double fkmax,
                             double eps) {
  SCL_PLAN(p);
  SCL_CHECK(check_dims(p, lx1, ly, lz));
  return by_prec(p->prec, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    SCL_CHECK(check_k2<R>(p, k2));
    void* s;
    SCL_CHECK(plan_scratch(p, 6, &s));
    cpx<R>* V = (cpx<R>*)s;
    const int64_t fs = p->fstride();
    cpx<R>* W = V + 3 * fs;
    const void* vi[3] = {vx, vy, vz};
    void* wi[3] = {wx, wy, wz};
    for (int f = 0; f < 3; ++f) SCL_CHECK(upload<R>(p, vi[f], V + f * fs));
    for (int f = 0; f < 3; ++f) SCL_CHECK(upload<R>(p, wi[f], W + f * fs));
    SCL_CHECK(Ops<R>::force_coef(*p, V, fs, fkmax, eps, 0.0, 1, p->coef));
    const R fk = (R)fkmax;
    SCL_CHECK(Ops<R>::force_apply(*p, W, V, fs, p->coef, fk * fk, 1));
    for (int f = 0; f < 3; ++f) SCL_CHECK(download<R>(p, W + f * fs, wi[f]));
    return SCL_OK;
  });
}

extern "C" int32_t scl_force_heli(scl_plan3d* p, void* fx, void* fy, void* fz, const void* vx, const void* vy,
                                  const void* vz, const void* wx, const void* wy, const void* wz, const void* k2,
                                  int32_t lx1, int32_t ly, int32_t lz, double fkmax, double eps, double eta) {
  SCL_PLAN(p);
  SCL_CHECK(check_dims(p, lx1, ly, lz));
  return by_prec(p->prec, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    SCL_CHECK(check_k2<R>(p, k2));
    void* s;
    SCL_CHECK(plan_scratch(p, 6, &s));
    cpx<R>* V = (cpx<R>*)s;
    const int64_t fs = p->fstride();
    cpx<R>* W = V + 3 * fs;
    const void* vi[3] = {vx, vy, vz};
    const void* wi[3] = {wx, wy, wz};
    void* fo[3] = {fx, fy, fz};
    for (int f = 0; f < 3; ++f) SCL_CHECK(upload<R>(p, vi[f], V + f * fs));
    for (int f = 0; f < 3; ++f) SCL_CHECK(upload<R>(p, wi[f], W + f * fs));
    SCL_CHECK(Ops<R>::force_coef(*p, V, fs, fkmax, eps, eta, 2, p->coef));
    const R fk = (R)fkmax;
    SCL_CHECK(Ops<R>::force_apply(*p, W, V, fs, p->coef, fk * fk, 2));
    for (int f = 0; f < 3; ++f) SCL_CHECK(download<R>(p, W + f * fs, fo[f]));
    return SCL_OK;
  });
}

// ---------------------------------------------------------------------------
// One RHS / one IF-RK4 step on device fields
template <class R>
static RhsParams<R> rhs_params(const scl_plan3d* p, double nu, double dt, int forced, double fkmax, int dealias) {
  RhsParams<R> rp;
  rp.dt = (R)dt;
  rp.nu = (R)nu;
  const R fk = (R)fkmax;
  rp.fk2 = fk * fk;
  // dns-deform.f90:75-76: kcut = 2*lx/3 (default real), kcut2 = kcut*kcut
  const float kcut = 2.0f * (float)p->g.mx / 3.0f;
  rp.kc2 = (R)kcut * (R)kcut;
  rp.forced = forced;
  rp.dealias = dealias;
  return rp;
}

// Event hooks for per-kernel timing (nullptr: none).
struct Timer {
  std::vector<hipEvent_t>* ev = nullptr;
  std::vector<int>* kinds = nullptr;
  hipStream_t s = nullptr;
  void mark(int kind) {
    if (!ev) return;
    hipEvent_t e;
    (void)hipEventCreate(&e);
    (void)hipEventRecord(e, s);
    ev->push_back(e);
    kinds->push_back(kind);
  }
};

// RHS of stage `mode` (0: N(st) into W[0..2]) with the fields of a solver
// (the fused five-pass DNS RHS; nc = 3).
template <class R>
static int32_t rk_stage(scl_plan3d* p, const DevFields& d, int mode, const RhsParams<R>& rp, double eps, double eta,
                        double fkmax, const cpx<R>* st, Timer* tm, bool primed = false, bool fuse = false) {
  const int64_t fs = d.fs;
  cpx<R>* V = d.V<R>();
  cpx<R>* DU = d.DU<R>();
  cpx<R>* S = d.S<R>();
  cpx<R>* W = d.W<R>();
  Timer none;
  if (!tm) tm = &none;
  if (rp.forced) {
    tm->mark(SCL_K_FORCE);
    SCL_CHECK(Ops<R>::force_coef(*p, st, fs, fkmax, eps, eta, rp.forced, d.coef));
  }
  if (!primed) {
    tm->mark(SCL_K_ZINV);
    SCL_CHECK(Ops<R>::zinv_vort(*p, st, fs, W, fs));
  }
  tm->mark(SCL_K_YINV);
  SCL_CHECK(Ops<R>::yinv_vort(*p, W, fs));
  tm->mark(SCL_K_XPASS);
  SCL_CHECK(Ops<R>::row_convec(*p, W, fs, reinterpret_cast<cpx<R>*>(d.phys_out)));
  tm->mark(SCL_K_YFWD);
  SCL_CHECK(Ops<R>::colfft(*p, W, fs, 3, 1, -1));
  tm->mark(fuse ? (mode >= 4 ? SCL_K_ZFWD4_ZINV : SCL_K_ZFWD_ZINV) : SCL_K_ZFWD);
  SCL_CHECK(Ops<R>::zfwd_rhs(*p, mode, W, fs, V, DU, S, st, d.coef, rp, fuse));
  // The kx = 0 Hermitian copy (symmetrize.f90:51-61) goes on N (mode 0),
  // on the collected stage states (s_out: the adjoint's tape), and on v at
  // the end of an unchained step.  Skipping it on du and the stage states of
  // stages 1..3, and on v when the last stage z-inverses it for the next
  // step (fuse), changes them only at rounding level: the kx = 0 plane of N
  // is Hermitian up to rounding (it is the transform of real products), and
  // c2r ignores the anti-Hermitian part of the kx = 0 lines.  Leaving v as
  // the chain transformed it also makes a dump of v the exact state the
  // chain continues from, so a restart reproduces the run bit for bit.
  if (((mode >= 1 && mode < 4) || (mode >= 4 && fuse)) && !d.s_out) {
    tm->mark(-1);
    return SCL_OK;
  }
  tm->mark(SCL_K_HERM);
  FieldPtrs<R> fp{};
  if (mode == 0) {
    for (int q = 0; q < 3; ++q) fp.p[q] = W + q * fs;
    fp.n = 3;
  } else if (mode < 4) {
    for (int q = 0; q < 3; ++q) {
      fp.p[q] = DU + q * fs;
      fp.p[3 + q] = S + q * fs;
    }
    fp.n = 6;
  } else {
    for (int q = 0; q < 3; ++q) fp.p[q] = V + q * fs;
    fp.n = 3;
  }
  SCL_CHECK(Ops<R>::herm_fix(*p, fp));
  tm->mark(-1);
  return SCL_OK;
}

// Stage chaining: the z-forward of stages 1..3 also z-inverses the next
// stage state into W[0..4] (k_zfwd_rhs FUSED), so stages 2..4 skip their
// z-inverse pass and the stage states are stored only on the forced modes.
// Not when the stage states are collected (s_out: the adjoint's tape).
template <class R>
static bool chain_stages(const scl_plan3d*, const DevFields& d) {
  return !d.s_out;
}

// One IF-RK4 step.  primed: W[0..4] already hold the z-inverse of v (the
// previous step's fused stage 4), so stage 1 runs no z-inverse pass;
// fuse_last: stage 4 z-inverses the new v into W[0..4] for the next step.
// The device solver sets both from its first step on, so every step after
// the first reads the fused z-inverse whatever the grouping of steps into
// calls or graphs: results do not depend on it.
// The truncation radius after a step: kcut > 0 from the parameters, else
// dns-deform.f90:75-76's 2 lx/3 (rhs_params).
template <class R>
static void apply_kcut(RhsParams<R>& rp, const scl_dns_params& prm) {
  if (prm.kcut > 0) rp.kc2 = (R)prm.kcut * (R)prm.kcut;
}

template <class R>
static int32_t rk4_step_dev(scl_plan3d* p, const DevFields& d, const scl_dns_params& prm, Timer* tm,
                            bool primed = false, bool fuse_last = false) {
  RhsParams<R> rp = rhs_params<R>(p, prm.nu, prm.dt, prm.forced, prm.fkmax, prm.dealias_rhs);
  apply_kcut<R>(rp, prm);
  rp.nonfinite = d.nonfinite;
  const bool chain = chain_stages<R>(p, d);
  for (int stage = 1; stage <= 4; ++stage) {
    const cpx<R>* st = stage == 1 ? d.V<R>() : d.S<R>();
    SCL_CHECK(rk_stage<R>(p, d, stage, rp, prm.eps, prm.eta, prm.fkmax, st, tm,
                          chain && (stage > 1 || primed), chain && (stage < 4 || fuse_last)));
  }
  return SCL_OK;
}

// One IF-AB2 step (spectral_dns_dealiased.f90:124-296 / specdns-forced-
// omp-ft.f90:224-274) of the DNS solver: the umax energy (device-guarded),
// dt and the weights (k_ab_begin), the RHS passes of one stage with the
// z-forward epilogue of mode 5 (Euler), 6 (leapfrog) or 7 (Adams-Bashforth),
// fused with the z-inverse of the new v for the next step, then the kx = 0
// Hermitian copy on v and the time update (k_ab_end).  istep: steps since
// the state was set.
// The step's dt and weights into ab (the umax energy first on refresh steps).
template <class R>
static int32_t ab_begin(scl_plan3d* p, const DevFields& d, const scl_dns_params& prm, double* ab, double* ab_part) {
  const Geom& g = p->g;
  AbConsts c;
  c.stepper = prm.stepper;
  c.ieout = prm.ieout;
  c.betadx = prm.beta * M_PI / g.mx;  // beta (pi/lx), spectral_dns_dealiased.f90:74-75
  c.dtw = prm.dtw;
  c.myeps = 2.0 * (double)std::numeric_limits<float>::epsilon();
  c.dt_fixed = prm.dt;
  if (prm.stepper == 1) {
    hipLaunchKernelGGL((k_ab_energy<R>), dim3(DIAG_BLOCKS), dim3(256), 0, p->stream, d.V<R>(), d.fs, g,
                       (const double*)ab, prm.ieout, ab_part);
    SCL_CHECK(check_launch("ab_energy"));
  }
  hipLaunchKernelGGL(k_ab_begin, dim3(1), dim3(64), 0, p->stream, ab, (const double*)ab_part, DIAG_BLOCKS, c);
  return check_launch("ab_begin");
}

static int32_t ab_end(scl_plan3d* p, double* ab) {
  hipLaunchKernelGGL(k_ab_end, dim3(1), dim3(64), 0, p->stream, ab);
  return check_launch("ab_end");
}

template <class R>
static int32_t ab2_step_dev(scl_plan3d* p, const DevFields& d, const scl_dns_params& prm, double* ab,
                            double* ab_part, int64_t istep, bool primed, Timer* tm) {
  RhsParams<R> rp = rhs_params<R>(p, prm.nu, 0.0, prm.forced, prm.fkmax, 0);
  apply_kcut<R>(rp, prm);
  rp.nonfinite = d.nonfinite;
  rp.ab = ab;
  Timer none;
  if (!tm) tm = &none;
  tm->mark(-1);
  SCL_CHECK(ab_begin<R>(p, d, prm, ab, ab_part));
  const int mode = istep == 0 ? 5 : istep == 1 ? 6 : 7;
  SCL_CHECK(rk_stage<R>(p, d, mode, rp, prm.eps, prm.eta, prm.fkmax, d.V<R>(), tm, primed, true));
  return ab_end(p, ab);
}

// The LES and/or scalar RHS (multi-pass) for stage `mode` (IF-RK4 1..4,
// IF-AB2 5..7 with the weights in ab).
template <class R>
static int32_t rk_stage_ext(scl_plan3d* p, const DevFields& d, int mode, const scl_dns_params& prm, const cpx<R>* st,
                            Timer* tm, const double* ab = nullptr) {
  const int64_t fs = d.fs;
  Timer none;
  if (!tm) tm = &none;
  ExtLayout L{prm.sgs ? 1 : 0, prm.scalar ? 1 : 0};
  ExtParams<R> ep;
  ep.r = rhs_params<R>(p, prm.nu, prm.dt, prm.forced, prm.fkmax, prm.dealias_rhs);
  apply_kcut<R>(ep.r, prm);
  ep.r.nonfinite = d.nonfinite;
  ep.r.ab = ab;
  ep.nu_phi = (R)(prm.nu / (prm.schmidt > 0 ? prm.schmidt : 1.0));
  ep.grad = (R)prm.grad;
  ep.sign_phi = 1;
  const double delta = prm.delta > 0 ? prm.delta : 2.0 * M_PI / p->g.nx;
  if (prm.forced) {
    tm->mark(SCL_K_FORCE);
    SCL_CHECK(Ops<R>::force_coef(*p, st, fs, prm.fkmax, prm.eps, prm.eta, prm.forced, d.coef));
  }
  tm->mark(SCL_K_ZFWD);
  SCL_CHECK(ExtOps<R>::rhs(*p, mode, L, d.W<R>(), fs, d.V<R>(), d.DU<R>(), d.S<R>(), st, d.coef, ep,
                           (R)(delta * delta), reinterpret_cast<cpx<R>*>(d.phys_out)));
  // the kx = 0 Hermitian copy: as in rk_stage
  if (mode >= 1 && mode < 4 && !d.s_out) {
    tm->mark(-1);
    return SCL_OK;
  }
  tm->mark(SCL_K_HERM);
  FieldPtrs<R> fp{};
  const int nc = d.nc;
  if (mode == 0) {
    for (int q = 0; q < nc; ++q) fp.p[q] = d.W<R>() + q * fs;
    fp.n = nc;
  } else if (mode < 4) {
    for (int q = 0; q < nc; ++q) {
      fp.p[q] = d.DU<R>() + q * fs;
      fp.p[nc + q] = d.S<R>() + q * fs;
    }
    fp.n = 2 * nc;
  } else {
    for (int q = 0; q < nc; ++q) fp.p[q] = d.V<R>() + q * fs;
    fp.n = nc;
  }
  SCL_CHECK(Ops<R>::herm_fix(*p, fp));
  tm->mark(-1);
  return SCL_OK;
}

template <class R>
int32_t scl::solver_stage(scl_plan3d* p, const DevFields& d, int mode, const scl_dns_params& prm, const cpx<R>* st) {
  if (prm.sgs || prm.scalar) return rk_stage_ext<R>(p, d, mode, prm, st, nullptr);
  const RhsParams<R> rp = rhs_params<R>(p, prm.nu, prm.dt, prm.forced, prm.fkmax, prm.dealias_rhs);
  return rk_stage<R>(p, d, mode, rp, prm.eps, prm.eta, prm.fkmax, st, nullptr);
}
template int32_t scl::solver_stage<float>(scl_plan3d*, const DevFields&, int, const scl_dns_params&,
                                          const cpx<float>*);
template int32_t scl::solver_stage<double>(scl_plan3d*, const DevFields&, int, const scl_dns_params&,
                                           const cpx<double>*);

extern "C" int32_t scl_ratesdudt(scl_plan3d* p, const void* vx, const void* vy, const void* vz, void* nx_, void* ny_,
                                 void* nz_, int32_t forced, double fkmax, double eps, double eta) {
  SCL_PLAN(p);
  if (forced < 0 || forced > 2) return set_error("forced must be 0, 1 or 2"), SCL_EINVAL;
  return by_prec(p->prec, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    void* s;
    SCL_CHECK(plan_scratch(p, 15, &s));
    DevFields d{s, p->fstride(), 3, p->coef};
    const void* vi[3] = {vx, vy, vz};
    for (int f = 0; f < 3; ++f) SCL_CHECK(upload<R>(p, vi[f], d.f<R>(6 + f)));
    const RhsParams<R> rp = rhs_params<R>(p, 0.0, 0.0, forced, fkmax, 0);
    SCL_CHECK(rk_stage<R>(p, d, 0, rp, eps, eta, fkmax, d.f<R>(6), nullptr));
    void* out[3] = {nx_, ny_, nz_};
    for (int f = 0; f < 3; ++f) SCL_CHECK(download<R>(p, d.f<R>(9 + f), out[f]));
    return SCL_OK;
  });
}

extern "C" int32_t scl_rk4step(scl_plan3d* p, void* vx, void* vy, void* vz, double nu, double dt, int32_t forced,
                               double fkmax, double eps, double eta) {
  SCL_PLAN(p);
  if (forced < 0 || forced > 2) return set_error("forced must be 0, 1 or 2"), SCL_EINVAL;
  return by_prec(p->prec, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    void* s;
    SCL_CHECK(plan_scratch(p, 15, &s));
    DevFields d{s, p->fstride(), 3, p->coef};
    void* vi[3] = {vx, vy, vz};
    for (int f = 0; f < 3; ++f) SCL_CHECK(upload<R>(p, vi[f], d.f<R>(f)));
    scl_dns_params prm{};
    prm.nu = nu;
    prm.dt = dt;
    prm.forced = forced;
    prm.fkmax = fkmax;
    prm.eps = eps;
    prm.eta = eta;
    SCL_CHECK(rk4_step_dev<R>(p, d, prm, nullptr));
    for (int f = 0; f < 3; ++f) SCL_CHECK(download<R>(p, d.f<R>(f), vi[f]));
    return SCL_OK;
  });
}

extern "C" int32_t scl_initialize_heli(int32_t prec, void* ux, void* uy, void* uz, int32_t nx, int32_t ny, int32_t nz,
                                       int32_t iseed, double hrel, double akp, double u0) {
  if (!ux || !uy || !uz || nx < 4 || ny < 2 || nz < 2 || (nx & 1) || (ny & 1) || (nz & 1) || !(hrel >= 0 && hrel <= 1))
    return set_error("bad initialize_heli arguments (even sizes, 0 <= hrel <= 1)"), SCL_EINVAL;
  if (prec == SCL_SP) return initialize_heli3d<float>((float*)ux, (float*)uy, (float*)uz, nx, ny, nz, iseed, hrel, akp, u0);
  return initialize_heli3d<double>((double*)ux, (double*)uy, (double*)uz, nx, ny, nz, iseed, hrel, akp, u0);
}

extern "C" int32_t scl_initialize3d(int32_t prec, void* vx, void* vy, void* vz, int32_t nx, int32_t ny, int32_t nz,
                                    int32_t iseed, double akp, double u0, double kcut) {
  if (!vx || !vy || !vz || nx < 4 || ny < 2 || nz < 2 || (nx & 1)) return set_error("bad initialize arguments"), SCL_EINVAL;
  if (prec == SCL_SP) return initialize3d<float>((float*)vx, (float*)vy, (float*)vz, nx, ny, nz, iseed, akp, u0, kcut);
  return initialize3d<double>((double*)vx, (double*)vy, (double*)vz, nx, ny, nz, iseed, akp, u0, kcut);
}

// ---------------------------------------------------------------------------
// Device-resident solver
static unsigned* nonfinite_flag(const scl_dns* d) { return reinterpret_cast<unsigned*>(d->coef + 8); }

static DevFields dev_fields(const scl_dns* d) {
  DevFields f{d->fields, d->plan->fstride(), d->nc, d->coef};
  f.nonfinite = nonfinite_flag(d);
  return f;
}

// Clears the sticky non-finite flag (a new state was set).
static int32_t clear_nonfinite(scl_dns* d) {
  SCL_HIP(hipMemsetAsync(nonfinite_flag(d), 0, sizeof(unsigned), d->plan->stream));
  return SCL_OK;
}

// SCL_ENONFINITE if a stage 4 since the last clear wrote a non-finite v
// (after the work queued so far has finished).
static int32_t read_nonfinite(scl_dns* d) {
  unsigned h = 0;
  SCL_HIP(hipMemcpyAsync(&h, nonfinite_flag(d), sizeof(h), hipMemcpyDeviceToHost, d->plan->stream));
  SCL_HIP(hipStreamSynchronize(d->plan->stream));
  if (h) {
    set_error("non-finite value in the solution (stepped from a non-finite or diverging state)");
    return SCL_ENONFINITE;
  }
  return SCL_OK;
}

static void free_dns(scl_dns* d) {
  if (d->exec) (void)hipGraphExecDestroy(d->exec);
  if (d->graph) (void)hipGraphDestroy(d->graph);
  if (d->fields) (void)hipFree(d->fields);
  if (d->dyn) (void)hipFree(d->dyn);
  if (d->gt) (void)hipFree(d->gt);
  if (d->part) (void)hipFree(d->part);
  if (d->coef) (void)hipFree(d->coef);
  if (d->ab) (void)hipFree(d->ab);
  if (d->ab_part) (void)hipFree(d->ab_part);
  delete d;
}

// IF-AB2: a new state restarts the bootstrap at the current time.
static int32_t ab_reset(scl_dns* d) {
  if (!d->ab) return SCL_OK;
  double h[AB_NSLOT];
  SCL_HIP(hipMemcpyAsync(h, d->ab, sizeof(h), hipMemcpyDeviceToHost, d->plan->stream));
  SCL_HIP(hipStreamSynchronize(d->plan->stream));
  const double t = h[AB_TIME];
  for (double& x : h) x = 0.0;
  h[AB_TIME] = t;
  h[AB_DTO] = d->p.dt_first > 0 ? d->p.dt_first : 1.0e30;
  h[AB_UMAX] = d->p.umax;
  h[AB_TWRITE] = t;
  SCL_HIP(hipMemcpyAsync(d->ab, h, sizeof(h), hipMemcpyHostToDevice, d->plan->stream));
  SCL_HIP(hipStreamSynchronize(d->plan->stream));  // h is on the host stack
  d->ab_steps = 0;
  return SCL_OK;
}

extern "C" int32_t scl_dns_create(scl_plan3d* p, const scl_dns_params* prm, scl_dns** out) {
  if (!out || !prm) return set_error("null argument"), SCL_EINVAL;
  *out = nullptr;
  SCL_PLAN(p);
  if (prm->forced < 0 || prm->forced > 2) return set_error("forced must be 0, 1 or 2"), SCL_EINVAL;
  if (prm->sgs < 0 || prm->sgs > 2) return set_error("sgs must be 0, 1 or 2"), SCL_EINVAL;
  if (prm->stepper < 0 || prm->stepper > 2) return set_error("stepper must be 0, 1 or 2"), SCL_EINVAL;
  if (prm->stepper == 1 && !(prm->beta > 0 && prm->umax > 0))
    return set_error("stepper 1 needs beta > 0 and umax > 0"), SCL_EINVAL;
  if (prm->stepper == 2 && !(prm->dt > 0)) return set_error("stepper 2 needs dt > 0"), SCL_EINVAL;
  if (prm->scalar && !(prm->schmidt > 0)) return set_error("scalar needs schmidt > 0"), SCL_EINVAL;
  auto* d = new scl_dns();
  d->plan = p;
  d->p = *prm;
  d->ext = prm->sgs != 0 || prm->scalar != 0;
  d->nc = prm->scalar ? 4 : 3;
  d->nw = d->ext ? ExtLayout{prm->sgs ? 1 : 0, prm->scalar ? 1 : 0}.nin() : 6;
  if (d->ext) {
    const ExtLayout L{prm->sgs ? 1 : 0, prm->scalar ? 1 : 0};
    d->nw = std::max(L.nin(), L.nout());
  }
  if (prm->sgs == 2) d->p.use_graph = 0;  // the update step differs from the others
  const size_t fb = (size_t)p->fstride() * p->esize();
  const size_t bytes = (size_t)(3 * d->nc + d->nw) * fb;
  auto fail = [&](size_t b) {
    (void)hipGetLastError();
    free_dns(d);
    set_error("device allocation of %zu bytes for the solver failed", b);
    return SCL_ENOMEM;
  };
  if (hipMalloc(&d->fields, bytes) != hipSuccess) return fail(bytes);
  if (hipMalloc((void**)&d->coef, 16 * sizeof(double)) != hipSuccess) return fail(128);
  if (prm->stepper) {
    if (hipMalloc((void**)&d->ab, AB_NSLOT * sizeof(double)) != hipSuccess) return fail(AB_NSLOT * 8);
    if (hipMalloc((void**)&d->ab_part, DIAG_BLOCKS * sizeof(double)) != hipSuccess) return fail(DIAG_BLOCKS * 8);
    if (hipMemsetAsync(d->ab, 0, AB_NSLOT * sizeof(double), p->stream) != hipSuccess) return fail(0);
  }
  if (prm->sgs == 2) {
    if (hipMalloc(&d->dyn, 24 * fb) != hipSuccess) return fail(24 * fb);
    if (hipMalloc(&d->gt, (size_t)p->fstride() * (p->esize() / 2)) != hipSuccess) return fail(fb / 2);
    if (hipMalloc((void**)&d->part, (2 * ExtOps<float>::DIAG_BLOCKS_DYN + 8) * sizeof(double)) != hipSuccess)
      return fail(8192);
  }
  double c0[8] = {0, 0, 0, 0, 0, prm->cs2, 0, 0};
  if (hipMemsetAsync(d->fields, 0, bytes, p->stream) != hipSuccess ||
      hipMemsetAsync(d->coef, 0, 16 * sizeof(double), p->stream) != hipSuccess ||
      hipMemcpyAsync(d->coef, c0, sizeof(c0), hipMemcpyHostToDevice, p->stream) != hipSuccess) {
    (void)hipGetLastError();
    free_dns(d);
    return set_error("solver initialisation failed"), SCL_EHIP;
  }
  if (prm->sgs == 2) {
    const double kc = 2.0 * (p->g.nx / 2) / 3.0;
    const double kt = prm->k_test > 0 ? prm->k_test : kc / 2.0;
    int32_t st = by_prec(p->prec, [&](auto zero) -> int32_t {
      using R = decltype(zero);
      return ExtOps<R>::gtest(*p, (R*)d->gt, (R)kt);
    });
    if (st != SCL_OK) {
      free_dns(d);
      return st;
    }
  }
  if (hipStreamSynchronize(p->stream) != hipSuccess) {
    free_dns(d);
    return set_error("solver initialisation failed"), SCL_EHIP;
  }
  if (const int32_t st = ab_reset(d); st != SCL_OK) {
    free_dns(d);
    return st;
  }
  *out = d;
  return SCL_OK;
}

extern "C" int32_t scl_dns_destroy(scl_dns* d) {
  if (!d) return SCL_OK;
  (void)hipStreamSynchronize(d->plan->stream);
  free_dns(d);
  return SCL_OK;
}

static int32_t dns_io(scl_dns* d, int first, int n, const void* const* in, void* const* out) {
  scl_plan3d* p = d->plan;
  return by_prec(p->prec, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    void* s;
    SCL_CHECK(plan_scratch(p, 0, &s));
    DevFields f = dev_fields(d);
    for (int q = 0; q < n; ++q) {
      if (in) SCL_CHECK(upload<R>(p, in[q], f.f<R>(first + q)));
      else SCL_CHECK(download<R>(p, f.f<R>(first + q), out[q]));
    }
    SCL_HIP(hipStreamSynchronize(p->stream));
    return SCL_OK;
  });
}

extern "C" int32_t scl_dns_set_velocity(scl_dns* d, const void* vx, const void* vy, const void* vz) {
  if (!d) return set_error("null solver"), SCL_EINVAL;
  const void* vi[3] = {vx, vy, vz};
  SCL_CHECK(clear_nonfinite(d));
  d->chained = false;
  SCL_CHECK(ab_reset(d));
  return dns_io(d, 0, 3, vi, nullptr);
}

extern "C" int32_t scl_dns_get_velocity(scl_dns* d, void* vx, void* vy, void* vz) {
  if (!d) return set_error("null solver"), SCL_EINVAL;
  void* vo[3] = {vx, vy, vz};
  return dns_io(d, 0, 3, nullptr, vo);
}

extern "C" int32_t scl_dns_set_scalar(scl_dns* d, const void* phi) {
  if (!d || d->nc < 4) return set_error("solver has no scalar"), SCL_EINVAL;
  const void* vi[1] = {phi};
  SCL_CHECK(clear_nonfinite(d));
  SCL_CHECK(ab_reset(d));
  return dns_io(d, 3, 1, vi, nullptr);
}

extern "C" int32_t scl_dns_get_scalar(scl_dns* d, void* phi) {
  if (!d || d->nc < 4) return set_error("solver has no scalar"), SCL_EINVAL;
  void* vo[1] = {phi};
  return dns_io(d, 3, 1, nullptr, vo);
}

extern "C" int32_t scl_dns_initialize(scl_dns* d, int32_t iseed, double akp, double u0) {
  if (!d) return set_error("null solver"), SCL_EINVAL;
  const Geom& g = d->plan->g;
  const size_t n = (size_t)(g.mx + 1) * g.nrow * 2;
  if (d->plan->prec == SCL_SP) {
    std::vector<float> a(n), b(n), c(n);
    SCL_CHECK(initialize3d<float>(a.data(), b.data(), c.data(), g.nx, g.ny, g.nz, iseed, akp, u0, -1.0));
    return scl_dns_set_velocity(d, a.data(), b.data(), c.data());
  }
  std::vector<double> a(n), b(n), c(n);
  SCL_CHECK(initialize3d<double>(a.data(), b.data(), c.data(), g.nx, g.ny, g.nz, iseed, akp, u0, -1.0));
  return scl_dns_set_velocity(d, a.data(), b.data(), c.data());
}

extern "C" int32_t scl_dns_initialize_heli(scl_dns* d, int32_t iseed, double hrel, double akp, double u0) {
  if (!d) return set_error("null solver"), SCL_EINVAL;
  const Geom& g = d->plan->g;
  const size_t n = (size_t)(g.mx + 1) * g.nrow * 2;
  const size_t es = d->plan->prec == SCL_SP ? sizeof(float) : sizeof(double);
  std::vector<char> a(n * es), b(n * es), c(n * es);
  SCL_CHECK(scl_initialize_heli(d->plan->prec, a.data(), b.data(), c.data(), g.nx, g.ny, g.nz, iseed, hrel, akp, u0));
  return scl_dns_set_velocity(d, a.data(), b.data(), c.data());
}

template <class R>
static int32_t dns_step_once(scl_dns* d, Timer* tm) {
  scl_plan3d* p = d->plan;
  DevFields f = dev_fields(d);
  if (!d->ext) {
    const bool primed = d->chained;
    d->chained = false;  // until the step is queued whole
    if (d->p.stepper) {
      SCL_CHECK(ab2_step_dev<R>(p, f, d->p, d->ab, d->ab_part, d->ab_steps, primed, tm));
      d->ab_steps += 1;
    } else {
      SCL_CHECK(rk4_step_dev<R>(p, f, d->p, tm, primed, true));
    }
    d->chained = true;
    return SCL_OK;
  }
  const scl_dns_params& prm = d->p;
  if (prm.sgs == 2 && d->steps % (prm.iupdate > 0 ? prm.iupdate : 1) == 0) {
    const double delta = prm.delta > 0 ? prm.delta : 2.0 * M_PI / p->g.nx;
    const double dtest = prm.delta_test > 0 ? prm.delta_test : 2.0 * delta;
    SCL_CHECK(ExtOps<R>::dyn_update(*p, f.V<R>(), (cpx<R>*)d->dyn, f.fs, (const R*)d->gt, (R)delta, (R)dtest, d->coef,
                                    d->part));
  }
  if (prm.stepper) {
    // IF-AB2 of spectral_les.f90 / spectral_dns_scalar.f90: one RHS of the
    // LES / scalar passes with the AB2 combine
    SCL_CHECK(ab_begin<R>(p, f, prm, d->ab, d->ab_part));
    const int mode = d->ab_steps == 0 ? 5 : d->ab_steps == 1 ? 6 : 7;
    SCL_CHECK(rk_stage_ext<R>(p, f, mode, prm, f.V<R>(), tm, d->ab));
    SCL_CHECK(ab_end(p, d->ab));
    d->ab_steps += 1;
    return SCL_OK;
  }
  for (int stage = 1; stage <= 4; ++stage) {
    const cpx<R>* st = stage == 1 ? f.V<R>() : f.S<R>();
    SCL_CHECK(rk_stage_ext<R>(p, f, stage, prm, st, tm));
  }
  return SCL_OK;
}

template <class R>
static int32_t dns_launch_steps(scl_dns* d, int nsteps, Timer* tm) {
  scl_plan3d* p = d->plan;
  if (d->p.use_graph && !tm) {
    // IF-AB2: the two bootstrap steps run eagerly; the captured step is the
    // Adams-Bashforth one
    while (d->p.stepper && d->ab_steps < 2 && nsteps > 0) {
      SCL_CHECK(dns_step_once<R>(d, nullptr));
      d->steps += 1;
      --nsteps;
    }
    if (nsteps == 0) return SCL_OK;
    if (!d->ext && !d->chained) {
      // the first step after the state was set: its z-inverse pass runs
      // here, so that the captured step is the chained one
      DevFields f = dev_fields(d);
      SCL_CHECK(Ops<R>::zinv_vort(*p, f.V<R>(), f.fs, f.W<R>(), f.fs));
      d->chained = true;
    }
    if (!d->exec) {
      SCL_HIP(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
      const int64_t keep = d->steps, keep_ab = d->ab_steps;
      int32_t st = dns_step_once<R>(d, nullptr);
      d->steps = keep;
      d->ab_steps = keep_ab;
      if (st != SCL_OK) d->chained = false;
      hipGraph_t gr = nullptr;
      hipError_t e = hipStreamEndCapture(p->stream, &gr);
      if (st != SCL_OK) {
        if (gr) (void)hipGraphDestroy(gr);
        return st;
      }
      if (e != hipSuccess) return set_error("graph capture failed: %s", hipGetErrorString(e)), SCL_EHIP;
      d->graph = gr;
      SCL_HIP(hipGraphInstantiate(&d->exec, d->graph, nullptr, nullptr, 0));
    }
    for (int i = 0; i < nsteps; ++i) {
      SCL_HIP(hipGraphLaunch(d->exec, p->stream));
      d->steps += 1;
      if (d->p.stepper) d->ab_steps += 1;
    }
  } else {
    for (int i = 0; i < nsteps; ++i) {
      SCL_CHECK(dns_step_once<R>(d, tm));
      d->steps += 1;
    }
  }
  d->time += nsteps * d->p.dt;
  return SCL_OK;
}

extern "C" int32_t scl_dns_step(scl_dns* d, int32_t nsteps) {
  if (!d || nsteps < 0) return set_error("bad arguments"), SCL_EINVAL;
  return by_prec(d->plan->prec, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    return dns_launch_steps<R>(d, nsteps, nullptr);
  });
}

extern "C" int32_t scl_dns_sync(scl_dns* d) {
  if (!d) return set_error("null solver"), SCL_EINVAL;
  SCL_HIP(hipStreamSynchronize(d->plan->stream));
  return read_nonfinite(d);
}

extern "C" double scl_dns_time(const scl_dns* d) {
  if (!d) return 0.0;
  if (!d->ab) return d->time;
  double t = 0.0;
  if (hipMemcpyAsync(&t, d->ab + AB_TIME, sizeof(t), hipMemcpyDeviceToHost, d->plan->stream) != hipSuccess ||
      hipStreamSynchronize(d->plan->stream) != hipSuccess) {
    (void)hipGetLastError();
    return std::numeric_limits<double>::quiet_NaN();
  }
  return t;
}

extern "C" int32_t scl_dns_step_info(scl_dns* d, double* time, double* dt_last, int64_t* steps) {
  if (!d) return set_error("null solver"), SCL_EINVAL;
  double t = d->time, dt = d->p.dt;
  if (d->ab) {
    double h[AB_NSLOT];
    SCL_HIP(hipMemcpyAsync(h, d->ab, sizeof(h), hipMemcpyDeviceToHost, d->plan->stream));
    SCL_HIP(hipStreamSynchronize(d->plan->stream));
    t = h[AB_TIME];
    dt = h[AB_DT];
  } else {
    SCL_HIP(hipStreamSynchronize(d->plan->stream));
  }
  if (time) *time = t;
  if (dt_last) *dt_last = dt;
  if (steps) *steps = d->ab ? d->ab_steps : d->steps;
  return SCL_OK;
}

extern "C" int32_t scl_dns_step_timed(scl_dns* d, int32_t nsteps, double* kind_ms, int32_t* kind_count,
                                      double* total_ms) {
  if (!d || nsteps < 1) return set_error("bad arguments"), SCL_EINVAL;
  scl_plan3d* p = d->plan;
  std::vector<hipEvent_t> ev;
  std::vector<int> kinds;
  Timer tm{&ev, &kinds, p->stream};
  int32_t st = by_prec(p->prec, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    return dns_launch_steps<R>(d, nsteps, &tm);
  });
  (void)hipStreamSynchronize(p->stream);
  if (st == SCL_OK) {
    for (int k = 0; k < SCL_NKIND; ++k) {
      if (kind_ms) kind_ms[k] = 0.0;
      if (kind_count) kind_count[k] = 0;
    }
    for (size_t i = 0; i + 1 < ev.size(); ++i) {
      if (kinds[i] < 0) continue;
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, ev[i], ev[i + 1]);
      if (kind_ms) kind_ms[kinds[i]] += ms;
      if (kind_count) kind_count[kinds[i]] += 1;
    }
    if (total_ms && ev.size() >= 2) {
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, ev.front(), ev.back());
      *total_ms = ms;
    }
  }
  for (auto e : ev) (void)hipEventDestroy(e);
  return st;
}

// ---------------------------------------------------------------------------
// Diagnostics
template <class R>
static int32_t diag_run(scl_plan3d* p, const cpx<R>* v, int nbins, std::vector<double>& res) {
  const int len = 2 * nbins + DIAG_TOT;
  void* buf;
  SCL_CHECK(plan_diag(p, (size_t)(diag_part_len(p->g.mx) + 2 * p->g.mx + DIAG_TOT) * sizeof(double) + 64, &buf));
  double* part = (double*)buf;
  double* fin = part + diag_part_len(nbins);
  SCL_CHECK(diag_launch<R>(p->g, v, p->fstride(), nbins, part, fin, p->stream));
  res.resize(len);
  SCL_HIP(hipMemcpyAsync(res.data(), fin, len * sizeof(double), hipMemcpyDeviceToHost, p->stream));
  SCL_HIP(hipStreamSynchronize(p->stream));
  return SCL_OK;
}

// Velocity-derivative skewness <d^3> / <d^2>^(3/2), d = du_x/dx in physical
// space (skewness.f90:1-48), FP64 sums, of the field vx (split layout).
template <class R>
static int32_t skewness_run(scl_plan3d* p, const cpx<R>* vx, double* out) {
  void* s;
  SCL_CHECK(plan_scratch(p, 1, &s));
  cpx<R>* f = (cpx<R>*)s;
  const Geom& g = p->g;
  hipLaunchKernelGGL((k_ddx_field<R>), dim3((unsigned)((g.ntot + 255) / 256)), dim3(256), 0, p->stream, vx, f, g);
  SCL_CHECK(check_launch("ddx_field"));
  SCL_CHECK(Ops<R>::c2r3d(*p, f, p->fstride(), 1));
  void* db;
  SCL_CHECK(plan_diag(p, (size_t)(diag_part_len(g.mx) + 2 * g.mx + DIAG_TOT) * sizeof(double) + 64, &db));
  double* part = (double*)db;
  double* sums = part + 2 * DIAG_BLOCKS;
  hipLaunchKernelGGL((k_moments23<R>), dim3(DIAG_BLOCKS), dim3(256), 0, p->stream, f, g.nmain, part);
  SCL_CHECK(check_launch("moments23"));
  hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(2), 0, p->stream, part, DIAG_BLOCKS, sums);
  SCL_CHECK(check_launch("moments_sum"));
  double h[2];
  SCL_HIP(hipMemcpyAsync(h, sums, sizeof(h), hipMemcpyDeviceToHost, p->stream));
  SCL_HIP(hipStreamSynchronize(p->stream));
  const double n = (double)g.nx * g.ny * g.nz;
  *out = (h[1] / n) / std::pow(h[0] / n, 1.5);
  return SCL_OK;
}

template <class R>
static int32_t divmax_run(scl_plan3d* p, const cpx<R>* v, double* out) {
  void* s;
  SCL_CHECK(plan_scratch(p, 1, &s));
  cpx<R>* f = (cpx<R>*)s;
  const Geom& g = p->g;
  hipLaunchKernelGGL((k_div_field<R>), dim3((unsigned)((g.ntot + 255) / 256)), dim3(256), 0, p->stream, v,
                     p->fstride(), f, g);
  SCL_CHECK(check_launch("div_field"));
  SCL_CHECK(Ops<R>::c2r3d(*p, f, p->fstride(), 1));
  void* db;
  SCL_CHECK(plan_diag(p, (size_t)(diag_part_len(g.mx) + 2 * g.mx + DIAG_TOT) * sizeof(double) + 64, &db));
  unsigned long long* mx = (unsigned long long*)((char*)db + p->diag_bytes - 64);
  SCL_HIP(hipMemsetAsync(mx, 0, sizeof(unsigned long long), p->stream));
  hipLaunchKernelGGL((k_absmax<R>), dim3(256), dim3(256), 0, p->stream, f, g.nmain, mx);
  SCL_CHECK(check_launch("absmax"));
  unsigned long long h = 0;
  SCL_HIP(hipMemcpyAsync(&h, mx, sizeof(h), hipMemcpyDeviceToHost, p->stream));
  SCL_HIP(hipStreamSynchronize(p->stream));
  double dv;
  std::memcpy(&dv, &h, sizeof(dv));
  *out = dv;
  return SCL_OK;
}

template <class R>
static int32_t scalar_diag_run(scl_plan3d* p, const cpx<R>* phi, double* var, double* kvar) {
  void* buf;
  SCL_CHECK(plan_diag(p, (size_t)(diag_part_len(p->g.mx) + 2 * p->g.mx + DIAG_TOT) * sizeof(double) + 64, &buf));
  double* part = (double*)buf;
  double* sums = part + 2 * DIAG_BLOCKS;
  hipLaunchKernelGGL((k_scalar_diag<R>), dim3(DIAG_BLOCKS), dim3(256), 0, p->stream, phi, p->g, part);
  SCL_CHECK(check_launch("scalar_diag"));
  hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(2), 0, p->stream, part, DIAG_BLOCKS, sums);
  SCL_CHECK(check_launch("scalar_sum"));
  double h[2];
  SCL_HIP(hipMemcpyAsync(h, sums, sizeof(h), hipMemcpyDeviceToHost, p->stream));
  SCL_HIP(hipStreamSynchronize(p->stream));
  *var = h[0];
  *kvar = h[1];
  return SCL_OK;
}

extern "C" int32_t scl_dns_spectrum(scl_dns* d, double* ek, int32_t nbins) {
  if (!d || !ek || nbins < 1 || nbins > d->plan->g.mx) return set_error("bad spectrum arguments"), SCL_EINVAL;
  scl_plan3d* p = d->plan;
  return by_prec(p->prec, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    std::vector<double> r;
    SCL_CHECK(diag_run<R>(p, (const cpx<R>*)d->fields, nbins, r));
    std::copy(r.begin(), r.begin() + nbins, ek);
    return SCL_OK;
  });
}

extern "C" int32_t scl_dns_helicity_spectrum(scl_dns* d, double* hk, int32_t nbins) {
  if (!d || !hk || nbins < 1 || nbins > d->plan->g.mx) return set_error("bad spectrum arguments"), SCL_EINVAL;
  scl_plan3d* p = d->plan;
  return by_prec(p->prec, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    std::vector<double> r;
    SCL_CHECK(diag_run<R>(p, (const cpx<R>*)d->fields, nbins, r));
    std::copy(r.begin() + nbins, r.begin() + 2 * nbins, hk);
    return SCL_OK;
  });
}

extern "C" int32_t scl_dns_diagnostics(scl_dns* d, scl_diag* out) {
  if (!d || !out) return set_error("null argument"), SCL_EINVAL;
  scl_plan3d* p = d->plan;
  return by_prec(p->prec, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    std::vector<double> r;
    const int nb = p->g.mx;
    const cpx<R>* v = (const cpx<R>*)d->fields;
    SCL_CHECK(diag_run<R>(p, v, nb, r));
    out->energy = r[2 * nb + 0];
    out->enstrophy = r[2 * nb + 1];
    out->dissipation = 2.0 * d->p.nu * r[2 * nb + 1];
    out->helicity = r[2 * nb + 2];
    out->helicity_dissipation = 2.0 * d->p.nu * r[2 * nb + 5];
    out->divrms_spec = std::sqrt(r[2 * nb + 3]);
    out->nonfinite = r[2 * nb + 4] > 0 ? 1 : 0;
    out->scalar_variance = 0;
    out->scalar_dissipation = 0;
    if (d->nc == 4) {
      double var, kvar;
      SCL_CHECK(scalar_diag_run<R>(p, v + 3 * p->fstride(), &var, &kvar));
      out->scalar_variance = var;
      out->scalar_dissipation = 2.0 * d->p.nu / d->p.schmidt * kvar;
    }
    double c[8];
    SCL_HIP(hipMemcpy(c, d->coef, sizeof(c), hipMemcpyDeviceToHost));
    out->cs2 = d->p.sgs ? c[5] : 0.0;
    SCL_CHECK(skewness_run<R>(p, v, &out->skewness));
    SCL_CHECK(divmax_run<R>(p, v, &out->divmax));
    // every field of *out is filled; the status reports a non-finite state
    const int32_t nf = read_nonfinite(d);
    if (nf == SCL_OK && out->nonfinite) {
      set_error("non-finite value in the solution");
      return SCL_ENONFINITE;
    }
    if (nf == SCL_ENONFINITE) out->nonfinite = 1;
    return nf;
  });
}

// ---------------------------------------------------------------------------
// SGS stress and scalar flux on host arrays
// Runs the LES part of the extended pipeline on v with only tau as output:
// tau_ij spectral in W[3..8] after prep -> c2r -> phys -> r2c.
template <class R>
static int32_t tau_static(scl_plan3d* p, cpx<R>* W, int64_t fs, double cs2, double delta, double* coefdev) {
  // W[0..2] holds v on entry (as the stage state of the prep kernel)
  const Geom& g = p->g;
  const ExtLayout L{1, 0};
  cpx<R>* st = W + 13 * fs;  // copy of v past the 13 work fields
  SCL_HIP(hipMemcpyAsync(st, W, 3 * fs * sizeof(cpx<R>), hipMemcpyDeviceToDevice, p->stream));
  double c0[8] = {0, 0, 0, 0, 0, cs2, 0, 0};
  SCL_HIP(hipMemcpyAsync(coefdev, c0, sizeof(c0), hipMemcpyHostToDevice, p->stream));
  hipLaunchKernelGGL((k_ext_prep<R>), ExtOps<R>::grid_all(g), dim3(256), 0, p->stream, st, W, fs, g, L);
  SCL_CHECK(check_launch("ext_prep"));
  SCL_CHECK(Ops<R>::c2r3d(*p, W, fs, L.nin()));
  const R scale = (R)1 / (R)((double)g.nx * g.ny * g.nz);
  SCL_CHECK(ExtOps<R>::phys(*p, W, fs, L, scale, (R)(delta * delta), coefdev + 5, nullptr));
  return Ops<R>::r2c3d(*p, W + L.tau_out() * fs, fs, 6);
}

extern "C" int32_t scl_smag(scl_plan3d* p, const void* vx, const void* vy, const void* vz, void* t11, void* t12,
                            void* t13, void* t22, void* t23, void* t33, const void* kx, const void* ky, const void* kz,
                            int32_t lx1, int32_t ly, int32_t lz, int32_t nx, int32_t ny, int32_t nz, double delta) {
  SCL_PLAN(p);
  SCL_CHECK(check_dims(p, lx1, ly, lz));
  if (nx != p->g.nx || ny != p->g.ny || nz != p->g.nz) return set_error("nx,ny,nz do not match the plan"), SCL_EINVAL;
  return by_prec(p->prec, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    SCL_CHECK(check_k1d<R>(p, kx, ky, kz));
    void* s;
    SCL_CHECK(plan_scratch(p, 16, &s));
    cpx<R>* W = (cpx<R>*)s;
    const int64_t fs = p->fstride();
    const void* vi[3] = {vx, vy, vz};
    for (int q = 0; q < 3; ++q) SCL_CHECK(upload<R>(p, vi[q], W + q * fs));
    // smag.f90:7: c1 = 0.026 (a real(sp) parameter)
    SCL_CHECK(tau_static<R>(p, W, fs, (double)(R)0.026f, delta, p->coef));
    void* to[6] = {t11, t12, t13, t22, t23, t33};
    for (int q = 0; q < 6; ++q) SCL_CHECK(download<R>(p, W + (3 + q) * fs, to[q]));
    return SCL_OK;
  });
}

// kx, ky, kz given as 3D arrays (dynsmag.f90:14): spot-check they are wavenumber()'s
template <class R>
static int32_t check_k3d(const scl_plan3d* p, const void* kx, const void* ky, const void* kz) {
  const R* a[3] = {(const R*)kx, (const R*)ky, (const R*)kz};
  if (!a[0] || !a[1] || !a[2]) return set_error("null wavenumber array"), SCL_EINVAL;
  const Geom& g = p->g;
  const int pts[5][3] = {{0, 0, 0}, {g.mx, g.ny - 1, g.nz - 1}, {1, g.ny / 2, 3 % g.nz}, {2, 1, g.nz / 2}, {g.mx / 3, 5 % g.ny, g.nz - 2}};
  for (auto& q : pts) {
    const int64_t o = ((int64_t)q[2] * g.ny + q[1]) * (g.mx + 1) + q[0];
    const R want[3] = {(R)q[0], (R)kwave(q[1], g.ny), (R)kwave(q[2], g.nz)};
    for (int c = 0; c < 3; ++c)
      if (a[c][o] != want[c]) return set_error("3D wavenumber array %d is not wavenumber()'s", c), SCL_EINVAL;
  }
  return SCL_OK;
}

extern "C" int32_t scl_dynsmag(scl_plan3d* p, const void* ux, const void* uy, const void* uz, void* t11, void* t12,
                               void* t13, void* t22, void* t23, void* t33, const void* kx, const void* ky,
                               const void* kz, const void* g_test, int32_t lx1, int32_t ly, int32_t lz, int32_t nx,
                               int32_t ny, int32_t nz, int32_t lx, double delta, double delta_test, double* mm,
                               double* lm, double* cs2, int32_t update) {
  SCL_PLAN(p);
  SCL_CHECK(check_dims(p, lx1, ly, lz));
  if (nx != p->g.nx || ny != p->g.ny || nz != p->g.nz || lx != p->g.mx)
    return set_error("nx,ny,nz,lx do not match the plan"), SCL_EINVAL;
  if (!cs2 || (update && (!mm || !lm))) return set_error("null mm/lm/cs2"), SCL_EINVAL;
  return by_prec(p->prec, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    SCL_CHECK(check_k3d<R>(p, kx, ky, kz));
    void* s;
    SCL_CHECK(plan_scratch(p, 40, &s));
    cpx<R>* W = (cpx<R>*)s;
    const int64_t fs = p->fstride();
    const void* vi[3] = {ux, uy, uz};
    for (int q = 0; q < 3; ++q) SCL_CHECK(upload<R>(p, vi[q], W + q * fs));
    if (update) {
      if (!g_test) return set_error("null g_test"), SCL_EINVAL;
      // scratch: velocity W[37..39], 24 dynamic fields W[13..36], g_test
      // staged in W[4] (reference layout) and W[3] (split layout), as reals
      cpx<R>* V = W + 37 * fs;
      cpx<R>* D = W + 13 * fs;
      R* gsplit = (R*)(W + 3 * fs);
      R* gref = (R*)(W + 4 * fs);
      SCL_HIP(hipMemcpyAsync(V, W, 3 * fs * sizeof(cpx<R>), hipMemcpyDeviceToDevice, p->stream));
      SCL_HIP(hipMemcpyAsync(gref, g_test, (size_t)(p->g.mx + 1) * p->g.nrow * sizeof(R), hipMemcpyHostToDevice,
                             p->stream));
      hipLaunchKernelGGL((k_ref_to_split_real<R>), dim3((unsigned)((p->g.nrow * (p->g.mx + 1) + 255) / 256)),
                         dim3(256), 0, p->stream, gref, gsplit, p->g);
      SCL_CHECK(check_launch("ref_to_split_real"));
      void* part;
      SCL_CHECK(plan_diag(p, (size_t)(diag_part_len(p->g.mx) + 2 * p->g.mx + DIAG_TOT) * sizeof(double) + 64, &part));
      SCL_CHECK(ExtOps<R>::dyn_update(*p, V, D, fs, gsplit, (R)delta, (R)delta_test, p->coef, (double*)part));
      // tau = -cs2 * FFT(2 D^2 |S| S_ij)/N^3 (dynsmag.f90:186-193), from D[0..5]
      hipLaunchKernelGGL((k_tau_from_s<R>), ExtOps<R>::grid_all(p->g), dim3(256), 0, p->stream, D, fs, p->g,
                         (const double*)p->coef);
      SCL_CHECK(check_launch("tau_from_s"));
      double c[8];
      SCL_HIP(hipMemcpyAsync(c, p->coef, sizeof(c), hipMemcpyDeviceToHost, p->stream));
      SCL_HIP(hipStreamSynchronize(p->stream));
      *cs2 = c[5];
      *mm = c[6];
      *lm = c[7];
      void* to[6] = {t11, t12, t13, t22, t23, t33};
      for (int q = 0; q < 6; ++q) SCL_CHECK(download<R>(p, D + q * fs, to[q]));
      return SCL_OK;
    }
    SCL_CHECK(tau_static<R>(p, W, fs, *cs2, delta, p->coef));
    void* to[6] = {t11, t12, t13, t22, t23, t33};
    for (int q = 0; q < 6; ++q) SCL_CHECK(download<R>(p, W + (3 + q) * fs, to[q]));
    return SCL_OK;
  });
}

extern "C" int32_t scl_convecphi_dns(scl_plan3d* p, const void* vx, const void* vy, const void* vz, const void* phi,
                                     void* wphi, const void* kx, const void* ky, const void* kz, int32_t lx1,
                                     int32_t ly, int32_t lz, int32_t nx, int32_t ny, int32_t nz) {
  SCL_PLAN(p);
  SCL_CHECK(check_dims(p, lx1, ly, lz));
  if (nx != p->g.nx || ny != p->g.ny || nz != p->g.nz) return set_error("nx,ny,nz do not match the plan"), SCL_EINVAL;
  return by_prec(p->prec, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    SCL_CHECK(check_k1d<R>(p, kx, ky, kz));
    void* s;
    SCL_CHECK(plan_scratch(p, 4, &s));
    cpx<R>* W = (cpx<R>*)s;
    const int64_t fs = p->fstride();
    const void* vi[4] = {vx, vy, vz, phi};
    for (int q = 0; q < 4; ++q) SCL_CHECK(upload<R>(p, vi[q], W + q * fs));
    SCL_CHECK(Ops<R>::c2r3d(*p, W, fs, 4));
    const Geom& g = p->g;
    hipLaunchKernelGGL((k_flux<R>), ExtOps<R>::grid_main(g), dim3(256), 0, p->stream, W, fs,
                       (R)1 / (R)((double)g.nx * g.ny * g.nz), g);
    SCL_CHECK(check_launch("flux"));
    SCL_CHECK(Ops<R>::r2c3d(*p, W, fs, 3));
    hipLaunchKernelGGL((k_divflux<R>), ExtOps<R>::grid_all(g), dim3(256), 0, p->stream, W, fs, g);
    SCL_CHECK(check_launch("divflux"));
    return download<R>(p, W, wphi);
  });
}

// ---------------------------------------------------------------------------
// 3/2-rule padded products (padd.f90, unpadd.f90, fftwplan_les.f90,
// smag_dealias.f90).  The padded grid here is the smallest power of two per
// axis that is at least 3n/2 (2n for the plan's powers of two): a product of
// two fields band-limited to |k_i| < n/2 is alias-free on every grid of at
// least 3n/2 points per axis, so the modes kept by unpadd come out as on the
// reference's 3n/2 grid, while the transforms stay radix-2/4/8/16.
static dim3 grid1(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

static int pad_size(int n) {
  int m = 1;
  while (m < 3 * n / 2) m <<= 1;
  return m;
}

static int32_t padded_plan(scl_plan3d* p, scl_plan3d** out) {
  if (!p->padded) {
    scl_plan3d* b = nullptr;
    SCL_CHECK(scl_plan3d_create(pad_size(p->g.nx), pad_size(p->g.ny), pad_size(p->g.nz), p->prec, &b));
    const int32_t st = scl_plan3d_set_stream(b, p->stream);
    if (st != SCL_OK) {
      scl_plan3d_destroy(b);
      return st;
    }
    p->padded = b;
  } else if (p->padded->stream != p->stream) {
    SCL_CHECK(scl_plan3d_set_stream(p->padded, p->stream));
  }
  *out = p->padded;
  return SCL_OK;
}

static int32_t check_pad_dims(const scl_plan3d* p, int lxb1, int lyb, int lzb, int nxb, int nyb, int nzb) {
  const Geom& g = p->g;
  if (lxb1 != nxb / 2 + 1 || lyb != nyb || lzb != nzb || 2 * nxb < 3 * g.nx || 2 * nyb < 3 * g.ny ||
      2 * nzb < 3 * g.nz)
    return set_error("padded sizes (%d,%d,%d) / (%d,%d,%d): need lxb1 = nxb/2+1, lyb = nyb, lzb = nzb, each at "
                     "least 3/2 of the plan's grid",
                     lxb1, lyb, lzb, nxb, nyb, nzb),
           SCL_EINVAL;
  return SCL_OK;
}

extern "C" int32_t scl_padd(int32_t prec, const void* u, void* ub, int32_t lx1, int32_t ly, int32_t lz, int32_t lxb1,
                            int32_t lyb, int32_t lzb) {
  if (!u || !ub || lx1 < 1 || ly < 2 || lz < 2 || lxb1 < lx1 || lyb < ly || lzb < lz || (ly & 1) || (lz & 1))
    return set_error("bad padd arguments"), SCL_EINVAL;
  const size_t es = prec == SCL_SP ? 8 : 16;
  const char* a = (const char*)u;
  char* b = (char*)ub;
  std::memset(b, 0, (size_t)lxb1 * lyb * lzb * es);
  for (int z = 0; z < lz; ++z) {
    const int zb = z < lz / 2 ? z : lzb - lz + z;
    for (int y = 0; y < ly; ++y) {
      const int yb = y < ly / 2 ? y : lyb - ly + y;
      std::memcpy(b + (((size_t)zb * lyb + yb) * lxb1) * es, a + (((size_t)z * ly + y) * lx1) * es, (size_t)lx1 * es);
    }
  }
  return SCL_OK;
}

extern "C" int32_t scl_unpadd(int32_t prec, const void* ub, void* u, int32_t lx1, int32_t ly, int32_t lz,
                              int32_t lxb1, int32_t lyb, int32_t lzb) {
  if (!u || !ub || lx1 < 1 || ly < 2 || lz < 2 || lxb1 < lx1 || lyb < ly || lzb < lz || (ly & 1) || (lz & 1))
    return set_error("bad unpadd arguments"), SCL_EINVAL;
  const size_t es = prec == SCL_SP ? 8 : 16;
  const char* b = (const char*)ub;
  char* a = (char*)u;
  for (int z = 0; z < lz; ++z) {
    const int zb = z < lz / 2 ? z : lzb - lz + z;
    for (int y = 0; y < ly; ++y) {
      const int yb = y < ly / 2 ? y : lyb - ly + y;
      std::memcpy(a + (((size_t)z * ly + y) * lx1) * es, b + (((size_t)zb * lyb + yb) * lxb1) * es, (size_t)lx1 * es);
    }
  }
  return SCL_OK;
}

extern "C" int32_t scl_convec_padded(scl_plan3d* p, const void* vx, const void* vy, const void* vz, void* wx,
                                     void* wy, void* wz, int32_t lx1, int32_t ly, int32_t lz, int32_t lxb1,
                                     int32_t lyb, int32_t lzb, int32_t nxb, int32_t nyb, int32_t nzb) {
  SCL_PLAN(p);
  SCL_CHECK(check_dims(p, lx1, ly, lz));
  SCL_CHECK(check_pad_dims(p, lxb1, lyb, lzb, nxb, nyb, nzb));
  scl_plan3d* pb;
  SCL_CHECK(padded_plan(p, &pb));
  return by_prec(p->prec, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    void *s, *sb;
    SCL_CHECK(plan_scratch(p, 1, &s));
    SCL_CHECK(plan_scratch(pb, 6, &sb));
    cpx<R>* U = (cpx<R>*)s;
    cpx<R>* B = (cpx<R>*)sb;
    const int64_t fsb = pb->fstride();
    const void* in[6] = {vx, vy, vz, wx, wy, wz};
    for (int f = 0; f < 6; ++f) {
      SCL_CHECK(upload<R>(p, in[f], U));
      hipLaunchKernelGGL((k_pad<R>), grid1(pb->g.ntot), dim3(256), 0, p->stream, U, p->g, B + f * fsb, pb->g);
      SCL_CHECK(check_launch("pad"));
    }
    SCL_CHECK(Ops<R>::colfft(*pb, B, fsb, 6, 2, +1));
    SCL_CHECK(Ops<R>::colfft(*pb, B, fsb, 6, 1, +1));
    SCL_CHECK(Ops<R>::row_convec(*pb, B, fsb));  // u x w / (Mx My Mz)
    SCL_CHECK(Ops<R>::colfft(*pb, B, fsb, 3, 1, -1));
    SCL_CHECK(Ops<R>::colfft(*pb, B, fsb, 3, 2, -1));
    void* out[3] = {wx, wy, wz};
    for (int f = 0; f < 3; ++f) {
      hipLaunchKernelGGL((k_unpad<R>), grid1(p->g.ntot), dim3(256), 0, p->stream, B + f * fsb, pb->g, U, p->g);
      SCL_CHECK(check_launch("unpad"));
      SCL_CHECK(download<R>(p, U, out[f]));
    }
    return SCL_OK;
  });
}

extern "C" int32_t scl_smag_dealias(scl_plan3d* p, const void* vx, const void* vy, const void* vz, void* t11,
                                    void* t12, void* t13, void* t22, void* t23, void* t33, const void* kx,
                                    const void* ky, const void* kz, int32_t lx1, int32_t ly, int32_t lz,
                                    int32_t lxb1, int32_t lyb, int32_t lzb, int32_t nxb, int32_t nyb, int32_t nzb,
                                    double delta) {
  SCL_PLAN(p);
  SCL_CHECK(check_dims(p, lx1, ly, lz));
  SCL_CHECK(check_pad_dims(p, lxb1, lyb, lzb, nxb, nyb, nzb));
  scl_plan3d* pb;
  SCL_CHECK(padded_plan(p, &pb));
  return by_prec(p->prec, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    SCL_CHECK(check_k3d<R>(p, kx, ky, kz));
    void *s, *sb;
    SCL_CHECK(plan_scratch(p, 9, &s));
    SCL_CHECK(plan_scratch(pb, 6, &sb));
    cpx<R>* V = (cpx<R>*)s;
    const int64_t fs = p->fstride(), fsb = pb->fstride();
    cpx<R>* S = V + 3 * fs;
    cpx<R>* B = (cpx<R>*)sb;
    const void* in[3] = {vx, vy, vz};
    for (int f = 0; f < 3; ++f) SCL_CHECK(upload<R>(p, in[f], V + f * fs));
    hipLaunchKernelGGL((k_strain_smagd<R>), grid1(p->g.ntot), dim3(256), 0, p->stream, V, S, fs, p->g);
    SCL_CHECK(check_launch("strain_smagd"));
    for (int q = 0; q < 6; ++q) {
      hipLaunchKernelGGL((k_pad<R>), grid1(pb->g.ntot), dim3(256), 0, p->stream, S + q * fs, p->g, B + q * fsb, pb->g);
      SCL_CHECK(check_launch("pad"));
    }
    SCL_CHECK(Ops<R>::c2r3d(*pb, B, fsb, 6));
    // smag_dealias.f90:7,44-57: c1 = 0.026 (real(sp)), const = 1/(padded points)
    const double c1 = (double)(float)0.026f;
    const double coef = -2.0 * c1 * delta * delta / ((double)pb->g.nx * pb->g.ny * pb->g.nz);
    hipLaunchKernelGGL((k_smagd_phys<R>), grid1(pb->g.nmain), dim3(256), 0, p->stream, B, fsb, pb->g.nmain, (R)coef);
    SCL_CHECK(check_launch("smagd_phys"));
    SCL_CHECK(Ops<R>::r2c3d(*pb, B, fsb, 6));
    void* out[6] = {t11, t12, t13, t22, t23, t33};
    for (int q = 0; q < 6; ++q) {
      hipLaunchKernelGGL((k_unpad<R>), grid1(p->g.ntot), dim3(256), 0, p->stream, B + q * fsb, pb->g, S, p->g);
      SCL_CHECK(check_launch("unpad"));
      SCL_CHECK(download<R>(p, S, out[q]));
    }
    return SCL_OK;
  });
}

extern "C" int32_t scl_skewness(scl_plan3d* p, const void* ux, const void* kx, double* s, int32_t nx, int32_t ny,
                                int32_t nz) {
  SCL_PLAN(p);
  if (nx != p->g.nx || ny != p->g.ny || nz != p->g.nz) return set_error("nx,ny,nz do not match the plan"), SCL_EINVAL;
  if (!s || !kx) return set_error("null kx or s"), SCL_EINVAL;
  return by_prec(p->prec, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    const R* k = (const R*)kx;
    for (int i = 0; i <= p->g.mx; ++i)
      if (k[i] != (R)i) return set_error("kx(%d) is not wavenumber()'s", i + 1), SCL_EINVAL;
    void* w;
    SCL_CHECK(plan_scratch(p, 2, &w));
    cpx<R>* u = (cpx<R>*)w + p->fstride();  // the second field: skewness_run uses the first
    SCL_CHECK(upload<R>(p, ux, u));
    return skewness_run<R>(p, u, s);
  });
}

// ---------------------------------------------------------------------------
// FFTW 2.x Fortran-77 execution entry points (convec_dns.f90:22-27,45-47),
// by the names unmodified reference code calls: in place on the global
// plan's grid; the plan token and the `ignore` argument are not used.
extern "C" void rfftwnd_f77_one_complex_to_real_(int64_t* /*plan*/, void* a, void* /*ignore*/) {
  const int32_t st = scl_c2r3d(nullptr, a);
  if (st != SCL_OK) {
    fprintf(stderr, "libscl: rfftwnd_f77_one_complex_to_real failed: %s\n", scl_last_error());
    abort();
  }
}

extern "C" void rfftwnd_f77_one_real_to_complex_(int64_t* /*plan*/, void* a, void* /*ignore*/) {
  const int32_t st = scl_r2c3d(nullptr, a);
  if (st != SCL_OK) {
    fprintf(stderr, "libscl: rfftwnd_f77_one_real_to_complex failed: %s\n", scl_last_error());
    abort();
  }
}
