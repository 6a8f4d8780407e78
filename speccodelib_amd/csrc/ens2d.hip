// C ABI of the batched 2D vorticity ensembles: size groups, per-member
// storage mapping, the IF-RK4 step as one hipGraph, the schedule autotuner.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <limits>
#include <vector>

#include "ic.hpp"
#include "plan.hpp"
#include "vort2d.hpp"

namespace scl {
int32_t check_launch(const char* what);
// exp(-2 pi i j/L), j < L, in FP64 (defined with the plan layer)
void twiddles(int L, std::vector<double>& re, std::vector<double>& im);
}

namespace {

using namespace scl;

template <int V>
using Int = std::integral_constant<int, V>;

// One batch of same-size members: six fields of Geom(n, n, nb) in one
// allocation: three work fields, the state w, the stage accumulator, the
// stage vorticity.
struct Group {
  int n = 0, nb = 0;
  Geom g{};
  const void* tw = nullptr;  // exp(-2 pi i j/n), j < n (owned by the ensemble)
  void* base = nullptr;
  int e = 8, cw = 8, row_grid = 0;
  hipStream_t stream = nullptr;  // branch stream (batched mode)
  hipEvent_t done = nullptr;
  int64_t fs() const { return g.ntot + SCL_FIELD_PAD; }  // padded as the 3D fields
  template <class R>
  cpx<R>* field(int i) const {
    return reinterpret_cast<cpx<R>*>(base) + i * fs();
  }
};

// Scheme 1 (turb_2d.f90) keeps the previous nonlinear term in F_ACC, the
// state before the first step in F_S and the unshifted half of the
// nonlinear term in a seventh field, F_NA.
enum { F_W = 0, F_WS = 3, F_ACC = 4, F_S = 5, F_NA = 6, NFIELDS = 6, NFIELDS_TURB = 7 };

// FP64 plans keep one schedule: E = 8 with 8 columns, 4 at 1024 points
// (8 would need 163 KB of LDS with the line's twiddles).
template <class R, int N, int E, int CW>
constexpr bool schedule_ok() {
  if constexpr (sizeof(R) == 8) return E == 8 && CW == (N >= 1024 ? 4 : 8) && Col2D<R, N, E, CW>::valid;
  else return Col2D<R, N, E, CW>::valid;
}

template <class R, class F>
int32_t with_col2d(int n, int e, int cw, F&& f) {
  auto go = [&](auto NC) -> int32_t {
    constexpr int N = decltype(NC)::value;
    auto pick = [&](auto EC, auto CC) -> int32_t {
      constexpr int E = decltype(EC)::value, CW = decltype(CC)::value;
      if constexpr (schedule_ok<R, N, E, CW>()) {
        return f(NC, EC, CC);
      } else {
        set_error("schedule e=%d cw=%d is not available for n=%d in this precision", E, CW, N);
        return SCL_EINVAL;
      }
    };
    if (e == 8 && cw == 4) return pick(Int<8>{}, Int<4>{});
    if (e == 8 && cw == 8) return pick(Int<8>{}, Int<8>{});
    if (e == 8 && cw == 16) return pick(Int<8>{}, Int<16>{});
    if (e == 16 && cw == 8) return pick(Int<16>{}, Int<8>{});
    if (e == 16 && cw == 16) return pick(Int<16>{}, Int<16>{});
    set_error("schedule e=%d cw=%d: e must be 8 or 16, cw 8 or 16 (or 4 with e = 8)", e, cw);
    return SCL_EINVAL;
  };
  switch (n) {
    case 16: return go(Int<16>{});
    case 32: return go(Int<32>{});
    case 64: return go(Int<64>{});
    case 128: return go(Int<128>{});
    case 256: return go(Int<256>{});
    case 512: return go(Int<512>{});
    case 1024: return go(Int<1024>{});
    default: break;
  }
  set_error("2D size %d is not a power of two in [16, 1024]", n);
  return SCL_EINVAL;
}

template <class R, class F>
int32_t with_row2d(int m, F&& f) {
  switch (m) {
    case 8: return f(Int<8>{});
    case 16: return f(Int<16>{});
    case 32: return f(Int<32>{});
    case 64: return f(Int<64>{});
    case 128: return f(Int<128>{});
    case 256: return f(Int<256>{});
    case 512: return f(Int<512>{});
    default: break;
  }
  set_error("2D size %d is not available", 2 * m);
  return SCL_EINVAL;
}

template <class R>
bool schedule_valid(int n, int e, int cw) {
  return with_col2d<R>(n, e, cw, [](auto, auto, auto) -> int32_t { return SCL_OK; }) == SCL_OK;
}

}  // namespace

struct scl_ens2d {
  int prec = SCL_SP;
  scl_ens2d_params p{};
  std::vector<int> sizes, members, first;  // user groups; first internal group of each
  std::vector<Group> groups;                // internal groups
  std::vector<std::pair<int, void*>> tw;    // twiddle table per size
  hipStream_t stream = nullptr;
  hipEvent_t fork = nullptr, t0 = nullptr, t1 = nullptr;
  hipGraph_t graph[2] = {nullptr, nullptr};      // one step; a chunk of steps
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  int chunk = 10;  // steps per chunk graph (batched mode; 1 per member)
  // One stream for all groups: 2.03 vs 2.20 ms/step for a graph branch per
  // size group (SCL_ENS2D_FORK=1), whose concurrent groups slow each other.
  bool fork_groups = false;
  double* diag = nullptr;
  int ncu = 256;
  int64_t steps = 0;
  double time = 0.0;
  int total_members() const {
    int t = 0;
    for (int m : members) t += m;
    return t;
  }
  size_t esize() const { return prec == SCL_SP ? 2 * sizeof(float) : 2 * sizeof(double); }
};

namespace {

template <class R>
Rhs2D<R> rhs_params(const scl_ens2d& e, int n) {
  Rhs2D<R> r;
  r.dt = (R)e.p.dt;
  r.fk2min = (R)(e.p.fkmin * e.p.fkmin);
  r.fk2max = (R)(e.p.fkmax * e.p.fkmax);
  r.eps = (R)e.p.eps;
  const double kc = 2.0 * (n / 2) / 3.0;
  r.kc2 = (R)(kc * kc);
  r.forced = e.p.forced;
  r.rate.c_hyper = (R)e.p.c_hyper;
  r.rate.c_hypo = (R)e.p.c_hypo;
  r.rate.damp = (R)e.p.damp;
  r.rate.n_hyper = e.p.n_hyper;
  r.rate.n_hypo = e.p.n_hypo;
  return r;
}

// Workgroups of the x-pass: persistent (as many as can be resident) or one
// per row group.
template <class R, int M>
int64_t row_grid_size(const scl_ens2d& e, const Group& gr) {
  using RC = RowCfg<R, M>;
  static int resident = 0;
  if (!resident) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_row_adv2d<R, M>, RC::BS, 0) != hipSuccess ||
        per_cu < 1)
      per_cu = 1;
    resident = per_cu;
  }
  const int64_t ngroups = gr.g.nrow / RC::RPW;
  return gr.row_grid ? ngroups : std::min<int64_t>(ngroups, (int64_t)resident * e.ncu);
}

template <class R>
int32_t yinv_launch(const Group& gr, const cpx<R>* src, int shift, hipStream_t st) {
  const cpx<R>* tw = reinterpret_cast<const cpx<R>*>(gr.tw);
  return with_col2d<R>(gr.n, gr.e, gr.cw, [&](auto NC, auto EC, auto CC) -> int32_t {
    constexpr int N = decltype(NC)::value, E = decltype(EC)::value, CW = decltype(CC)::value;
    hipLaunchKernelGGL((k_yinv2d<R, N, E, CW>), dim3(col_grid<CW, 1>(gr.g)), dim3(Col2D<R, N, E, CW>::BS), 0, st,
                       src, gr.field<R>(F_W), gr.fs(), gr.g, tw, 1, shift);
    return check_launch("yinv2d");
  });
}

template <class R>
int32_t row_launch(const scl_ens2d& e, const Group& gr, hipStream_t st) {
  const cpx<R>* tw = reinterpret_cast<const cpx<R>*>(gr.tw);
  return with_row2d<R>(gr.n / 2, [&](auto MC) -> int32_t {
    constexpr int M = decltype(MC)::value;
    const R scale = (R)1 / (R)((double)gr.n * gr.n);
    hipLaunchKernelGGL((k_row_adv2d<R, M>), dim3((unsigned)row_grid_size<R, M>(e, gr)), dim3(RowCfg<R, M>::BS), 0,
                       st, gr.field<R>(F_W), gr.fs(), gr.g, scale, tw, 2, 1);
    return check_launch("row_adv2d");
  });
}

// One RHS evaluation (mode 0) or IF-RK4 stage (modes 1..4) of a group.
// primed: W[0..2] already hold the y-inverse of the stage vorticity (the
// previous stage was fused); fuse: the y-forward also y-inverses the new
// stage vorticity into W[0..2] (k_yfwd2d FUSED), for the next stage.
template <class R>
int32_t stage(const scl_ens2d& e, const Group& gr, int mode, hipStream_t st, bool primed = false, bool fuse = false) {
  using C = cpx<R>;
  C* W = gr.field<R>(F_W);
  C* WS = gr.field<R>(F_WS);
  C* ACC = gr.field<R>(F_ACC);
  C* S = gr.field<R>(F_S);
  const C* sin = mode <= 1 ? WS : S;
  const C* tw = reinterpret_cast<const C*>(gr.tw);
  const int64_t fs = gr.fs();
  const Rhs2D<R> rp = rhs_params<R>(e, gr.n);
  if (fuse && mode < 1) return set_error("fused 2D stage: mode %d is not a stage 1..4", mode), SCL_EINVAL;
  if (!primed) SCL_CHECK(yinv_launch<R>(gr, sin, 0, st));
  SCL_CHECK(row_launch<R>(e, gr, st));
  return with_col2d<R>(gr.n, gr.e, gr.cw, [&](auto NC, auto EC, auto CC) -> int32_t {
    constexpr int N = decltype(NC)::value, E = decltype(EC)::value, CW = decltype(CC)::value;
    const dim3 grd(col_grid<CW, 1>(gr.g)), blk(Col2D<R, N, E, CW>::BS);
    if (fuse)
      hipLaunchKernelGGL((k_yfwd2d<R, N, E, CW, true>), grd, blk, 0, st, W, fs, WS, ACC, S, sin, mode, rp, gr.g, tw, 1);
    else
      hipLaunchKernelGGL((k_yfwd2d<R, N, E, CW, false>), grd, blk, 0, st, W, fs, WS, ACC, S, sin, mode, rp, gr.g, tw,
                         1);
    return check_launch("yfwd2d");
  });
}

template <class R>
Turb2DP<R> turb_params(const scl_ens2d& e, int n) {
  Turb2DP<R> t;
  t.dt = (R)e.p.dt;
  t.sigma = (R)e.p.sigma;
  t.nu = (R)e.p.nu;
  const R etak = (R)e.p.eta / (R)104;  // turb_2d.f90:111
  t.etak4 = etak / (R)4;
  t.kf = (R)e.p.kf;
  // dealiasing(): ass = 2.0/9.0*(my*my) in single precision (turb_2d.f90:519)
  t.kc2 = (R)((2.0f / 9.0f) * (float)((double)n * n));
  t.forced = e.p.forced;
  return t;
}

// One turb_2d.f90 step (mode TURB_STEP0, TURB_STEP1, TURB_AB2) or nonlinear
// term (TURB_RHS) of a group: the unshifted half, then the shifted half and
// the update.
template <class R>
int32_t turb_stage(const scl_ens2d& e, const Group& gr, int mode, hipStream_t st) {
  using C = cpx<R>;
  const C* tw = reinterpret_cast<const C*>(gr.tw);
  const Turb2DP<R> tp = turb_params<R>(e, gr.n);
  for (int half = 0; half < 2; ++half) {
    SCL_CHECK(yinv_launch<R>(gr, gr.field<R>(F_WS), half, st));
    SCL_CHECK(row_launch<R>(e, gr, st));
    SCL_CHECK(with_col2d<R>(gr.n, gr.e, gr.cw, [&](auto NC, auto EC, auto CC) -> int32_t {
      constexpr int N = decltype(NC)::value, E = decltype(EC)::value, CW = decltype(CC)::value;
      hipLaunchKernelGGL((k_yfwd2d_turb<R, N, E, CW>), dim3(col_grid<CW, 1>(gr.g)), dim3(Col2D<R, N, E, CW>::BS), 0,
                         st, gr.field<R>(F_W), gr.fs(), gr.field<R>(F_NA), gr.field<R>(F_WS), gr.field<R>(F_ACC),
                         gr.field<R>(F_S), half ? mode : (int)TURB_HALF, tp, gr.g, tw, 1);
      return check_launch("yfwd2d_turb");
    }));
  }
  return SCL_OK;
}

int turb_mode(int64_t step) { return step == 0 ? TURB_STEP0 : step == 1 ? TURB_STEP1 : TURB_AB2; }

// nsteps steps of a group as one stage chain: the first stage runs its own
// y-inverse, every later one gets it from the previous stage's fused
// y-forward, and the last stage is not fused (the chain leaves no state in
// W).  The kx = 0 column is made Hermitian in the same workgroup, so the
// chain gives the same results as separate stages.
template <class R>
int32_t group_steps(const scl_ens2d& e, const Group& gr, hipStream_t st, int nsteps, int64_t first) {
  if (e.p.scheme == 1) {
    for (int s = 0; s < nsteps; ++s) SCL_CHECK(turb_stage<R>(e, gr, turb_mode(first + s), st));
    return SCL_OK;
  }
  for (int s = 0; s < nsteps; ++s)
    for (int mode = 1; mode <= 4; ++mode)
      SCL_CHECK(stage<R>(e, gr, mode, st, s > 0 || mode > 1, s < nsteps - 1 || mode < 4));
  return SCL_OK;
}

// nsteps steps of every group, each group's nsteps steps as one chain, the
// groups one after the other on the ensemble's stream.  With fork_groups
// (batched mode) each group runs on its own branch instead (fork/join
// through events, which graph capture turns into graph edges) and the
// branches join once, at the end.
// first: the index of the first step (scheme 1 bootstraps on steps 0, 1).
int32_t launch_steps(scl_ens2d& e, int nsteps, int64_t first) {
  const bool fork = e.fork_groups && !e.p.per_member && e.groups.size() > 1;
  if (fork) SCL_HIP(hipEventRecord(e.fork, e.stream));
  for (auto& gr : e.groups) {
    hipStream_t st = fork ? gr.stream : e.stream;
    if (fork) SCL_HIP(hipStreamWaitEvent(st, e.fork, 0));
    SCL_CHECK(e.prec == SCL_SP ? group_steps<float>(e, gr, st, nsteps, first)
                               : group_steps<double>(e, gr, st, nsteps, first));
    if (fork) {
      SCL_HIP(hipEventRecord(gr.done, st));
      SCL_HIP(hipStreamWaitEvent(e.stream, gr.done, 0));
    }
  }
  return SCL_OK;
}

void drop_graph(scl_ens2d& e) {
  for (int i = 0; i < 2; ++i) {
    if (e.exec[i]) (void)hipGraphExecDestroy(e.exec[i]);
    if (e.graph[i]) (void)hipGraphDestroy(e.graph[i]);
    e.exec[i] = nullptr;
    e.graph[i] = nullptr;
  }
}

// Graph i of e (0: one step, 1: a chunk of e.chunk steps), built on first
// use; scheme 1 graphs hold IF-AB2 steps only.
int32_t graph_exec(scl_ens2d& e, int i, hipGraphExec_t* out) {
  if (!e.exec[i]) {
    SCL_HIP(hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal));
    const int32_t st = launch_steps(e, i ? e.chunk : 1, 2);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(e.stream, &g);
    if (st != SCL_OK) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    SCL_HIP(ce);
    e.graph[i] = g;
    SCL_HIP(hipGraphInstantiate(&e.exec[i], e.graph[i], nullptr, nullptr, 0));
  }
  *out = e.exec[i];
  return SCL_OK;
}

void advance_clock(scl_ens2d& e, int nsteps) {
  for (int s = 0; s < nsteps; ++s, ++e.steps)
    e.time += (e.p.scheme == 1 && e.steps < 2) ? 0.5 * e.p.dt : e.p.dt;  // turb_2d.f90:457,467,481
}

int32_t run_steps(scl_ens2d& e, int nsteps) {
  if (nsteps < 0) return set_error("nsteps < 0"), SCL_EINVAL;
  // scheme 1: the two bootstrap steps run eagerly
  for (; nsteps > 0 && e.p.scheme == 1 && e.steps < 2; --nsteps) {
    SCL_CHECK(launch_steps(e, 1, e.steps));
    advance_clock(e, 1);
  }
  int done = 0;
  if (e.p.use_graph) {
    hipGraphExec_t x;
    if (e.chunk > 1 && nsteps >= e.chunk) {
      SCL_CHECK(graph_exec(e, 1, &x));
      for (; done + e.chunk <= nsteps; done += e.chunk) SCL_HIP(hipGraphLaunch(x, e.stream));
    }
    if (done < nsteps) {
      SCL_CHECK(graph_exec(e, 0, &x));
      for (; done < nsteps; ++done) SCL_HIP(hipGraphLaunch(x, e.stream));
    }
  } else {
    SCL_CHECK(launch_steps(e, nsteps, e.steps));
    done = nsteps;
  }
  advance_clock(e, done);
  return SCL_OK;
}

// (user group, member) -> (internal group, member index in it)
int32_t locate(const scl_ens2d& e, int ug, int j, const Group** gr, int* b) {
  if (ug < 0 || ug >= (int)e.sizes.size()) return set_error("group %d out of range", ug), SCL_EINVAL;
  if (j < 0 || j >= e.members[ug]) return set_error("member %d out of range for group %d", j, ug), SCL_EINVAL;
  if (e.p.per_member) {
    *gr = &e.groups[e.first[ug] + j];
    *b = 0;
  } else {
    *gr = &e.groups[e.first[ug]];
    *b = j;
  }
  return SCL_OK;
}

// Host reference layout (lx1, ly) <-> split layout of member b, field f.
int32_t copy_member(const scl_ens2d& e, const Group& gr, int b, int f, void* host, bool to_device) {
  const size_t es = e.esize();
  const int n = gr.n, mx = n / 2, lx1 = mx + 1;
  char* dbase = reinterpret_cast<char*>(gr.base) + (size_t)f * gr.fs() * es;
  char* dmain = dbase + (size_t)b * n * mx * es;
  char* dnyq = dbase + ((size_t)gr.g.nmain + (size_t)b * n) * es;
  char* h = reinterpret_cast<char*>(host);
  if (to_device) {
    SCL_HIP(hipMemcpy2DAsync(dmain, mx * es, h, lx1 * es, mx * es, n, hipMemcpyHostToDevice, e.stream));
    SCL_HIP(hipMemcpy2DAsync(dnyq, es, h + mx * es, lx1 * es, es, n, hipMemcpyHostToDevice, e.stream));
  } else {
    SCL_HIP(hipMemcpy2DAsync(h, lx1 * es, dmain, mx * es, mx * es, n, hipMemcpyDeviceToHost, e.stream));
    SCL_HIP(hipMemcpy2DAsync(h + mx * es, lx1 * es, dnyq, es, es, n, hipMemcpyDeviceToHost, e.stream));
  }
  SCL_HIP(hipStreamSynchronize(e.stream));
  return SCL_OK;
}

// turb_2d.f90's wz(my, mx2) (C = 2 conj(D), see Turb2DP) <-> the host
// reference layout (lx1, ly) of the normalised coefficients D.  On the way
// in, symmetrize (turb_2d.f90:569-590): row ky = -n/2 and k = 0 zeroed, the
// kx = 0 column averaged with its mirror for 1 <= y <= n/2 - 2.
template <class R>
void turb_in(const cpx<R>* wz, int n, std::vector<cpx<R>>& h) {
  const int mx2 = n / 2, lx1 = mx2 + 1;
  h.assign((size_t)lx1 * n, mk<R>(0, 0));
  for (int x = 0; x < mx2; ++x)
    for (int y = 0; y < n; ++y) h[(size_t)y * lx1 + x] = (R)0.5 * conj(wz[(size_t)x * n + y]);
  for (int x = 0; x < lx1; ++x) h[(size_t)(n / 2) * lx1 + x] = mk<R>(0, 0);
  h[0] = mk<R>(0, 0);
  for (int y = 1; y <= n / 2 - 2; ++y)
    h[(size_t)y * lx1] = (R)0.5 * (h[(size_t)y * lx1] + conj(h[(size_t)(n - y) * lx1]));
  for (int y = 1; y <= n / 2 - 2; ++y) h[(size_t)(n - y) * lx1] = conj(h[(size_t)y * lx1]);
}

template <class R>
void turb_out(const std::vector<cpx<R>>& h, int n, cpx<R>* wz) {
  const int mx2 = n / 2, lx1 = mx2 + 1;
  for (int x = 0; x < mx2; ++x)
    for (int y = 0; y < n; ++y) wz[(size_t)x * n + y] = (R)2 * conj(h[(size_t)y * lx1 + x]);
}

template <class F>
int32_t with_prec(int prec, F&& f) {
  return prec == SCL_SP ? f(float{}) : f(double{});
}

// Member copies in the layout of the ensemble's scheme.
int32_t put_member(const scl_ens2d& e, const Group& gr, int b, int f, const void* w) {
  if (e.p.scheme != 1) return copy_member(e, gr, b, f, const_cast<void*>(w), true);
  return with_prec(e.prec, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    std::vector<cpx<R>> h;
    turb_in<R>(static_cast<const cpx<R>*>(w), gr.n, h);
    return copy_member(e, gr, b, f, h.data(), true);
  });
}

int32_t take_member(const scl_ens2d& e, const Group& gr, int b, int f, void* w) {
  if (e.p.scheme != 1) return copy_member(e, gr, b, f, w, false);
  return with_prec(e.prec, [&](auto zero) -> int32_t {
    using R = decltype(zero);
    std::vector<cpx<R>> h((size_t)(gr.n / 2 + 1) * gr.n);
    SCL_CHECK(copy_member(e, gr, b, f, h.data(), false));
    turb_out<R>(h, gr.n, static_cast<cpx<R>*>(w));
    return SCL_OK;
  });
}

// Seeded vorticity of one member, computed in FP64 (reference layout).
void initialize2d(std::vector<double>& w, int n, int32_t seed, double k0, double energy) {
  const int lx1 = n / 2 + 1, ly = n;
  w.assign(2 * (size_t)lx1 * ly, 0.0);
  const double pi = 3.1415926536;  // the reference's literal (constant_sp.f90:20)
  Ran1 ran(seed);
  for (int jj = 0; jj < ly; ++jj)
    for (int ii = 0; ii < lx1; ++ii) {
      const double t1 = ran.uniform<double>();
      double t2 = ran.uniform<double>();
      t2 = 2.0 * pi * t2;
      gasdev<double>(t1, t2, w[2 * ((size_t)jj * lx1 + ii)], w[2 * ((size_t)jj * lx1 + ii) + 1]);
    }
  auto at = [&](int x, int y) -> double* { return &w[2 * ((size_t)y * lx1 + x)]; };
  // symmetrize2d (symmetrize.f90:17-27)
  for (int x = 0; x < lx1; ++x) at(x, ly / 2)[0] = at(x, ly / 2)[1] = 0;
  for (int y = 0; y < ly; ++y) at(lx1 - 1, y)[0] = at(lx1 - 1, y)[1] = 0;
  at(0, 0)[0] = at(0, 0)[1] = 0;
  for (int y = ly / 2 + 1; y < ly; ++y) {
    at(0, y)[0] = at(0, ly - y)[0];
    at(0, y)[1] = -at(0, ly - y)[1];
  }
  const double kc = 2.0 * (n / 2) / 3.0, kc2 = kc * kc;
  double esum = 0.0;
  for (int y = 0; y < ly; ++y) {
    const int ky = kwave(y, ly);
    for (int x = 0; x < lx1; ++x) {
      const int ik2 = x * x + ky * ky;
      const double k2 = ik2 == 0 ? std::numeric_limits<double>::min() : (double)ik2;
      double* c = at(x, y);
      if (k2 >= kc2) {
        c[0] = c[1] = 0;
        continue;
      }
      const double k = std::sqrt(k2), s = std::exp(-k2 / (2 * k0 * k0));
      c[0] = c[0] * k * s;
      c[1] = c[1] * k * s;
      const double a = (c[0] * c[0] + c[1] * c[1]) / k2;
      esum += x == 0 ? 0.5 * a : a;
    }
  }
  const double f = esum > 0 ? std::sqrt(energy / esum) : 0.0;
  for (double& v : w) v *= f;
}

template <class R>
int32_t upload_member_f64(const scl_ens2d& e, const Group& gr, int b, const std::vector<double>& w) {
  std::vector<R> h(w.size());
  for (size_t i = 0; i < w.size(); ++i) h[i] = (R)w[i];
  return copy_member(e, gr, b, F_WS, h.data(), true);
}

void destroy_groups(scl_ens2d& e) {
  for (auto& gr : e.groups) {
    if (gr.base) (void)hipFree(gr.base);
    if (gr.stream) (void)hipStreamDestroy(gr.stream);
    if (gr.done) (void)hipEventDestroy(gr.done);
  }
  e.groups.clear();
  for (auto& t : e.tw) (void)hipFree(t.second);
  e.tw.clear();
}

template <class R>
int32_t upload_twiddles(int n, void** out) {
  std::vector<double> re, im;
  twiddles(n, re, im);
  std::vector<cpx<R>> h(n);
  for (int j = 0; j < n; ++j) h[j] = mk<R>((R)re[j], (R)im[j]);
  SCL_HIP(hipMalloc(out, n * sizeof(cpx<R>)));
  SCL_HIP(hipMemcpy(*out, h.data(), n * sizeof(cpx<R>), hipMemcpyHostToDevice));
  return SCL_OK;
}

}  // namespace

extern "C" int32_t scl_ens2d_create(int32_t prec, int32_t ngroups, const int32_t* sizes, const int32_t* members,
                                    const scl_ens2d_params* params, scl_ens2d** out) {
  if (!out || !sizes || !members || !params) return set_error("null argument"), SCL_EINVAL;
  *out = nullptr;
  if (prec != SCL_SP && prec != SCL_DP) return set_error("bad precision %d", prec), SCL_EINVAL;
  if (ngroups < 1) return set_error("ngroups must be >= 1"), SCL_EINVAL;
  for (int g = 0; g < ngroups; ++g) {
    const int n = sizes[g];
    // n >= 32: the x-pass packs 1024/n rows per workgroup and needs n^2 >= 1024
    if (n < 32 || n > 1024 || (n & (n - 1)))
      return set_error("group %d: size %d must be a power of two in [32, 1024]", g, n), SCL_EINVAL;
    if (members[g] < 1) return set_error("group %d: members must be >= 1", g), SCL_EINVAL;
  }
  if (!(params->dt > 0)) return set_error("dt must be > 0"), SCL_EINVAL;
  if (params->n_hyper < 0 || params->n_hypo < 0) return set_error("negative viscosity order"), SCL_EINVAL;
  if (params->scheme != 0 && params->scheme != 1) return set_error("scheme %d: 0 or 1", params->scheme), SCL_EINVAL;
  if (params->scheme == 1 && !(std::isfinite(params->sigma) && std::isfinite(params->nu) && params->nu != 0 &&
                               std::isfinite(params->eta) && std::isfinite(params->kf)))
    return set_error("scheme 1: sigma, nu (nonzero), eta and kf must be finite"), SCL_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    (void)hipGetLastError();
    return set_error("no HIP device"), SCL_ENODEV;
  }
  auto* e = new scl_ens2d();
  e->prec = prec;
  e->p = *params;
  // per-member mode keeps one step per graph: a chunk would hold
  // chunk x members x 12 nodes
  e->chunk = params->per_member ? 1 : 10;
  if (const char* f = std::getenv("SCL_ENS2D_FORK")) e->fork_groups = std::atoi(f) != 0;
  auto fail = [&](int32_t st) {
    scl_ens2d_destroy(e);
    return st;
  };
  {
    int dev = 0, ncu = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0)
      e->ncu = ncu;
  }
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) != hipSuccess || hipEventCreate(&e->t0) != hipSuccess ||
      hipEventCreate(&e->t1) != hipSuccess)
    return set_error("stream/event creation failed"), fail(SCL_EHIP);
  const size_t es = e->esize();
  for (int ug = 0; ug < ngroups; ++ug) {
    const int n = sizes[ug];
    e->sizes.push_back(n);
    e->members.push_back(members[ug]);
    e->first.push_back((int)e->groups.size());
    const void* tw = nullptr;
    for (auto& t : e->tw)
      if (t.first == n) tw = t.second;
    if (!tw) {
      void* d = nullptr;
      const int32_t st = prec == SCL_SP ? upload_twiddles<float>(n, &d) : upload_twiddles<double>(n, &d);
      if (st != SCL_OK) return fail(st);
      e->tw.push_back({n, d});
      tw = d;
    }
    const int copies = params->per_member ? members[ug] : 1;
    const int nb = params->per_member ? 1 : members[ug];
    for (int c = 0; c < copies; ++c) {
      Group gr;
      gr.n = n;
      gr.nb = nb;
      gr.g = make_geom(n, n, nb);
      gr.tw = tw;
      gr.e = 8;  // at 1024 points: one 1024-thread workgroup per CU, 4 waves per SIMD
      gr.cw = prec == SCL_DP && n >= 1024 ? 4 : 8;  // FP64 1024: 4 columns (LDS)
      const size_t bytes = (size_t)(params->scheme == 1 ? NFIELDS_TURB : NFIELDS) * gr.fs() * es;
      if (hipMalloc(&gr.base, bytes) != hipSuccess) {
        (void)hipGetLastError();
        e->groups.push_back(gr);
        return set_error("device allocation of %zu bytes failed", bytes), fail(SCL_ENOMEM);
      }
      if (hipMemsetAsync(gr.base, 0, bytes, e->stream) != hipSuccess) {
        e->groups.push_back(gr);
        return set_error("memset failed"), fail(SCL_EHIP);
      }
      if (!params->per_member && ngroups > 1) {
        if (hipStreamCreateWithFlags(&gr.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&gr.done, hipEventDisableTiming) != hipSuccess) {
          e->groups.push_back(gr);
          return set_error("stream creation failed"), fail(SCL_EHIP);
        }
      }
      e->groups.push_back(gr);
    }
  }
  if (hipMalloc((void**)&e->diag, 2 * sizeof(double) * e->total_members()) != hipSuccess)
    return set_error("device allocation failed"), fail(SCL_ENOMEM);
  if (hipStreamSynchronize(e->stream) != hipSuccess) return set_error("stream sync failed"), fail(SCL_EHIP);
  *out = e;
  return SCL_OK;
}

extern "C" int32_t scl_ens2d_destroy(scl_ens2d* e) {
  if (!e) return SCL_OK;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  drop_graph(*e);
  destroy_groups(*e);
  if (e->diag) (void)hipFree(e->diag);
  if (e->fork) (void)hipEventDestroy(e->fork);
  if (e->t0) (void)hipEventDestroy(e->t0);
  if (e->t1) (void)hipEventDestroy(e->t1);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  return SCL_OK;
}

extern "C" int32_t scl_ens2d_set_member(scl_ens2d* e, int32_t group, int32_t member, const void* w) {
  if (!e || !w) return set_error("null argument"), SCL_EINVAL;
  const Group* gr;
  int b;
  SCL_CHECK(locate(*e, group, member, &gr, &b));
  SCL_CHECK(put_member(*e, *gr, b, F_WS, w));
  // scheme 1: a new state starts turb_2d.f90's bootstrap again (its restart
  // reads wz and runs istep = 0, 1 first)
  if (e->p.scheme == 1) e->steps = 0;
  return SCL_OK;
}

extern "C" int32_t scl_ens2d_get_member(scl_ens2d* e, int32_t group, int32_t member, void* w) {
  if (!e || !w) return set_error("null argument"), SCL_EINVAL;
  const Group* gr;
  int b;
  SCL_CHECK(locate(*e, group, member, &gr, &b));
  return take_member(*e, *gr, b, F_WS, w);
}

extern "C" int32_t scl_ens2d_initialize(scl_ens2d* e, int32_t seed0, double k0, double energy) {
  if (!e) return set_error("null argument"), SCL_EINVAL;
  if (!(k0 > 0) || !(energy > 0)) return set_error("k0 and energy must be > 0"), SCL_EINVAL;
  std::vector<double> w;
  int q = 0;
  for (int ug = 0; ug < (int)e->sizes.size(); ++ug)
    for (int j = 0; j < e->members[ug]; ++j, ++q) {
      initialize2d(w, e->sizes[ug], seed0 - q, k0, energy);
      const Group* gr;
      int b;
      SCL_CHECK(locate(*e, ug, j, &gr, &b));
      SCL_CHECK(e->prec == SCL_SP ? upload_member_f64<float>(*e, *gr, b, w) : upload_member_f64<double>(*e, *gr, b, w));
    }
  e->steps = 0;
  e->time = 0.0;
  return SCL_OK;
}

extern "C" int32_t scl_ens2d_rhs(scl_ens2d* e, int32_t group, int32_t member, const void* w, void* rhs) {
  if (!e || !w || !rhs) return set_error("null argument"), SCL_EINVAL;
  const Group* gr;
  int b;
  SCL_CHECK(locate(*e, group, member, &gr, &b));
  SCL_CHECK(put_member(*e, *gr, b, F_WS, w));
  if (e->p.scheme == 1) {
    e->steps = 0;
    SCL_CHECK(e->prec == SCL_SP ? turb_stage<float>(*e, *gr, TURB_RHS, e->stream)
                                : turb_stage<double>(*e, *gr, TURB_RHS, e->stream));
    return take_member(*e, *gr, b, F_NA, rhs);
  }
  SCL_CHECK(e->prec == SCL_SP ? stage<float>(*e, *gr, 0, e->stream) : stage<double>(*e, *gr, 0, e->stream));
  return copy_member(*e, *gr, b, F_W, rhs, false);
}

extern "C" int32_t scl_ens2d_step(scl_ens2d* e, int32_t nsteps) {
  if (!e) return set_error("null argument"), SCL_EINVAL;
  return run_steps(*e, nsteps);
}

extern "C" int32_t scl_ens2d_sync(scl_ens2d* e) {
  if (!e) return set_error("null argument"), SCL_EINVAL;
  SCL_HIP(hipStreamSynchronize(e->stream));
  return SCL_OK;
}

extern "C" int32_t scl_ens2d_step_timed(scl_ens2d* e, int32_t nsteps, double* ms) {
  if (!e || !ms) return set_error("null argument"), SCL_EINVAL;
  SCL_HIP(hipStreamSynchronize(e->stream));
  SCL_HIP(hipEventRecord(e->t0, e->stream));
  SCL_CHECK(run_steps(*e, nsteps));
  SCL_HIP(hipEventRecord(e->t1, e->stream));
  SCL_HIP(hipEventSynchronize(e->t1));
  float t = 0;
  SCL_HIP(hipEventElapsedTime(&t, e->t0, e->t1));
  *ms = t;
  return SCL_OK;
}

extern "C" int32_t scl_ens2d_time(scl_ens2d* e, double* time) {
  if (!e || !time) return set_error("null argument"), SCL_EINVAL;
  *time = e->time;
  return SCL_OK;
}

extern "C" int32_t scl_ens2d_diagnostics(scl_ens2d* e, double* out) {
  if (!e || !out) return set_error("null argument"), SCL_EINVAL;
  int q = 0;
  for (int ug = 0; ug < (int)e->sizes.size(); ++ug) {
    const int ng = e->p.per_member ? e->members[ug] : 1;
    for (int c = 0; c < ng; ++c) {
      const Group& gr = e->groups[e->first[ug] + c];
      if (e->prec == SCL_SP)
        hipLaunchKernelGGL((k_ens_diag<float>), dim3(gr.nb), dim3(256), 0, e->stream,
                           gr.field<float>(F_WS), gr.g, e->diag + 2 * q);
      else
        hipLaunchKernelGGL((k_ens_diag<double>), dim3(gr.nb), dim3(256), 0, e->stream,
                           gr.field<double>(F_WS), gr.g, e->diag + 2 * q);
      SCL_CHECK(check_launch("ens_diag"));
      q += gr.nb;
    }
  }
  SCL_HIP(hipMemcpyAsync(out, e->diag, 2 * sizeof(double) * q, hipMemcpyDeviceToHost, e->stream));
  SCL_HIP(hipStreamSynchronize(e->stream));
  return SCL_OK;
}

static int32_t apply_schedule(scl_ens2d* e, int ug, int ee, int cw, int row_grid) {
  const int n = e->sizes[ug];
  const bool ok = e->prec == SCL_SP ? schedule_valid<float>(n, ee, cw) : schedule_valid<double>(n, ee, cw);
  if (!ok) return SCL_EINVAL;
  const int ng = e->p.per_member ? e->members[ug] : 1;
  for (int c = 0; c < ng; ++c) {
    Group& gr = e->groups[e->first[ug] + c];
    gr.e = ee;
    gr.cw = cw;
    gr.row_grid = row_grid ? 1 : 0;
  }
  drop_graph(*e);
  return SCL_OK;
}

extern "C" int32_t scl_ens2d_set_schedule(scl_ens2d* e, int32_t group, int32_t ee, int32_t cw, int32_t row_grid) {
  if (!e) return set_error("null argument"), SCL_EINVAL;
  if (group < 0 || group >= (int)e->sizes.size()) return set_error("group %d out of range", group), SCL_EINVAL;
  SCL_HIP(hipStreamSynchronize(e->stream));
  return apply_schedule(e, group, ee, cw, row_grid);
}

extern "C" int32_t scl_ens2d_get_schedule(scl_ens2d* e, int32_t group, int32_t* ee, int32_t* cw, int32_t* row_grid) {
  if (!e || !ee || !cw || !row_grid) return set_error("null argument"), SCL_EINVAL;
  if (group < 0 || group >= (int)e->sizes.size()) return set_error("group %d out of range", group), SCL_EINVAL;
  const Group& gr = e->groups[e->first[group]];
  *ee = gr.e;
  *cw = gr.cw;
  *row_grid = gr.row_grid;
  return SCL_OK;
}

// The tuner times whole steps (the chain of fused stages the runs use; an
// RHS alone ranks the schedules differently at 1024 points), so it saves
// the fields a step changes for the next one (the state, and for scheme 1
// the previous nonlinear term) and puts them back after each candidate.
extern "C" int32_t scl_ens2d_tune(scl_ens2d* e, int32_t reps, double* best_ms) {
  if (!e) return set_error("null argument"), SCL_EINVAL;
  if (reps < 1) return set_error("reps must be >= 1"), SCL_EINVAL;
  SCL_HIP(hipStreamSynchronize(e->stream));
  const size_t es = e->esize();
  size_t save_bytes = 0;
  for (const Group& gr : e->groups) save_bytes = std::max(save_bytes, 2 * gr.fs() * es);
  void* save = nullptr;
  if (hipMalloc(&save, save_bytes) != hipSuccess) {
    (void)hipGetLastError();
    return set_error("tuner: device allocation of %zu bytes failed", save_bytes), SCL_ENOMEM;
  }
  // fields 0: state, 1: previous nonlinear term (scheme 1)
  auto copy = [&](const Group& gr, bool to_save) -> int32_t {
    const size_t fb = gr.fs() * es;
    char* b = static_cast<char*>(gr.base);
    char* sv = static_cast<char*>(save);
    const int fields[2] = {F_WS, F_ACC};
    for (int k = 0; k < (e->p.scheme == 1 ? 2 : 1); ++k) {
      char* dev = b + fields[k] * fb;
      SCL_HIP(hipMemcpyAsync(to_save ? sv + k * fb : dev, to_save ? dev : sv + k * fb, fb, hipMemcpyDeviceToDevice,
                             e->stream));
    }
    return SCL_OK;
  };
  const int64_t first = e->p.scheme == 1 ? std::max<int64_t>(e->steps, 2) : 0;
  int32_t st = SCL_OK;
  for (int ug = 0; ug < (int)e->sizes.size() && st == SCL_OK; ++ug) {
    const int n = e->sizes[ug];
    const int ng = e->p.per_member ? e->members[ug] : 1;
    double best = 1e300;
    int be = -1, bc = -1, br = -1;
    for (int ee : {8, 16})
      for (int cw : {4, 8, 16})
        for (int rg : {0, 1}) {
          if (st != SCL_OK || apply_schedule(e, ug, ee, cw, rg) != SCL_OK) continue;
          double per = 0.0;
          for (int c = 0; c < ng && st == SCL_OK; ++c) {
            const Group& gr = e->groups[e->first[ug] + c];
            auto steps = [&](int k) -> int32_t {
              return e->prec == SCL_SP ? group_steps<float>(*e, gr, e->stream, k, first)
                                       : group_steps<double>(*e, gr, e->stream, k, first);
            };
            float t = 0;
            auto timed = [&]() -> int32_t {
              SCL_CHECK(copy(gr, true));
              SCL_CHECK(steps(1));  // warm-up
              SCL_HIP(hipEventRecord(e->t0, e->stream));
              SCL_CHECK(steps(reps));
              SCL_HIP(hipEventRecord(e->t1, e->stream));
              SCL_CHECK(copy(gr, false));
              SCL_HIP(hipEventSynchronize(e->t1));
              SCL_HIP(hipEventElapsedTime(&t, e->t0, e->t1));
              return SCL_OK;
            };
            st = timed();
            per += (double)t / reps;
          }
          if (st == SCL_OK && per < best) {
            best = per;
            be = ee;
            bc = cw;
            br = rg;
          }
        }
    if (st != SCL_OK) break;
    if (be < 0) {
      st = (set_error("no valid schedule for size %d", n), SCL_EINVAL);
      break;
    }
    st = apply_schedule(e, ug, be, bc, br);
    if (best_ms) best_ms[ug] = best;
  }
  (void)hipStreamSynchronize(e->stream);
  (void)hipFree(save);
  return st;
}
