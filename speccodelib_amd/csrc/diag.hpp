// FP64 diagnostics: shell spectra E(ll), H(ll) (spec.f90:66-89,
// helispec.f90:67-91), totals, and the moments of the velocity-derivative
// skewness (skewness.f90:1-48).  Every sum runs in an order fixed by the
// grid alone (no atomics), so a diagnostic of a given field is the same bits
// on every call.
#pragma once

#include <climits>

#include "kernels3d.hpp"

namespace scl {

int32_t check_launch(const char* what);

constexpr int DIAG_TOT = 8;     // E, sum k2 E, H, |k.v|^2, nonfinite, sum k2 H, spare
constexpr int DIAG_NT = 6;      // totals accumulated per thread
constexpr int DIAG_BLOCKS = 512;
constexpr int DIAG_WAVES = 4;   // wavefronts per workgroup (256 threads)

// Doubles of the per-workgroup partials of k_diag_partial for nbins shells.
inline int64_t diag_part_len(int nbins) { return (int64_t)DIAG_BLOCKS * (2 * nbins + DIAG_TOT); }
// LDS bytes of k_diag_partial.
inline size_t diag_lds_bytes(int nbins) { return (size_t)(DIAG_WAVES * 2 * nbins + 256 * DIAG_NT) * sizeof(double); }

// Shell sums by wavefront shuffles.  A wavefront walks whole rows (y, z),
// 64 consecutive kx per pass (kx = nx/2 is the row's nyq entry, x = mx).
// Along a row ll = floor(|k| + 1/2) never decreases with kx, so the entries
// of one shell in a pass are a contiguous run of lanes: a segmented
// inclusive scan (shuffles up by 1, 2, ..., 32, adding only from lanes of
// the same shell) leaves the run's sum in its last lane, which adds it to
// its wavefront's own LDS histogram.  Lanes of one instruction write
// distinct shells, passes follow each other, and the four histograms and
// the per-thread totals are then combined in a fixed order.
template <class R>
__global__ __launch_bounds__(256) void k_diag_partial(const cpx<R>* __restrict__ v, int64_t fstride, Geom g, int nbins,
                                                      double* __restrict__ part) {
  extern __shared__ double sh[];  // DIAG_WAVES x (2 nbins) histograms, then 256 x DIAG_NT totals
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  for (int i = tid; i < DIAG_WAVES * 2 * nbins; i += 256) sh[i] = 0.0;
  __syncthreads();
  double* hE = sh + wave * 2 * nbins;
  double* hH = hE + nbins;
  double tE = 0, tK = 0, tH = 0, tD = 0, tN = 0, tHK = 0;
  const int64_t wstride = (int64_t)gridDim.x * DIAG_WAVES;
  for (int64_t row = (int64_t)blockIdx.x * DIAG_WAVES + wave; row < g.nrow; row += wstride) {
    int y, z;
    g.yz(row, y, z);
    const int ky = kwave(y, g.ny), kz = kwave(z, g.nz);
    for (int x0 = 0; x0 <= g.mx; x0 += 64) {
      const int x = x0 + lane;
      double e = 0, h = 0;
      int ll = INT_MAX;  // lanes past the row end: a shell of their own, never written
      if (x <= g.mx) {
        const int64_t i = x < g.mx ? row * g.mx + x : g.nmain + row;
        const int kx = x;
        const double k2 = (double)kx * kx + (double)ky * ky + (double)kz * kz;
        const cpx<R> a = v[i], b = v[i + fstride], c = v[i + 2 * fstride];
        const double ar = a.x, ai = a.y, br = b.x, bi = b.y, cr = c.x, ci = c.y;
        const double w = x == 0 ? 0.5 : 1.0;
        e = w * (ar * ar + ai * ai + br * br + bi * bi + cr * cr + ci * ci);
        // w = i k x v; h = 2 Re(w . conj v) (helispec.f90:77-82)
        const double wxr = -(ky * ci - kz * bi), wxi = ky * cr - kz * br;
        const double wyr = -(kz * ai - kx * ci), wyi = kz * ar - kx * cr;
        const double wzr = -(kx * bi - ky * ai), wzi = kx * br - ky * ar;
        h = w * 2.0 * (wxr * ar + wxi * ai + wyr * br + wyi * bi + wzr * cr + wzi * ci);
        const double dr = kx * ar + ky * br + kz * cr, di = kx * ai + ky * bi + kz * ci;
        tE += e;
        tK += k2 * e;
        tH += h;
        tHK += k2 * h;
        tD += w * (dr * dr + di * di);
        if (!(isfinite(ar) && isfinite(ai) && isfinite(br) && isfinite(bi) && isfinite(cr) && isfinite(ci))) tN += 1.0;
        ll = (int)floor(sqrt(k2) + 0.5);
      }
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const double pe = __shfl_up(e, o), ph = __shfl_up(h, o);
        const int pl = __shfl_up(ll, o);
        if (lane >= o && pl == ll) {
          e += pe;
          h += ph;
        }
      }
      const int nl = __shfl_down(ll, 1);
      if ((lane == 63 || nl != ll) && ll >= 1 && ll <= nbins) {
        hE[ll - 1] += e;
        hH[ll - 1] += h;
      }
    }
  }
  double* tot = sh + DIAG_WAVES * 2 * nbins;
  tot[tid * DIAG_NT + 0] = tE;
  tot[tid * DIAG_NT + 1] = tK;
  tot[tid * DIAG_NT + 2] = tH;
  tot[tid * DIAG_NT + 3] = tD;
  tot[tid * DIAG_NT + 4] = tN;
  tot[tid * DIAG_NT + 5] = tHK;
  __syncthreads();
  const int stride = 2 * nbins + DIAG_TOT;
  double* out = part + (int64_t)blockIdx.x * stride;
  for (int j = tid; j < 2 * nbins; j += 256) {
    double s = 0;
    for (int wv = 0; wv < DIAG_WAVES; ++wv) s += sh[wv * 2 * nbins + j];
    out[j] = s;
  }
  if (tid < DIAG_NT) {
    double s = 0;
    for (int t = 0; t < 256; ++t) s += tot[t * DIAG_NT + tid];
    out[2 * nbins + tid] = s;
  }
  if (tid >= DIAG_NT && tid < DIAG_TOT) out[2 * nbins + tid] = 0.0;
}

// out[j] = sum over the nblocks partials (block order)
static __global__ __launch_bounds__(256) void k_diag_final(const double* __restrict__ part, int nblocks, int len,
                                                    double* __restrict__ out) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= len) return;
  double s = 0;
  for (int b = 0; b < nblocks; ++b) s += part[(int64_t)b * len + j];
  out[j] = s;
}

// Shell spectra and totals of the 3 fields at v into out[0 .. 2 nbins +
// DIAG_TOT): E(1..nbins), H(1..nbins), then the DIAG_TOT totals.  part holds
// diag_part_len(nbins) doubles, every one of which is written before it is
// read (nothing needs clearing).  nbins in [1, mx].
template <class R>
inline int32_t diag_launch(const Geom& g, const cpx<R>* v, int64_t fstride, int nbins, double* part, double* out,
                           hipStream_t s) {
  const int len = 2 * nbins + DIAG_TOT;
  hipLaunchKernelGGL((k_diag_partial<R>), dim3(DIAG_BLOCKS), dim3(256), diag_lds_bytes(nbins), s, v, fstride, g,
                     nbins, part);
  SCL_CHECK(check_launch("diag_partial"));
  hipLaunchKernelGGL(k_diag_final, dim3((len + 255) / 256), dim3(256), 0, s, part, DIAG_BLOCKS, len, out);
  return check_launch("diag_final");
}

// d = i kx v (the x derivative, skewness.f90:18-21) into f (split layout)
template <class R>
__global__ __launch_bounds__(256) void k_ddx_field(const cpx<R>* __restrict__ v, cpx<R>* __restrict__ f, Geom g) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.ntot) return;
  const R kx = i < g.nmain ? (R)(i % g.mx) : (R)g.mx;
  f[i] = mul_i(kx * v[i]);
}

// sum x^2 and sum x^3 over the packed physical reals of the main rows
// (skewness.f90:27-45), per-workgroup partials part[2b], part[2b + 1]
template <class R>
__global__ __launch_bounds__(256) void k_moments23(const cpx<R>* __restrict__ f, int64_t n, double* __restrict__ part) {
  double s2 = 0, s3 = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const cpx<R> a = f[i];
    const double x = a.x, y = a.y;
    s2 += x * x + y * y;
    s3 += x * x * x + y * y * y;
  }
  __shared__ double red[2][256];
  red[0][threadIdx.x] = s2;
  red[1][threadIdx.x] = s3;
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

// d = i k.v into f (split layout) for the physical divergence
template <class R>
__global__ __launch_bounds__(256) void k_div_field(const cpx<R>* __restrict__ v, int64_t fstride, cpx<R>* __restrict__ f,
                                                   Geom g) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.ntot) return;
  int x, y, z;
  split_xyz(g, i, x, y, z);
  const R kx = (R)x, ky = (R)kwave(y, g.ny), kz = (R)kwave(z, g.nz);
  const cpx<R> a = v[i], b = v[i + fstride], c = v[i + 2 * fstride];
  f[i] = mul_i(mk<R>(kx * a.x + ky * b.x + kz * c.x, kx * a.y + ky * b.y + kz * c.y));
}

// max |x| over the packed real rows (main part), as an order-free max of
// the float bit patterns of |x| (non-negative floats order like unsigned).
template <class R>
__global__ __launch_bounds__(256) void k_absmax(const cpx<R>* __restrict__ f, int64_t n, unsigned long long* out) {
  double m = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const cpx<R> a = f[i];
    // NaN counts as +inf (fmax would drop it)
    const double ax = fabs((double)a.x), ay = fabs((double)a.y);
    constexpr double big = std::numeric_limits<double>::max(), inf = std::numeric_limits<double>::infinity();
    m = fmax(m, fmax(ax <= big ? ax : inf, ay <= big ? ay : inf));
  }
  __shared__ double red[256];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, (unsigned long long)__double_as_longlong(red[0]));
}

}  // namespace scl
