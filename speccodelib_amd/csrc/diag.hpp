// FP64 diagnostics: shell spectra E(ll), H(ll) (spec.f90:66-89,
// helispec.f90:67-91) and totals, as per-block partial sums combined in a
// fixed block order.  Within a block each wavefront owns its own LDS
// histogram.
#pragma once

#include "kernels3d.hpp"

namespace scl {

constexpr int DIAG_TOT = 8;  // E, sum k2 E, H, |k.v|^2, nonfinite, sum k2 H, spare
constexpr int DIAG_NT = 6;   // totals accumulated per thread
constexpr int DIAG_BLOCKS = 512;

template <class R>
__global__ __launch_bounds__(256) void k_diag_partial(const cpx<R>* __restrict__ v, int64_t fstride, Geom g, int nbins,
                                                      double* __restrict__ part) {
  extern __shared__ double sh[];  // 4 waves x (2 nbins) + 256 x DIAG_NT totals
  const int tid = threadIdx.x, wave = tid / 64;
  for (int i = tid; i < 4 * 2 * nbins; i += 256) sh[i] = 0.0;
  __syncthreads();
  double* hE = sh + wave * 2 * nbins;
  double* hH = hE + nbins;
  double tE = 0, tK = 0, tH = 0, tD = 0, tN = 0, tHK = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + tid; i < g.ntot; i += (int64_t)gridDim.x * 256) {
    int x, y, z;
    split_xyz(g, i, x, y, z);
    const int kx = x, ky = kwave(y, g.ny), kz = kwave(z, g.nz);
    const double k2 = (double)kx * kx + (double)ky * ky + (double)kz * kz;
    const cpx<R> a = v[i], b = v[i + fstride], c = v[i + 2 * fstride];
    const double ar = a.x, ai = a.y, br = b.x, bi = b.y, cr = c.x, ci = c.y;
    const double w = x == 0 ? 0.5 : 1.0;
    const double e = w * (ar * ar + ai * ai + br * br + bi * bi + cr * cr + ci * ci);
    // w = i k x v; h = 2 Re(w . conj v) (helispec.f90:77-82)
    const double wxr = -(ky * ci - kz * bi), wxi = ky * cr - kz * br;
    const double wyr = -(kz * ai - kx * ci), wyi = kz * ar - kx * cr;
    const double wzr = -(kx * bi - ky * ai), wzi = kx * br - ky * ar;
    const double h = w * 2.0 * (wxr * ar + wxi * ai + wyr * br + wyi * bi + wzr * cr + wzi * ci);
    const double dr = kx * ar + ky * br + kz * cr, di = kx * ai + ky * bi + kz * ci;
    tE += e;
    tK += k2 * e;
    tH += h;
    tHK += k2 * h;
    tD += w * (dr * dr + di * di);
    if (!(isfinite(ar) && isfinite(ai) && isfinite(br) && isfinite(bi) && isfinite(cr) && isfinite(ci))) tN += 1.0;
    const int ll = (int)floor(sqrt(k2) + 0.5);
    if (ll >= 1 && ll <= nbins) {
      atomicAdd(&hE[ll - 1], e);
      atomicAdd(&hH[ll - 1], h);
    }
  }
  double* tot = sh + 4 * 2 * nbins;
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
    for (int wv = 0; wv < 4; ++wv) s += sh[wv * 2 * nbins + j];
    out[j] = s;
  }
  if (tid < DIAG_NT) {
    double s = 0;
    for (int t = 0; t < 256; ++t) s += tot[t * DIAG_NT + tid];
    out[2 * nbins + tid] = s;
  }
}

__global__ __launch_bounds__(256) void k_diag_final(const double* __restrict__ part, int nblocks, int len,
                                                    double* __restrict__ out) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= len) return;
  double s = 0;
  for (int b = 0; b < nblocks; ++b) s += part[(int64_t)b * len + j];
  out[j] = s;
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
