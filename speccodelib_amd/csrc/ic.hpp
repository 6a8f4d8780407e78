// Seeded initial velocity field (host code).
//
// initialize.f90:373-463 (initialize3dvellowmemdealiased) fixes the draw
// order: modes in (kk, jj, ii) loop order, x fastest; per mode three complex
// Gaussians, each from two ran1 draws (t1, t2 = 2 pi t2) through gasdev
// (:507-513).  The projection onto k-normal planes uses the complex k.v of
// initialize3dvelhighmemdealiased (:328-332): the low-memory routine keeps
// only the real part of k.v (it is declared real, :389,418), which leaves
// the imaginary part divergent.  Then symmetrize (:429-431), the 2/3 cut
// (:433-437), the k^4 exp(-2k^2/akp^2) shape (:441) and the rescale to
// sum' |v|^2 = 1.5 u0^2 (:454-460), with that sum accumulated in FP64.
//
// ran1 (:516-548) is the Numerical Recipes generator: Park-Miller minimal
// standard LCG in Schrage form behind a 32-entry Bays-Durham shuffle.  Each
// call here starts from a fresh generator state (the reference keeps the
// state in SAVEd module variables); as in the reference, a positive seed on
// a fresh state is replaced by 1 (:529-530), so distinct fields need
// distinct negative seeds.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

#include "scl_core.hpp"

namespace scl {

struct Ran1 {
  int32_t idum, iy = 0;
  int32_t iv[32] = {0};
  explicit Ran1(int32_t seed) : idum(seed) {}
  int32_t next() {
    constexpr int32_t IA = 16807, IM = 2147483647, IQ = 127773, IR = 2836, NTAB = 32;
    constexpr int32_t NDIV = 1 + (IM - 1) / NTAB;
    if (idum <= 0 || iy == 0) {
      idum = (-idum > 1) ? -idum : 1;
      for (int j = NTAB + 8; j >= 1; --j) {
        const int32_t k = idum / IQ;
        idum = IA * (idum - k * IQ) - IR * k;
        if (idum < 0) idum += IM;
        if (j <= NTAB) iv[j - 1] = idum;
      }
      iy = iv[0];
    }
    const int32_t k = idum / IQ;
    idum = IA * (idum - k * IQ) - IR * k;
    if (idum < 0) idum += IM;
    const int32_t j = 1 + iy / NDIV;
    iy = iv[j - 1];
    iv[j - 1] = idum;
    return iy;
  }
  template <class R>
  R uniform() {  // ran1 = min(am*iy, rnmx)
    const R am = (R)1 / (R)2147483647, rnmx = (R)1 - (R)1.2e-7;
    const R u = am * (R)next();
    return u < rnmx ? u : rnmx;
  }
};

// gasdev(t1, t2) = sqrt(-2 log t1) * (cos t2, sin t2) / sqrt(2)
template <class R>
inline void gasdev(R t1, R t2, R& re, R& im) {
  const R r = std::sqrt((R)-2 * std::log(t1));
  const R s2 = std::sqrt((R)2);
  re = r * std::cos(t2) / s2;
  im = r * std::sin(t2) / s2;
}

// Hermitian copy and plane zeroing of symmetrize3d (symmetrize.f90:41-61)
// on a host array in the reference layout.
template <class R>
inline void symmetrize_host(R* c, int lx1, int ly, int lz) {
  auto at = [&](int x, int y, int z) -> R* { return c + 2 * (((int64_t)z * ly + y) * lx1 + x); };
  for (int z = 0; z < lz; ++z)
    for (int x = 0; x < lx1; ++x) at(x, ly / 2, z)[0] = at(x, ly / 2, z)[1] = 0;
  for (int y = 0; y < ly; ++y)
    for (int x = 0; x < lx1; ++x) at(x, y, lz / 2)[0] = at(x, y, lz / 2)[1] = 0;
  for (int z = 0; z < lz; ++z)
    for (int y = 0; y < ly; ++y) at(lx1 - 1, y, z)[0] = at(lx1 - 1, y, z)[1] = 0;
  at(0, 0, 0)[0] = at(0, 0, 0)[1] = 0;
  for (int y = 0; y < ly; ++y)
    for (int z = lz / 2 + 1; z < lz; ++z) {
      const R* s = at(0, (ly - y) % ly, lz - z);
      R* d = at(0, y, z);
      d[0] = s[0];
      d[1] = -s[1];
    }
  for (int y = ly / 2 + 1; y < ly; ++y) {
    const R* s = at(0, ly - y, 0);
    R* d = at(0, y, 0);
    d[0] = s[0];
    d[1] = -s[1];
  }
}

template <class R>
int32_t initialize3d(R* vx, R* vy, R* vz, int nx, int ny, int nz, int32_t iseed, double akp_d, double u0_d,
                     double kcut_d) {
  const int lx1 = nx / 2 + 1, ly = ny, lz = nz;
  const R pi = (R)3.1415926536;  // the reference's literal (constant_sp.f90:20)
  const R tiny = std::numeric_limits<R>::min();
  Ran1 ran(iseed);
  R* v[3] = {vx, vy, vz};
  const R kcut = kcut_d > 0 ? (R)kcut_d : (R)(2.0f * (float)(nx / 2) / 3.0f);
  const R kcut2 = kcut * kcut;
  for (int kk = 0; kk < lz; ++kk) {
    const R kz = (R)kwave(kk, lz);
    for (int jj = 0; jj < ly; ++jj) {
      const R ky = (R)kwave(jj, ly);
      for (int ii = 0; ii < lx1; ++ii) {
        const R kx = (R)ii;
        const int ik2 = ii * ii + kwave(jj, ly) * kwave(jj, ly) + kwave(kk, lz) * kwave(kk, lz);
        const int64_t o = 2 * (((int64_t)kk * ly + jj) * lx1 + ii);
        if (ik2 > 0 && (R)ik2 >= kcut2) {
          // the 2/3 cut below zeroes this mode (and its kx = 0 mirror, at
          // the same |k|): draw its six numbers to keep the stream, skip
          // the Gaussians
          for (int q = 0; q < 6; ++q) ran.next();
          for (int q = 0; q < 3; ++q) v[q][o] = v[q][o + 1] = 0;
          continue;
        }
        R g[3][2];
        for (int q = 0; q < 3; ++q) {
          const R t1 = ran.uniform<R>();
          R t2 = ran.uniform<R>();
          t2 = (R)2 * pi * t2;
          gasdev<R>(t1, t2, g[q][0], g[q][1]);
        }
        const R den = kx * kx + ky * ky + kz * kz + tiny;
        const R tr = (kx * g[0][0] + ky * g[1][0] + kz * g[2][0]) / den;
        const R ti = (kx * g[0][1] + ky * g[1][1] + kz * g[2][1]) / den;
        const R k[3] = {kx, ky, kz};
        for (int q = 0; q < 3; ++q) {
          v[q][o] = g[q][0] - k[q] * tr;
          v[q][o + 1] = g[q][1] - k[q] * ti;
        }
      }
    }
  }
  for (int q = 0; q < 3; ++q) symmetrize_host<R>(v[q], lx1, ly, lz);
  const R akp = (R)akp_d, u0 = (R)u0_d;
  const R pref = u0 * std::sqrt((R)8 * std::sqrt((R)2 / pi) / ((R)3 * pi * akp * akp * akp * akp * akp));
  double esum = 0.0;
  for (int kk = 0; kk < lz; ++kk) {
    const int kz = kwave(kk, lz);
    for (int jj = 0; jj < ly; ++jj) {
      const int ky = kwave(jj, ly);
      for (int ii = 0; ii < lx1; ++ii) {
        const int ik2 = ii * ii + ky * ky + kz * kz;
        const R k2 = ik2 == 0 ? tiny : (R)ik2;
        const int64_t o = 2 * (((int64_t)kk * ly + jj) * lx1 + ii);
        R amp = pref * std::sqrt(k2) * std::exp(-k2 / (akp * akp));
        if (k2 >= kcut2) amp = 0;
        double e = 0.0;
        for (int q = 0; q < 3; ++q) {
          v[q][o] = amp == 0 ? (R)0 : v[q][o] * amp;
          v[q][o + 1] = amp == 0 ? (R)0 : v[q][o + 1] * amp;
          e += (double)v[q][o] * v[q][o] + (double)v[q][o + 1] * v[q][o + 1];
        }
        esum += ii == 0 ? 0.5 * e : e;
      }
    }
  }
  if (!(esum > 0)) {
    set_error("initial field has zero energy (grid too small for kcut/akp)");
    return SCL_EINVAL;
  }
  const R s = std::sqrt((R)1.5 * u0 * u0 / (R)esum);
  const int64_t n = 2 * (int64_t)lx1 * ly * lz;
  for (int q = 0; q < 3; ++q)
    for (int64_t i = 0; i < n; ++i) v[q][i] = s * v[q][i];
  return SCL_OK;
}

// Helical initial field of initialize_heli.f90:1-109 (spectral_heli_dns.f90's
// start).  Per mode, in (kk, jj, ii) order with x fastest, two complex
// Gaussians am, then ap, from ran1 pairs (:22-35); am rescaled on the shells
// |k - ii| < 1/2 by sqrt((1 - hrel) P/M) with P, M the sums of |ap|^2, |am|^2
// over |k - ii - oneless/2| < 1/2 (:37-44, the two masks as written there;
// the sums in FP64); u = sqrt(3/2) (ap h+ + am h-) A(k), A(k) = u0 sqrt(8
// sqrt(2/pi)/(3 pi akp^5)) k exp(-k^2/akp^2), with the helical basis h+ of
// :46-88 (i k x h+ = |k| h+, h- = conj h+; its kx-axis form for ky = kz = 0)
// and h- = conj(h+); then symmetrize and the rescale to sum' |u|^2 =
// 1.5 u0^2 (:90-101).  At k = 0 the file's basis is 0/0; the mean is zero.
template <class R>
int32_t initialize_heli3d(R* ux, R* uy, R* uz, int nx, int ny, int nz, int32_t iseed, double hrel_d, double akp_d,
                          double u0_d) {
  const int lx1 = nx / 2 + 1, ly = ny, lz = nz;
  const int64_t nm = (int64_t)lx1 * ly * lz;
  const R pi = (R)3.1415926536;  // the reference's literal (constant_sp.f90:20)
  const R smallest = (R)2 * (R)std::numeric_limits<float>::epsilon();  // constant_sp.f90:7
  const R oneless = (R)1 - smallest;
  std::vector<R> am(2 * nm), ap(2 * nm), km(nm);
  Ran1 ran(iseed);
  for (int64_t o = 0; o < nm; ++o) {
    R t1 = ran.uniform<R>(), t2 = (R)2 * pi * ran.uniform<R>();
    gasdev<R>(t1, t2, am[2 * o], am[2 * o + 1]);
    t1 = ran.uniform<R>();
    t2 = (R)2 * pi * ran.uniform<R>();
    gasdev<R>(t1, t2, ap[2 * o], ap[2 * o + 1]);
  }
  for (int kk = 0; kk < lz; ++kk)
    for (int jj = 0; jj < ly; ++jj)
      for (int ii = 0; ii < lx1; ++ii) {
        const int ik2 = ii * ii + kwave(jj, ly) * kwave(jj, ly) + kwave(kk, lz) * kwave(kk, lz);
        km[((int64_t)kk * ly + jj) * lx1 + ii] = std::sqrt((R)ik2);
      }
  const R hrel = (R)hrel_d;
  for (int sh = 1; sh <= lx1 - 1; ++sh) {
    double P = 0.0, M = 0.0;
    for (int64_t o = 0; o < nm; ++o)
      if (std::fabs(km[o] - (R)sh - (R)0.5 * oneless) < (R)0.5) {
        P += (double)ap[2 * o] * ap[2 * o] + (double)ap[2 * o + 1] * ap[2 * o + 1];
        M += (double)am[2 * o] * am[2 * o] + (double)am[2 * o + 1] * am[2 * o + 1];
      }
    const R f = M > 0 ? (R)std::sqrt(((double)1 - (double)hrel) * P / M) : (R)0;
    for (int64_t o = 0; o < nm; ++o)
      if (std::fabs(km[o] - (R)sh) < (R)0.5) {
        am[2 * o] *= f;
        am[2 * o + 1] *= f;
      }
  }
  const R akp = (R)akp_d, u0 = (R)u0_d;
  const R pref = u0 * std::sqrt((R)8 * std::sqrt((R)2 / pi) / ((R)3 * pi * akp * akp * akp * akp * akp));
  const R h = std::sqrt((R)2) / (R)2, s32 = std::sqrt((R)3 / (R)2);
  R* u[3] = {ux, uy, uz};
  for (int kk = 0; kk < lz; ++kk) {
    const R kz = (R)kwave(kk, lz);
    for (int jj = 0; jj < ly; ++jj) {
      const R ky = (R)kwave(jj, ly);
      for (int ii = 0; ii < lx1; ++ii) {
        const R kx = (R)ii;
        const int64_t o = ((int64_t)kk * ly + jj) * lx1 + ii;
        const R k = km[o];
        if (o == 0) {
          for (int q = 0; q < 3; ++q) u[q][0] = u[q][1] = 0;
          continue;
        }
        const R a = pref * k * std::exp(-k * k / (akp * akp));
        // h+ = (re, im) per component
        R hp[3][2];
        if (jj == 0 && kk == 0) {  // on the kx axis (:50-58)
          const R k2d = std::max(std::sqrt(kx * kx + ky * ky), smallest);
          hp[0][0] = h * (-ky / k2d);
          hp[0][1] = h * (-kx * kz / k / k2d);
          hp[1][0] = h * (kx / k2d);
          hp[1][1] = h * (-ky * kz / k / k2d);
          hp[2][0] = 0;
          hp[2][1] = h * (ky * ky + kx * kx) / k / k2d;
        } else {  // (:70-78)
          const R k2d = std::sqrt(kz * kz + ky * ky);
          hp[0][0] = 0;
          hp[0][1] = h * (ky * ky + kz * kz) / k / k2d;
          hp[1][0] = h * (-kz / k2d);
          hp[1][1] = h * (-kx * ky / k / k2d);
          hp[2][0] = h * (ky / k2d);
          hp[2][1] = h * (-kz * kx / k / k2d);
        }
        const R apr = ap[2 * o], api = ap[2 * o + 1], amr = am[2 * o], ami = am[2 * o + 1];
        for (int q = 0; q < 3; ++q) {
          const R hr = hp[q][0], hi = hp[q][1];  // h- = (hr, -hi)
          const R re = (apr * hr - api * hi) + (amr * hr + ami * hi);
          const R im = (apr * hi + api * hr) + (ami * hr - amr * hi);
          u[q][2 * o] = s32 * re * a;
          u[q][2 * o + 1] = s32 * im * a;
        }
      }
    }
  }
  for (int q = 0; q < 3; ++q) symmetrize_host<R>(u[q], lx1, ly, lz);
  double esum = 0.0;
  for (int64_t o = 0; o < nm; ++o) {
    double e = 0.0;
    for (int q = 0; q < 3; ++q) e += (double)u[q][2 * o] * u[q][2 * o] + (double)u[q][2 * o + 1] * u[q][2 * o + 1];
    esum += (o % lx1) == 0 ? 0.5 * e : e;
  }
  if (!(esum > 0)) {
    set_error("helical initial field has zero energy");
    return SCL_EINVAL;
  }
  const R sc = std::sqrt((R)1.5 * u0 * u0 / (R)esum);
  for (int q = 0; q < 3; ++q)
    for (int64_t i = 0; i < 2 * nm; ++i) u[q][i] = sc * u[q][i];
  return SCL_OK;
}

}  // namespace scl
