/*
 * Oracle (TEST INFRASTRUCTURE ONLY): the sequential random stream that the
 * reference's initial-condition routines draw from, restated in C so that the
 * numpy oracle can build 64^3..128^3 initial fields in well under a second.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this file's library, as the checker.  The product never links it.
 *
 * ran1   : initialize.f90:516-548  (Numerical Recipes ran1: Park-Miller
 *          minimal standard LCG, Schrage factorisation, Bays-Durham shuffle
 *          with a 32-entry table; state is fresh per oracle call)
 * gasdev : initialize.f90:507-513  (sqrt(-2 log t1) (cos t2, sin t2)/sqrt 2)
 * draw order: initialize.f90:397-417 (kk, jj, ii loops; x fastest; per mode
 *          ran1 pairs for vx, vy, vz; t2 = 2*pi*ran1 with the reference's
 *          pi literal 3.1415926536).
 */
#include <math.h>
#include <stdint.h>

typedef struct {
  int32_t idum, iy;
  int32_t iv[32];
} ran1_state;

enum { IA = 16807, IM = 2147483647, IQ = 127773, IR = 2836, NTAB = 32 };
#define NDIV (1 + (IM - 1) / NTAB)

static void ran1_init(ran1_state *s, int32_t seed) {
  s->idum = seed;
  s->iy = 0;
  for (int j = 0; j < NTAB; ++j) s->iv[j] = 0;
}

/* Returns the integer iy that the reference scales by am = 1/im. */
static int32_t ran1_next(ran1_state *s) {
  if (s->idum <= 0 || s->iy == 0) {
    s->idum = (-s->idum > 1) ? -s->idum : 1;
    for (int j = NTAB + 8; j >= 1; --j) {
      int32_t k = s->idum / IQ;
      s->idum = IA * (s->idum - k * IQ) - IR * k;
      if (s->idum < 0) s->idum += IM;
      if (j <= NTAB) s->iv[j - 1] = s->idum;
    }
    s->iy = s->iv[0];
  }
  int32_t k = s->idum / IQ;
  s->idum = IA * (s->idum - k * IQ) - IR * k;
  if (s->idum < 0) s->idum += IM;
  int32_t j = 1 + s->iy / NDIV;
  s->iy = s->iv[j - 1];
  s->iv[j - 1] = s->idum;
  return s->iy;
}

/* Raw integer stream: out[i] = iy of the i-th ran1 call. */
void oracle_ran1_stream(int32_t seed, int64_t n, int32_t *out) {
  ran1_state s;
  ran1_init(&s, seed);
  for (int64_t i = 0; i < n; ++i) out[i] = ran1_next(&s);
}

/* Gaussian draws of initialize.f90:397-417 in FP64 (the constant_dp build):
 * g[3*m + c] = gasdev for component c of mode m, m in x-fastest order. */
void oracle_gauss3_f64(int32_t seed, int64_t nmodes, double *g /* 2*3*nmodes */) {
  const double am = 1.0 / (double)IM, rnmx = 1.0 - 1.2e-7, pi = 3.1415926536;
  ran1_state s;
  ran1_init(&s, seed);
  for (int64_t m = 0; m < nmodes; ++m) {
    for (int c = 0; c < 3; ++c) {
      double t1 = am * ran1_next(&s);
      if (t1 > rnmx) t1 = rnmx;
      double t2 = am * ran1_next(&s);
      if (t2 > rnmx) t2 = rnmx;
      t2 = 2.0 * pi * t2;
      double r = sqrt(-2.0 * log(t1));
      g[2 * (3 * m + c) + 0] = r * cos(t2) / sqrt(2.0);
      g[2 * (3 * m + c) + 1] = r * sin(t2) / sqrt(2.0);
    }
  }
}

/* Same draws in FP32 arithmetic (the constant_sp build). */
void oracle_gauss3_f32(int32_t seed, int64_t nmodes, float *g) {
  const float am = 1.0f / (float)IM, rnmx = 1.0f - 1.2e-7f, pi = 3.1415926536f;
  ran1_state s;
  ran1_init(&s, seed);
  for (int64_t m = 0; m < nmodes; ++m) {
    for (int c = 0; c < 3; ++c) {
      float t1 = am * (float)ran1_next(&s);
      if (t1 > rnmx) t1 = rnmx;
      float t2 = am * (float)ran1_next(&s);
      if (t2 > rnmx) t2 = rnmx;
      t2 = 2.0f * pi * t2;
      float r = sqrtf(-2.0f * logf(t1));
      g[2 * (3 * m + c) + 0] = r * cosf(t2) / sqrtf(2.0f);
      g[2 * (3 * m + c) + 1] = r * sinf(t2) / sqrtf(2.0f);
    }
  }
}
