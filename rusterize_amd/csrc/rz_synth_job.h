// rz_synth_job.h -- the seeded job generator of the batched workload
// (4,096 independent jobs), shared by the host generator (tests, oracle
// inputs) and the device generator (benchmarks): one source, so both produce
// the same coordinates bit for bit.  Only IEEE add / mul / div and exact
// scaling are used (no libm), and both sides compile with -ffp-contract=off.
//
// Job k of seed s:
//   side      = round(256 * 2^(5u))           square extent, 256..8192 px
//   n_poly    = round(10 * 2^(u * log2 5000))   10..50,000 polygons
//   fun       = floor(7u)                       one of the seven functions
// Polygon p of job k: star-shaped ring with
//   n_vert    = round(4 * 2^(10u))             4..4,096 vertices
//   centre    = U(0, side)^2, radius R = 2^(1 + u*L), L = floor(log2 side) - 4
//   vertex j  = centre + r_j (cos a_j, sin a_j), a_j = 2 pi (j + u_j) / n_vert,
//               r_j = R * (0.3 + 0.7 u'_j); the ring is closed (n_vert + 1
//               coordinates).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define RZ_HD __host__ __device__
#else
#define RZ_HD
#endif

namespace rzsynth {

RZ_HD inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Uniform [0, 1) from a 64-bit hash (53 bits).
RZ_HD inline double unit(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

// Stream value i of a (seed, a, b, c) key.
RZ_HD inline double draw(uint64_t seed, uint64_t a, uint64_t b, uint64_t c, uint64_t i) {
  return unit(mix64(mix64(mix64(mix64(seed ^ 0x5851F42D4C957F2Dull) + a) + b) + c + 0x632BE59BD9B4E019ull * i));
}

// 2^x for x in [0, 64): exact integer part, polynomial (Taylor, 13 terms) for
// the fraction.
RZ_HD inline double exp2_det(double x) {
  int n = (int)x;
  double f = (x - (double)n) * 0.6931471805599453;  // f * ln 2
  double term = 1.0, sum = 1.0;
  for (int k = 1; k <= 13; ++k) {
    term = term * f / (double)k;
    sum = sum + term;
  }
  double p = 1.0;
  for (int k = 0; k < n; ++k) p = p * 2.0;
  return sum * p;
}

// sin and cos of a in [0, 2 pi) by quadrant reduction and Taylor series.
RZ_HD inline void sincos_det(double a, double* s, double* c) {
  const double half_pi = 1.5707963267948966;
  int q = (int)(a / half_pi);
  if (q > 3) q = 3;
  const double r = a - (double)q * half_pi;  // [0, pi/2)
  const double r2 = r * r;
  double ts = r, ss = r, tc = 1.0, sc = 1.0;
  for (int k = 1; k <= 11; ++k) {
    ts = -ts * r2 / (double)((2 * k) * (2 * k + 1));
    ss = ss + ts;
    tc = -tc * r2 / (double)((2 * k - 1) * (2 * k));
    sc = sc + tc;
  }
  switch (q) {
    case 0: *s = ss; *c = sc; break;
    case 1: *s = sc; *c = -ss; break;
    case 2: *s = -ss; *c = -sc; break;
    default: *s = -sc; *c = ss; break;
  }
}

struct JobSpec {
  int32_t side, n_poly, fun;
};

RZ_HD inline JobSpec job_spec(uint64_t seed, uint64_t k) {
  JobSpec j;
  j.side = (int32_t)(256.0 * exp2_det(5.0 * draw(seed, 1, k, 0, 0)) + 0.5);
  j.n_poly = (int32_t)(10.0 * exp2_det(12.287712379549449 * draw(seed, 1, k, 0, 1)) + 0.5);
  j.fun = (int32_t)(7.0 * draw(seed, 1, k, 0, 2));
  if (j.fun > 6) j.fun = 6;
  return j;
}

struct PolySpec {
  int32_t n_vert;
  double cx, cy, R;
};

RZ_HD inline int32_t floor_log2(int32_t v) {
  int32_t l = 0;
  while ((v >> (l + 1)) > 0) ++l;
  return l;
}

RZ_HD inline PolySpec poly_spec(uint64_t seed, uint64_t k, uint64_t p, int32_t side) {
  PolySpec s;
  s.n_vert = (int32_t)(4.0 * exp2_det(10.0 * draw(seed, 2, k, p, 0)) + 0.5);
  s.cx = (double)side * draw(seed, 2, k, p, 1);
  s.cy = (double)side * draw(seed, 2, k, p, 2);
  int32_t L = floor_log2(side) - 4;
  if (L < 1) L = 1;
  s.R = exp2_det(1.0 + (double)L * draw(seed, 2, k, p, 3));
  return s;
}

// Vertex j (0 <= j <= n_vert; j == n_vert closes the ring) of polygon p.
RZ_HD inline void vertex(uint64_t seed, uint64_t k, uint64_t p, const PolySpec& s, int32_t j,
                         double* x, double* y) {
  if (j == s.n_vert) j = 0;
  const double a = 6.283185307179586 * ((double)j + draw(seed, 3, k, p, 2 * (uint64_t)j)) / (double)s.n_vert;
  const double r = s.R * (0.3 + 0.7 * draw(seed, 3, k, p, 2 * (uint64_t)j + 1));
  double sn, cs;
  sincos_det(a, &sn, &cs);
  *x = s.cx + r * cs;
  *y = s.cy + r * sn;
}

// Field value of polygon p: U[0, 1).
RZ_HD inline double field(uint64_t seed, uint64_t k, uint64_t p) { return draw(seed, 4, k, p, 0); }

}  // namespace rzsynth
