/*
 * workload_rsa.c -- seeded random sequential addition (RSA) of spheres in a
 * box, used to synthesise the particle inputs of the benchmark and test
 * workloads (uniform random, non-overlapping).  Host code only; not part of
 * the time step.
 *
 * Spheres are placed in the order given.  Each sphere draws candidate centres
 * uniformly in [lo + r, hi - r] from a splitmix64 stream seeded by `seed`
 * until one overlaps no placed sphere (cell list, cell = 2 * r_max).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

static inline uint64_t splitmix64(uint64_t* s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static inline double u01(uint64_t* s) { return (double)(splitmix64(s) >> 11) * 0x1.0p-53; }

/* returns the number of spheres placed (n on success); pos is (3, n) */
int64_t mfixdem_rsa_place(int64_t n, const double* radius, const double* lo, const double* hi,
                          uint64_t seed, int64_t max_tries, double* pos) {
  double rmax = 0.0;
  for (int64_t i = 0; i < n; i++) rmax = radius[i] > rmax ? radius[i] : rmax;
  if (n == 0 || rmax <= 0.0) return 0;
  double cw = 2.0 * rmax;
  int nc[3];
  for (int c = 0; c < 3; c++) {
    nc[c] = (int)((hi[c] - lo[c]) / cw);
    if (nc[c] < 1) nc[c] = 1;
  }
  int64_t ncell = (int64_t)nc[0] * nc[1] * nc[2];
  int64_t* head = malloc(sizeof(int64_t) * ncell);
  int64_t* next = malloc(sizeof(int64_t) * n);
  if (!head || !next) { free(head); free(next); return -1; }
  for (int64_t c = 0; c < ncell; c++) head[c] = -1;
  uint64_t st = seed * 0x2545F4914F6CDD1Dull + 1;
  int64_t placed = 0;
  for (int64_t i = 0; i < n; i++) {
    double r = radius[i];
    int ok = 0;
    double x[3];
    for (int64_t t = 0; t < max_tries && !ok; t++) {
      for (int c = 0; c < 3; c++) x[c] = lo[c] + r + u01(&st) * (hi[c] - lo[c] - 2.0 * r);
      int ci[3];
      for (int c = 0; c < 3; c++) {
        ci[c] = (int)((x[c] - lo[c]) / cw);
        if (ci[c] >= nc[c]) ci[c] = nc[c] - 1;
        if (ci[c] < 0) ci[c] = 0;
      }
      ok = 1;
      for (int dz = -1; dz <= 1 && ok; dz++)
        for (int dy = -1; dy <= 1 && ok; dy++)
          for (int dx = -1; dx <= 1 && ok; dx++) {
            int a = ci[0] + dx, b = ci[1] + dy, cc = ci[2] + dz;
            if (a < 0 || b < 0 || cc < 0 || a >= nc[0] || b >= nc[1] || cc >= nc[2]) continue;
            for (int64_t j = head[((int64_t)cc * nc[1] + b) * nc[0] + a]; j >= 0; j = next[j]) {
              double d0 = x[0] - pos[j], d1 = x[1] - pos[n + j], d2 = x[2] - pos[2 * n + j];
              double rr = r + radius[j];
              if (d0 * d0 + d1 * d1 + d2 * d2 < rr * rr) { ok = 0; break; }
            }
          }
    }
    if (!ok) break;
    pos[i] = x[0];
    pos[n + i] = x[1];
    pos[2 * n + i] = x[2];
    int ci[3];
    for (int c = 0; c < 3; c++) {
      ci[c] = (int)((x[c] - lo[c]) / cw);
      if (ci[c] >= nc[c]) ci[c] = nc[c] - 1;
    }
    int64_t cell = ((int64_t)ci[2] * nc[1] + ci[1]) * nc[0] + ci[0];
    next[i] = head[cell];
    head[cell] = i;
    placed++;
  }
  free(head);
  free(next);
  return placed;
}
