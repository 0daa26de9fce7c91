/*
 * cast_oracle.c — CPU restatement of the reference CPU gate path.
 * TEST INFRASTRUCTURE ONLY (parity checker + timed CPU baseline); see
 * cast_oracle.h for what each function restates and where it is pinned.
 *
 * Built with -ffp-contract=off so the scalar path performs exactly the IEEE
 * operations of the reference's num-complex arithmetic (no fused
 * multiply-adds): (a+bi)(c+di) = (ac - bd) + (ad + bc)i, then acc += that.
 */
#define _GNU_SOURCE
#include "cast_oracle.h"

#include <pthread.h>
#include <sched.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

uint64_t orc_pdep(uint64_t src, uint64_t mask) {
  uint64_t out = 0;
  for (uint64_t m = mask; m != 0; m &= m - 1) {
    if (src & 1u)
      out |= m & (~m + 1);
    src >>= 1;
  }
  return out;
}

static uint64_t full_mask(unsigned n) { return n >= 64 ? ~(uint64_t)0 : (((uint64_t)1 << n) - 1); }

/* src/cpu/statevector.rs:469-516 */
void orc_apply_gate_f64(double *sv, unsigned n, const double *mat, const uint32_t *qubits,
                        unsigned k) {
  uint64_t tmask = 0;
  for (unsigned b = 0; b < k; ++b)
    tmask |= (uint64_t)1 << qubits[b];
  const uint64_t task_mask = full_mask(n) ^ tmask;
  const uint64_t n_tasks = (uint64_t)1 << (n - k);
  const unsigned K = 1u << k;
  uint64_t delta[64], idx[64];
  double nr[64], ni[64];
  for (unsigned i = 0; i < K; ++i)
    delta[i] = orc_pdep(i, tmask);
  for (uint64_t t = 0; t < n_tasks; ++t) {
    const uint64_t dep = orc_pdep(t, task_mask);
    for (unsigned i = 0; i < K; ++i)
      idx[i] = dep | delta[i];
    for (unsigned r = 0; r < K; ++r) {
      double ar = 0.0, ai = 0.0;
      for (unsigned c = 0; c < K; ++c) {
        const double a = mat[2 * (r * K + c)], b = mat[2 * (r * K + c) + 1];
        const double x = sv[2 * idx[c]], y = sv[2 * idx[c] + 1];
        const double pr = a * x - b * y;
        const double pi = a * y + b * x;
        ar = ar + pr;
        ai = ai + pi;
      }
      nr[r] = ar;
      ni[r] = ai;
    }
    for (unsigned r = 0; r < K; ++r) {
      sv[2 * idx[r]] = nr[r];
      sv[2 * idx[r] + 1] = ni[r];
    }
  }
}

void orc_apply_gate_f32(float *sv, unsigned n, const double *mat, const uint32_t *qubits,
                        unsigned k) {
  uint64_t tmask = 0;
  for (unsigned b = 0; b < k; ++b)
    tmask |= (uint64_t)1 << qubits[b];
  const uint64_t task_mask = full_mask(n) ^ tmask;
  const uint64_t n_tasks = (uint64_t)1 << (n - k);
  const unsigned K = 1u << k;
  uint64_t delta[64], idx[64];
  double nr[64], ni[64];
  for (unsigned i = 0; i < K; ++i)
    delta[i] = orc_pdep(i, tmask);
  for (uint64_t t = 0; t < n_tasks; ++t) {
    const uint64_t dep = orc_pdep(t, task_mask);
    for (unsigned i = 0; i < K; ++i)
      idx[i] = dep | delta[i];
    for (unsigned r = 0; r < K; ++r) {
      double ar = 0.0, ai = 0.0;
      for (unsigned c = 0; c < K; ++c) {
        const double a = mat[2 * (r * K + c)], b = mat[2 * (r * K + c) + 1];
        const double x = (double)sv[2 * idx[c]], y = (double)sv[2 * idx[c] + 1];
        const double pr = a * x - b * y;
        const double pi = a * y + b * x;
        ar = ar + pr;
        ai = ai + pi;
      }
      nr[r] = ar;
      ni[r] = ai;
    }
    for (unsigned r = 0; r < K; ++r) {
      sv[2 * idx[r]] = (float)nr[r];
      sv[2 * idx[r] + 1] = (float)ni[r];
    }
  }
}

/* statevector.rs:438-448 */
double orc_norm_squared_f64(const double *x, size_t n) {
  double s = 0.0;
  for (size_t i = 0; i < n; ++i)
    s = s + x[i] * x[i];
  return s;
}

double orc_norm_squared_f32(const float *x, size_t n) {
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) {
    const double v = (double)x[i];
    s = s + v * v;
  }
  return s;
}

/* statevector.rs:450-457 */
void orc_scale_f64(double *x, size_t n, double factor) {
  for (size_t i = 0; i < n; ++i)
    x[i] = x[i] * factor;
}

void orc_scale_f32(float *x, size_t n, double factor) {
  for (size_t i = 0; i < n; ++i)
    x[i] = (float)((double)x[i] * factor);
}

void orc_marginal_f64(const double *sv, unsigned n, const uint32_t *pos, unsigned m,
                      double *out) {
  const uint64_t N = (uint64_t)1 << n;
  for (uint64_t b = 0; b < ((uint64_t)1 << m); ++b)
    out[b] = 0.0;
  for (uint64_t j = 0; j < N; ++j) {
    uint64_t b = 0;
    for (unsigned t = 0; t < m; ++t)
      b |= ((j >> pos[t]) & 1u) << t;
    out[b] += sv[2 * j] * sv[2 * j] + sv[2 * j + 1] * sv[2 * j + 1];
  }
}

void orc_dm_populations_f64(const double *rho, unsigned n, double *out) {
  for (uint64_t i = 0; i < ((uint64_t)1 << n); ++i)
    out[i] = rho[2 * (i | (i << n))];
}

void orc_expect_zz_f64(const double *sv, unsigned n, const uint32_t *pairs, size_t n_pairs,
                       double *out) {
  const uint64_t N = (uint64_t)1 << n;
  for (size_t p = 0; p < n_pairs; ++p) {
    const unsigned a = pairs[2 * p], b = pairs[2 * p + 1];
    double s = 0.0;
    for (uint64_t j = 0; j < N; ++j) {
      const double pr = sv[2 * j] * sv[2 * j] + sv[2 * j + 1] * sv[2 * j + 1];
      s += (((j >> a) ^ (j >> b)) & 1u) ? -pr : pr;
    }
    out[p] = s;
  }
}
