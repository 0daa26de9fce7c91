/*
 * cast_oracle_simd.c — the timed CPU baseline: a restatement of the
 * reference's CPU SIMD kernel family (split re/im layout, simd lanes over the
 * lowest non-target qubits, task ranges split over OS threads).
 * TEST INFRASTRUCTURE ONLY; see cast_oracle.h.
 *
 * Built with -O3 and fp contraction allowed (the reference JIT compiles its
 * kernels under fast-math, src/cpp/cpu/cpu_gen.cpp:149), and with
 * target_clones so one binary uses AVX-512 where the host has it.
 */
#define _GNU_SOURCE
#include "cast_oracle.h"

#include <pthread.h>
#include <sched.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* ── SIMD split layout ──────────────────────────────────────────────────── */

static uint64_t insert_zero_to_bit(uint64_t x, unsigned bit) {
  const uint64_t lo = ((uint64_t)1 << bit) - 1;
  return (x & lo) + ((x & ~lo) << 1);
}

size_t orc_simd_len(unsigned n, unsigned s) { return (size_t)2 << (n > s ? n : s); }

void orc_simd_pack_f64(double *split, const double *inter, unsigned n, unsigned s) {
  const uint64_t N = (uint64_t)1 << n;
  memset(split, 0, orc_simd_len(n, s) * sizeof(double));
  for (uint64_t j = 0; j < N; ++j) {
    const uint64_t p = insert_zero_to_bit(j, s);
    split[p] = inter[2 * j];
    split[p | ((uint64_t)1 << s)] = inter[2 * j + 1];
  }
}

void orc_simd_unpack_f64(double *inter, const double *split, unsigned n, unsigned s) {
  const uint64_t N = (uint64_t)1 << n;
  for (uint64_t j = 0; j < N; ++j) {
    const uint64_t p = insert_zero_to_bit(j, s);
    inter[2 * j] = split[p];
    inter[2 * j + 1] = split[p | ((uint64_t)1 << s)];
  }
}

void orc_simd_pack_f32(float *split, const double *inter, unsigned n, unsigned s) {
  const uint64_t N = (uint64_t)1 << n;
  memset(split, 0, orc_simd_len(n, s) * sizeof(float));
  for (uint64_t j = 0; j < N; ++j) {
    const uint64_t p = insert_zero_to_bit(j, s);
    split[p] = (float)inter[2 * j];
    split[p | ((uint64_t)1 << s)] = (float)inter[2 * j + 1];
  }
}

void orc_simd_unpack_f32(double *inter, const float *split, unsigned n, unsigned s) {
  const uint64_t N = (uint64_t)1 << n;
  for (uint64_t j = 0; j < N; ++j) {
    const uint64_t p = insert_zero_to_bit(j, s);
    inter[2 * j] = (double)split[p];
    inter[2 * j + 1] = (double)split[p | ((uint64_t)1 << s)];
  }
}

unsigned orc_num_cpus(void) {
  cpu_set_t set;
  if (sched_getaffinity(0, sizeof(set), &set) == 0) {
    const int c = CPU_COUNT(&set);
    if (c > 0)
      return (unsigned)c;
  }
  const long n = sysconf(_SC_NPROCESSORS_ONLN);
  return n > 0 ? (unsigned)n : 1u;
}

/* Task geometry shared by the worker threads.  Amplitude bits split into
 * targets, simd bits (the lowest s non-target positions, bit_layout.cpp:5-37)
 * and task bits (the rest).  The split-layout position is linear in the
 * amplitude bits, so pos = pos(task) + pos(target state) + pos(lane). */
typedef struct {
  void *sv;
  unsigned n, s, k;
  const double *mat64;
  const float *mat32;
  uint32_t ins[64]; /* ascending positions of target and simd bits */
  unsigned n_ins;
  uint64_t pt[64];  /* split position of target state i   */
  uint64_t pl[16];  /* split position of lane l           */
  int contiguous;   /* pl[l] == l                          */
  uint64_t t0, t1;
} simd_job_t;

static uint64_t task_pos(const simd_job_t *j, uint64_t t) {
  for (unsigned b = 0; b < j->n_ins; ++b)
    t = insert_zero_to_bit(t, j->ins[b]);
  return insert_zero_to_bit(t, j->s);
}

#define CAST_ORC_CLONES __attribute__((target_clones("avx512f", "avx2", "default")))

CAST_ORC_CLONES static void simd_range_f64(const simd_job_t *j) {
  enum { S = 8 };
  const unsigned K = 1u << j->k;
  double *sv = (double *)j->sv;
  double re[64][S], im[64][S];
  for (uint64_t t = j->t0; t < j->t1; ++t) {
    const uint64_t base = task_pos(j, t);
    for (unsigned i = 0; i < K; ++i) {
      const double *p = sv + base + j->pt[i];
      if (j->contiguous) {
        for (unsigned l = 0; l < S; ++l) {
          re[i][l] = p[l];
          im[i][l] = p[l + S];
        }
      } else {
        for (unsigned l = 0; l < S; ++l) {
          re[i][l] = p[j->pl[l]];
          im[i][l] = p[j->pl[l] + S];
        }
      }
    }
    for (unsigned r = 0; r < K; ++r) {
      double ar[S], ai[S];
      for (unsigned l = 0; l < S; ++l)
        ar[l] = ai[l] = 0.0;
      for (unsigned c = 0; c < K; ++c) {
        const double mr = j->mat64[2 * (r * K + c)], mi = j->mat64[2 * (r * K + c) + 1];
        for (unsigned l = 0; l < S; ++l) {
          ar[l] += mr * re[c][l] - mi * im[c][l];
          ai[l] += mr * im[c][l] + mi * re[c][l];
        }
      }
      double *p = sv + base + j->pt[r];
      if (j->contiguous) {
        for (unsigned l = 0; l < S; ++l) {
          p[l] = ar[l];
          p[l + S] = ai[l];
        }
      } else {
        for (unsigned l = 0; l < S; ++l) {
          p[j->pl[l]] = ar[l];
          p[j->pl[l] + S] = ai[l];
        }
      }
    }
  }
}

CAST_ORC_CLONES static void simd_range_f32(const simd_job_t *j) {
  enum { S = 16 };
  const unsigned K = 1u << j->k;
  float *sv = (float *)j->sv;
  float re[64][S], im[64][S];
  for (uint64_t t = j->t0; t < j->t1; ++t) {
    const uint64_t base = task_pos(j, t);
    for (unsigned i = 0; i < K; ++i) {
      const float *p = sv + base + j->pt[i];
      if (j->contiguous) {
        for (unsigned l = 0; l < S; ++l) {
          re[i][l] = p[l];
          im[i][l] = p[l + S];
        }
      } else {
        for (unsigned l = 0; l < S; ++l) {
          re[i][l] = p[j->pl[l]];
          im[i][l] = p[j->pl[l] + S];
        }
      }
    }
    for (unsigned r = 0; r < K; ++r) {
      float ar[S], ai[S];
      for (unsigned l = 0; l < S; ++l)
        ar[l] = ai[l] = 0.0f;
      for (unsigned c = 0; c < K; ++c) {
        const float mr = j->mat32[2 * (r * K + c)], mi = j->mat32[2 * (r * K + c) + 1];
        for (unsigned l = 0; l < S; ++l) {
          ar[l] += mr * re[c][l] - mi * im[c][l];
          ai[l] += mr * im[c][l] + mi * re[c][l];
        }
      }
      float *p = sv + base + j->pt[r];
      if (j->contiguous) {
        for (unsigned l = 0; l < S; ++l) {
          p[l] = ar[l];
          p[l + S] = ai[l];
        }
      } else {
        for (unsigned l = 0; l < S; ++l) {
          p[j->pl[l]] = ar[l];
          p[j->pl[l] + S] = ai[l];
        }
      }
    }
  }
}

static void *simd_worker_f64(void *arg) {
  simd_range_f64((const simd_job_t *)arg);
  return NULL;
}

static void *simd_worker_f32(void *arg) {
  simd_range_f32((const simd_job_t *)arg);
  return NULL;
}

/* Split-layout scalar position of amplitude index a. */
static uint64_t split_pos(uint64_t a, unsigned s) { return insert_zero_to_bit(a, s); }

static int simd_apply(void *sv, int f32, unsigned n, unsigned s, const double *mat,
                      const uint32_t *qubits, unsigned k, unsigned n_threads) {
  const unsigned S = 1u << s;
  if ((f32 && S != 16) || (!f32 && S != 8))
    return 1; /* the baseline is built for 512-bit vectors: s = 3 (f64), 4 (f32) */
  if (k == 0 || k > 6 || n < k + s || n > 40)
    return 1;
  for (unsigned b = 1; b < k; ++b)
    if (qubits[b] <= qubits[b - 1])
      return 1;
  if (qubits[k - 1] >= n)
    return 1;

  simd_job_t g;
  memset(&g, 0, sizeof(g));
  g.sv = sv;
  g.n = n;
  g.s = s;
  g.k = k;
  uint64_t tmask = 0;
  for (unsigned b = 0; b < k; ++b)
    tmask |= (uint64_t)1 << qubits[b];
  /* simd bits: the lowest s non-target positions */
  uint64_t smask = 0;
  unsigned found = 0;
  for (unsigned q = 0; found < s; ++q)
    if (!((tmask >> q) & 1u)) {
      smask |= (uint64_t)1 << q;
      ++found;
    }
  const uint64_t ins_mask = tmask | smask;
  g.n_ins = 0;
  for (unsigned q = 0; q < 64; ++q)
    if ((ins_mask >> q) & 1u)
      g.ins[g.n_ins++] = q;
  /* task_pos() inserts the zeros of ins[] into the task index, then the
   * split layout's re/im bit at s; lanes and targets are added afterwards. */
  const unsigned K = 1u << k;
  for (unsigned i = 0; i < K; ++i)
    g.pt[i] = split_pos(orc_pdep(i, tmask), s);
  g.contiguous = 1;
  for (unsigned l = 0; l < S; ++l) {
    g.pl[l] = split_pos(orc_pdep(l, smask), s);
    if (g.pl[l] != l)
      g.contiguous = 0;
  }

  float *m32 = NULL;
  if (f32) {
    m32 = (float *)malloc(sizeof(float) * 2 * K * K);
    if (!m32)
      return 1;
    for (unsigned i = 0; i < 2 * K * K; ++i)
      m32[i] = (float)mat[i];
    g.mat32 = m32;
  } else {
    g.mat64 = mat;
  }

  const uint64_t n_tasks = (uint64_t)1 << (n - k - s);
  unsigned nt = n_threads ? n_threads : orc_num_cpus();
  if (nt > 256)
    nt = 256;
  if ((uint64_t)nt > n_tasks)
    nt = (unsigned)n_tasks;
  simd_job_t jobs[256];
  pthread_t th[256];
  int started[256];
  for (unsigned i = 0; i < nt; ++i) {
    jobs[i] = g;
    jobs[i].t0 = n_tasks * i / nt;
    jobs[i].t1 = n_tasks * (i + 1) / nt;
    started[i] = 0;
  }
  int rc = 0;
  for (unsigned i = 1; i < nt; ++i) {
    if (pthread_create(&th[i], NULL, f32 ? simd_worker_f32 : simd_worker_f64, &jobs[i]) == 0)
      started[i] = 1;
    else
      rc = 1;
  }
  if (f32)
    simd_range_f32(&jobs[0]);
  else
    simd_range_f64(&jobs[0]);
  for (unsigned i = 1; i < nt; ++i) {
    if (started[i])
      pthread_join(th[i], NULL);
    else if (f32)
      simd_range_f32(&jobs[i]); /* thread could not start: run inline */
    else
      simd_range_f64(&jobs[i]);
  }
  free(m32);
  (void)rc;
  return 0;
}

int orc_simd_apply_f64(double *split, unsigned n, unsigned s, const double *mat,
                       const uint32_t *qubits, unsigned k, unsigned n_threads) {
  return simd_apply(split, 0, n, s, mat, qubits, k, n_threads);
}

int orc_simd_apply_f32(float *split, unsigned n, unsigned s, const double *mat,
                       const uint32_t *qubits, unsigned k, unsigned n_threads) {
  return simd_apply(split, 1, n, s, mat, qubits, k, n_threads);
}
