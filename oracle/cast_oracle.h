/*
 * cast_oracle.h — CPU restatement of the reference's CPU gate path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the
 * gfx950 backend and the timed CPU baseline of bench.py.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product (libcast_hip.so, cast_simulator_amd) never links or calls it.
 *
 * What it restates (reference files, read not compiled):
 *   orc_apply_gate_*      src/cpu/statevector.rs:469-516 (scalar bit-deposit
 *                         apply, f64 complex MACs in column order) with
 *                         pdep_usize :592-606.  Storage here is interleaved
 *                         (re, im) instead of the SIMD split layout; values
 *                         are identical because the scalar path only reads
 *                         and writes whole amplitudes through amp()/set_amp().
 *                         An F32 statevector is widened to f64 per read and
 *                         rounded once per write (CpuScalar::from_f64,
 *                         statevector.rs:11-31).
 *   orc_norm_squared_*    statevector.rs:438-448 (sum of squared scalars in f64)
 *   orc_scale_*           statevector.rs:450-457 (x = from_f64(x * factor))
 *   orc_simd_*            the CPU SIMD kernel family used as the timed
 *                         baseline: split re/im layout
 *                         (statevector.rs:172-186, 556-587), simd bits = the
 *                         lowest s non-target qubits (src/cpp/cpu/internal/
 *                         bit_layout.cpp:5-37), per task a K x 2^s tile
 *                         updated with reContrib = ReM*Rev - ImM*Imv
 *                         (src/cpp/cpu/internal/kernel_codegen.cpp:110-118),
 *                         task ranges split evenly over OS threads spawned per
 *                         apply (src/cpu/kernel.rs:371-432), runtime matrix
 *                         as in StackLoad mode (matrix_data.cpp:21-26).
 *
 * Parity pins: see tests/test_oracle.py (reference known answers from
 * src/cuda/tests.rs:79-187 and tests/cpu_cuda_compare.rs, brute-force
 * Kronecker expansion, DFT closed form for the QFT, norm known answer
 * from src/simulator/tests.rs:320-339).
 */
#ifndef CAST_ORACLE_H
#define CAST_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Bit-deposit of src into the set bits of mask (statevector.rs:592-606). */
uint64_t orc_pdep(uint64_t src, uint64_t mask);

/* Scalar apply.  sv: interleaved (re, im), 2 * 2^n scalars.  mat: row-major
 * K x K complex as 2*K*K doubles (re, im).  qubits: k ascending targets. */
void orc_apply_gate_f64(double *sv, unsigned n, const double *mat, const uint32_t *qubits,
                        unsigned k);
void orc_apply_gate_f32(float *sv, unsigned n, const double *mat, const uint32_t *qubits,
                        unsigned k);

double orc_norm_squared_f64(const double *x, size_t n_elements);
double orc_norm_squared_f32(const float *x, size_t n_elements);
void orc_scale_f64(double *x, size_t n_elements, double factor);
void orc_scale_f32(float *x, size_t n_elements, double factor);

/* <Z_a Z_b> = sum_j |amp_j|^2 * (-1)^(bit_a(j) xor bit_b(j)), in f64. */
void orc_expect_zz_f64(const double *sv, unsigned n, const uint32_t *pairs, size_t n_pairs,
                       double *out);

/* Marginal probabilities over measured qubits pos[0..m): out[b] (2^m entries,
 * zeroed here) += |amp_j|^2 at b = compress_bits(j, pos), visiting j in
 * ascending order (src/simulator/measure.rs:15-25, compress_bits
 * src/types/gate.rs:636-643). */
void orc_marginal_f64(const double *sv, unsigned n, const uint32_t *pos, unsigned m,
                      double *out);

/* Density-matrix populations: out[i] = Re of element i | (i << n) of the
 * 2n-qubit vectorised rho, i < 2^n (src/simulator/mod.rs:351-366). */
void orc_dm_populations_f64(const double *rho, unsigned n, double *out);

/* ── SIMD split-layout baseline ──────────────────────────────────────────── */

/* Scalars needed for an n-qubit statevector in split layout with 2^s lanes
 * (statevector.rs:556-561: 2^(max(n, s) + 1)). */
size_t orc_simd_len(unsigned n, unsigned s);
/* Interleaved <-> split layout conversion (precision of the split buffer =
 * f64 or f32; the interleaved side is always f64). */
void orc_simd_pack_f64(double *split, const double *inter, unsigned n, unsigned s);
void orc_simd_unpack_f64(double *inter, const double *split, unsigned n, unsigned s);
void orc_simd_pack_f32(float *split, const double *inter, unsigned n, unsigned s);
void orc_simd_unpack_f32(double *inter, const float *split, unsigned n, unsigned s);
/* Apply one gate with n_threads OS threads (0 = all logical CPUs).
 * Requires n >= k + s (src/cpu/kernel.rs:381-383).  Returns 0 on success. */
int orc_simd_apply_f64(double *split, unsigned n, unsigned s, const double *mat,
                       const uint32_t *qubits, unsigned k, unsigned n_threads);
int orc_simd_apply_f32(float *split, unsigned n, unsigned s, const double *mat,
                       const uint32_t *qubits, unsigned k, unsigned n_threads);
/* Number of logical CPUs the process may run on. */
unsigned orc_num_cpus(void);

#ifdef __cplusplus
}
#endif

#endif /* CAST_ORACLE_H */
