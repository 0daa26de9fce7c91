#ifndef CAST_HIP_H
#define CAST_HIP_H

/*
 * cast_hip.h — C ABI of the MI355X (gfx950) gate-application backend.
 *
 * This header is the drop-in boundary that replaces the reference's CUDA
 * bridge (src/cpp/include/ffi_cuda.h + ffi_types.h) for the
 * "apply a fused k-qubit complex matrix to a statevector" path.  A Rust
 * binding (the equivalent of src/ffi/cuda.rs) binds exactly these symbols.
 *
 * Conventions (kept from the reference, src/cpp/cuda/ffi_cuda.cpp:1-13):
 *   - int-returning calls return 0 on success and 1 on failure;
 *     constructors return a null handle (or 0 device pointer) on failure.
 *   - Every call that can fail takes a caller-owned (err_buf, err_buf_len).
 *     On failure a NUL-terminated, truncated message is written there;
 *     on success err_buf[0] is set to '\0' (src/cpp/internal/err_buf.hpp:12-23).
 *   - No C++ exception ever crosses this boundary.
 *   - Device memory is a raw device address carried as uint64_t; streams,
 *     events and gate plans are opaque void* handles.
 *   - Precision is a uint8_t: 0 = F32, 1 = F64 (same values as
 *     cast_precision_t).
 *   - n_elements always counts SCALARS: a statevector of n qubits has
 *     n_elements = 2 * 2^n (interleaved re, im), as in
 *     src/cuda/statevector.rs:77.
 *   - The host side of sv_upload / sv_download is always f64 pairs; an F32
 *     statevector is narrowed / widened inside the library
 *     (src/cpp/cuda/cuda_exec.cpp:481-533).
 *
 * Replaced reference machinery: the per-gate LLVM IR -> PTX -> cubin ->
 * CUmodule -> CUfunction chain (cast_cuda_compile_gate_ptx,
 * cast_cuda_ptx_to_cubin, cast_cuda_cubin_load, cast_cuda_module_get_function,
 * ffi_cuda.h:27-62) becomes ONE host-side "gate plan": the matrix is
 * classified (zero / +-1 snapping with the reference's ztol/K, otol/K rule),
 * converted to device precision, uploaded once, and a pre-compiled gfx950
 * kernel variant is chosen.  No code generation happens at run time.
 */

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ── Shared types (src/cpp/include/ffi_types.h:13-21) ─────────────────────── */

typedef enum cast_precision_t {
  CAST_PRECISION_F32 = 0,
  CAST_PRECISION_F64 = 1,
} cast_precision_t;

typedef struct cast_complex64_t {
  double re;
  double im;
} cast_complex64_t;

/* ── Gate-plan spec (replaces cast_cuda_kernel_gen_spec_t, ffi_cuda.h:16-23) ─
 * ztol / otol keep the reference meaning (src/cuda/types.rs:12-56): a matrix
 * scalar with |v| < ztol/K becomes exactly 0 and one with |v -+ 1| < otol/K
 * becomes exactly +-1 (src/cpp/cuda/cuda_gen.cpp:67-112).  sm_major/sm_minor
 * and maxnreg have no meaning on gfx950; `variant` may force a kernel family
 * (CAST_HIP_VARIANT_*), 0 lets the plan builder choose. */
typedef struct cast_hip_kernel_gen_spec_t {
  cast_precision_t precision;
  double ztol;
  double otol;
  uint32_t variant;
  uint32_t flags; /* CAST_HIP_FLAG_* */
} cast_hip_kernel_gen_spec_t;

/* Visit wave tiles (and matrix-core tiles: the low 10 index bits) in
 * bit-reversed order of their low index bits.  Which order streams faster
 * depends on how the gate's target bits fall on the HBM channel/bank
 * mapping; the host autotuner tries both. */
#define CAST_HIP_FLAG_TILE_SWIZZLE 1u
/* Wave tiles of k = 4, 5 in the three-product form (3 K instead of 4 K FMAs
 * per output: Re = Mr vr - Mi vi, Im = (Mr + Mi)(vr + vi) - Mr vr - Mi vi),
 * for gates where the FMA rate, not HBM, limits the wave tiles. */
#define CAST_HIP_FLAG_TILE_GAUSS 2u
/* Visit wave tiles so that the blocks dispatched to each of the 8 XCDs
 * (block b runs on XCD b mod 8) cover one contiguous eighth of the tiles,
 * instead of every XCD interleaving with the others.  The matrix-core
 * kernels (grid-stride loops) take the same flag: each XCD's workgroups then
 * cover a contiguous eighth of every grid-stride step.  Faster for most
 * target sets and slower for some; the host autotuner tries it.  With
 * variant AUTO (no order flag) the library uses it for wave tiles, for F64
 * matrix-core launches and for F32 split-product launches from 32 qubits
 * on; an explicitly requested variant without this flag keeps launch
 * order. */
#define CAST_HIP_FLAG_TILE_XCD 4u
/* Dense k = 6 at F32 on the matrix cores: through bf16 split products
 * (every f32 operand split into three bf16 parts, six partial products per
 * real product on v_mfma_f32_16x16x32_bf16, f32 accumulation), or on the
 * f32 matrix cores (v_mfma_f32_16x16x4_f32).  At most one of the two; with
 * neither the library picks.  Ignored elsewhere. */
#define CAST_HIP_FLAG_MFMA_SPLIT 8u
#define CAST_HIP_FLAG_MFMA_F32 16u

/* AUTO: monomial when every row has <= 1 nonzero, else the dense family, whose
 * kernel (direct / wave tile / matrix core) is picked per launch from k, the
 * statevector size and the precision.  A forced variant falls back to the
 * direct kernel where its shape requirement is not met. */
#define CAST_HIP_VARIANT_AUTO 0u
#define CAST_HIP_VARIANT_DENSE 1u    /* dense, one combo per lane, K amplitudes in registers    */
#define CAST_HIP_VARIANT_MONOMIAL 2u /* <=1 nonzero per row (diag / permutation with phases)    */
#define CAST_HIP_VARIANT_TILE 3u     /* dense, coalesced 64*K wave tiles; low-order targets      */
                                     /* exchanged across lanes (k <= 5, n >= k + 6)            */
#define CAST_HIP_VARIANT_MFMA 4u     /* dense k = 4..6 on matrix cores, three real products     */

/* Read-only description of a built plan (no reference counterpart: the
 * reference exposes this through PTX text, src/cuda/kernel.rs:847-882). */
typedef struct cast_hip_gate_plan_info_t {
  uint32_t n_qubits;   /* k                                              */
  uint32_t precision;  /* 0 = F32, 1 = F64                               */
  uint32_t variant;    /* CAST_HIP_VARIANT_* actually chosen             */
  uint32_t nnz_rows;   /* rows that differ from the identity row         */
  uint64_t scalar_nnz; /* nonzero re/im scalars after snapping           */
  uint32_t qubits[8];  /* ascending target qubits                        */
} cast_hip_gate_plan_info_t;

/* Device description (replaces cast_cuda_device_sm, ffi_cuda.h:36). */
typedef struct cast_hip_device_info_t {
  char name[128];
  char arch[64];          /* e.g. "gfx950:sramecc+:xnack-"            */
  uint32_t n_cu;
  uint32_t wavefront_size;
  uint32_t lds_per_block; /* bytes                                     */
  uint32_t l2_bytes;
  uint32_t clock_khz;
  uint32_t mem_clock_khz;
  uint32_t mem_bus_width; /* bits                                      */
  uint32_t reserved;
  uint64_t total_mem;     /* bytes                                     */
} cast_hip_device_info_t;

/* ── Device queries (ffi_cuda.h:36-39) ───────────────────────────────────── */

int cast_hip_device_info(cast_hip_device_info_t *out, char *err_buf, size_t err_buf_len);
/* Replaces cast_cuda_free_memory (ffi_cuda.h:38, cuda_exec.cpp:133-153). */
int cast_hip_free_memory(uint64_t *out_free_bytes, uint64_t *out_total_bytes, char *err_buf,
                         size_t err_buf_len);

/* ── Streams (ffi_cuda.h:43-45) ──────────────────────────────────────────── */

void *cast_hip_stream_create(char *err_buf, size_t err_buf_len);
void cast_hip_stream_destroy(void *stream);
int cast_hip_stream_sync(void *stream, char *err_buf, size_t err_buf_len);

/* ── Events (ffi_cuda.h:66-71) ───────────────────────────────────────────── */

/* A stream that does not synchronise with the null stream, so work queued on
 * it overlaps with synchronous null-stream calls made by other host threads
 * (plan uploads while requests are in flight).  No reference counterpart. */
void *cast_hip_stream_create_nonblocking(char *err_buf, size_t err_buf_len);
void *cast_hip_event_create(char *err_buf, size_t err_buf_len);
/* *out_done = 1 when all work captured by the event has completed, else 0;
 * never blocks.  No reference counterpart (the reference only synchronises). */
int cast_hip_event_query(void *event, int *out_done, char *err_buf, size_t err_buf_len);
int cast_hip_event_record(void *event, void *stream, char *err_buf, size_t err_buf_len);
int cast_hip_event_synchronize(void *event, char *err_buf, size_t err_buf_len);
void cast_hip_event_destroy(void *event);
int cast_hip_event_elapsed_ms(void *start_event, void *end_event, float *out_ms, char *err_buf,
                              size_t err_buf_len);

/* ── Device-to-device async copy (ffi_cuda.h:75-76) ─────────────────────── */

int cast_hip_memcpy_dtod_async(uint64_t dst, uint64_t src, size_t n_bytes, void *stream,
                               char *err_buf, size_t err_buf_len);

/* ── Raw device memory (no reference counterpart beyond cuMemAlloc inside
 *    cast_cuda_sv_alloc, cuda_exec.cpp:429-449) — for host engines that lay
 *    out task tables and matrix pools in device memory. ───────────────────── */

uint64_t cast_hip_malloc(size_t n_bytes, char *err_buf, size_t err_buf_len); /* 0 on failure */
void cast_hip_free(uint64_t dptr);
/* Host -> device copy on `stream`; the host buffer must stay valid until the
 * stream reaches the copy (synchronise before reusing it). */
int cast_hip_memcpy_htod_async(uint64_t dst, const void *src, size_t n_bytes, void *stream,
                               char *err_buf, size_t err_buf_len);
/* Device -> host copy; returns after completion. */
int cast_hip_memcpy_dtoh(void *dst, uint64_t src, size_t n_bytes, void *stream, char *err_buf,
                         size_t err_buf_len);
/* Page-locked host memory, for copies that overlap with device work. */
void *cast_hip_host_alloc(size_t n_bytes, char *err_buf, size_t err_buf_len); /* NULL on failure */
void cast_hip_host_free(void *ptr);
/* Device -> host copy queued on `stream`; dst must be page-locked
 * (cast_hip_host_alloc) and stays undefined until the stream reaches it. */
int cast_hip_memcpy_dtoh_async(void *dst, uint64_t src, size_t n_bytes, void *stream,
                               char *err_buf, size_t err_buf_len);

/* ── Statevector memory (ffi_cuda.h:80-88) ───────────────────────────────── */

/* Returns 0 on failure (e.g. out of device memory). */
uint64_t cast_hip_sv_alloc(size_t n_elements, uint8_t precision, char *err_buf,
                           size_t err_buf_len);
void cast_hip_sv_free(uint64_t dptr);
/* Sets |0...0>: every scalar 0, re(amp[0]) = 1 (cuda_exec.cpp:452-479).
 * Runs on the null stream and returns after completion. */
int cast_hip_sv_zero(uint64_t dptr, size_t n_elements, uint8_t precision, char *err_buf,
                     size_t err_buf_len);
/* |0...0> queued on `stream` (non-blocking). */
int cast_hip_sv_zero_async(uint64_t dptr, size_t n_elements, uint8_t precision, void *stream,
                           char *err_buf, size_t err_buf_len);
int cast_hip_sv_upload(uint64_t dptr, const double *host_data, size_t n_elements,
                       uint8_t precision, char *err_buf, size_t err_buf_len);
int cast_hip_sv_download(uint64_t dptr, double *host_data, size_t n_elements, uint8_t precision,
                         char *err_buf, size_t err_buf_len);

/* ── Reductions (ffi_cuda.h:94-99; cuda_kernels.cu:123-196) ─────────────── */

/* Sum of squares of all n_elements scalars, accumulated in f64 for both
 * precisions.  Synchronous (null stream). */
int cast_hip_norm_squared(uint64_t dptr, size_t n_elements, uint8_t precision,
                          double *out_norm_sq, char *err_buf, size_t err_buf_len);
/* x *= scale for all n_elements scalars (scale applied in f64, rounded once).
 * Synchronous (null stream). */
int cast_hip_scale(uint64_t dptr, size_t n_elements, uint8_t precision, double scale,
                   char *err_buf, size_t err_buf_len);

/* ── Gate plans (replace ffi_cuda.h:27-32, 49-62) ────────────────────────── */

/* Build a plan for a k-qubit gate.  matrix is row-major K x K (K = 2^k,
 * matrix_len == K*K); qubits must be strictly ascending (canonical order,
 * src/types/gate.rs:483-521) — local basis bit b addresses global qubit
 * qubits[b] (cuda_gen.cpp:241-246).  1 <= k <= 6. */
int cast_hip_gate_plan_create(const cast_hip_kernel_gen_spec_t *spec,
                              const cast_complex64_t *matrix, size_t matrix_len,
                              const uint32_t *qubits, size_t n_qubits, void **out_plan,
                              char *err_buf, size_t err_buf_len);
void cast_hip_gate_plan_destroy(void *plan);
int cast_hip_gate_plan_info(const void *plan, cast_hip_gate_plan_info_t *out, char *err_buf,
                            size_t err_buf_len);

/* Kernel family (CAST_HIP_VARIANT_DENSE, _MONOMIAL, _TILE or _MFMA) that
 * cast_hip_gate_apply launches for this plan on a statevector of sv_n_qubits
 * qubits.  No reference counterpart (diagnostics and benchmarks). */
int cast_hip_gate_plan_family(const void *plan, uint32_t sv_n_qubits, uint32_t *out_variant,
                              char *err_buf, size_t err_buf_len);

/* Enqueue the gate on `stream` (NULL = null stream) against the statevector
 * at sv_dptr holding 2^sv_n_qubits amplitudes of the plan's precision.
 * Replaces cast_cuda_kernel_launch (ffi_cuda.h:59-61, cuda_exec.cpp:302-335).
 * Non-blocking. */
int cast_hip_gate_apply(const void *plan, void *stream, uint64_t sv_dptr, uint32_t sv_n_qubits,
                        char *err_buf, size_t err_buf_len);

/* Enqueue n_items independent (plan, statevector) pairs as ONE launch (64
 * items per launch, more in several launches): item i applies plans[i] to
 * svs[i] (2^sv_n_qubits[i] amplitudes, sv_n_qubits[i] <= 40) with the direct
 * kernel.  All plans must share precision and k.  For callers that hold many
 * small statevectors (one launch instead of one per statevector).  No
 * reference counterpart (the reference launches one kernel per apply). */
int cast_hip_gate_apply_batch(const void *const *plans, const uint64_t *svs,
                              const uint32_t *sv_n_qubits, size_t n_items, void *stream,
                              char *err_buf, size_t err_buf_len);

/* ⟨Z_a Z_b⟩ for n_pairs qubit pairs (pairs = [a0,b0,a1,b1,...]), accumulated
 * in f64 over |amp|^2; no normalisation is applied.  Synchronous on `stream`.
 * No reference counterpart (the reference downloads the whole statevector,
 * src/simulator/mod.rs:282-291). */
int cast_hip_expect_zz(uint64_t sv_dptr, uint32_t sv_n_qubits, uint8_t precision,
                       const uint32_t *pairs, size_t n_pairs, double *out, void *stream,
                       char *err_buf, size_t err_buf_len);

/* Queued on `stream`: out[0] = sum |amp|^2 and out[1 + j] = sum |amp|^2
 * (-1)^(bit a_j xor bit b_j), f64, into device memory `out` (1 + n_pairs
 * doubles, n_pairs <= 32), using device `scratch` of
 * cast_hip_observe_scratch_bytes(n_pairs) bytes that no other queued work
 * uses until this completes.  Deterministic (fixed reduction order).  No
 * reference counterpart. */
size_t cast_hip_observe_scratch_bytes(uint32_t n_pairs);
int cast_hip_observe_async(uint64_t sv_dptr, uint32_t sv_n_qubits, uint8_t precision,
                           const uint32_t *pairs, size_t n_pairs, uint64_t out, uint64_t scratch,
                           void *stream, char *err_buf, size_t err_buf_len);

/* cast_hip_observe_async without the scratch argument: the reduction
 * partials live in a static device array of the library (528 KiB per GPU,
 * held until the process ends; the copy on the device of `stream`, or of
 * the current device for the null stream), so the call allocates nothing (no hipMalloc,
 * no pinned memory) and only enqueues two kernels on `stream`.  The caller
 * passes the statevector and the device output (1 + n_pairs doubles);
 * n_pairs above 32 is rejected.  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (other streams, other threads) race
 * and must be serialised by the caller.  No counterpart in the reference:
 * it is for callers that cannot allocate (code under stream capture, tight
 * serving loops, harnesses that forbid hipMalloc). */
int cast_hip_observe_noalloc(uint64_t sv_dptr, uint32_t sv_n_qubits, uint8_t precision,
                             const uint32_t *pairs, size_t n_pairs, uint64_t out, void *stream,
                             char *err_buf, size_t err_buf_len);

/* Marginal measurement probabilities over `measured_qubits` (distinct, any
 * order): out[b] = sum of |amp_j|^2 over every j whose bit measured_qubits[i]
 * equals bit i of b, for b < 2^n_measured, accumulated in f64 on the device;
 * only the 2^n_measured results cross to the host (`out`, host memory).
 * Synchronous on `stream`.  Replaces the download-and-accumulate of
 * Backend::marginal_probabilities for Cuda (src/simulator/mod.rs:282-291,
 * src/simulator/measure.rs:15-25). */
int cast_hip_marginal_probabilities(uint64_t sv_dptr, uint32_t sv_n_qubits, uint8_t precision,
                                    const uint32_t *measured_qubits, size_t n_measured,
                                    double *out, void *stream, char *err_buf, size_t err_buf_len);

/* Populations of a density matrix held as a 2 n_physical-qubit vectorised
 * statevector: out[i] = Re rho[i][i], the element at index i | (i << n),
 * for i < 2^n_physical (host memory), f64.  Synchronous on `stream`.
 * Replaces the download in QuantumState<Cuda>::populations / trace
 * (src/simulator/mod.rs:388-414). */
int cast_hip_dm_populations(uint64_t sv_dptr, uint32_t n_physical, uint8_t precision, double *out,
                            void *stream, char *err_buf, size_t err_buf_len);

/* <a|b> = sum_j conj(a_j) b_j over two statevectors of sv_n_qubits qubits,
 * each F32 or F64, accumulated in f64; out_re_im = {Re, Im}.  Synchronous on
 * `stream`.  Fidelity |<a|b>|^2 / (<a|a><b|b>) of a reduced-precision result
 * against an f64 one without copying either to the host.  No reference
 * counterpart. */
int cast_hip_inner_product(uint64_t a, uint8_t precision_a, uint64_t b, uint8_t precision_b,
                           uint32_t sv_n_qubits, double *out_re_im, void *stream,
                           char *err_buf, size_t err_buf_len);

/* The same queued on `stream` (returns without waiting): device out[0..1] =
 * {Re, Im} of <a|b>, using device `scratch` of
 * cast_hip_inner_product_scratch_bytes() bytes that no other queued work uses
 * until this completes.  Deterministic (fixed reduction order).  No reference
 * counterpart (serving's f32-against-f64 check runs it between requests
 * without waiting for the device). */
size_t cast_hip_inner_product_scratch_bytes(void);
int cast_hip_inner_product_async(uint64_t a, uint8_t precision_a, uint64_t b, uint8_t precision_b,
                                 uint32_t sv_n_qubits, uint64_t out, uint64_t scratch,
                                 void *stream, char *err_buf, size_t err_buf_len);

/* ── Batched trajectories in one persistent launch ───────────────────────────
 * No reference counterpart (the reference branches deterministic ensembles
 * on the host, src/simulator/trajectory.rs:180-265).  Each workgroup owns
 * one statevector slot and takes whole tasks from an atomic counter until the
 * task list is exhausted (no workgroup ever waits on another).  A task starts
 * from a checkpoint statevector, applies its gate list (k <= 4 at F64, <= 6
 * at F32, k = 6 only on statevectors of >= 12 qubits) — the first gate reads
 * the checkpoint and writes the slot — and writes f64 sums of |amp|^2 and of
 * |amp|^2 (-1)^(bit a xor bit b) for each requested pair.  Matrix pool
 * entries (run precision): k = 1..3, the K x K complex matrix, row-major
 * interleaved (K^2 complex slots); k = 4..6, three real K x K matrices Re M,
 * Im M, Re M + Im M one after the other (3 K^2 / 2 complex slots), for the
 * three-product update. */
typedef struct cast_hip_traj_gate_t {
  uint32_t matrix_offset; /* complex elements into the matrix pool      */
  uint8_t k;              /* 1..4 (F64), 1..6 (F32)                     */
  uint8_t q[6];           /* ascending target qubits                    */
  uint8_t pad;
} cast_hip_traj_gate_t;

typedef struct cast_hip_traj_task_t {
  uint32_t first_gate; /* into the gate table                        */
  uint32_t n_gates;
  uint32_t checkpoint; /* into the checkpoint table                  */
  uint32_t pad;
} cast_hip_traj_task_t;

typedef struct cast_hip_traj_launch_t {
  uint32_t precision;     /* 0 = F32, 1 = F64                           */
  uint32_t n_qubits;
  uint64_t tasks;         /* device cast_hip_traj_task_t[n_tasks]         */
  uint64_t order;         /* device uint32_t[n_tasks]: task visit order   */
  uint32_t n_tasks;
  uint32_t n_gates;       /* size of the gate table                     */
  uint64_t gates;         /* device cast_hip_traj_gate_t[n_gates]         */
  uint64_t matrices;      /* device complex pool (run precision)        */
  uint64_t n_matrix_elems;/* complex elements in the pool               */
  uint64_t checkpoints;   /* device: n_checkpoints statevectors of 2^n  */
                          /* amplitudes each, contiguous                */
  uint32_t n_checkpoints;
  uint32_t n_slots;       /* workgroups = statevector slots             */
  uint64_t slots;         /* device: n_slots * 2^n amplitudes           */
  uint32_t block;         /* threads per workgroup: 64, 128, 256 or 512 */
  uint32_t n_pairs;       /* <= 32                                      */
  uint32_t pairs[64];     /* a0, b0, a1, b1, ...                        */
  uint64_t out;           /* device double[n_tasks * (1 + n_pairs)]     */
  uint64_t dump;          /* 0, or device buffer for final states:      */
                          /* n_tasks * 2^n amplitudes, indexed by task  */
  uint64_t work;          /* device scratch: 2 * uint32_t (counter, error flag) */
  uint32_t waves_per_simd;/* occupancy the kernel is built for: 2 (256 VGPRs, */
                          /* no spills) or 4 (128 VGPRs, two 512-thread     */
                          /* workgroups per CU); 0 = 2                      */
  uint32_t flags;         /* CAST_HIP_TRAJ_FLAG_*, 0 = defaults          */
} cast_hip_traj_launch_t;

/* k = 5, 6 gates (F32) on the f32 matrix cores (v_mfma_f32_16x16x4_f32)
 * instead of the default bf16 split products (see CAST_HIP_FLAG_MFMA_SPLIT). */
#define CAST_HIP_TRAJ_FLAG_MFMA_F32 1u

/* Enqueue the persistent launch on `stream` (returns without waiting).  A
 * gate descriptor that is invalid on the device (k outside 1..4 at F64 or
 * 1..6 at F32, k = 6 on fewer than 12 qubits, a qubit >= n, a matrix outside
 * the pool, a checkpoint outside the table) is skipped and sets the error
 * flag (second uint32_t of `work`), which the caller checks after the stream
 * completes. */
int cast_hip_traj_run(const cast_hip_traj_launch_t *args, void *stream, char *err_buf,
                      size_t err_buf_len);

/* Library version string (static storage). */
const char *cast_hip_version(void);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* CAST_HIP_H */
