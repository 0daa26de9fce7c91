// Shared helpers of the gfx950 backend library (libcast_hip.so).
#pragma once

#include "cast_hip.h"

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>

namespace cast_hip {

// Error carried from the implementation to the C boundary, where it is
// turned into (return 1, err_buf) — the convention of
// src/cpp/cuda/ffi_cuda.cpp:1-13 and src/cpp/internal/err_buf.hpp:12-23.
struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline void write_err(char *buf, size_t len, const std::string &msg) {
  if (buf == nullptr || len == 0)
    return;
  const size_t n = msg.size() < len - 1 ? msg.size() : len - 1;
  std::memcpy(buf, msg.data(), n);
  buf[n] = '\0';
}

inline void clear_err(char *buf, size_t len) {
  if (buf != nullptr && len > 0)
    buf[0] = '\0';
}

inline void check(hipError_t rc, const char *what) {
  if (rc != hipSuccess)
    throw Error(std::string(what) + " failed: " + hipGetErrorString(rc));
}

// Throws unless a HIP device is present; selects nothing (the caller's
// current device is used, device 0 by default).
void ensure_device();

// Cached properties of the current device.
struct DeviceProps {
  int device = 0;
  int n_cu = 0;
  int warp = 64;
  size_t lds_per_block = 0;
  int l2_bytes = 0;
  std::string arch;
  std::string name;
};
const DeviceProps &device_props();

inline size_t scalar_bytes(uint8_t precision) { return precision == 0 ? 4 : 8; }

inline void check_precision(uint8_t precision) {
  if (precision > 1)
    throw Error("precision must be 0 (F32) or 1 (F64), got " + std::to_string(precision));
}

// Geometry of one gate on one statevector.  q[] holds the ascending target
// qubits; combos = 2^(n-k) is the number of K-amplitude groups.
struct GateGeom {
  uint64_t combos;
  uint32_t q[6];
  uint32_t k;
  // Matrix-core kernels: 0 = waves visit tiles in launch order; 1 = the
  // grid (a multiple of 8 workgroups) is renumbered so that the workgroups
  // dispatched to one XCD (every 8th) cover a contiguous 1/8 of each
  // grid-stride step.
  uint32_t xcd;
  // Matrix-core kernels: 0, or the number of low tile-index bits visited in
  // bit-reversed order (n_tiles is a power of two >= 2^swz).
  uint32_t swz;
  uint32_t pad;
};

// Plan built by cast_hip_gate_plan_create.
struct GatePlan {
  uint32_t k = 0;
  uint8_t precision = 1;
  uint32_t variant = 0;  // CAST_HIP_VARIANT_* (DENSE = dense family, launcher picks)
  bool forced = false;   // variant requested explicitly by the spec
  uint32_t flags = 0;    // CAST_HIP_FLAG_*
  uint32_t q[8] = {};
  int device = 0;
  void *d_mat = nullptr;       // K*K complex, interleaved, device precision
  void *d_mat3 = nullptr;      // [Re M | Im M | Re M + Im M], 3*K*K real (k = 4, 5)
  size_t mat_bytes = 0, mat3_bytes = 0;
  double *h_mat = nullptr;     // snapped matrix in f64 (host copy, 2*K*K)
  uint64_t scalar_nnz = 0;
  // Monomial form: rows that differ from identity, each with one source col.
  uint32_t n_active = 0;
  uint8_t act_row[64] = {};
  uint8_t act_col[64] = {};
  double act_re[64] = {};
  double act_im[64] = {};
};

// Launch entry points implemented in gate_kernels.hip.
void launch_gate(const GatePlan &plan, hipStream_t stream, void *sv, uint32_t n_qubits);
// Kernel family launch_gate uses for `plan` on an n-qubit statevector.
uint32_t plan_family(const GatePlan &plan, uint32_t n_qubits);
struct BatchItem {
  void *sv;
  const void *mat;
  uint64_t combos;
  uint8_t q[6];
  uint8_t k;
  uint8_t pad;
};
void launch_gate_batch(const BatchItem *items, size_t n, uint8_t precision, uint32_t k,
                       hipStream_t stream);

// Statevector utility kernels implemented in sv_kernels.hip.
void launch_sv_zero(void *sv, size_t n_elements, uint8_t precision, hipStream_t stream);
double norm_squared(const void *sv, size_t n_elements, uint8_t precision, hipStream_t stream);
void launch_scale(void *sv, size_t n_elements, uint8_t precision, double scale,
                  hipStream_t stream);
void expect_zz(const void *sv, uint32_t n_qubits, uint8_t precision, const uint32_t *pairs,
               size_t n_pairs, double *out, hipStream_t stream);

// <a|b> in f64 over two statevectors of n qubits, any precisions; synchronous.
void inner_product(const void *a, uint8_t prec_a, const void *b, uint8_t prec_b,
                   uint32_t n_qubits, double *out, hipStream_t stream);
// The same queued on `stream` into device out[0..1] = {Re, Im}, with device
// scratch of inner_scratch_bytes().
size_t inner_scratch_bytes();
void launch_inner(const void *a, uint8_t prec_a, const void *b, uint8_t prec_b, uint32_t n_qubits,
                  double *out, double *scratch, hipStream_t stream);

// Marginal probabilities over `m` measured qubits (bin bit i = address bit
// qubits[i]) and the density-matrix diagonal, f64, into host memory;
// synchronous on `stream` (sv_kernels.hip).
void marginal_probabilities(const void *sv, uint32_t n_qubits, uint8_t precision,
                            const uint32_t *qubits, uint32_t m, double *out_host,
                            hipStream_t stream);
void dm_populations(const void *sv, uint32_t n_physical, uint8_t precision, double *out_host,
                    hipStream_t stream);

// Deterministic f64 norm + pair sums into device `out` (sv_kernels.hip).
size_t observe_scratch_bytes(uint32_t n_pairs);
void launch_observe(const void *sv, uint32_t n_qubits, uint8_t precision, const uint32_t *pairs,
                    size_t n_pairs, double *out, double *scratch, hipStream_t stream);
// The same with the library's static device scratch (no allocation; at most
// observe_noalloc_max_pairs() pairs).
size_t observe_noalloc_max_pairs();
void launch_observe_noalloc(const void *sv, uint32_t n_qubits, uint8_t precision,
                            const uint32_t *pairs, size_t n_pairs, double *out,
                            hipStream_t stream);

// Device blocks for small long-lived objects (gate matrices), cached by
// power-of-two size class: a released block is kept for reuse rather than
// returned with hipFree, which waits for the whole device.
void *cache_alloc(size_t bytes);
void cache_release(void *p, size_t bytes);

// Persistent batched-trajectory launch (traj_kernels.hip); validates the
// descriptor and enqueues the counter reset and the kernel.
void launch_traj(const cast_hip_traj_launch_t &args, hipStream_t stream);

} // namespace cast_hip
