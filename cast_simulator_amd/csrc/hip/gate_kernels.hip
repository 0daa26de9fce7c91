// Gate-application kernels for gfx950 (MI355X).
//
// A k-qubit gate with ascending targets q[0..k) updates, for every combo
// c in [0, 2^(n-k)), the K = 2^k amplitudes at base(c) | delta(i), where
// base(c) inserts a zero bit at each target position of c (the bit-deposit
// of src/cpp/cuda/cuda_gen.cpp:174-211) and delta(i) sets bit q[b] for each
// set bit b of i (cuda_gen.cpp:241-246).  new[r] = sum_c M[r][c] * old[c]
// (cuda_gen.cpp:255-283), stored in place.
//
// Statevector layout: interleaved complex, amplitude j at scalars 2j, 2j+1
// (src/cuda/statevector.rs:46-57), so each amplitude is one 16-byte (f64)
// or 8-byte (f32) vector access.
//
// Kernel families:
//   k_dense      one combo per lane, K amplitudes in VGPRs, matrix read with
//                wave-uniform addresses (scalar loads) — coalesced whenever
//                the lowest target is >= 3, since consecutive lanes then
//                touch consecutive amplitudes.
//   k_monomial   <= 1 nonzero per row: touches only rows that differ from the
//                identity row, so e.g. CZ / CP move 1/4 of the bytes.
//   k_dense_batch  k_dense over many (matrix, statevector) items, item =
//                blockIdx.y, for batched trajectories and serving.

#include "common.hpp"

#include <algorithm>
#include <cstdlib>

namespace cast_hip {
namespace {

template <typename T> struct V2;
template <> struct V2<float> { using type = float2; };
template <> struct V2<double> { using type = double2; };

__device__ __forceinline__ uint64_t insert_zero(uint64_t x, uint32_t q) {
  const uint64_t lo_mask = (uint64_t(1) << q) - 1;
  return ((x & ~lo_mask) << 1) | (x & lo_mask);
}

// Deposit the combo index into the non-target bit positions.  Targets are
// ascending, so inserting zeros from the lowest position up is exact.
template <int KL>
__device__ __forceinline__ uint64_t combo_base(uint64_t c, const uint32_t *q) {
#pragma unroll
  for (int b = 0; b < KL; ++b)
    c = insert_zero(c, q[b]);
  return c;
}

template <int KL>
__device__ __forceinline__ uint64_t local_delta(uint32_t i, const uint32_t *q) {
  uint64_t d = 0;
#pragma unroll
  for (int b = 0; b < KL; ++b)
    d |= uint64_t((i >> b) & 1u) << q[b];
  return d;
}

// ── Dense: one combo per lane ───────────────────────────────────────────────

template <typename T, int KL>
__device__ __forceinline__ void dense_update(typename V2<T>::type *__restrict__ amp,
                                             const T *__restrict__ mat, uint64_t base,
                                             const uint64_t *delta) {
  constexpr int K = 1 << KL;
  using VT = typename V2<T>::type;
  T vr[K], vi[K];
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const VT x = amp[base + delta[i]];
    vr[i] = x.x;
    vi[i] = x.y;
  }
#pragma unroll
  for (int r = 0; r < K; ++r) {
    T ar0 = T(0), ar1 = T(0), ai = T(0);
#pragma unroll
    for (int c = 0; c < K; ++c) {
      const T mr = mat[2 * (r * K + c)];
      const T mi = mat[2 * (r * K + c) + 1];
      ar0 = __builtin_fma(mr, vr[c], ar0);
      ar1 = __builtin_fma(mi, vi[c], ar1);
      ai = __builtin_fma(mr, vi[c], ai);
      ai = __builtin_fma(mi, vr[c], ai);
    }
    VT out;
    out.x = ar0 - ar1;
    out.y = ai;
    amp[base + delta[r]] = out;
  }
}

template <typename T, int KL>
__global__ void __launch_bounds__(256) k_dense(T *__restrict__ sv, const T *__restrict__ mat,
                                               GateGeom g) {
  constexpr int K = 1 << KL;
  using VT = typename V2<T>::type;
  VT *__restrict__ amp = reinterpret_cast<VT *>(sv);
  uint64_t delta[K];
#pragma unroll
  for (int i = 0; i < K; ++i)
    delta[i] = local_delta<KL>(uint32_t(i), g.q);
  // One combo per lane and no loop: the matrix entries are then read once
  // per wave through scalar loads instead of being hoisted into VGPRs.
  const uint64_t c = (uint64_t(blockIdx.y) * gridDim.x + blockIdx.x) * blockDim.x + threadIdx.x;
  if (c < g.combos)
    dense_update<T, KL>(amp, mat, combo_base<KL>(c, g.q), delta);
}

// ── Monomial: at most one nonzero per row ───────────────────────────────────

template <typename T, int KL> struct MonoArgs {
  uint32_t n_active;
  uint8_t row[1 << KL];
  uint8_t col[1 << KL];
  T re[1 << KL];
  T im[1 << KL];
};

template <typename T, int KL>
__global__ void __launch_bounds__(256) k_monomial(T *__restrict__ sv, MonoArgs<T, KL> m,
                                                  GateGeom g) {
  constexpr int K = 1 << KL;
  using VT = typename V2<T>::type;
  VT *__restrict__ amp = reinterpret_cast<VT *>(sv);
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t c = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; c < g.combos; c += stride) {
    const uint64_t base = combo_base<KL>(c, g.q);
    VT x[K];
#pragma unroll
    for (int j = 0; j < K; ++j)
      if (uint32_t(j) < m.n_active)
        x[j] = amp[base + local_delta<KL>(m.col[j], g.q)];
#pragma unroll
    for (int j = 0; j < K; ++j)
      if (uint32_t(j) < m.n_active) {
        VT o;
        o.x = m.re[j] * x[j].x - m.im[j] * x[j].y;
        o.y = m.re[j] * x[j].y + m.im[j] * x[j].x;
        amp[base + local_delta<KL>(m.row[j], g.q)] = o;
      }
  }
}

// ── Batched dense: item = blockIdx.y ────────────────────────────────────────

constexpr int kMaxBatch = 64;
struct BatchArgs {
  BatchItem it[kMaxBatch];
};

template <typename T, int KL>
__global__ void __launch_bounds__(256) k_dense_batch(BatchArgs a) {
  constexpr int K = 1 << KL;
  using VT = typename V2<T>::type;
  const BatchItem &it = a.it[blockIdx.y];
  VT *__restrict__ amp = reinterpret_cast<VT *>(it.sv);
  const T *__restrict__ mat = reinterpret_cast<const T *>(it.mat);
  uint32_t q[6];
#pragma unroll
  for (int b = 0; b < 6; ++b)
    q[b] = it.q[b];
  uint64_t delta[K];
#pragma unroll
  for (int i = 0; i < K; ++i)
    delta[i] = local_delta<KL>(uint32_t(i), q);
  const uint64_t c = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c < it.combos)
    dense_update<T, KL>(amp, mat, combo_base<KL>(c, q), delta);
}

// ── Launch helpers ──────────────────────────────────────────────────────────

constexpr unsigned kBlock = 256;

unsigned grid_for(uint64_t combos) {
  static const unsigned cap = [] {
    const char *s = std::getenv("CAST_HIP_GRID_PER_CU");
    long per_cu = 16;
    if (s && *s) {
      const long v = std::strtol(s, nullptr, 10);
      if (v > 0 && v <= 4096)
        per_cu = v;
    }
    return unsigned(std::max(1, device_props().n_cu) * per_cu);
  }();
  const uint64_t need = (combos + kBlock - 1) / kBlock;
  return unsigned(std::max<uint64_t>(1, std::min<uint64_t>(need, cap)));
}

// Full grid, one combo per lane; blocks beyond 2^20 spill into grid.y.
dim3 full_grid(uint64_t combos) {
  const uint64_t blocks = (combos + kBlock - 1) / kBlock;
  constexpr uint64_t kMaxX = uint64_t(1) << 20;
  if (blocks <= kMaxX)
    return dim3(unsigned(blocks), 1, 1);
  return dim3(unsigned(kMaxX), unsigned((blocks + kMaxX - 1) / kMaxX), 1);
}

template <typename T, int KL>
void launch_dense_t(const GatePlan &p, hipStream_t s, void *sv, const GateGeom &g) {
  hipLaunchKernelGGL((k_dense<T, KL>), full_grid(g.combos), dim3(kBlock), 0, s,
                     static_cast<T *>(sv), static_cast<const T *>(p.d_mat), g);
  check(hipGetLastError(), "hipLaunchKernelGGL(k_dense)");
}

template <typename T, int KL>
void launch_mono_t(const GatePlan &p, hipStream_t s, void *sv, const GateGeom &g) {
  MonoArgs<T, KL> m{};
  m.n_active = p.n_active;
  for (uint32_t j = 0; j < p.n_active; ++j) {
    m.row[j] = p.act_row[j];
    m.col[j] = p.act_col[j];
    m.re[j] = T(p.act_re[j]);
    m.im[j] = T(p.act_im[j]);
  }
  hipLaunchKernelGGL((k_monomial<T, KL>), dim3(grid_for(g.combos)), dim3(kBlock), 0, s,
                     static_cast<T *>(sv), m, g);
  check(hipGetLastError(), "hipLaunchKernelGGL(k_monomial)");
}

template <typename T>
void launch_by_k(const GatePlan &p, hipStream_t s, void *sv, const GateGeom &g) {
  const bool mono = p.variant == 2;
#define CAST_HIP_CASE(KL)                                                                        \
  case KL:                                                                                       \
    if (mono)                                                                                    \
      launch_mono_t<T, KL>(p, s, sv, g);                                                         \
    else                                                                                         \
      launch_dense_t<T, KL>(p, s, sv, g);                                                        \
    return;
  switch (p.k) {
    CAST_HIP_CASE(1)
    CAST_HIP_CASE(2)
    CAST_HIP_CASE(3)
    CAST_HIP_CASE(4)
    CAST_HIP_CASE(5)
    CAST_HIP_CASE(6)
  default:
    throw Error("unsupported gate size k=" + std::to_string(p.k));
  }
#undef CAST_HIP_CASE
}

template <typename T>
void launch_batch_by_k(const BatchArgs &a, unsigned n, uint32_t k, uint64_t max_combos,
                       hipStream_t s) {
  const uint64_t blocks = (max_combos + kBlock - 1) / kBlock;
  if (blocks > 0x7fffffffULL)
    throw Error("batched statevector too large");
  const dim3 grid(unsigned(blocks), n);
#define CAST_HIP_BCASE(KL)                                                                       \
  case KL:                                                                                       \
    hipLaunchKernelGGL((k_dense_batch<T, KL>), grid, dim3(kBlock), 0, s, a);                    \
    break;
  switch (k) {
    CAST_HIP_BCASE(1)
    CAST_HIP_BCASE(2)
    CAST_HIP_BCASE(3)
    CAST_HIP_BCASE(4)
    CAST_HIP_BCASE(5)
    CAST_HIP_BCASE(6)
  default:
    throw Error("unsupported gate size k=" + std::to_string(k));
  }
#undef CAST_HIP_BCASE
  check(hipGetLastError(), "hipLaunchKernelGGL(k_dense_batch)");
}

} // namespace

void launch_gate(const GatePlan &plan, hipStream_t stream, void *sv, uint32_t n_qubits) {
  if (n_qubits < plan.k)
    throw Error("statevector has fewer qubits than the gate requires");
  if (plan.q[plan.k - 1] >= n_qubits)
    throw Error("gate acts on qubit " + std::to_string(plan.q[plan.k - 1]) +
                " outside a " + std::to_string(n_qubits) + "-qubit statevector");
  if (n_qubits > 40)
    throw Error("statevector qubit count " + std::to_string(n_qubits) + " is out of range");
  GateGeom g{};
  g.combos = uint64_t(1) << (n_qubits - plan.k);
  g.k = plan.k;
  for (uint32_t b = 0; b < plan.k; ++b)
    g.q[b] = plan.q[b];
  if (plan.variant == 2 && plan.n_active == 0)
    return; // identity: nothing to move
  if (plan.precision == 0)
    launch_by_k<float>(plan, stream, sv, g);
  else
    launch_by_k<double>(plan, stream, sv, g);
}

void launch_gate_batch(const BatchItem *items, size_t n, uint8_t precision, uint32_t k,
                       hipStream_t stream) {
  for (size_t off = 0; off < n; off += kMaxBatch) {
    const unsigned m = unsigned(std::min<size_t>(kMaxBatch, n - off));
    BatchArgs a{};
    uint64_t max_combos = 0;
    for (unsigned i = 0; i < m; ++i) {
      a.it[i] = items[off + i];
      max_combos = std::max(max_combos, items[off + i].combos);
    }
    if (precision == 0)
      launch_batch_by_k<float>(a, m, k, max_combos, stream);
    else
      launch_batch_by_k<double>(a, m, k, max_combos, stream);
  }
}

} // namespace cast_hip
