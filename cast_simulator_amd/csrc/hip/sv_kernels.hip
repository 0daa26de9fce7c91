// Statevector utility kernels: |0> fill, f64-accumulated norm^2, scale and
// <Z_a Z_b> expectations.
//
// Counterparts in the reference: cuda_exec.cpp:452-479 (memset + one H->D
// scalar for |0>), cuda_kernels.cu:123-196 (normSqPass1/2 with f64
// accumulation, scaleKernel).  Here |0> is one fill kernel (no memset node,
// no host copy), and every reduction accumulates in f64 whatever the
// statevector precision.

#include "common.hpp"

#include <algorithm>
#include <atomic>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

namespace cast_hip {
namespace {

constexpr unsigned kRedBlock = 256;
constexpr unsigned kMaxRedGrid = 2048;
constexpr unsigned kMaxPairs = 32;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    v += __shfl_xor(v, off, 64);
  return v;
}

// Block-wide sum; every thread gets the result.
__device__ __forceinline__ double block_sum(double v, double *lds) {
  v = wave_sum(v);
  const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0)
    lds[wave] = v;
  __syncthreads();
  double t = 0.0;
  const unsigned nw = (blockDim.x + 63u) >> 6;
  for (unsigned w = 0; w < nw; ++w)
    t += lds[w];
  return t;
}

template <typename T>
__global__ void __launch_bounds__(kRedBlock) k_zero(T *__restrict__ x, size_t n) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    x[i] = (i == 0) ? T(1) : T(0);
}

template <typename T>
__global__ void __launch_bounds__(kRedBlock) k_norm_pass1(const T *__restrict__ x, size_t n,
                                                          double *__restrict__ partial) {
  __shared__ double lds[kRedBlock / 64];
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  const size_t tid = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  double acc = 0.0;
  constexpr int VW = 16 / sizeof(T);
  using VT = typename std::conditional<sizeof(T) == 8, double2, float4>::type;
  const size_t nv = n / VW;
  const VT *xv = reinterpret_cast<const VT *>(x);
  for (size_t i = tid; i < nv; i += stride) {
    const VT v = xv[i];
    const T *e = reinterpret_cast<const T *>(&v);
#pragma unroll
    for (int j = 0; j < VW; ++j) {
      const double d = double(e[j]);
      acc = __builtin_fma(d, d, acc);
    }
  }
  for (size_t i = nv * VW + tid; i < n; i += stride) {
    const double d = double(x[i]);
    acc = __builtin_fma(d, d, acc);
  }
  const double s = block_sum(acc, lds);
  if (threadIdx.x == 0)
    partial[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kRedBlock) k_sum_partials(const double *__restrict__ partial,
                                                            unsigned n, unsigned width,
                                                            double *__restrict__ out) {
  // Reduces column `blockIdx.x` of a row-major (n x width) array.
  __shared__ double lds[kRedBlock / 64];
  double acc = 0.0;
  for (unsigned i = threadIdx.x; i < n; i += blockDim.x)
    acc += partial[size_t(i) * width + blockIdx.x];
  const double s = block_sum(acc, lds);
  if (threadIdx.x == 0)
    out[blockIdx.x] = s;
}

// In-place scale: one 16-byte vector per thread over a grid that covers the
// whole statevector (block b, thread t: vector b * blockDim + t), the access
// pattern of a one-tile-per-wave gate launch; a scalar tail, if any, is done
// by the first threads.  With `xcd` (block count a multiple of 8) block b
// takes the vectors of block (b mod 8) * (blocks / 8) + b / 8, so the
// workgroups of one XCD (every 8th) stream one contiguous eighth: 6.58 TB/s
// against 6.47 in launch order over 16 GiB of random f64 data (6.72 / 6.58
// over zeros).
template <typename T>
__global__ void __launch_bounds__(kRedBlock) k_scale(T *__restrict__ x, size_t n, double scale,
                                                     uint32_t xcd) {
  constexpr int VW = 16 / sizeof(T);
  using VT = T __attribute__((ext_vector_type(16 / sizeof(T))));
  size_t b = size_t(blockIdx.y) * gridDim.x + blockIdx.x;
  if (xcd) {
    const size_t nb = size_t(gridDim.x) * gridDim.y;
    b = (b & 7u) * (nb >> 3) + (b >> 3);
  }
  const size_t i = b * blockDim.x + threadIdx.x;
  const size_t nv = n / VW;
  if (i < nv) {
    VT *xv = reinterpret_cast<VT *>(x);
    VT v = __builtin_nontemporal_load(xv + i);
#pragma unroll
    for (int j = 0; j < VW; ++j)
      v[j] = T(double(v[j]) * scale);
    __builtin_nontemporal_store(v, xv + i);
  }
  if (i < n - nv * VW)
    x[nv * VW + i] = T(double(x[nv * VW + i]) * scale);
}

struct PairArgs {
  uint32_t n_pairs;
  uint32_t a[kMaxPairs];
  uint32_t b[kMaxPairs];
};

template <typename T>
__global__ void __launch_bounds__(kRedBlock) k_expect_zz(const T *__restrict__ sv, uint64_t n_amp,
                                                         PairArgs pa, double *__restrict__ partial) {
  __shared__ double lds[kRedBlock / 64];
  using VT = typename std::conditional<sizeof(T) == 8, double2, float2>::type;
  const VT *amp = reinterpret_cast<const VT *>(sv);
  double acc[kMaxPairs];
#pragma unroll
  for (unsigned p = 0; p < kMaxPairs; ++p)
    acc[p] = 0.0;
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < n_amp; j += stride) {
    const VT v = amp[j];
    const double re = double(v.x), im = double(v.y);
    const double pr = __builtin_fma(re, re, im * im);
#pragma unroll
    for (unsigned p = 0; p < kMaxPairs; ++p)
      if (p < pa.n_pairs) {
        const uint64_t par = ((j >> pa.a[p]) ^ (j >> pa.b[p])) & 1u;
        acc[p] += par ? -pr : pr;
      }
  }
#pragma unroll
  for (unsigned p = 0; p < kMaxPairs; ++p)
    if (p < pa.n_pairs) {
      const double s = block_sum(acc[p], lds);
      if (threadIdx.x == 0)
        partial[size_t(blockIdx.x) * pa.n_pairs + p] = s;
    }
}

// Norm (slot 0) and pair sums (slots 1..P) per block, written to
// partial[block * (1 + P) + slot].
template <typename T>
__global__ void __launch_bounds__(kRedBlock) k_observe(const T *__restrict__ sv, uint64_t n_amp,
                                                       PairArgs pa, double *__restrict__ partial) {
  __shared__ double lds[kRedBlock / 64];
  using VT = typename std::conditional<sizeof(T) == 8, double2, float2>::type;
  const VT *amp = reinterpret_cast<const VT *>(sv);
  double acc[1 + kMaxPairs];
#pragma unroll
  for (unsigned p = 0; p < 1 + kMaxPairs; ++p)
    acc[p] = 0.0;
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < n_amp; j += stride) {
    const VT v = amp[j];
    const double re = double(v.x), im = double(v.y);
    const double pr = __builtin_fma(re, re, im * im);
    acc[0] += pr;
#pragma unroll
    for (unsigned p = 0; p < kMaxPairs; ++p) {
      const uint64_t par = ((j >> pa.a[p]) ^ (j >> pa.b[p])) & 1u;
      acc[1 + p] += par ? -pr : pr; // unused slots (a = b = 0) are never read
    }
  }
  const unsigned w = 1 + pa.n_pairs;
#pragma unroll
  for (unsigned p = 0; p < 1 + kMaxPairs; ++p)
    if (p < w) {
      const double s = block_sum(acc[p], lds);
      if (threadIdx.x == 0)
        partial[size_t(blockIdx.x) * w + p] = s;
    }
}

// <a|b> = sum conj(a_j) b_j in f64, per-block partials (re, im) at
// partial[2 * block + {0, 1}].
template <typename TA, typename TB>
__global__ void __launch_bounds__(kRedBlock) k_inner(const TA *__restrict__ a,
                                                     const TB *__restrict__ b, uint64_t n_amp,
                                                     double *__restrict__ partial) {
  __shared__ double lds[kRedBlock / 64];
  double re = 0.0, im = 0.0;
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < n_amp; j += stride) {
    const double ar = double(a[2 * j]), ai = double(a[2 * j + 1]);
    const double br = double(b[2 * j]), bi = double(b[2 * j + 1]);
    re = __builtin_fma(ar, br, __builtin_fma(ai, bi, re));
    im = __builtin_fma(ar, bi, __builtin_fma(-ai, br, im));
  }
  const double sr = block_sum(re, lds);
  const double si = block_sum(im, lds);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = sr;
    partial[2 * blockIdx.x + 1] = si;
  }
}

// Marginal probabilities over measured qubits pos[0..m): bin b collects
// |amp_j|^2 for every j whose bit pos[i] equals bit i of b (the reference's
// compress_bits, src/types/gate.rs:636-643, summed as in
// src/simulator/measure.rs:15-25).  Amplitude (b, r) of bin b and rest index
// r is j = scatter(b) | pdep(r, ~M), M the measured mask.
constexpr unsigned kMaxMarginalQubits = 40;
struct MarginalArgs {
  uint64_t mask;                  // M
  uint32_t m;                     // measured qubits
  uint32_t pos[kMaxMarginalQubits];
};

__device__ __forceinline__ uint64_t insert_zero_u64(uint64_t x, uint32_t p) {
  const uint64_t lo = (uint64_t(1) << p) - 1;
  return ((x & ~lo) << 1) | (x & lo);
}

__device__ __forceinline__ uint64_t scatter_bin(uint64_t b, const MarginalArgs &a) {
  uint64_t base = 0;
#pragma unroll
  for (unsigned i = 0; i < kMaxMarginalQubits; ++i)
    if (i < a.m)
      base |= ((b >> i) & 1u) << a.pos[i];
  return base;
}

__device__ __forceinline__ uint64_t deposit_rest(uint64_t r, uint64_t mask) {
  for (uint64_t mm = mask; mm; mm &= mm - 1) // ascending measured positions
    r = insert_zero_u64(r, uint32_t(__builtin_ctzll(mm)));
  return r;
}

template <typename T> __device__ __forceinline__ double amp_prob(const T *sv, uint64_t j) {
  using VT = typename std::conditional<sizeof(T) == 8, double2, float2>::type;
  const VT v = reinterpret_cast<const VT *>(sv)[j];
  const double re = double(v.x), im = double(v.y);
  return __builtin_fma(re, re, im * im);
}

// Address-major (n >= 6): every wave reads 64 consecutive amplitudes, lane =
// address bits 0..5, whether those bits are measured or not.  Workgroup
// w = c * B_hi + h fixes the measured bits >= 6 to pattern h (deposited in
// ascending position order) and sums chunk c (chunk_len indices r over the
// unmeasured bits >= 6, waves interleaved).  Each lane's f64 sum is then
// reduced over the unmeasured lane bits with xor shuffles, summed over the
// waves in LDS, and lanes whose unmeasured bits are zero write the bin of
// their measured low bits and h into partial[c * B + bin].
template <typename T>
__global__ void __launch_bounds__(kRedBlock) k_marginal_rest(const T *__restrict__ sv,
                                                             MarginalArgs a, uint64_t chunk_len,
                                                             double *__restrict__ partial) {
  __shared__ double lds[kRedBlock];
  const uint64_t hi_mask = a.mask & ~uint64_t(63);
  const uint32_t lo_mask = uint32_t(a.mask & 63u);
  const uint32_t n_hi = uint32_t(__builtin_popcountll(hi_mask));
  const uint64_t w = uint64_t(blockIdx.y) * gridDim.x + blockIdx.x;
  const uint64_t h = w & ((uint64_t(1) << n_hi) - 1), c = w >> n_hi;
  uint64_t base = 0; // h deposited at the measured positions >= 6
  {
    uint64_t mm = hi_mask;
    for (uint32_t i = 0; mm; ++i, mm &= mm - 1)
      base |= ((h >> i) & 1u) << __builtin_ctzll(mm);
  }
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, n_waves = blockDim.x >> 6;
  double acc = 0.0;
  const uint64_t r0 = c * chunk_len;
  for (uint64_t r = r0 + wave; r < r0 + chunk_len; r += n_waves)
    acc += amp_prob(sv, deposit_rest(r << 6, hi_mask) | base | lane);
  for (uint32_t t = 0; t < 6; ++t)
    if (!((lo_mask >> t) & 1u))
      acc += __shfl_xor(acc, 1 << t, 64);
  lds[threadIdx.x] = acc;
  __syncthreads();
  if (wave == 0 && (lane & ~lo_mask) == 0) {
    double s = 0.0;
    for (uint32_t v = 0; v < n_waves; ++v)
      s += lds[v * 64 + lane];
    const uint64_t j = base | lane;
    uint64_t bin = 0;
#pragma unroll
    for (unsigned i = 0; i < kMaxMarginalQubits; ++i)
      if (i < a.m)
        bin |= ((j >> a.pos[i]) & 1u) << i;
    partial[(c << a.m) + bin] = s;
  }
}

// Bin-major (few amplitudes per bin): one thread per bin, rest indices in
// ascending order.
template <typename T>
__global__ void __launch_bounds__(kRedBlock) k_marginal_bin(const T *__restrict__ sv,
                                                            MarginalArgs a, uint64_t n_rest,
                                                            double *__restrict__ out) {
  const uint64_t n_bins = uint64_t(1) << a.m;
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t b = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; b < n_bins; b += stride) {
    const uint64_t base = scatter_bin(b, a);
    double acc = 0.0;
    for (uint64_t r = 0; r < n_rest; ++r)
      acc += amp_prob(sv, deposit_rest(r, a.mask) | base);
    out[b] = acc;
  }
}

// Density-matrix diagonal: the 2n-qubit vectorised rho holds rho[i][i] at
// index i | (i << n) (src/simulator/mod.rs:349-363); out[i] = Re of it.
template <typename T>
__global__ void __launch_bounds__(kRedBlock) k_dm_diag(const T *__restrict__ sv, uint32_t n,
                                                       double *__restrict__ out) {
  const uint64_t dim = uint64_t(1) << n;
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < dim; i += stride)
    out[i] = double(sv[2 * (i | (i << n))]);
}

unsigned red_grid(uint64_t work) {
  const uint64_t need = (work + kRedBlock - 1) / kRedBlock;
  const uint64_t cap = std::min<uint64_t>(kMaxRedGrid, uint64_t(device_props().n_cu) * 4);
  return unsigned(std::max<uint64_t>(1, std::min(need, cap)));
}

// Reduction scratch (device partials + pinned host result), one per process
// and guarded by a mutex so concurrent callers are serialised.
struct RedScratch {
  std::mutex mu;
  double *d_partial = nullptr; // kMaxRedGrid * kMaxPairs
  double *d_out = nullptr;     // kMaxPairs
  double *h_out = nullptr;     // pinned, kMaxPairs
  void ensure() {
    if (d_partial)
      return;
    check(hipMalloc(&d_partial, sizeof(double) * kMaxRedGrid * kMaxPairs), "hipMalloc(partial)");
    check(hipMalloc(&d_out, sizeof(double) * kMaxPairs), "hipMalloc(out)");
    check(hipHostMalloc(&h_out, sizeof(double) * kMaxPairs, hipHostMallocDefault),
          "hipHostMalloc(out)");
  }
};
RedScratch &red_scratch() {
  static RedScratch s;
  return s;
}

} // namespace

void launch_sv_zero(void *sv, size_t n_elements, uint8_t precision, hipStream_t stream) {
  const unsigned grid = red_grid(n_elements) * 4;
  if (precision == 0)
    hipLaunchKernelGGL(k_zero<float>, dim3(grid), dim3(kRedBlock), 0, stream,
                       static_cast<float *>(sv), n_elements);
  else
    hipLaunchKernelGGL(k_zero<double>, dim3(grid), dim3(kRedBlock), 0, stream,
                       static_cast<double *>(sv), n_elements);
  check(hipGetLastError(), "hipLaunchKernelGGL(k_zero)");
}

double norm_squared(const void *sv, size_t n_elements, uint8_t precision, hipStream_t stream) {
  RedScratch &s = red_scratch();
  std::lock_guard<std::mutex> lock(s.mu);
  s.ensure();
  const unsigned grid = red_grid(n_elements / (precision == 0 ? 4 : 2));
  if (precision == 0)
    hipLaunchKernelGGL(k_norm_pass1<float>, dim3(grid), dim3(kRedBlock), 0, stream,
                       static_cast<const float *>(sv), n_elements, s.d_partial);
  else
    hipLaunchKernelGGL(k_norm_pass1<double>, dim3(grid), dim3(kRedBlock), 0, stream,
                       static_cast<const double *>(sv), n_elements, s.d_partial);
  check(hipGetLastError(), "hipLaunchKernelGGL(k_norm_pass1)");
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kRedBlock), 0, stream, s.d_partial, grid, 1u,
                     s.d_out);
  check(hipGetLastError(), "hipLaunchKernelGGL(k_sum_partials)");
  check(hipMemcpyAsync(s.h_out, s.d_out, sizeof(double), hipMemcpyDeviceToHost, stream),
        "hipMemcpyAsync(norm)");
  check(hipStreamSynchronize(stream), "hipStreamSynchronize(norm)");
  return s.h_out[0];
}

void launch_scale(void *sv, size_t n_elements, uint8_t precision, double scale,
                  hipStream_t stream) {
  const size_t vw = precision == 0 ? 4 : 2;
  const uint64_t threads = std::max<uint64_t>(n_elements / vw, n_elements % vw);
  const uint64_t blocks = (threads + kRedBlock - 1) / kRedBlock;
  const uint64_t cap = uint64_t(1) << 20;
  const dim3 grid = blocks <= cap ? dim3(unsigned(blocks))
                                  : dim3(unsigned(cap), unsigned((blocks + cap - 1) / cap));
  const uint64_t n_blocks = uint64_t(grid.x) * grid.y;
  const uint32_t xcd = n_blocks >= 8 && n_blocks % 8 == 0;
  if (precision == 0)
    hipLaunchKernelGGL(k_scale<float>, grid, dim3(kRedBlock), 0, stream,
                       static_cast<float *>(sv), n_elements, scale, xcd);
  else
    hipLaunchKernelGGL(k_scale<double>, grid, dim3(kRedBlock), 0, stream,
                       static_cast<double *>(sv), n_elements, scale, xcd);
  check(hipGetLastError(), "hipLaunchKernelGGL(k_scale)");
}

void expect_zz(const void *sv, uint32_t n_qubits, uint8_t precision, const uint32_t *pairs,
               size_t n_pairs, double *out, hipStream_t stream) {
  RedScratch &s = red_scratch();
  std::lock_guard<std::mutex> lock(s.mu);
  s.ensure();
  const uint64_t n_amp = uint64_t(1) << n_qubits;
  for (size_t off = 0; off < n_pairs; off += kMaxPairs) {
    PairArgs pa{};
    pa.n_pairs = uint32_t(std::min<size_t>(kMaxPairs, n_pairs - off));
    for (uint32_t p = 0; p < pa.n_pairs; ++p) {
      pa.a[p] = pairs[2 * (off + p)];
      pa.b[p] = pairs[2 * (off + p) + 1];
    }
    const unsigned grid = red_grid(n_amp);
    if (precision == 0)
      hipLaunchKernelGGL(k_expect_zz<float>, dim3(grid), dim3(kRedBlock), 0, stream,
                         static_cast<const float *>(sv), n_amp, pa, s.d_partial);
    else
      hipLaunchKernelGGL(k_expect_zz<double>, dim3(grid), dim3(kRedBlock), 0, stream,
                         static_cast<const double *>(sv), n_amp, pa, s.d_partial);
    check(hipGetLastError(), "hipLaunchKernelGGL(k_expect_zz)");
    hipLaunchKernelGGL(k_sum_partials, dim3(pa.n_pairs), dim3(kRedBlock), 0, stream, s.d_partial,
                       grid, pa.n_pairs, s.d_out);
    check(hipGetLastError(), "hipLaunchKernelGGL(k_sum_partials)");
    check(hipMemcpyAsync(s.h_out, s.d_out, sizeof(double) * pa.n_pairs, hipMemcpyDeviceToHost,
                         stream),
          "hipMemcpyAsync(zz)");
    check(hipStreamSynchronize(stream), "hipStreamSynchronize(zz)");
    for (uint32_t p = 0; p < pa.n_pairs; ++p)
      out[off + p] = s.h_out[p];
  }
}

size_t observe_scratch_bytes(uint32_t n_pairs) {
  return sizeof(double) * kMaxRedGrid * (1 + std::min<uint32_t>(n_pairs, kMaxPairs));
}

namespace {
// Reduction partials of the allocation-free observe call: room for the
// largest launch_observe (kMaxRedGrid blocks x (1 + kMaxPairs) slots, 528
// KiB).  Each GPU the library runs on holds its own copy until the process
// ends.  Every slot a call's sum kernel reads is first written by the same
// call's k_observe, so nothing needs clearing; calls that overlap on one GPU
// share it and race.
static __device__ __align__(256) double scratch[size_t(kMaxRedGrid) * (1 + kMaxPairs)];

constexpr int kMaxDevices = 64;
std::atomic<double *> g_scratch_ptr[kMaxDevices];
std::mutex g_scratch_mu;

// Device address of `scratch` on the device `stream` belongs to (the current
// device for the null stream), looked up once per device with that device
// current (the host-side name is not a device address and must never be
// passed to a kernel).
double *device_scratch(hipStream_t stream) {
  int cur = 0;
  check(hipGetDevice(&cur), "hipGetDevice");
  int dev = cur;
  if (stream) {
    hipDevice_t sd = 0;
    check(hipStreamGetDevice(stream, &sd), "hipStreamGetDevice");
    dev = int(sd);
  }
  if (dev < 0 || dev >= kMaxDevices)
    throw Error("device ordinal " + std::to_string(dev) + " out of range");
  double *p = g_scratch_ptr[dev].load(std::memory_order_acquire);
  if (p)
    return p;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  p = g_scratch_ptr[dev].load(std::memory_order_relaxed);
  if (!p) {
    if (dev != cur)
      check(hipSetDevice(dev), "hipSetDevice");
    void *addr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
    if (dev != cur)
      check(hipSetDevice(cur), "hipSetDevice");
    check(e, "hipGetSymbolAddress(scratch)");
    if (!addr)
      throw Error("hipGetSymbolAddress(scratch) returned a null address");
    p = static_cast<double *>(addr);
    g_scratch_ptr[dev].store(p, std::memory_order_release);
  }
  return p;
}
} // namespace

size_t observe_noalloc_max_pairs() { return kMaxPairs; }

void launch_observe_noalloc(const void *sv, uint32_t n_qubits, uint8_t precision,
                            const uint32_t *pairs, size_t n_pairs, double *out,
                            hipStream_t stream) {
  if (n_pairs > kMaxPairs)
    throw Error("observe (no allocation): at most " + std::to_string(kMaxPairs) +
                " qubit pairs, the size of the device scratch; got " + std::to_string(n_pairs));
  static_assert(sizeof(scratch) >= sizeof(double) * kMaxRedGrid * (1 + kMaxPairs),
                "scratch must hold the largest launch_observe");
  launch_observe(sv, n_qubits, precision, pairs, n_pairs, out, device_scratch(stream), stream);
}

void launch_observe(const void *sv, uint32_t n_qubits, uint8_t precision, const uint32_t *pairs,
                    size_t n_pairs, double *out, double *scratch, hipStream_t stream) {
  if (n_pairs > kMaxPairs)
    throw Error("observe: at most 32 qubit pairs");
  PairArgs pa{};
  pa.n_pairs = uint32_t(n_pairs);
  for (uint32_t p = 0; p < pa.n_pairs; ++p) {
    pa.a[p] = pairs[2 * p];
    pa.b[p] = pairs[2 * p + 1];
  }
  const uint64_t n_amp = uint64_t(1) << n_qubits;
  const unsigned grid = red_grid(n_amp);
  if (precision == 0)
    hipLaunchKernelGGL(k_observe<float>, dim3(grid), dim3(kRedBlock), 0, stream,
                       static_cast<const float *>(sv), n_amp, pa, scratch);
  else
    hipLaunchKernelGGL(k_observe<double>, dim3(grid), dim3(kRedBlock), 0, stream,
                       static_cast<const double *>(sv), n_amp, pa, scratch);
  check(hipGetLastError(), "hipLaunchKernelGGL(k_observe)");
  hipLaunchKernelGGL(k_sum_partials, dim3(1 + pa.n_pairs), dim3(kRedBlock), 0, stream, scratch,
                     grid, 1 + pa.n_pairs, out);
  check(hipGetLastError(), "hipLaunchKernelGGL(k_sum_partials)");
}

size_t inner_scratch_bytes() { return sizeof(double) * kMaxRedGrid * 2; }

void launch_inner(const void *a, uint8_t prec_a, const void *b, uint8_t prec_b, uint32_t n_qubits,
                  double *out, double *scratch, hipStream_t stream) {
  const uint64_t n_amp = uint64_t(1) << n_qubits;
  const unsigned grid = red_grid(n_amp);
  auto go = [&](auto *pa, auto *pb) {
    using A = std::remove_const_t<std::remove_pointer_t<decltype(pa)>>;
    using B = std::remove_const_t<std::remove_pointer_t<decltype(pb)>>;
    hipLaunchKernelGGL((k_inner<A, B>), dim3(grid), dim3(kRedBlock), 0, stream, pa, pb, n_amp,
                       scratch);
  };
  if (prec_a == 0 && prec_b == 0)
    go(static_cast<const float *>(a), static_cast<const float *>(b));
  else if (prec_a == 0)
    go(static_cast<const float *>(a), static_cast<const double *>(b));
  else if (prec_b == 0)
    go(static_cast<const double *>(a), static_cast<const float *>(b));
  else
    go(static_cast<const double *>(a), static_cast<const double *>(b));
  check(hipGetLastError(), "hipLaunchKernelGGL(k_inner)");
  hipLaunchKernelGGL(k_sum_partials, dim3(2), dim3(kRedBlock), 0, stream, scratch, grid, 2u, out);
  check(hipGetLastError(), "hipLaunchKernelGGL(k_sum_partials)");
}

void inner_product(const void *a, uint8_t prec_a, const void *b, uint8_t prec_b,
                   uint32_t n_qubits, double *out, hipStream_t stream) {
  RedScratch &s = red_scratch();
  std::lock_guard<std::mutex> lock(s.mu);
  s.ensure();
  launch_inner(a, prec_a, b, prec_b, n_qubits, s.d_out, s.d_partial, stream);
  check(hipMemcpyAsync(s.h_out, s.d_out, 2 * sizeof(double), hipMemcpyDeviceToHost, stream),
        "hipMemcpyAsync(inner)");
  check(hipStreamSynchronize(stream), "hipStreamSynchronize(inner)");
  out[0] = s.h_out[0];
  out[1] = s.h_out[1];
}

} // namespace cast_hip

namespace cast_hip {

namespace {
// Device buffer for one synchronous call: small sizes from the block cache,
// large ones straight from the allocator.
struct CallBuffer {
  void *p = nullptr;
  size_t bytes = 0;
  explicit CallBuffer(size_t b) : bytes(b) {
    if (bytes <= (size_t(16) << 20))
      p = cache_alloc(bytes);
    else
      check(hipMalloc(&p, bytes), "hipMalloc(marginal buffer)");
  }
  ~CallBuffer() {
    if (!p)
      return;
    if (bytes <= (size_t(16) << 20))
      cache_release(p, bytes);
    else
      (void)hipFree(p);
  }
  CallBuffer(const CallBuffer &) = delete;
  CallBuffer &operator=(const CallBuffer &) = delete;
};

dim3 wide_grid(uint64_t blocks) {
  const uint64_t cap = 1u << 20;
  if (blocks <= cap)
    return dim3(unsigned(blocks));
  return dim3(unsigned(cap), unsigned(blocks / cap));
}
} // namespace

void marginal_probabilities(const void *sv, uint32_t n_qubits, uint8_t precision,
                            const uint32_t *qubits, uint32_t m, double *out_host,
                            hipStream_t stream) {
  if (m > kMaxMarginalQubits || m > n_qubits)
    throw Error("marginal: more measured qubits than the statevector has");
  MarginalArgs a{};
  a.m = m;
  for (uint32_t i = 0; i < m; ++i) {
    if (qubits[i] >= n_qubits)
      throw Error("marginal: measured qubit " + std::to_string(qubits[i]) + " out of range");
    if (a.mask >> qubits[i] & 1u)
      throw Error("marginal: measured qubit " + std::to_string(qubits[i]) + " repeated");
    a.mask |= uint64_t(1) << qubits[i];
    a.pos[i] = qubits[i];
  }
  const uint64_t n_bins = uint64_t(1) << m, n_rest = uint64_t(1) << (n_qubits - m);
  CallBuffer out(sizeof(double) * n_bins);
  double *d_out = static_cast<double *>(out.p);
  const uint32_t n_hi = uint32_t(__builtin_popcountll(a.mask & ~uint64_t(63)));
  // Indices over the unmeasured positions >= 6 (64 amplitudes each).
  const uint64_t n_rest_hi = n_qubits >= 6 ? uint64_t(1) << (n_qubits - 6 - n_hi) : 0;
  if (n_rest_hi >= 16) {
    const uint64_t chunk = std::min<uint64_t>(
        n_rest_hi, std::max<uint64_t>(16, (uint64_t(1) << n_qubits) >> 20));
    const uint64_t n_chunks = n_rest_hi / chunk, blocks = (uint64_t(1) << n_hi) * n_chunks;
    std::unique_ptr<CallBuffer> part;
    double *d_part = d_out;
    if (n_chunks > 1) {
      part = std::make_unique<CallBuffer>(sizeof(double) * n_bins * n_chunks);
      d_part = static_cast<double *>(part->p);
    }
    if (precision == 0)
      hipLaunchKernelGGL(k_marginal_rest<float>, wide_grid(blocks), dim3(kRedBlock), 0, stream,
                         static_cast<const float *>(sv), a, chunk, d_part);
    else
      hipLaunchKernelGGL(k_marginal_rest<double>, wide_grid(blocks), dim3(kRedBlock), 0, stream,
                         static_cast<const double *>(sv), a, chunk, d_part);
    check(hipGetLastError(), "hipLaunchKernelGGL(k_marginal_rest)");
    if (n_chunks > 1) {
      hipLaunchKernelGGL(k_sum_partials, dim3(unsigned(n_bins)), dim3(kRedBlock), 0, stream,
                         d_part, unsigned(n_chunks), unsigned(n_bins), d_out);
      check(hipGetLastError(), "hipLaunchKernelGGL(k_sum_partials)");
    }
    check(hipMemcpyAsync(out_host, d_out, sizeof(double) * n_bins, hipMemcpyDeviceToHost, stream),
          "hipMemcpyAsync(marginal)");
    check(hipStreamSynchronize(stream), "hipStreamSynchronize(marginal)");
    return;
  }
  const unsigned grid = unsigned(std::min<uint64_t>((n_bins + kRedBlock - 1) / kRedBlock,
                                                    uint64_t(device_props().n_cu) * 16));
  if (precision == 0)
    hipLaunchKernelGGL(k_marginal_bin<float>, dim3(grid), dim3(kRedBlock), 0, stream,
                       static_cast<const float *>(sv), a, n_rest, d_out);
  else
    hipLaunchKernelGGL(k_marginal_bin<double>, dim3(grid), dim3(kRedBlock), 0, stream,
                       static_cast<const double *>(sv), a, n_rest, d_out);
  check(hipGetLastError(), "hipLaunchKernelGGL(k_marginal_bin)");
  check(hipMemcpyAsync(out_host, d_out, sizeof(double) * n_bins, hipMemcpyDeviceToHost, stream),
        "hipMemcpyAsync(marginal)");
  check(hipStreamSynchronize(stream), "hipStreamSynchronize(marginal)");
}

void dm_populations(const void *sv, uint32_t n_physical, uint8_t precision, double *out_host,
                    hipStream_t stream) {
  const uint64_t dim = uint64_t(1) << n_physical;
  CallBuffer out(sizeof(double) * dim);
  const unsigned grid = unsigned(std::min<uint64_t>((dim + kRedBlock - 1) / kRedBlock,
                                                    uint64_t(device_props().n_cu) * 16));
  if (precision == 0)
    hipLaunchKernelGGL(k_dm_diag<float>, dim3(grid), dim3(kRedBlock), 0, stream,
                       static_cast<const float *>(sv), n_physical, static_cast<double *>(out.p));
  else
    hipLaunchKernelGGL(k_dm_diag<double>, dim3(grid), dim3(kRedBlock), 0, stream,
                       static_cast<const double *>(sv), n_physical, static_cast<double *>(out.p));
  check(hipGetLastError(), "hipLaunchKernelGGL(k_dm_diag)");
  check(hipMemcpyAsync(out_host, out.p, sizeof(double) * dim, hipMemcpyDeviceToHost, stream),
        "hipMemcpyAsync(dm populations)");
  check(hipStreamSynchronize(stream), "hipStreamSynchronize(dm populations)");
}

} // namespace cast_hip
