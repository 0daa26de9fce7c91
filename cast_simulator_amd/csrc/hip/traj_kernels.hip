// Persistent batched-trajectory kernel for gfx950.
//
// One workgroup owns one statevector slot and takes whole tasks from a
// global counter until the list is exhausted; no workgroup ever waits on
// another, so any grid size drains.  A task is (checkpoint, gate list): the
// first applied gate reads the checkpoint statevector and writes the slot,
// every later gate updates the slot in place, with a workgroup barrier
// between gates.  Gates use wave tiles (coalesced 64-amplitude rows, cross-
// lane exchange of targets below bit 6) when n >= k + 6, else one combo per
// thread.  A wave keeps 16 amplitude loads in flight per lane whatever k is
// (16 / K tiles per step for k <= 4), since one workgroup alone must cover
// the memory latency of its statevector.  Gates of k = 4 to 6 (k = 5, 6 in
// f32 only) use the three-product form, the pool holding [Re M | Im M | Re M
// + Im M]: k = 4 on the VALU, k = 5 and 6 on the matrix cores.  At the end the workgroup reduces, in f64, the
// squared norm and sum |amp|^2 (-1)^(bit a xor bit b) for each requested
// pair (a, b), and can copy the final state out.
//
// The gate update is the one of src/cpp/cuda/cuda_gen.cpp:241-283 (local
// basis i sets bit q[b] for bit b of i; new[r] = sum_c M[r][c] old[c]).

#include "cast_hip.h"
#include "common.hpp"
#include "device_core.hpp"

#include <string>

namespace cast_hip {
namespace {

using dev::combo_base;
using dev::insert_zero;
using dev::local_delta;
using dev::V2;

constexpr int kMaxPairs = 32;
constexpr int kMaxBlock = 512;
constexpr int kMaxWaves = kMaxBlock / 64;
constexpr uint32_t kMaxTrajGateQubits = 6; // 4 at f64

// Layouts the host bindings rely on (no padding surprises across compilers).
static_assert(sizeof(cast_hip_traj_gate_t) == 12, "cast_hip_traj_gate_t layout");
static_assert(sizeof(cast_hip_traj_task_t) == 16, "cast_hip_traj_task_t layout");
static_assert(sizeof(cast_hip_traj_launch_t) == 376, "cast_hip_traj_launch_t layout");

struct TrajScalars {
  uint64_t n_matrix_elems;
  uint32_t n_tasks, n_gates, n_checkpoints, n_qubits, n_pairs, log2_block, flags;
  uint32_t pairs[2 * kMaxPairs];
};

// Gate operands staged in LDS by the k = 6 (f32 products) and the split-
// product gates, one buffer shared by both: 48 KiB at f32, none at f64.
template <typename T> constexpr size_t kGateLdsBytes = sizeof(T) == 4 ? 48 * 1024 : 16;

__global__ void k_fill_u32(uint32_t *__restrict__ p, uint32_t n, uint32_t v) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    p[i] = v;
}

using dev::opaque_u32;

// One gate over the whole statevector by the workgroup, wave tiles.  q[] are
// the ascending targets, of which exactly L are below 6.
template <typename T, int KL, int L>
__device__ __forceinline__ void wg_tile_gate(const T *src, T *dst, const T *__restrict__ mat,
                                             const uint32_t *q, uint32_t n) {
  constexpr int U = KL >= 4 ? 1 : (16 >> KL); // tiles per step
  // Tile layout: register bit j < L -> the j-th non-target position >= 6,
  // register bit j >= L -> target q[j] (targets >= 6); the tile index is
  // deposited around the ascending set of those positions.
  uint64_t tmask = 0;
#pragma unroll
  for (int b = 0; b < KL; ++b)
    tmask |= uint64_t(1) << q[b];
  uint32_t reg_pos[KL], low_t[L > 0 ? L : 1], ins[KL];
  uint64_t f = ~tmask & ~uint64_t(63), fmask = 0;
#pragma unroll
  for (int j = 0; j < L; ++j) {
    low_t[j] = q[j];
    reg_pos[j] = uint32_t(__builtin_ctzll(f));
    fmask |= uint64_t(1) << reg_pos[j];
    f &= f - 1;
  }
#pragma unroll
  for (int j = L; j < KL; ++j)
    reg_pos[j] = q[j];
  uint64_t im = fmask | (tmask & ~uint64_t(63));
#pragma unroll
  for (int b = 0; b < KL; ++b) {
    ins[b] = uint32_t(__builtin_ctzll(im)) - 6;
    im &= im - 1;
  }

  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, n_waves = blockDim.x >> 6;
  const uint64_t n_tiles = uint64_t(1) << (n - 6 - KL);
  auto tile_ptr = [&](uint64_t tile) {
#pragma unroll
    for (int b = 0; b < KL; ++b)
      tile = insert_zero(tile, ins[b]);
    return 2 * ((tile << 6) | lane);
  };
  if (n_tiles >= uint64_t(U)) {
    // n_tiles and U are powers of two, so every step is whole.
    for (uint64_t t0 = uint64_t(wave) * U; t0 < n_tiles; t0 += uint64_t(n_waves) * U) {
      // Re-derived each step so the K*K matrix entries are read by scalar
      // loads inside the loop rather than hoisted into VGPRs.
      const T *m = mat + opaque_u32(0);
      const T *s[U];
      T *d[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint64_t off = tile_ptr(t0 + u);
        s[u] = src + off;
        d[u] = dst + off;
      }
      dev::tile_step<T, KL, L, U, (KL >= 4)>(s, d, m, reg_pos, low_t, lane);
    }
  } else {
    for (uint64_t t = wave; t < n_tiles; t += n_waves) {
      const T *m = mat + opaque_u32(0);
      const uint64_t off = tile_ptr(t);
      dev::tile_update<T, KL, L, (KL >= 4)>(src + off, dst + off, m, reg_pos, low_t, lane);
    }
  }
}

// Small statevectors (n < k + 6): one combo per thread.
template <typename T, int KL>
__device__ __forceinline__ void wg_dense_gate(const T *src, T *dst, const T *__restrict__ mat,
                                              const uint32_t *q, uint32_t n) {
  constexpr int K = 1 << KL;
  using VT = typename V2<T>::type;
  uint64_t delta[K];
#pragma unroll
  for (int i = 0; i < K; ++i)
    delta[i] = local_delta<KL>(uint32_t(i), q);
  const uint64_t combos = uint64_t(1) << (n - KL);
  for (uint64_t c = threadIdx.x; c < combos; c += blockDim.x) {
    const T *m = mat + opaque_u32(0);
    const VT *in = reinterpret_cast<const VT *>(src);
    VT *out = reinterpret_cast<VT *>(dst);
    const uint64_t base = combo_base<KL>(c, q);
    if constexpr (KL >= 4) { // three real matrices [Re | Im | Re + Im]
      T vr[K], vi[K];
#pragma unroll
      for (int i = 0; i < K; ++i) {
        const VT x = in[base + delta[i]];
        vr[i] = x.x;
        vi[i] = x.y;
      }
#pragma unroll
      for (int r = 0; r < K; ++r) {
        T t1 = T(0), t2 = T(0), t3 = T(0);
#pragma unroll
        for (int cc = 0; cc < K; ++cc) {
          t1 = dev::fma_t(m[r * K + cc], vr[cc], t1);
          t2 = dev::fma_t(m[K * K + r * K + cc], vi[cc], t2);
          t3 = dev::fma_t(m[2 * K * K + r * K + cc], vr[cc] + vi[cc], t3);
        }
        VT o;
        o.x = t1 - t2;
        o.y = t3 - t1 - t2;
        out[base + delta[r]] = o;
      }
    } else {
      dev::dense_combo<T, KL>(in, out, m, base, delta);
    }
  }
}

// k = 5 at f32 on the matrix cores: the three-product contraction of the
// plain-path matrix-core kernel (Out = M V over tiles of 16 combos with
// v_mfma_f32_16x16x4), run by the workgroup's waves over the statevector.
// Lane l owns combo l % 16 of a tile and loads inputs 4 s + l / 16; the
// three real K x K matrices ([Re | Im | Re + Im] from the pool) stay in
// registers for the whole gate.  Two tiles per step keep 16 loads per lane
// in flight.  Any target set: addresses come from the combo's bit deposit.
// (The VALU three-product update of k = 5 costs 12 flops per byte at f32,
// above the 9.8 flop/B VALU ridge of the chip; here it is MFMA work.)
template <int KL>
__device__ __forceinline__ void wg_mfma_gate(const float *src, float *dst,
                                             const float *__restrict__ m3, const uint32_t *q,
                                             uint32_t n) {
  constexpr int K = 1 << KL, RB = K / 16, KS = K / 4, U = 2;
  using VT = typename V2<float>::type;
  using AV = dev::Acc4<float>::type;
  const VT *in = reinterpret_cast<const VT *>(src);
  VT *out = reinterpret_cast<VT *>(dst);
  const uint32_t lane = threadIdx.x & 63u, j = lane >> 4, col = lane & 15u;
  const uint32_t wave = threadIdx.x >> 6, n_waves = blockDim.x >> 6;
  float a[3][RB][KS];
#pragma unroll
  for (int m = 0; m < 3; ++m)
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
      for (int s = 0; s < KS; ++s)
        a[m][rb][s] = m3[m * K * K + (16 * rb + col) * K + 4 * s + j];
  uint64_t din[KS];
#pragma unroll
  for (int s = 0; s < KS; ++s)
    din[s] = local_delta<KL>(4 * s + j, q);
  // accumulator register i of lane l holds output row 16 rb + 4 (l / 16) + i
  uint64_t dout[RB][4];
#pragma unroll
  for (int rb = 0; rb < RB; ++rb)
#pragma unroll
    for (int i = 0; i < 4; ++i)
      dout[rb][i] = local_delta<KL>(16 * rb + 4 * j + i, q);
  // n >= KL + 5, so n_tiles = 2^(n - KL - 4) is a multiple of U
  const uint64_t n_tiles = uint64_t(1) << (n - KL - 4);
  for (uint64_t t0 = uint64_t(wave) * U; t0 < n_tiles; t0 += uint64_t(n_waves) * U) {
    uint64_t base[U];
    float br[U][KS], bi[U][KS];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      base[u] = combo_base<KL>((t0 + u) * 16 + col, q);
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        const VT x = in[base[u] + din[s]];
        br[u][s] = x.x;
        bi[u][s] = x.y;
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int rb = 0; rb < RB; ++rb) {
        AV t1 = {0, 0, 0, 0}, t2 = {0, 0, 0, 0}, t3 = {0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < KS; ++s) {
          t1 = dev::mfma16(a[0][rb][s], br[u][s], t1);
          t2 = dev::mfma16(a[1][rb][s], bi[u][s], t2);
          t3 = dev::mfma16(a[2][rb][s], br[u][s] + bi[u][s], t3);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          VT o;
          o.x = t1[i] - t2[i];
          o.y = t3[i] - t1[i] - t2[i];
          out[base[u] + dout[rb][i]] = o;
        }
      }
  }
}

// k = 6 at f32 on the matrix cores, as the plain path's k = 6 kernel: Re M
// and Im M (32 KiB) are staged in LDS once per gate, laid out so that the
// operand (m, rb, s) of lane l is element ((m * 4 + rb) * 16 + s) * 64 + l
// (every read is 64 consecutive values); Re M + Im M is formed on the fly.
// A lane owns combo l % 16 of a 16-combo tile and keeps its 16 inputs in
// registers while two 16-row output blocks are produced per pass (six
// independent accumulator chains).  Needs n >= 12 (checked by the caller).
// The caller's barrier after every gate keeps the next gate's staging from
// overwriting operands still in use.
__device__ __forceinline__ void wg_mfma6_gate(const float *src, float *dst,
                                              const float *__restrict__ m3, const uint32_t *q,
                                              uint32_t n, float *a_lds) {
  constexpr int KL = 6, K = 64, RB = 4, KS = 16;
  using VT = typename V2<float>::type;
  using AV = dev::Acc4<float>::type;
  for (uint32_t i = threadIdx.x; i < 2 * RB * KS * 64; i += blockDim.x) {
    const uint32_t l = i & 63u, s = (i >> 6) & (KS - 1), rb = (i >> 10) & (RB - 1), m = i >> 12;
    a_lds[i] = m3[m * K * K + (16 * rb + (l & 15u)) * K + 4 * s + (l >> 4)];
  }
  __syncthreads();
  const VT *in = reinterpret_cast<const VT *>(src);
  VT *out = reinterpret_cast<VT *>(dst);
  const uint32_t lane = threadIdx.x & 63u, j = lane >> 4, col = lane & 15u;
  const uint32_t wave = threadIdx.x >> 6, n_waves = blockDim.x >> 6;
  uint64_t din[KS];
#pragma unroll
  for (int s = 0; s < KS; ++s)
    din[s] = local_delta<KL>(4 * s + j, q);
  const uint64_t n_tiles = uint64_t(1) << (n - KL - 4);
  for (uint64_t t = wave; t < n_tiles; t += n_waves) {
    const uint64_t base = combo_base<KL>(t * 16 + col, q);
    float br[KS], bi[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const VT x = in[base + din[s]];
      br[s] = x.x;
      bi[s] = x.y;
    }
#pragma unroll
    for (int rp = 0; rp < RB; rp += 2) {
      // Re-read per tile (not hoisted out of the tile loop into registers).
      const float *a = a_lds + rp * KS * 64 + lane + opaque_u32(0);
      AV t1[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}}, t2[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}},
         t3[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
#pragma unroll
      for (int s = 0; s < KS; ++s)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const float ar = a[h * KS * 64 + s * 64];
          const float ai = a[RB * KS * 64 + h * KS * 64 + s * 64];
          t1[h] = dev::mfma16(ar, br[s], t1[h]);
          t2[h] = dev::mfma16(ai, bi[s], t2[h]);
          t3[h] = dev::mfma16(ar + ai, br[s] + bi[s], t3[h]);
        }
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          VT o;
          o.x = t1[h][i] - t2[h][i];
          o.y = t3[h][i] - t1[h][i] - t2[h][i];
          out[base + local_delta<KL>(16 * (rp + h) + 4 * j + i, q)] = o;
        }
    }
  }
}

// k = 5, 6 at f32 through bf16 split products (dev::split_tile): the splits
// of Re M and Im M are staged in LDS once per gate (k = 6: 48 KiB, k = 5:
// 12 KiB of the shared gate buffer); the workgroup's waves take 16-combo
// tiles in turn.  Needs n >= KL + 4.  The caller's barrier after every gate
// keeps the next gate's staging from overwriting operands still in use.
template <int KL>
__device__ __forceinline__ void wg_split_gate(const float *src, float *dst,
                                              const float *__restrict__ m3, const uint32_t *q,
                                              uint32_t n, dev::bf16x8_t *a_lds) {
  // k = 5: two tiles per step, so that each lane keeps 16 loads in flight as
  // at k = 6 (one workgroup streams the whole statevector).
  constexpr int U = KL == 5 ? 2 : 1;
  dev::split_stage<KL>(m3, a_lds);
  __syncthreads();
  const dev::SplitLane<KL> sl(q);
  const uint32_t wave = threadIdx.x >> 6, n_waves = blockDim.x >> 6;
  // n >= KL + 5 for k = 5 (checked by the caller), so n_tiles is a multiple of U
  const uint64_t n_tiles = uint64_t(1) << (n - KL - 4);
  for (uint64_t t = uint64_t(wave) * U; t < n_tiles; t += uint64_t(n_waves) * U)
    dev::split_tile<KL, U>(reinterpret_cast<const float2 *>(src),
                           reinterpret_cast<float2 *>(dst), a_lds, q, sl, t);
}

template <typename T, int KL>
__device__ __forceinline__ void wg_gate_k(const T *src, T *dst, const T *__restrict__ mat,
                                          const uint32_t *q, uint32_t n) {
  if (n < uint32_t(KL) + 6)
    return wg_dense_gate<T, KL>(src, dst, mat, q, n);
  uint32_t n_low = 0;
#pragma unroll
  for (int b = 0; b < KL; ++b)
    n_low += q[b] < 6;
  if constexpr (KL == 5) { // f32 only: matrix cores, any targets
    if constexpr (sizeof(T) == 4)
      wg_mfma_gate<KL>(src, dst, mat, q, n);
  } else {
    switch (n_low) {
    case 0:
      return wg_tile_gate<T, KL, 0>(src, dst, mat, q, n);
    case 1:
      if constexpr (KL >= 1)
        return wg_tile_gate<T, KL, 1>(src, dst, mat, q, n);
      break;
    case 2:
      if constexpr (KL >= 2)
        return wg_tile_gate<T, KL, 2>(src, dst, mat, q, n);
      break;
    case 3:
      if constexpr (KL == 3 || KL == 4)
        return wg_tile_gate<T, KL, 3>(src, dst, mat, q, n);
      break;
    case 4:
      if constexpr (KL == 4)
        return wg_tile_gate<T, KL, 4>(src, dst, mat, q, n);
      break;
    default:
      break;
    }
  }
}

// Squared norm and the pair sums of the workgroup's final state, written to
// out_row[0 .. 1 + n_pairs).  Amplitude i = tid + B r (B = blockDim.x = 2^bB).
// The sign of pair (a, b) at i factors into a per-thread part (positions
// below bB), applied once at the end, and a part that depends only on the row
// r: the parity of r & m for the pair's row mask m, computed with scalar
// instructions since r is uniform across the workgroup.  The inner loop then
// does one f64 FMA per pair with a uniform +-1; 16 rows are loaded at a time.
template <typename T>
__device__ __forceinline__ void wg_observe(const T *sv, uint32_t n, uint32_t n_pairs, uint32_t bB,
                           const uint32_t *s_pairs, double *out_row, double *red) {
  using VT = typename V2<T>::type;
  constexpr int R = 16;
  const VT *amp = reinterpret_cast<const VT *>(sv);
  const uint32_t tid = threadIdx.x;
  uint32_t rmask[kMaxPairs];
  uint32_t tneg = 0; // bit j: the per-thread sign of pair j is -1
#pragma unroll
  for (int j = 0; j < kMaxPairs; ++j) {
    rmask[j] = 0;
    if (uint32_t(j) < n_pairs) {
      const uint32_t qa = __builtin_amdgcn_readfirstlane(s_pairs[2 * j]);
      const uint32_t qb = __builtin_amdgcn_readfirstlane(s_pairs[2 * j + 1]);
      rmask[j] = (qa >= bB ? 1u << (qa - bB) : 0u) ^ (qb >= bB ? 1u << (qb - bB) : 0u);
      const uint32_t tb = (qa < bB ? (tid >> qa) & 1u : 0u) ^ (qb < bB ? (tid >> qb) & 1u : 0u);
      tneg |= tb << j;
    }
  }
  double acc[1 + kMaxPairs];
#pragma unroll
  for (int j = 0; j < 1 + kMaxPairs; ++j)
    acc[j] = 0.0;
  const uint64_t len = uint64_t(1) << n;
  const uint32_t rows = n > bB ? (1u << (n - bB)) : 1u;
  for (uint32_t r0 = 0; r0 < rows; r0 += R) {
    double p[R];
#pragma unroll
    for (int u = 0; u < R; ++u) {
      const uint64_t i = tid + (uint64_t(r0 + u) << bB);
      p[u] = 0.0;
      if (r0 + u < rows && i < len) {
        const VT x = amp[i];
        const double re = double(x.x), im = double(x.y);
        p[u] = re * re + im * im;
      }
    }
#pragma unroll
    for (int u = 0; u < R; ++u) {
      const uint32_t r = r0 + u;
      acc[0] += p[u];
      // All 32 slots are accumulated (unused ones have mask 0 and are never
      // read): no per-pair guard for the compiler to turn into selects.
#pragma unroll
      for (int j = 0; j < kMaxPairs; ++j) {
        const double sg = (__builtin_popcount(r & rmask[j]) & 1) ? -1.0 : 1.0;
        acc[1 + j] = __builtin_fma(sg, p[u], acc[1 + j]);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < kMaxPairs; ++j)
    if ((tneg >> j) & 1u)
      acc[1 + j] = -acc[1 + j];

  // Workgroup reduction in a fixed order: xor-shuffle within each wave, then
  // thread j sums entry j over the waves.
  const uint32_t lane = tid & 63u, wave = tid >> 6, n_waves = blockDim.x >> 6;
  const uint32_t cnt = 1 + n_pairs;
#pragma unroll
  for (int j = 0; j < 1 + kMaxPairs; ++j) {
    if (uint32_t(j) < cnt) {
      double x = acc[j];
#pragma unroll
      for (int m = 32; m >= 1; m >>= 1)
        x += __shfl_xor(x, m, 64);
      if (lane == 0)
        red[wave * (1 + kMaxPairs) + j] = x;
    }
  }
  __syncthreads();
  if (tid < cnt) {
    double s = 0.0;
    for (uint32_t w = 0; w < n_waves; ++w)
      s += red[w * (1 + kMaxPairs) + tid];
    out_row[tid] = s;
  }
  __syncthreads();
}

// Built for OCC waves per SIMD (second launch-bound argument): 2 allows 256
// VGPRs (no spills, one 512-thread workgroup per CU), 4 allows 128 (spills,
// but two workgroups per CU keep twice the loads in flight).
template <typename T, int OCC>
__global__ void __launch_bounds__(kMaxBlock, OCC)
    k_traj(const cast_hip_traj_task_t *__restrict__ tasks, const uint32_t *__restrict__ order,
           const cast_hip_traj_gate_t *__restrict__ gates, const T *__restrict__ matrices,
           const T *__restrict__ checkpoints, T *__restrict__ slots, double *__restrict__ out,
           T *__restrict__ dump, uint32_t *__restrict__ work, TrajScalars a) {
  __shared__ uint32_t s_task;
  __shared__ uint32_t s_pairs[2 * kMaxPairs];
  __shared__ double red[kMaxWaves * (1 + kMaxPairs)];
  __shared__ __align__(16) unsigned char gate_lds[kGateLdsBytes<T>];
  const bool f32_products = a.flags & CAST_HIP_TRAJ_FLAG_MFMA_F32;
  if (threadIdx.x < 2 * a.n_pairs)
    s_pairs[threadIdx.x] = a.pairs[threadIdx.x];
  for (;;) {
    // The leading barrier keeps every path into the lane-0 fetch below
    // convergent: without it the compiler may thread lanes that skipped a
    // lane-0-only branch straight to the next barrier, into a loop that waits
    // for a task index lane 0 never gets to fetch.
    __syncthreads();
    if (threadIdx.x == 0)
      s_task = atomicAdd(&work[0], 1u);
    __syncthreads();
    const uint32_t t = __builtin_amdgcn_readfirstlane(s_task);
    __syncthreads();
    if (t >= a.n_tasks)
      break;
    const uint32_t n = opaque_u32(a.n_qubits);
    const uint64_t len = uint64_t(1) << n;
    T *slot = slots + 2 * len * blockIdx.x;
    const uint32_t ti = __builtin_amdgcn_readfirstlane(order[t]);
    bool task_ok = ti < a.n_tasks;
    cast_hip_traj_task_t task{};
    if (task_ok) {
      const cast_hip_traj_task_t tk = tasks[ti];
      task.first_gate = __builtin_amdgcn_readfirstlane(tk.first_gate);
      task.n_gates = __builtin_amdgcn_readfirstlane(tk.n_gates);
      task.checkpoint = __builtin_amdgcn_readfirstlane(tk.checkpoint);
      task_ok = task.checkpoint < a.n_checkpoints && task.first_gate <= a.n_gates &&
                task.n_gates <= a.n_gates - task.first_gate;
    }
    if (!task_ok) {
      atomicOr(&work[1], 1u); // every lane: no lane-dependent branch here
      continue;
    }
    const T *cur = checkpoints + 2 * len * task.checkpoint; // state the next gate reads
    for (uint32_t g = 0; g < task.n_gates; ++g) {
      // The descriptor is the same for the whole workgroup; readfirstlane
      // states that, so geometry and matrix address live in SGPRs.
      const cast_hip_traj_gate_t d = gates[task.first_gate + g];
      const uint32_t k = __builtin_amdgcn_readfirstlane(d.k);
      const uint32_t moff = __builtin_amdgcn_readfirstlane(d.matrix_offset);
      uint32_t q[kMaxTrajGateQubits];
      bool ok = k >= 1 && k <= (sizeof(T) == 4 ? kMaxTrajGateQubits : 4u);
      ok = ok && (k < 6 || n >= 12); // the k = 6 path needs 64 tiles of 16 combos
#pragma unroll
      for (uint32_t b = 0; b < kMaxTrajGateQubits; ++b) {
        q[b] = __builtin_amdgcn_readfirstlane(d.q[b]);
        if (b < k)
          ok = ok && q[b] < n && (b == 0 || q[b] > q[b - 1]);
      }
      // k = 4, 5 entries hold three real K x K matrices: 3 K^2 / 2 complex slots.
      const uint64_t slots = k >= 4 ? uint64_t(3) << (2 * k - 1) : (uint64_t(1) << (2 * k));
      ok = ok && uint64_t(moff) + slots <= a.n_matrix_elems;
      if (!ok) {
        atomicOr(&work[1], 1u);
        continue;
      }
      const T *mat = matrices + 2 * uint64_t(moff);
      switch (k) {
      case 1:
        wg_gate_k<T, 1>(cur, slot, mat, q, n);
        break;
      case 2:
        wg_gate_k<T, 2>(cur, slot, mat, q, n);
        break;
      case 3:
        wg_gate_k<T, 3>(cur, slot, mat, q, n);
        break;
      case 4:
        wg_gate_k<T, 4>(cur, slot, mat, q, n);
        break;
      case 5: // f32 only (checked above)
        if constexpr (sizeof(T) == 4) {
          if (f32_products || n < 11)
            wg_gate_k<T, 5>(cur, slot, mat, q, n);
          else
            wg_split_gate<5>(cur, slot, mat, q, n, reinterpret_cast<dev::bf16x8_t *>(gate_lds));
        }
        break;
      default: // k = 6, f32 only, n >= 12 (checked above)
        if constexpr (sizeof(T) == 4) {
          if (f32_products)
            wg_mfma6_gate(cur, slot, mat, q, n, reinterpret_cast<float *>(gate_lds));
          else
            wg_split_gate<6>(cur, slot, mat, q, n, reinterpret_cast<dev::bf16x8_t *>(gate_lds));
        }
        break;
      }
      cur = slot;
      __syncthreads();
    }
    wg_observe<T>(cur, n, a.n_pairs, a.log2_block, s_pairs, out + uint64_t(ti) * (1 + a.n_pairs),
                  red);
    if (dump) {
      using VT = typename V2<T>::type;
      const VT *src = reinterpret_cast<const VT *>(cur);
      VT *dst = reinterpret_cast<VT *>(dump + 2 * len * ti);
      for (uint64_t i0 = threadIdx.x; i0 < len; i0 += uint64_t(blockDim.x) * 8) {
        VT x[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const uint64_t i = i0 + uint64_t(u) * blockDim.x;
          if (i < len)
            x[u] = src[i];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const uint64_t i = i0 + uint64_t(u) * blockDim.x;
          if (i < len)
            dst[i] = x[u];
        }
      }
      __syncthreads();
    }
  }
}

template <typename T, int OCC>
void launch_traj_t(const cast_hip_traj_launch_t &L, const TrajScalars &a, hipStream_t s) {
  hipLaunchKernelGGL((k_traj<T, OCC>), dim3(L.n_slots), dim3(L.block), 0, s,
                     reinterpret_cast<const cast_hip_traj_task_t *>(L.tasks),
                     reinterpret_cast<const uint32_t *>(L.order),
                     reinterpret_cast<const cast_hip_traj_gate_t *>(L.gates),
                     reinterpret_cast<const T *>(L.matrices),
                     reinterpret_cast<const T *>(L.checkpoints), reinterpret_cast<T *>(L.slots),
                     reinterpret_cast<double *>(L.out), reinterpret_cast<T *>(L.dump),
                     reinterpret_cast<uint32_t *>(L.work), a);
  check(hipGetLastError(), "hipLaunchKernelGGL(k_traj)");
}

} // namespace

void launch_traj(const cast_hip_traj_launch_t &L, hipStream_t stream) {
  if (L.precision > 1)
    throw Error("trajectory launch: precision must be 0 (F32) or 1 (F64)");
  if (L.n_qubits < 1 || L.n_qubits > 34)
    throw Error("trajectory launch: n_qubits must be in 1..34");
  if (L.n_slots < 1 || L.n_slots > (1u << 20))
    throw Error("trajectory launch: n_slots must be in 1..2^20");
  if (L.block != 64 && L.block != 128 && L.block != 256 && L.block != 512)
    throw Error("trajectory launch: block must be 64, 128, 256 or 512");
  if (L.n_pairs > kMaxPairs)
    throw Error("trajectory launch: at most 32 qubit pairs");
  if (L.waves_per_simd != 0 && L.waves_per_simd != 2 && L.waves_per_simd != 4)
    throw Error("trajectory launch: waves_per_simd must be 0, 2 or 4");
  for (uint32_t j = 0; j < L.n_pairs; ++j) {
    const uint32_t qa = L.pairs[2 * j], qb = L.pairs[2 * j + 1];
    if (qa >= L.n_qubits || qb >= L.n_qubits || qa == qb)
      throw Error("trajectory launch: pair " + std::to_string(j) + " = (" + std::to_string(qa) +
                  ", " + std::to_string(qb) + ") is not two distinct qubits below n");
  }
  if (!L.work)
    throw Error("trajectory launch: work buffer is null");
  if (L.n_tasks > 0 && (!L.tasks || !L.order || !L.out || !L.slots || !L.checkpoints ||
                        L.n_checkpoints == 0))
    throw Error("trajectory launch: null task, order, output, slot or checkpoint buffer");
  if (L.n_gates > 0 && (!L.gates || !L.matrices))
    throw Error("trajectory launch: null gate table or matrix pool");
  uint32_t *work = reinterpret_cast<uint32_t *>(L.work);
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, stream, work, 2u, 0u);
  check(hipGetLastError(), "hipLaunchKernelGGL(k_fill_u32)");
  if (L.n_tasks == 0)
    return;
  TrajScalars a{};
  a.n_matrix_elems = L.n_matrix_elems;
  a.n_tasks = L.n_tasks;
  a.n_gates = L.n_gates;
  a.n_checkpoints = L.n_checkpoints;
  a.n_qubits = L.n_qubits;
  a.n_pairs = L.n_pairs;
  a.log2_block = uint32_t(__builtin_ctz(L.block));
  if (L.flags & ~CAST_HIP_TRAJ_FLAG_MFMA_F32)
    throw Error("trajectory launch: unknown flags " + std::to_string(L.flags));
  a.flags = L.flags;
  for (uint32_t j = 0; j < 2 * L.n_pairs; ++j)
    a.pairs[j] = L.pairs[j];
  const bool occ4 = L.waves_per_simd == 4;
  if (L.precision == 0)
    occ4 ? launch_traj_t<float, 4>(L, a, stream) : launch_traj_t<float, 2>(L, a, stream);
  else
    occ4 ? launch_traj_t<double, 4>(L, a, stream) : launch_traj_t<double, 2>(L, a, stream);
}

} // namespace cast_hip
