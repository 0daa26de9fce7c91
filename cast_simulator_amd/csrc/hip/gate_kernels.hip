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
//                blockIdx.z (cast_hip_gate_apply_batch).
//   k_tile, k_tile3, k_mfma, k_mfma6, k_mfma_split: see their sections.
//
// Tile orders (wave tiles and matrix-core tiles): launch order, the low
// tile-index bits bit-reversed, XCD-contiguous (workgroups are dispatched
// to the 8 XCDs round-robin, so renumbering them gives each XCD one
// contiguous eighth of the tiles), or both.  Plan flags select them; the
// host autotuner tries each per gate; launches chosen automatically use the
// XCD order for wave tiles, f64 matrix cores and f32 split products from 32
// qubits on.

#include "cast_hip.h"
#include "common.hpp"
#include "device_core.hpp"

#include <algorithm>
#include <cstdlib>
#include <string>

namespace cast_hip {
namespace {

using dev::combo_base;
using dev::insert_zero;
using dev::ld_amp;
using dev::ld_stream;
using dev::local_delta;
using dev::NV;
using dev::st_amp;
using dev::st_stream;
using dev::V2;

// ── Dense: one combo per lane ───────────────────────────────────────────────

template <typename T, int KL, bool NTS>
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
    dev::dense_combo<T, KL, NTS>(amp, amp, mat, combo_base<KL>(c, g.q), delta);
}

// ── Wave tiles with cross-lane exchange of low-order targets ───────────────
//
// A wave owns a tile of 64 * K amplitudes: the 6 lowest address bits (one per
// lane), plus K register slots whose bits are, during load and store, the
// first L free (non-target) positions >= 6 ("F") followed by the targets
// >= 6.  Every load and store instruction therefore moves 64 consecutive
// amplitudes (1 KiB at f64) and the streaming hint is safe.  The L targets
// below 6 live in lane bits after the load; each is exchanged with register
// bit j (xor-shuffle across lanes, one exchange per register pair) so that
// register bits then enumerate exactly the K target states of one combo per
// lane.  Outputs are produced 2^L rows at a time, exchanged back, and stored.

struct TileGeom {
  uint64_t n_tiles;    // 2^(n - 6 - k)
  uint32_t ins[6];     // ascending insert positions (minus 6) of F and targets >= 6
  uint32_t reg_pos[6]; // address bit of register bit b during load/store
  uint32_t low_t[6];   // targets below 6, ascending
  uint32_t swz;        // 0, or the number of low tile-index bits visited bit-reversed
  uint32_t xcd_bits;   // 0, or log2(blocks): block b runs the tiles of block
                       // (b mod 8) * (blocks / 8) + b / 8 (contiguous per XCD)
};

// Tile visited by the wave `tile` (its launch-order index) under the
// geometry's orders; both are bijections on [0, n_tiles).
__device__ __forceinline__ uint64_t visit_order(uint64_t tile, const TileGeom &g) {
  if (g.xcd_bits) { // 4 waves per block, blocks = 2^xcd_bits >= 8
    const uint64_t b = tile >> 2;
    const uint64_t rb = ((b & 7u) << (g.xcd_bits - 3)) | (b >> 3);
    tile = (rb << 2) | (tile & 3u);
  }
  if (g.swz) { // n_tiles is a power of two >= 2^swz
    const uint64_t lo = tile & ((uint64_t(1) << g.swz) - 1);
    tile = (tile ^ lo) | (__brevll(lo) >> (64 - g.swz));
  }
  return tile;
}

template <typename T, int KL, int L>
__global__ void __launch_bounds__(256) k_tile(T *__restrict__ sv, const T *__restrict__ mat,
                                              TileGeom g) {
  const uint32_t lane = threadIdx.x & 63u;
  uint64_t tile =
      ((uint64_t(blockIdx.y) * gridDim.x + blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  if (tile >= g.n_tiles)
    return; // whole wave
  tile = visit_order(tile, g);
  uint64_t tb = tile;
#pragma unroll
  for (int b = 0; b < KL; ++b)
    tb = insert_zero(tb, g.ins[b]);
  T *__restrict__ p = sv + 2 * ((tb << 6) | lane);

  dev::tile_update<T, KL, L>(p, p, mat, g.reg_pos, g.low_t, lane);
}

// Three-product wave tiles (k = 4, 5): the same tile update with the real
// matrices [Re | Im | Re + Im].  At f64 and k = 5 the third input vector
// pushes the update past 256 VGPRs; the occupancy bound (two workgroups of
// four waves per CU, i.e. two waves per SIMD) keeps it at 256.
template <typename T, int KL, int L>
__global__ void __launch_bounds__(256, 2) k_tile3(T *__restrict__ sv, const T *__restrict__ m3,
                                                  TileGeom g) {
  const uint32_t lane = threadIdx.x & 63u;
  uint64_t tile =
      ((uint64_t(blockIdx.y) * gridDim.x + blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  if (tile >= g.n_tiles)
    return; // whole wave
  tile = visit_order(tile, g);
  uint64_t tb = tile;
#pragma unroll
  for (int b = 0; b < KL; ++b)
    tb = insert_zero(tb, g.ins[b]);
  T *__restrict__ p = sv + 2 * ((tb << 6) | lane);
  dev::tile_update<T, KL, L, true>(p, p, m3, g.reg_pos, g.low_t, lane);
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

// ── Matrix cores: complex matvec as three real MFMA products ───────────────
//
// For a tile of 16 consecutive combos the update is the complex product
// Out(K x 16) = M(K x K) * V(K x 16).  It is computed with three real
// products (Gauss):  t1 = Mr Vr,  t2 = Mi Vi,  t3 = (Mr + Mi)(Vr + Vi),
// Re Out = t1 - t2,  Im Out = t3 - t1 - t2  — 3 K^2 instead of 4 K^2
// multiply-adds.  Each product runs on v_mfma_*_16x16x4: lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16] and receives, in accumulator
// register i, D[4 (l / 16) + i][l % 16] (f32) or D[l / 16 + 4 i][l % 16]
// (f64) — layouts confirmed on gfx950 with a one-hot probe.  So lane l owns
// combo (l % 16) of the tile, loads the input amplitudes c = 4 s + l / 16
// (s = 0 .. K/4-1) and stores the rows its accumulators hold.  The three real
// matrices are loaded into registers once per wave and reused for every tile
// the wave processes.

using dev::Acc4;
using dev::mfma16;

template <typename T, int KL>
__global__ void __launch_bounds__(256) k_mfma(T *__restrict__ sv, const T *__restrict__ m3,
                                              GateGeom g, uint64_t n_tiles) {
  constexpr int K = 1 << KL, RB = K / 16, KS = K / 4;
  static_assert(RB >= 1, "MFMA path needs k >= 4");
  using VT = typename V2<T>::type;
  using AV = typename Acc4<T>::type;
  VT *__restrict__ amp = reinterpret_cast<VT *>(sv);
  const unsigned lane = threadIdx.x & 63u, j = lane >> 4, col = lane & 15u;

  T a[3][RB][KS];
#pragma unroll
  for (int m = 0; m < 3; ++m)
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
      for (int s = 0; s < KS; ++s)
        a[m][rb][s] = m3[m * K * K + (16 * rb + col) * K + 4 * s + j];
  uint64_t din[KS], dout[RB][4];
#pragma unroll
  for (int s = 0; s < KS; ++s)
    din[s] = local_delta<KL>(4 * s + j, g.q);
  // Accumulator register i of lane l holds row 4 (l / 16) + i for f32 and
  // row (l / 16) + 4 i for f64 (v_mfma_f64_16x16x4 interleaves its rows).
#pragma unroll
  for (int rb = 0; rb < RB; ++rb)
#pragma unroll
    for (int i = 0; i < 4; ++i)
      dout[rb][i] = local_delta<KL>(16 * rb + (sizeof(T) == 8 ? j + 4 * i : 4 * j + i), g.q);

  // Loads of a tile are issued after the previous tile's stores: issuing the
  // next tile's loads during this tile's products (software pipelining, 244
  // registers, still 2 waves per SIMD) measured 2.3 % slower at f64 k = 5.
  const uint64_t n_waves = (uint64_t(gridDim.x) * blockDim.x) >> 6;
  for (uint64_t t = dev::grid_wave(g.xcd); t < n_tiles; t += n_waves) {
    const uint64_t base = combo_base<KL>(dev::swizzled_tile(t, g.swz) * 16 + col, g.q);
    T br[KS], bi[KS], bs[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const VT x = ld_amp(amp + base + din[s]);
      br[s] = x.x;
      bi[s] = x.y;
      bs[s] = x.x + x.y;
    }
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
      AV t1 = {0, 0, 0, 0}, t2 = {0, 0, 0, 0}, t3 = {0, 0, 0, 0};
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        t1 = mfma16(a[0][rb][s], br[s], t1);
        t2 = mfma16(a[1][rb][s], bi[s], t2);
        t3 = mfma16(a[2][rb][s], bs[s], t3);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        VT o;
        o.x = t1[i] - t2[i];
        o.y = t3[i] - t1[i] - t2[i];
        st_amp(amp + base + dout[rb][i], o);
      }
    }
  }
}

// ── Matrix cores, k = 6: real matrices in LDS ──────────────────────────────
//
// Same three-product contraction as k_mfma for K = 64, where the A operands
// (3 x 4 x 16 values per lane) no longer fit in registers.  Re M and Im M are
// staged once per workgroup in LDS, laid out so that operand (m, rb, s) of
// lane l is element ((m * 4 + rb) * 16 + s) * 64 + l: every read is 64
// consecutive values (no bank conflicts).  Re M + Im M is formed on the fly,
// which keeps the f64 staging area at 64 KiB.  The 16 input amplitudes of a
// lane stay in registers while the four 16-row output blocks are produced.

constexpr unsigned kMfma6Block = 512;

template <typename T>
__global__ void __launch_bounds__(kMfma6Block) k_mfma6(T *__restrict__ sv, const T *__restrict__ m3,
                                                       GateGeom g, uint64_t n_tiles) {
  constexpr int KL = 6, K = 64, RB = 4, KS = 16;
  using VT = typename V2<T>::type;
  using AV = typename Acc4<T>::type;
  __shared__ T a_lds[2 * RB * KS * 64];
  for (unsigned i = threadIdx.x; i < 2 * RB * KS * 64; i += blockDim.x) {
    const unsigned l = i & 63u, s = (i >> 6) & (KS - 1), rb = (i >> 10) & (RB - 1), m = i >> 12;
    a_lds[i] = m3[m * K * K + (16 * rb + (l & 15u)) * K + 4 * s + (l >> 4)];
  }
  __syncthreads();
  VT *__restrict__ amp = reinterpret_cast<VT *>(sv);
  const unsigned lane = threadIdx.x & 63u, j = lane >> 4, col = lane & 15u;
  uint64_t din[KS];
#pragma unroll
  for (int s = 0; s < KS; ++s)
    din[s] = local_delta<KL>(4 * s + j, g.q);
  const uint64_t n_waves = (uint64_t(gridDim.x) * blockDim.x) >> 6;
  for (uint64_t t = dev::grid_wave(g.xcd); t < n_tiles;
       t += n_waves) {
    const uint64_t base = combo_base<KL>(dev::swizzled_tile(t, g.swz) * 16 + col, g.q);
    T br[KS], bi[KS], bs[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const VT x = ld_amp(amp + base + din[s]);
      br[s] = x.x;
      bi[s] = x.y;
      bs[s] = x.x + x.y;
    }
    // Two 16-row output blocks at a time: six independent accumulator
    // chains keep the matrix cores busy between dependent MFMAs.
#pragma unroll
    for (int rp = 0; rp < RB; rp += 2) {
      // Re-read per tile (not hoisted out of the tile loop into 128+ VGPRs).
      const T *a = a_lds + rp * KS * 64 + lane + dev::opaque_u32(0);
      AV t1[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}}, t2[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}},
         t3[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
#pragma unroll
      for (int s = 0; s < KS; ++s) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const T ar = a[h * KS * 64 + s * 64];
          const T ai = a[RB * KS * 64 + h * KS * 64 + s * 64];
          t1[h] = mfma16(ar, br[s], t1[h]);
          t2[h] = mfma16(ai, bi[s], t2[h]);
          t3[h] = mfma16(ar + ai, bs[s], t3[h]);
        }
      }
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          VT o;
          o.x = t1[h][i] - t2[h][i];
          o.y = t3[h][i] - t1[h][i] - t2[h][i];
          const uint32_t row = 16 * (rp + h) + (sizeof(T) == 8 ? j + 4 * i : 4 * j + i);
          st_amp(amp + base + local_delta<KL>(row, g.q), o);
        }
    }
  }
}

// ── Matrix cores, f32 k = 5, 6 through bf16 split products ────────────────
//
// On the f32 matrix cores (157 TFLOP/s) a dense 6-qubit f32 gate is
// compute-bound: AI 32 flop/B against a ridge of 19.7.  The bf16 matrix
// cores run 16x the flops; each f32 operand is split into three bf16 parts
// and each real product takes six partial products (dev::split_tile), which
// keeps f32 accuracy.  One wave per 16-combo tile in a grid-stride loop;
// the splits of Re M and Im M are staged in LDS once per workgroup.

constexpr unsigned kSplitBlock = 256;

// One tile per wave step (two per step at k = 5 measured -0.6 % at 30
// qubits and +1.8 % at 34 on the config-3 circuit's gates).
template <int KL>
__global__ void __launch_bounds__(kSplitBlock) k_mfma_split(float *__restrict__ sv,
                                                          const float *__restrict__ m3,
                                                          GateGeom g, uint64_t n_tiles) {
  __shared__ dev::bf16x8_t a_lds[dev::SplitOperands<KL>::kFragments];
  dev::split_stage<KL>(m3, a_lds);
  __syncthreads();
  float2 *__restrict__ amp = reinterpret_cast<float2 *>(sv);
  const dev::SplitLane<KL> sl(g.q);
  const uint64_t n_waves = (uint64_t(gridDim.x) * blockDim.x) >> 6;
  for (uint64_t t = dev::grid_wave(g.xcd); t < n_tiles;
       t += n_waves)
    dev::split_tile<KL>(amp, amp, a_lds, g.q, sl, dev::swizzled_tile(t, g.swz));
}

// ── Batched dense: item = blockIdx.z ────────────────────────────────────────

constexpr int kMaxBatch = 64;
struct BatchArgs {
  BatchItem it[kMaxBatch];
};

template <typename T, int KL>
__global__ void __launch_bounds__(256) k_dense_batch(BatchArgs a) {
  constexpr int K = 1 << KL;
  using VT = typename V2<T>::type;
  const BatchItem &it = a.it[blockIdx.z];
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
  const uint64_t c = (uint64_t(blockIdx.y) * gridDim.x + blockIdx.x) * blockDim.x + threadIdx.x;
  if (c < it.combos)
    dev::dense_combo<T, KL>(amp, amp, mat, combo_base<KL>(c, q), delta);
}

// ── Launch helpers ──────────────────────────────────────────────────────────

constexpr unsigned kBlock = 256;

bool tile_gauss(const GatePlan &p); // defined with the family rule below
bool auto_choice(const GatePlan &p);

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
  // Nontemporal stores at f64 when consecutive lanes store whole 128-byte
  // lines: no target below bit 3 (measured on k = 5 gates: up to 17 % faster
  // there, up to 2x slower with a target at bit 0 or 1).
  uint32_t qmin = 64;
  for (uint32_t b = 0; b < g.k; ++b)
    qmin = std::min(qmin, g.q[b]);
  if (sizeof(T) == 8 && qmin >= 3)
    hipLaunchKernelGGL((k_dense<T, KL, true>), full_grid(g.combos), dim3(kBlock), 0, s,
                       static_cast<T *>(sv), static_cast<const T *>(p.d_mat), g);
  else
    hipLaunchKernelGGL((k_dense<T, KL, false>), full_grid(g.combos), dim3(kBlock), 0, s,
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

bool use_split(const GatePlan &p);
template <int KL>
void launch_mfma_split(const GatePlan &p, hipStream_t s, void *sv, const GateGeom &g);

// Tile order of a matrix-core launch over n_tiles (a power of two) tiles:
// XCD-contiguous (GateGeom::xcd) when the plan asks for the XCD order
// (CAST_HIP_FLAG_TILE_XCD), when `xcd_auto` is set and the family was chosen
// automatically, or, for measurements, with CAST_HIP_MFMA_XCD=1; the grid is
// then rounded down to a multiple of 8 (the kernels' grid-stride loops
// cover every tile).  The low 10 tile-index bits bit-reversed (GateGeom::swz)
// when the plan asks for CAST_HIP_FLAG_TILE_SWIZZLE.
unsigned mfma_order(const GatePlan &p, unsigned grid, uint64_t n_tiles, GateGeom &g,
                    bool xcd_auto = false) {
  static const bool env = [] {
    const char *e = std::getenv("CAST_HIP_MFMA_XCD");
    return e && std::string(e) == "1";
  }();
  g.xcd = 0;
  g.swz = 0;
  const bool xcd = env || (p.flags & CAST_HIP_FLAG_TILE_XCD) || (xcd_auto && auto_choice(p));
  if (xcd && grid >= 8) {
    g.xcd = 1;
    grid &= ~7u;
  }
  if ((p.flags & CAST_HIP_FLAG_TILE_SWIZZLE) && n_tiles >= 2) {
    uint32_t m = 0;
    while ((uint64_t(2) << m) <= n_tiles)
      ++m; // n_tiles = 2^m
    g.swz = std::min<uint32_t>(10, m);
  }
  return grid;
}

template <typename T, int KL>
void launch_mfma_t(const GatePlan &p, hipStream_t s, void *sv, const GateGeom &g0) {
  GateGeom g = g0;
  if constexpr (sizeof(T) == 4 && KL == 5) {
    if (use_split(p))
      return launch_mfma_split<5>(p, s, sv, g);
  }
  // Grid: the three real matrices are loaded into registers once per wave,
  // and long grid-stride loops stream worse than full grids.  At f64 the
  // grid gives each wave about 8 tiles, between 8 and 256 workgroups per
  // CU; at f32, 8 per CU.  Measured on the 29 k = 5 gates of the 30-qubit
  // f64 headline circuit: 4898 GB/s aggregate at 8 per CU (256 tiles per
  // wave), 5118 at 32, 5255 at 64, 5232-5271 at 128, 5312 at 192, 5372 at
  // 256 (8 tiles per wave), 5286 at 384, 5049 at 512; the serving workload
  // (12-28 qubits, f32 and f64): 79.3 circuits/s with 8 per CU, 72.3 with
  // 128.
  static const unsigned per_cu_env = [] {
    const char *e = std::getenv("CAST_HIP_MFMA_BLOCKS_PER_CU");
    long v = e && *e ? std::strtol(e, nullptr, 10) : 0;
    return unsigned(v > 0 && v <= 1024 ? v : 0);
  }();
  const uint64_t tiles = g.combos / 16;
  const uint64_t need = (tiles + 3) / 4; // 4 waves per block
  const uint64_t n_cu = uint64_t(std::max(1, device_props().n_cu));
  uint64_t cap = n_cu * 8;
  if (per_cu_env)
    cap = n_cu * per_cu_env;
  else if (sizeof(T) == 8)
    cap = std::min(n_cu * 256, std::max(cap, tiles / 32));
  // f64 launches chosen automatically visit tiles in XCD order: on the
  // config-3 circuit's k = 5 gates 5970 -> 6189 GB/s at 33 qubits, and
  // +0.5-0.7 % on the 30-qubit headline circuit's k = 5 gates.
  const unsigned grid = mfma_order(p, unsigned(std::max<uint64_t>(1, std::min(need, cap))), tiles,
                                   g, sizeof(T) == 8);
  hipLaunchKernelGGL((k_mfma<T, KL>), dim3(grid), dim3(kBlock), 0, s, static_cast<T *>(sv),
                     static_cast<const T *>(p.d_mat3), g, tiles);
  check(hipGetLastError(), "hipLaunchKernelGGL(k_mfma)");
}

// k = 5, 6 at f32 on the matrix cores: bf16 split products (k_mfma_split)
// or f32 products (k_mfma<float, 5>, k_mfma6<float>).  A plan flag forces
// one; otherwise CAST_HIP_K6_F32 / CAST_HIP_K5_F32 ("split" or "f32", for
// measurements) or the default below.
bool use_split(const GatePlan &p) {
  if (p.flags & CAST_HIP_FLAG_MFMA_SPLIT)
    return true;
  if (p.flags & CAST_HIP_FLAG_MFMA_F32)
    return false;
  auto read = [](const char *name) {
    const char *e = std::getenv(name);
    if (e && std::string(e) == "split")
      return 1;
    if (e && std::string(e) == "f32")
      return 0;
    return -1;
  };
  static const int env6 = read("CAST_HIP_K6_F32"), env5 = read("CAST_HIP_K5_F32");
  const int env = p.k == 6 ? env6 : env5;
  // Default: the split products, which measured both more accurate (rel L2
  // 6-9e-8 per gate against 1.6-2.1e-7) and faster on MI355X: over five
  // target sets, k = 6 at 34 qubits 5017 GB/s against 4508 (the f32 products
  // are matrix-core bound at 159 algorithmic TFLOP/s), k = 5 at 34 qubits
  // 5286 against 4957, and at 30 qubits 4351 / 4180 (k = 6) and 4489 / 4405
  // (k = 5).  High target sets can run faster with f32 products; the
  // per-gate autotuner tries both.
  return env >= 0 ? env == 1 : true;
}

template <int KL>
void launch_mfma_split(const GatePlan &p, hipStream_t s, void *sv, const GateGeom &g0) {
  GateGeom g = g0;
  const uint64_t tiles = g.combos / 16;
  const uint64_t need = (tiles + kSplitBlock / 64 - 1) / (kSplitBlock / 64);
  // Grid (each workgroup stages the split matrices once):
  //  * k = 6: at most 48 workgroups per CU (48 KiB of LDS per workgroup, 3
  //    resident per CU).  At 30 qubits on the config-3 circuit's 9 k = 6
  //    gates, caps of 3, 8, 16, 48, 128, 256, 512 per CU gave 4541 / 4544 /
  //    4588 / 4633 / 4446 / 4316 / 4020 GB/s.
  //  * k = 5: about 4 tiles per wave, between 48 and 512 workgroups per CU.
  //    Caps of 48, 128, 256, 512, 1024 and 2048 (one tile per wave) per CU
  //    gave 4828 / 4886 / 4953 / 5009-5027 / 5057 / 4830 GB/s on the 29 k = 5
  //    gates of the 30-qubit headline circuit at f32, and 5313-5489 at 512
  //    against 5346 at 48 on the 34-qubit config-3 gates.
  // CAST_HIP_SPLIT_BLOCKS_PER_CU fixes the cap for measurements.
  static const unsigned per_cu_env = [] {
    const char *e = std::getenv("CAST_HIP_SPLIT_BLOCKS_PER_CU");
    long v = e && *e ? std::strtol(e, nullptr, 10) : 0;
    return unsigned(v > 0 && v <= 4096 ? v : 0);
  }();
  const uint64_t n_cu = uint64_t(std::max(1, device_props().n_cu));
  uint64_t cap = n_cu * 48;
  if (per_cu_env)
    cap = n_cu * per_cu_env;
  else if (KL == 5)
    cap = std::min(n_cu * 512, std::max(cap, tiles / 16));
  // Automatically chosen launches on 32 qubits or more visit tiles in XCD
  // order: on the config-3 circuit's gates at 34 qubits k = 5 5345 -> 5419
  // GB/s, k = 6 4906 -> 4933; at 30 qubits it cost 1 % (k = 5).
  const bool large = tiles >= (uint64_t(1) << (32 - KL - 4));
  const unsigned grid = mfma_order(p, unsigned(std::max<uint64_t>(1, std::min(need, cap))), tiles,
                                   g, large);
  hipLaunchKernelGGL((k_mfma_split<KL>), dim3(grid), dim3(kSplitBlock), 0, s,
                     static_cast<float *>(sv), static_cast<const float *>(p.d_mat3), g, tiles);
  check(hipGetLastError(), "hipLaunchKernelGGL(k_mfma_split)");
}

template <typename T>
void launch_mfma6_t(const GatePlan &p, hipStream_t s, void *sv, const GateGeom &g0) {
  GateGeom g = g0;
  if constexpr (sizeof(T) == 4) {
    if (use_split(p))
      return launch_mfma_split<6>(p, s, sv, g);
  }
  const uint64_t tiles = g.combos / 16;
  const uint64_t need = (tiles + kMfma6Block / 64 - 1) / (kMfma6Block / 64);
  // 2 workgroups per CU are resident (LDS).  The grid gives each wave about
  // 256 tiles, between 2 and 32 workgroups per CU (each workgroup stages
  // the matrices in LDS once).  Measured: k = 6 at 34 qubits f32 136.3 ->
  // 138.5 TFLOP/s and at 33 qubits f64 74.8 -> 75.4 going from 2 to 32 per
  // CU, while the serving workload (12-28 qubits) served 79.4 circuits/s
  // with 2 per CU and 77.4 with 32.  CAST_HIP_MFMA6_BLOCKS_PER_CU fixes the
  // cap for measurements.
  static const unsigned per_cu_env = [] {
    const char *e = std::getenv("CAST_HIP_MFMA6_BLOCKS_PER_CU");
    long v = e && *e ? std::strtol(e, nullptr, 10) : 0;
    return unsigned(v > 0 && v <= 1024 ? v : 0);
  }();
  const uint64_t n_cu = uint64_t(std::max(1, device_props().n_cu));
  const uint64_t cap = per_cu_env ? n_cu * per_cu_env
                                  : std::min(n_cu * 32, std::max(n_cu * 2, tiles / 2048));
  // f64 launches chosen automatically visit tiles in XCD order: on the
  // config-3 circuit's k = 6 gates 4319 -> 4687 GB/s at 28 qubits, 4633 ->
  // 4920 at 30, unchanged at 26 and 33 (matrix-core bound there); f32 (the
  // f32-product form) keeps launch order.
  const unsigned grid = mfma_order(p, unsigned(std::max<uint64_t>(1, std::min(need, cap))), tiles,
                                   g, sizeof(T) == 8);
  hipLaunchKernelGGL((k_mfma6<T>), dim3(grid), dim3(kMfma6Block), 0, s, static_cast<T *>(sv),
                     static_cast<const T *>(p.d_mat3), g, tiles);
  check(hipGetLastError(), "hipLaunchKernelGGL(k_mfma6)");
}

template <typename T, int KL, int L>
void launch_tile_kl(const GatePlan &p, hipStream_t s, void *sv, const TileGeom &tg) {
  const uint64_t blocks = (tg.n_tiles + 3) / 4; // 4 waves per block, one tile per wave
  // Three-product form for k = 4, 5 when asked for (needs the real matrices).
  if constexpr (KL == 4 || KL == 5) {
    if (tile_gauss(p)) {
      hipLaunchKernelGGL((k_tile3<T, KL, L>), full_grid(blocks * kBlock), dim3(kBlock), 0, s,
                         static_cast<T *>(sv), static_cast<const T *>(p.d_mat3), tg);
      check(hipGetLastError(), "hipLaunchKernelGGL(k_tile, three products)");
      return;
    }
  }
  hipLaunchKernelGGL((k_tile<T, KL, L>), full_grid(blocks * kBlock), dim3(kBlock), 0, s,
                     static_cast<T *>(sv), static_cast<const T *>(p.d_mat), tg);
  check(hipGetLastError(), "hipLaunchKernelGGL(k_tile)");
}

template <typename T, int KL>
void launch_tile_k(const GatePlan &p, hipStream_t s, void *sv, const TileGeom &tg, uint32_t L) {
  switch (L) {
  case 0:
    return launch_tile_kl<T, KL, 0>(p, s, sv, tg);
  case 1:
    if constexpr (KL >= 1)
      return launch_tile_kl<T, KL, 1>(p, s, sv, tg);
    break;
  case 2:
    if constexpr (KL >= 2)
      return launch_tile_kl<T, KL, 2>(p, s, sv, tg);
    break;
  case 3:
    if constexpr (KL >= 3)
      return launch_tile_kl<T, KL, 3>(p, s, sv, tg);
    break;
  case 4:
    if constexpr (KL >= 4)
      return launch_tile_kl<T, KL, 4>(p, s, sv, tg);
    break;
  case 5:
    if constexpr (KL >= 5)
      return launch_tile_kl<T, KL, 5>(p, s, sv, tg);
    break;
  default:
    break;
  }
  throw Error("wave-tile kernel: invalid low-target count");
}

// Wave-tile geometry for a gate on an n-qubit statevector (n >= k + 6).
TileGeom tile_geom(const GatePlan &p, uint32_t n, uint32_t *n_low) {
  TileGeom tg{};
  tg.n_tiles = uint64_t(1) << (n - 6 - p.k);
  const dev::TileLayout t = dev::tile_layout(p.q, p.k);
  for (int b = 0; b < 6; ++b) {
    tg.ins[b] = t.ins[b];
    tg.reg_pos[b] = t.reg_pos[b];
    tg.low_t[b] = t.low_t[b];
  }
  const uint32_t L = t.n_low;
  if (p.flags & CAST_HIP_FLAG_TILE_SWIZZLE)
    tg.swz = std::min<uint32_t>(10, n - 6 - p.k);
  // blocks = n_tiles / 4 (4 waves per block, one tile each), a power of two.
  // Automatically chosen tiles (no order flag) use the XCD order too: on the
  // config-3 circuit's gates forced to wave tiles it streamed 5504 -> 6360
  // GB/s (k = 3) and 5524 -> 6316 (k = 4) at 33 qubits f64, 5507 -> 6474
  // (k = 3) at 34 qubits f32, and +1.4 to +5.7 % on the 30-qubit headline
  // circuit's k = 3, 4 gates; the per-gate autotuner still tries all three.
  const bool xcd = (p.flags & CAST_HIP_FLAG_TILE_XCD) ||
                   (auto_choice(p) && !(p.flags & CAST_HIP_FLAG_TILE_SWIZZLE));
  if (xcd && tg.n_tiles >= 32)
    tg.xcd_bits = n - 6 - p.k - 2;
  *n_low = L;
  return tg;
}

template <typename T>
void launch_tile(const GatePlan &p, hipStream_t s, void *sv, uint32_t n) {
  uint32_t L = 0;
  const TileGeom tg = tile_geom(p, n, &L);
  switch (p.k) {
  case 1:
    return launch_tile_k<T, 1>(p, s, sv, tg, L);
  case 2:
    return launch_tile_k<T, 2>(p, s, sv, tg, L);
  case 3:
    return launch_tile_k<T, 3>(p, s, sv, tg, L);
  case 4:
    return launch_tile_k<T, 4>(p, s, sv, tg, L);
  case 5:
    return launch_tile_k<T, 5>(p, s, sv, tg, L);
  default:
    throw Error("wave-tile kernel supports k <= 5");
  }
}

// Kernel family for a dense plan on an n-qubit statevector.  A forced family
// falls back to the direct kernel where its shape requirement is not met
// (wave tiles need n >= k + 6, matrix-core tiles k in {4, 5, 6} and n >= k + 4).
// Whether an automatic choice may pick the family (variant AUTO, or DENSE
// not forced).
bool auto_choice(const GatePlan &p) { return p.variant == 0 || (p.variant == 1 && !p.forced); }

uint32_t count_low(const GatePlan &p) {
  uint32_t n_low = 0;
  for (uint32_t b = 0; b < p.k; ++b)
    n_low += p.q[b] < 6;
  return n_low;
}

uint32_t dense_family(const GatePlan &p, uint32_t n) {
  const bool tile_ok = p.k <= 5 && n >= p.k + 6;
  const bool mfma_ok = p.d_mat3 != nullptr && n >= p.k + 4;
  const uint32_t n_low = count_low(p);
  uint32_t v = p.variant;
  if (auto_choice(p)) {
    // Automatic choice, from measurements on MI355X (the per-gate autotuner
    // refines it where it runs):
    //  * wave tiles stream full rows and win up to 3 low-order targets for
    //    k <= 4; beyond that the exchanged register groups cost more than
    //    the matrix-core kernel's 16-combo tiles;
    //  * k = 5: the matrix cores (f32: bf16 split products) for any
    //    targets: on the config-3 circuit's k = 5 gates without autotuning,
    //    wave tiles ran 4979 GB/s at 33 qubits f64 against 6089 on the
    //    matrix cores, and 4849 at 34 qubits f32 against 5214; on the
    //    30-qubit headline circuit the autotuner keeps the matrix cores for
    //    most k = 5 gates (19 of 29 at f64, 21 of 29 at f32);
    //  * k = 6: only the matrix-core kernel keeps its operands out of
    //    scratch memory.
    if (p.k >= 5 && mfma_ok)
      v = 4;
    else if (tile_ok && n_low <= 3)
      v = 3;
    else if (mfma_ok)
      v = 4;
    else
      v = tile_ok ? 3 : 1;
  }
  if (v == 3 && !tile_ok)
    v = 1;
  if (v == 4 && !mfma_ok)
    v = tile_ok ? 3 : 1;
  return v;
}

// Three-product form of the wave tiles: asked for by flag, or chosen
// automatically where it measured faster (k = 4; f32 k = 5).
bool tile_gauss(const GatePlan &p) {
  if (!p.d_mat3 || (p.k != 4 && p.k != 5))
    return false;
  if (p.flags & CAST_HIP_FLAG_TILE_GAUSS)
    return true;
  return auto_choice(p) && (p.k == 4 || p.precision == 0);
}

template <typename T>
void launch_by_k(const GatePlan &p, hipStream_t s, void *sv, const GateGeom &g) {
  const bool mono = p.variant == 2;
  if (!mono) {
    const uint32_t n = p.k + uint32_t(63 - __builtin_clzll(g.combos));
    const uint32_t fam = dense_family(p, n);
    if (fam == 4) {
      if (p.k == 4)
        return launch_mfma_t<T, 4>(p, s, sv, g);
      if (p.k == 5)
        return launch_mfma_t<T, 5>(p, s, sv, g);
      return launch_mfma6_t<T>(p, s, sv, g);
    }
    if (fam == 3)
      return launch_tile<T>(p, s, sv, n);
  }
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
  // Combo blocks over x and y as for a single gate (full_grid), item in z.
  dim3 grid = full_grid(max_combos);
  grid.z = n;
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

uint32_t plan_family(const GatePlan &plan, uint32_t n_qubits) {
  if (plan.variant == 2)
    return 2;
  return dense_family(plan, n_qubits);
}

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
