// Device-side building blocks shared by the gate kernels and the persistent
// trajectory kernel: bit-insert addressing and the wave-tile gate update.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cast_hip {
namespace dev {

// Index of the calling wave in a 1-D grid of workgroups.  With `xcd` set
// (grid a multiple of 8) workgroup b is renumbered (b mod 8) * (grid / 8) +
// b / 8: workgroups are dispatched to the 8 XCDs round-robin, so each XCD
// then owns a contiguous range of wave indices.  A bijection on the waves.
__device__ __forceinline__ uint64_t grid_wave(uint32_t xcd) {
  uint64_t b = blockIdx.x;
  if (xcd)
    b = (b & 7u) * (gridDim.x >> 3) + (b >> 3);
  return (b * blockDim.x + threadIdx.x) >> 6;
}

// Tile visited at loop index t: the low `swz` bits of t bit-reversed (0 =
// launch order).  A bijection on [0, 2^m) for every m >= swz.
__device__ __forceinline__ uint64_t swizzled_tile(uint64_t t, uint32_t swz) {
  if (!swz)
    return t;
  const uint64_t lo = t & ((uint64_t(1) << swz) - 1);
  return (t ^ lo) | (__brevll(lo) >> (64 - swz));
}


template <typename T> struct V2;
template <> struct V2<float> {
  using type = float2;
};
template <> struct V2<double> {
  using type = double2;
};

typedef float f2v __attribute__((ext_vector_type(2)));
typedef double d2v __attribute__((ext_vector_type(2)));
template <typename T> struct NV;
template <> struct NV<float> {
  using type = f2v;
};
template <> struct NV<double> {
  using type = d2v;
};

template <typename T>
__device__ __forceinline__ typename NV<T>::type ld_stream(const T *p) {
  return __builtin_nontemporal_load(reinterpret_cast<const typename NV<T>::type *>(p));
}
template <typename T>
__device__ __forceinline__ void st_stream(T *p, typename NV<T>::type v) {
  __builtin_nontemporal_store(v, reinterpret_cast<typename NV<T>::type *>(p));
}

// One interleaved amplitude of a gather kernel (dense and matrix-core
// paths): nontemporal at f64, plain at f32.  Every gate touches each
// amplitude once per pass; measured on the MI355X, nontemporal accesses make
// the f64 k = 5 matrix-core and dense kernels 3-5 % faster, but the f32
// matrix-core kernels, whose 16-combo rows are single 128-byte lines, up to
// 27 % slower.
__device__ __forceinline__ float2 ld_amp(const float2 *p) { return *p; }
__device__ __forceinline__ double2 ld_amp(const double2 *p) {
  const d2v v = __builtin_nontemporal_load(reinterpret_cast<const d2v *>(p));
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ void st_amp(float2 *p, float2 v) { *p = v; }
__device__ __forceinline__ void st_amp(double2 *p, double2 v) {
  const d2v w = {v.x, v.y};
  __builtin_nontemporal_store(w, reinterpret_cast<d2v *>(p));
}

// Fused multiply-add in the operand precision (__builtin_fma is the double
// builtin: float operands would be promoted and rounded back).
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) {
  return __builtin_fma(a, b, c);
}

// A wave-uniform value the compiler cannot see through (kept in an SGPR).
// Adding opaque_u32(0) to a pointer keeps its provenance (so loads through it
// stay scalar loads) while stopping the compiler from hoisting or batching
// the loads across the point where it is taken.
__device__ __forceinline__ uint32_t opaque_u32(uint32_t v) {
  asm volatile("" : "+s"(v));
  return v;
}

__device__ __forceinline__ uint64_t insert_zero(uint64_t x, uint32_t q) {
  const uint64_t lo_mask = (uint64_t(1) << q) - 1;
  return ((x & ~lo_mask) << 1) | (x & lo_mask);
}

// Deposit the combo index into the non-target bit positions (the
// bit-deposit of src/cpp/cuda/cuda_gen.cpp:174-211).  Targets are ascending,
// so inserting zeros from the lowest position up is exact.
template <int KL>
__device__ __forceinline__ uint64_t combo_base(uint64_t c, const uint32_t *q) {
#pragma unroll
  for (int b = 0; b < KL; ++b)
    c = insert_zero(c, q[b]);
  return c;
}

// Address offset of local basis state i: bit b of i sets bit q[b]
// (cuda_gen.cpp:241-246).
template <int KL>
__device__ __forceinline__ uint64_t local_delta(uint32_t i, const uint32_t *q) {
  uint64_t d = 0;
#pragma unroll
  for (int b = 0; b < KL; ++b)
    d |= uint64_t((i >> b) & 1u) << q[b];
  return d;
}

// Dense update of one combo: K amplitudes read from `src`, K written to
// `dst` (src == dst for in-place), new[r] = sum_c M[r][c] old[c].  NTS:
// nontemporal stores at f64, which pay only when a wave's stores cover whole
// 128-byte lines (no target below bit 3); otherwise plain stores.
template <typename T, int KL, bool NTS = false>
__device__ __forceinline__ void dense_combo(const typename V2<T>::type *src,
                                            typename V2<T>::type *dst, const T *__restrict__ mat,
                                            uint64_t base, const uint64_t *delta) {
  constexpr int K = 1 << KL;
  using VT = typename V2<T>::type;
  T vr[K], vi[K];
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const VT x = ld_amp(src + base + delta[i]);
    vr[i] = x.x;
    vi[i] = x.y;
  }
#pragma unroll
  for (int r = 0; r < K; ++r) {
    T ar0 = T(0), ar1 = T(0), ai0 = T(0), ai1 = T(0);
#pragma unroll
    for (int c = 0; c < K; ++c) {
      const T mr = mat[2 * (r * K + c)];
      const T mi = mat[2 * (r * K + c) + 1];
      ar0 = fma_t(mr, vr[c], ar0);
      ar1 = fma_t(mi, vi[c], ar1);
      ai0 = fma_t(mr, vi[c], ai0);
      ai1 = fma_t(mi, vr[c], ai1);
    }
    VT out;
    out.x = ar0 - ar1;
    out.y = ai0 + ai1;
    if constexpr (NTS)
      st_amp(dst + base + delta[r], out);
    else
      dst[base + delta[r]] = out;
  }
}

// Wave tiles.  A wave owns 64 * K amplitudes: the 6 lowest address bits
// (one per lane) and K register slots.  Register bit b lives at address bit
// reg_pos[b] during load and store, so every load and store instruction moves
// 64 consecutive amplitudes.  The L targets below bit 6 (low_t) live in lane
// bits after the load; each is exchanged with register bit j by xor shuffles
// so that the registers of a lane then hold the K target states of one combo.
// Outputs are produced 2^L rows at a time, exchanged back, and stored.

// U wave tiles per step (U = 1 for the one-tile-per-wave gate kernel; more
// in loops where one workgroup must keep enough loads in flight on its own).
// src[u] / dst[u] point at scalar 2 * (tile_base | lane) of tile u (equal for
// in place).  All loads of the step are issued before any store.  Rows are the
// outer loop and tiles the inner one, so one matrix row (2 K scalars, read by
// scalar loads) is live at a time whatever U is.  Register arrays are local
// and indexed by constants only, so they stay in VGPRs.
//
// G3 selects the three-product form: `mat` then holds three real K x K
// matrices [Re M | Im M | Re M + Im M] and each output needs 3 K FMAs instead
// of 4 K (Re = Mr vr - Mi vi, Im = (Mr + Mi)(vr + vi) - Mr vr - Mi vi).
template <typename T, int KL, int L, int U, bool G3 = false>
__device__ __forceinline__ void tile_step(const T *const *src, T *const *dst,
                                          const T *__restrict__ mat, const uint32_t *reg_pos,
                                          const uint32_t *low_t, uint32_t lane) {
  constexpr int K = 1 << KL, G = 1 << L;
  T vr[U][K], vi[U][K];
#pragma unroll
  for (int u = 0; u < U; ++u)
#pragma unroll
    for (int i = 0; i < K; ++i) {
      const auto x = ld_stream(src[u] + 2 * local_delta<KL>(uint32_t(i), reg_pos));
      vr[u][i] = x.x;
      vi[u][i] = x.y;
    }
  // lane bit low_t[j]  <->  register bit j
#pragma unroll
  for (int j = 0; j < L; ++j) {
    const int m = 1 << low_t[j];
    const bool up = (lane >> low_t[j]) & 1u;
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int r = 0; r < K; ++r) {
        if (r & (1 << j))
          continue;
        const int r1 = r | (1 << j);
        const T sr = up ? vr[u][r] : vr[u][r1], si = up ? vi[u][r] : vi[u][r1];
        const T gr = __shfl_xor(sr, m, 64), gi = __shfl_xor(si, m, 64);
        if (up) {
          vr[u][r] = gr;
          vi[u][r] = gi;
        } else {
          vr[u][r1] = gr;
          vi[u][r1] = gi;
        }
      }
  }
  T vs[G3 ? U : 1][G3 ? K : 1];
  if constexpr (G3) {
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int c = 0; c < K; ++c)
        vs[u][c] = vr[u][c] + vi[u][c];
  }
#pragma unroll
  for (int h = 0; h < K / G; ++h) {
    T orr[U][G], oi[U][G];
#pragma unroll
    for (int q = 0; q < G; ++q) {
      const int r = h * G + q;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if constexpr (G3) {
          T t1 = T(0), t2 = T(0), t3 = T(0);
#pragma unroll
          for (int c = 0; c < K; ++c) {
            t1 = fma_t(mat[r * K + c], vr[u][c], t1);
            t2 = fma_t(mat[K * K + r * K + c], vi[u][c], t2);
            t3 = fma_t(mat[2 * K * K + r * K + c], vs[u][c], t3);
          }
          orr[u][q] = t1 - t2;
          oi[u][q] = t3 - t1 - t2;
        } else {
          // Four independent accumulators: dependent FMA chains of length K.
          T ar0 = T(0), ar1 = T(0), ai0 = T(0), ai1 = T(0);
#pragma unroll
          for (int c = 0; c < K; ++c) {
            const T mr = mat[2 * (r * K + c)];
            const T mi = mat[2 * (r * K + c) + 1];
            ar0 = fma_t(mr, vr[u][c], ar0);
            ar1 = fma_t(mi, vi[u][c], ar1);
            ai0 = fma_t(mr, vi[u][c], ai0);
            ai1 = fma_t(mi, vr[u][c], ai1);
          }
          orr[u][q] = ar0 - ar1;
          oi[u][q] = ai0 + ai1;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < L; ++j) {
      const int m = 1 << low_t[j];
      const bool up = (lane >> low_t[j]) & 1u;
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int q = 0; q < G; ++q) {
          if (q & (1 << j))
            continue;
          const int q1 = q | (1 << j);
          const T sr = up ? orr[u][q] : orr[u][q1], si = up ? oi[u][q] : oi[u][q1];
          const T gr = __shfl_xor(sr, m, 64), gi = __shfl_xor(si, m, 64);
          if (up) {
            orr[u][q] = gr;
            oi[u][q] = gi;
          } else {
            orr[u][q1] = gr;
            oi[u][q1] = gi;
          }
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int q = 0; q < G; ++q) {
        typename NV<T>::type o;
        o.x = orr[u][q];
        o.y = oi[u][q];
        st_stream(dst[u] + 2 * local_delta<KL>(uint32_t(h * G + q), reg_pos), o);
      }
  }
}

// One wave-tile update (src == dst for in place).
template <typename T, int KL, int L, bool G3 = false>
__device__ __forceinline__ void tile_update(const T *src, T *dst, const T *__restrict__ mat,
                                            const uint32_t *reg_pos, const uint32_t *low_t,
                                            uint32_t lane) {
  const T *const s[1] = {src};
  T *const d[1] = {dst};
  tile_step<T, KL, L, 1, G3>(s, d, mat, reg_pos, low_t, lane);
}

// Wave-tile geometry of a gate with ascending targets q[0..k) on an n-qubit
// statevector (n >= k + 6).  F = the first L non-target positions >= 6.
struct TileLayout {
  uint32_t ins[6];     // ascending insert positions (minus 6) of F and targets >= 6
  uint32_t reg_pos[6]; // address bit of register bit b during load/store
  uint32_t low_t[6];   // targets below 6
  uint32_t n_low;
};

__host__ __device__ inline TileLayout tile_layout(const uint32_t *q, uint32_t k) {
  TileLayout t{};
  uint32_t L = 0, h = 0, high[6] = {};
  for (uint32_t b = 0; b < k; ++b) {
    if (q[b] < 6)
      t.low_t[L++] = q[b];
    else
      high[h++] = q[b];
  }
  uint32_t nf = 0;
  for (uint32_t pos = 6; nf < L; ++pos) {
    bool is_t = false;
    for (uint32_t b = 0; b < k; ++b)
      is_t = is_t || q[b] == pos;
    if (!is_t)
      t.reg_pos[nf++] = pos;
  }
  for (uint32_t j = 0; j < h; ++j)
    t.reg_pos[L + j] = high[j];
  uint32_t ins[6];
  for (uint32_t b = 0; b < k; ++b)
    ins[b] = t.reg_pos[b];
  for (uint32_t a = 1; a < k; ++a) // insertion sort (k <= 6)
    for (uint32_t b = a; b > 0 && ins[b - 1] > ins[b]; --b) {
      const uint32_t x = ins[b];
      ins[b] = ins[b - 1];
      ins[b - 1] = x;
    }
  for (uint32_t b = 0; b < k; ++b)
    t.ins[b] = ins[b] - 6;
  t.n_low = L;
  return t;
}

// v_mfma_{f32,f64}_16x16x4: lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16]; accumulator register i of lane l holds
// D[4 (l / 16) + i][l % 16] (f32) or D[l / 16 + 4 i][l % 16] (f64).
template <typename T> struct Acc4;
template <> struct Acc4<float> {
  using type = float __attribute__((ext_vector_type(4)));
};
template <> struct Acc4<double> {
  using type = double __attribute__((ext_vector_type(4)));
};

__device__ __forceinline__ Acc4<float>::type mfma16(float a, float b, Acc4<float>::type c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ Acc4<double>::type mfma16(double a, double b, Acc4<double>::type c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// ── bf16 split products (f32 dense k = 5, 6 on the bf16 matrix cores) ─────
//
// Every f32 operand x is split into three bf16 parts x = x0 + x1 + x2 + r,
// |r| <= 2^-27 |x| (each part the round-to-nearest bf16 of what the earlier
// parts leave; those differences are exact in f32).  A product a b is the
// sum of the six partial products a_s b_t with s + t <= 2 (the dropped ones
// are below 2^-24 |a b|) on v_mfma_f32_16x16x32_bf16 with f32 accumulation,
// smallest first.  The complex product is four real products (Re M vr,
// Im M vi, and Re M vi + Im M vr into one accumulator), so only the splits
// of Re M and Im M are staged in LDS (k = 5: 12 KiB, k = 6: 48 KiB).
//
// v_mfma_f32_16x16x32_bf16: lane l supplies A[l % 16][8 (l / 16) + e] and
// B[8 (l / 16) + e][l % 16] as element e of an 8-vector and receives, in
// accumulator register i, D[4 (l / 16) + i][l % 16].  Lane l owns combo
// l % 16 of a 16-combo tile, loads input rows 32 ks + 8 (l / 16) + e and
// stores output rows 16 rb + 4 (l / 16) + i.

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4_t mfma_bf16(bf16x8_t a, bf16x8_t b, f32x4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

struct Split3 {
  __bf16 p[3];
};

__device__ __forceinline__ Split3 split3(float x) {
  Split3 s;
  s.p[0] = __bf16(x);
  const float r1 = x - float(s.p[0]);
  s.p[1] = __bf16(r1);
  const float r2 = r1 - float(s.p[1]);
  s.p[2] = __bf16(r2);
  return s;
}

template <int KL> struct SplitOperands {
  static_assert(KL == 5 || KL == 6, "bf16 split products for k = 5, 6");
  static constexpr int K = 1 << KL, RB = K / 16, KS = K / 32;
  // fragment (matrix m: 0 = Re, 1 = Im; split s; row block rb; k step ks)
  // of lane l at index (((m * 3 + s) * RB + rb) * KS + ks) * 64 + l: every
  // read is 64 consecutive 16-byte vectors
  static constexpr int kFragments = 2 * 3 * RB * KS * 64;
  __device__ static constexpr int at(int m, int s, int rb, int ks) {
    return (((m * 3 + s) * RB + rb) * KS + ks) * 64;
  }
};

// Stage the splits of Re M and Im M (from the three real matrices
// [Re | Im | Re + Im], K x K row-major each) into a_lds; every thread of the
// workgroup takes part, and the caller synchronises before use.
template <int KL>
__device__ __forceinline__ void split_stage(const float *__restrict__ m3, bf16x8_t *a_lds) {
  using SO = SplitOperands<KL>;
  constexpr int K = SO::K, RB = SO::RB, KS = SO::KS;
  for (unsigned i = threadIdx.x; i < 2 * RB * KS * 64; i += blockDim.x) {
    const unsigned l = i & 63u, ks = (i >> 6) % KS, rb = (i >> 6) / KS % RB,
                   m = (i >> 6) / (KS * RB);
    const float *row = m3 + m * K * K + (16 * rb + (l & 15u)) * K + 32 * ks + 8 * (l >> 4);
    bf16x8_t p0, p1, p2;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const Split3 sp = split3(row[e]);
      p0[e] = sp.p[0];
      p1[e] = sp.p[1];
      p2[e] = sp.p[2];
    }
    a_lds[SO::at(m, 0, rb, ks) + l] = p0;
    a_lds[SO::at(m, 1, rb, ks) + l] = p1;
    a_lds[SO::at(m, 2, rb, ks) + l] = p2;
  }
}

// Per-lane address parts of a split-product tile: input row 32 ks + 8 j + e
// has bits 0-2 (e) -> q0..q2, 3-4 (j) -> q3, q4, 5 (ks) -> q5; output row
// 16 rb + 4 j + i has bits 0-1 (i) -> q0, q1, 2-3 (j) -> q2, q3,
// 4-5 (rb) -> q4, q5 (j = lane / 16).
template <int KL> struct SplitLane {
  uint32_t lane, col;
  uint64_t din_j, dout_j;
  __device__ explicit SplitLane(const uint32_t *q) {
    const uint64_t one = 1;
    lane = threadIdx.x & 63u;
    col = lane & 15u;
    const uint32_t j = lane >> 4;
    din_j = ((j & 1u) ? one << q[3] : 0) | ((j & 2u) ? one << q[4] : 0);
    dout_j = ((j & 1u) ? one << q[2] : 0) | ((j & 2u) ? one << q[3] : 0);
  }
};

// U consecutive 16-combo tiles t0 .. t0 + U - 1: reads `in`, writes `out`
// (equal for in place).  All loads of the U tiles are issued before any
// product, so one wave keeps U times the loads in flight (for callers with
// few waves per CU).
template <int KL, int U = 1>
__device__ __forceinline__ void split_tile(const float2 *in, float2 *out, const bf16x8_t *a_lds,
                                           const uint32_t *q, const SplitLane<KL> &sl,
                                           uint64_t t0) {
  using SO = SplitOperands<KL>;
  constexpr int RB = SO::RB, KS = SO::KS;
  const uint64_t one = 1;
  uint64_t base[U];
  float2 x[U][KS][8];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    base[u] = combo_base<KL>((t0 + u) * 16 + sl.col, q);
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const uint64_t d = sl.din_j | local_delta<3>(uint32_t(e), q) | (ks ? one << q[5] : 0);
        x[u][ks][e] = in[base[u] + d];
      }
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    bf16x8_t br[3][KS], bi[3][KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const Split3 sr = split3(x[u][ks][e].x), si = split3(x[u][ks][e].y);
#pragma unroll
        for (int sp = 0; sp < 3; ++sp) {
          br[sp][ks][e] = sr.p[sp];
          bi[sp][ks][e] = si.p[sp];
        }
      }
#pragma unroll
    for (int rp = 0; rp < RB; rp += 2) {
      f32x4_t t1[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}}, t2[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}},
              t3[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
      // Re-read per tile (not hoisted out of the caller's tile loop into VGPRs).
      const bf16x8_t *a = a_lds + sl.lane + opaque_u32(0);
#pragma unroll
      for (int sa = 2; sa >= 0; --sa)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const bf16x8_t ar = a[SO::at(0, sa, rp + h, ks)];
            const bf16x8_t ai = a[SO::at(1, sa, rp + h, ks)];
#pragma unroll
            for (int sb = 2 - sa; sb >= 0; --sb) {
              t1[h] = mfma_bf16(ar, br[sb][ks], t1[h]);
              t3[h] = mfma_bf16(ar, bi[sb][ks], t3[h]);
              t2[h] = mfma_bf16(ai, bi[sb][ks], t2[h]);
              t3[h] = mfma_bf16(ai, br[sb][ks], t3[h]);
            }
          }
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const uint64_t d = sl.dout_j | local_delta<2>(uint32_t(i), q) |
                             (((rp + h) & 1) ? one << q[4] : 0) |
                             (((rp + h) & 2) ? one << q[5] : 0);
          out[base[u] + d] = make_float2(t1[h][i] - t2[h][i], t3[h][i]);
        }
    }
  }
}

} // namespace dev
} // namespace cast_hip
