// tile_common.h -- cell layout, lane maps and small helpers shared by the
// tiled gridder and degridder.
//
// LDS tiles are stored cell-major: a cell (u, v) holds the 2*npol real
// components of all its polarisations contiguously (32 bytes for four FP32
// polarisations).  Lanes move one 16-byte chunk each (ds_read_b128 /
// ds_write_b128), and eight consecutive lanes cover 128 contiguous bytes, so
// every wavefront access is free of bank conflicts.
//
// Float atomics on LDS run at about 1/25 of a plain read-modify-write on
// gfx950, so no kernel here uses them: a tile row band is only ever updated
// by the one wavefront that owns it.
#pragma once

#include "avgrid_internal.h"

namespace av {

template <typename R, int N>
struct alignas(sizeof(R) * N) Vec {
  R v[N];
};

template <typename R>
struct alignas(2 * sizeof(R)) Cplx2 {
  R re, im;
};

// Components of one cell and how they split into 16-byte chunks.
template <typename R, int NPOL>
struct CellLayout {
  static constexpr int kC = 2 * NPOL;                       // real components per cell
  static constexpr int kPerChunk = 16 / (int)sizeof(R);     // components in 16 bytes
  static constexpr int kCC = kC < kPerChunk ? kC : kPerChunk;  // components per lane chunk
  static constexpr int kH = kC / kCC;                       // chunks per cell
};

// Sum over each group of DVW lanes spaced H apart (lane = ... + dv*H + h);
// the total lands in the group's last v-tap lane (dv = DVW - 1).  Shifts
// inside a 16-lane DPP row use row_shr; wider ones use a lane shuffle.
__device__ inline float dpp_shr_f(float v, int n) {
  int iv = __builtin_bit_cast(int, v);
  int r;
  switch (n) {
    case 1: r = __builtin_amdgcn_update_dpp(0, iv, 0x111, 0xF, 0xF, true); break;
    case 2: r = __builtin_amdgcn_update_dpp(0, iv, 0x112, 0xF, 0xF, true); break;
    case 4: r = __builtin_amdgcn_update_dpp(0, iv, 0x114, 0xF, 0xF, true); break;
    default: r = __builtin_amdgcn_update_dpp(0, iv, 0x118, 0xF, 0xF, true); break;
  }
  return __builtin_bit_cast(float, r);
}
__device__ inline float shr_any(float v, int n) {
  if (n < 16) return dpp_shr_f(v, n);
  return __shfl_up(v, n);
}
__device__ inline double shr_any(double v, int n) {
  if (n < 16) {
    long long iv = __builtin_bit_cast(long long, v);
    int lo = (int)(iv & 0xffffffffll), hi = (int)(iv >> 32);
    float flo = __builtin_bit_cast(float, lo), fhi = __builtin_bit_cast(float, hi);
    int rl = __builtin_bit_cast(int, dpp_shr_f(flo, n));
    int rh = __builtin_bit_cast(int, dpp_shr_f(fhi, n));
    long long r = ((long long)(unsigned)rl) | ((long long)rh << 32);
    return __builtin_bit_cast(double, r);
  }
  return __shfl_up(v, n);
}

template <int H, int DVW, typename R>
__device__ inline R sum_over_dv(R v) {
#pragma unroll
  for (int off = H; off < H * DVW; off *= 2) v += shr_any(v, off);
  return v;
}

// Wave-local ordering point: LDS operations of one wavefront execute in
// order, so this only has to stop the compiler from moving LDS accesses
// across it.
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Workgroups are dealt to the chip's 8 XCDs (each with its own L2) round
// robin.  XcdGroups renumbers them so that each XCD receives whole groups of
// G consecutive workgroups, the groups themselves dealt out round robin:
// units of neighbouring channels (which share the cache lines of the
// visibilities, weights and flags of a row) then run under one L2, while the
// dense and sparse parts of the grid stay spread over all XCDs.  The grid
// size must be a multiple of 8 * G; the map is then a permutation of
// [0, gridDim.x).
constexpr int kXcds = 8;
__host__ __device__ inline uint32_t xcd_grouped_block(uint32_t b, uint32_t g) {
  const uint32_t k = b / kXcds;  // position in this XCD's sequence
  return ((k / g) * kXcds + b % kXcds) * g + k % g;
}
__host__ __device__ inline int64_t xcd_grouped_grid(int64_t blocks, int64_t g) {
  return (blocks + kXcds * g - 1) / (kXcds * g) * (kXcds * g);
}

// |ov*(tap - sc) + off| into the half-kernel table (standard_grid.py:384-396)
__device__ inline int tab_index(int ov, int tap, int sc, int off) {
  int k = ov * (tap - sc) + off;
  return k < 0 ? -k : k;
}

// Kernel weight of tap t at offset off, or 0 when t is outside [0, S).  The
// table read uses the clamped tap, so it is always in bounds, and the result
// is multiplied by a 0/1 mask: every lane issues the read, and the reads of
// several taps can be in flight together instead of each sitting behind a
// branch.
template <typename R>
__device__ inline R tap_weight(const R *tab, int ov, int t, int S, int sc, int off) {
  const int tc = t < 0 ? 0 : (t >= S ? S - 1 : t);
  const R w = tab[tab_index(ov, tc, sc, off)];
  const R in = (t >= 0 && t < S) ? (R)1 : (R)0;
  return w * in;
}

// Zero-padded kernel-weight table.  Entry (t, off), for taps t in [-1, S] and
// offsets off in [-half, half], holds tab[|ov*(t - sc) + off|] for
// 0 <= t < S and 0 on the padding rows t = -1 and t = S, so clamping any tap
// into [-1, S] gives its weight (or 0 outside the support) with one read.
struct WTab {
  int nw, half, S;
  __host__ __device__ WTab(int ov, int support)
      : nw(2 * (ov / 2 + 2) + 1), half(ov / 2 + 2), S(support) {}
  __host__ __device__ int size() const { return (S + 2) * nw; }
  __device__ int index(int t, int off) const {
    const int tc = t < -1 ? -1 : (t > S ? S : t);
    return (tc + 1) * nw + off + half;
  }
};

// Copy the call's zero-padded weight table (built once per call by
// prepare_tables) into LDS with every thread of the workgroup: four loads per
// thread in flight at a time (one round trip for the common tables), each
// from an in-bounds index so that no branch separates a load from its use.
template <typename R>
__device__ inline void copy_wtab(R *w, const WTab &W, const PlanDev &P) {
  constexpr int K = 4;
  const R *src;
  if constexpr (sizeof(R) == 4)
    src = P.wtab32;
  else
    src = P.wtab64;
  const int n = W.size();
  for (int i0 = threadIdx.x; i0 < n; i0 += K * blockDim.x) {
    R v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = src[min(i0 + k * (int)blockDim.x, n - 1)];
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (i0 + k * (int)blockDim.x < n) w[i0 + k * blockDim.x] = v[k];
  }
}

// ---- split-FP16 matrix-core operands ------------------------------------
// An FP32 operand x is carried as two FP16 halves, hi = fp16(x) and lo =
// fp16(x - hi); a product then takes three v_mfma_f32_16x16x32_f16 (hi*hi +
// hi*lo + lo*hi) with FP32 accumulation, and keeps 22 bits of each operand.
typedef _Float16 H3Half8 __attribute__((ext_vector_type(8)));
typedef float H3Float4 __attribute__((ext_vector_type(4)));
typedef uint32_t H3U4 __attribute__((ext_vector_type(4)));

// Eight FP32 values a_j as FP16 halves, two per register: hi_j = fp16(a_j)
// (v_cvt_pk_f16_f32, round to nearest even) and lo_j = fp16(a_j - hi_j).
// v_fma_mix forms its FMA in FP32 and rounds the result to FP16; mixlo
// writes bits 0-15 of its register and mixhi bits 16-31, keeping the other
// half.  The closing s_nop gives the wait states a matrix-core read of a
// VALU result needs (the compiler cannot see which instruction of the block
// writes which register).
__device__ inline void h3_split8(const float *a, H3U4 &hi, H3U4 &lo) {
  uint32_t h0, h1, h2, h3, l0, l1, l2, l3;
  asm("v_cvt_pk_f16_f32 %0, %8, %9\n\t"
      "v_cvt_pk_f16_f32 %1, %10, %11\n\t"
      "v_cvt_pk_f16_f32 %2, %12, %13\n\t"
      "v_cvt_pk_f16_f32 %3, %14, %15\n\t"
      "v_fma_mixlo_f16 %4, %8, 1.0, -%0 op_sel_hi:[0,0,1]\n\t"
      "v_fma_mixhi_f16 %4, %9, 1.0, -%0 op_sel:[0,0,1] op_sel_hi:[0,0,1]\n\t"
      "v_fma_mixlo_f16 %5, %10, 1.0, -%1 op_sel_hi:[0,0,1]\n\t"
      "v_fma_mixhi_f16 %5, %11, 1.0, -%1 op_sel:[0,0,1] op_sel_hi:[0,0,1]\n\t"
      "v_fma_mixlo_f16 %6, %12, 1.0, -%2 op_sel_hi:[0,0,1]\n\t"
      "v_fma_mixhi_f16 %6, %13, 1.0, -%2 op_sel:[0,0,1] op_sel_hi:[0,0,1]\n\t"
      "v_fma_mixlo_f16 %7, %14, 1.0, -%3 op_sel_hi:[0,0,1]\n\t"
      "v_fma_mixhi_f16 %7, %15, 1.0, -%3 op_sel:[0,0,1] op_sel_hi:[0,0,1]\n\t"
      "s_nop 1"
      : "=&v"(h0), "=&v"(h1), "=&v"(h2), "=&v"(h3), "=&v"(l0), "=&v"(l1), "=&v"(l2), "=&v"(l3)
      : "v"(a[0]), "v"(a[1]), "v"(a[2]), "v"(a[3]), "v"(a[4]), "v"(a[5]), "v"(a[6]),
        "v"(a[7]));
  hi = H3U4{h0, h1, h2, h3};
  lo = H3U4{l0, l1, l2, l3};
}

// Eight FP32 products x_j = w_j * d_j as FP16 halves, two per register:
// hi_j = fp16(x_j) and lo_j = fp16(x_j - hi_j), two v_fma_mix each.
__device__ inline void h3_split_products8(const float *w, const float *d, H3U4 &hi, H3U4 &lo) {
  uint32_t h0, h1, h2, h3, l0, l1, l2, l3;
#define AV_H3_PAIR(H, L, W0, D0, W1, D1)                                      \
  "v_fma_mixlo_f16 " H ", " W0 ", " D0 ", 0\n\t"                            \
  "v_fma_mixhi_f16 " H ", " W1 ", " D1 ", 0\n\t"                            \
  "v_fma_mixlo_f16 " L ", " W0 ", " D0 ", -" H " op_sel_hi:[0,0,1]\n\t"     \
  "v_fma_mixhi_f16 " L ", " W1 ", " D1 ", -" H " op_sel:[0,0,1] op_sel_hi:[0,0,1]\n\t"
  asm(AV_H3_PAIR("%0", "%4", "%8", "%16", "%9", "%17")
      AV_H3_PAIR("%1", "%5", "%10", "%18", "%11", "%19")
      AV_H3_PAIR("%2", "%6", "%12", "%20", "%13", "%21")
      AV_H3_PAIR("%3", "%7", "%14", "%22", "%15", "%23")
      "s_nop 1"
      : "=&v"(h0), "=&v"(h1), "=&v"(h2), "=&v"(h3), "=&v"(l0), "=&v"(l1), "=&v"(l2), "=&v"(l3)
      : "v"(w[0]), "v"(w[1]), "v"(w[2]), "v"(w[3]), "v"(w[4]), "v"(w[5]), "v"(w[6]),
        "v"(w[7]), "v"(d[0]), "v"(d[1]), "v"(d[2]), "v"(d[3]), "v"(d[4]), "v"(d[5]),
        "v"(d[6]), "v"(d[7]));
#undef AV_H3_PAIR
  hi = H3U4{h0, h1, h2, h3};
  lo = H3U4{l0, l1, l2, l3};
}

}  // namespace av
