// tile_common.h -- lane layout and LDS helpers shared by the tiled gridder
// and degridder.
//
// A wavefront works on whole samples.  Inside one sample the 64 lanes are
// laid out as (component q, v-tap dv): q = 2*pol + {re,im} picks an LDS
// plane and dv a column of the support footprint, and the u-taps are a loop.
// For a fixed u-tap the lanes of one sample touch consecutive words of up to
// four planes whose strides are = DVW (mod 32), so every LDS access of a
// sample is free of bank conflicts.
#pragma once

#include "avgrid_internal.h"

namespace av {

template <int NPOL, int DVW>
struct LaneMap {
  static constexpr int kQ = 2 * NPOL;                        // components per sample
  static constexpr int kLpi = kQ * DVW;                      // lanes wanted per sample
  static constexpr int kLanesPerItem = kLpi > 64 ? 64 : kLpi;
  static constexpr int kItemsPerIter = 64 / kLanesPerItem;   // samples per wave step
  static constexpr int kQPerLane = kLpi > 64 ? kLpi / 64 : 1;  // components per lane
  static constexpr int kQStride = kLanesPerItem / DVW;       // q distance between them
};

// plane stride: >= n and congruent to DVW modulo 32 (see above)
__host__ __device__ inline int plane_stride(int n, int dvw) {
  int r = n % 32;
  int s = n - r + dvw;
  if (s < n) s += 32;
  return s;
}

// Wave-local ordering point: LDS operations of one wavefront execute in
// order, so this only has to stop the compiler from moving LDS accesses
// across it.
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sum over each group of DVW consecutive lanes; the total lands in the last
// lane of the group.  Row shifts stay inside 16-lane DPP rows.
__device__ inline float dpp_shr(float v, int n) {
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
__device__ inline double dpp_shr(double v, int n) {
  long long iv = __builtin_bit_cast(long long, v);
  int lo = (int)(iv & 0xffffffffll), hi = (int)(iv >> 32);
  int rl, rh;
  switch (n) {
    case 1:
      rl = __builtin_amdgcn_update_dpp(0, lo, 0x111, 0xF, 0xF, true);
      rh = __builtin_amdgcn_update_dpp(0, hi, 0x111, 0xF, 0xF, true);
      break;
    case 2:
      rl = __builtin_amdgcn_update_dpp(0, lo, 0x112, 0xF, 0xF, true);
      rh = __builtin_amdgcn_update_dpp(0, hi, 0x112, 0xF, 0xF, true);
      break;
    case 4:
      rl = __builtin_amdgcn_update_dpp(0, lo, 0x114, 0xF, 0xF, true);
      rh = __builtin_amdgcn_update_dpp(0, hi, 0x114, 0xF, 0xF, true);
      break;
    default:
      rl = __builtin_amdgcn_update_dpp(0, lo, 0x118, 0xF, 0xF, true);
      rh = __builtin_amdgcn_update_dpp(0, hi, 0x118, 0xF, 0xF, true);
      break;
  }
  long long r = ((long long)(unsigned)rl) | ((long long)rh << 32);
  return __builtin_bit_cast(double, r);
}

template <int DVW, typename R>
__device__ inline R group_sum_to_last(R v) {
  if (DVW >= 16) v += dpp_shr(v, 8);
  v += dpp_shr(v, 4);
  v += dpp_shr(v, 2);
  v += dpp_shr(v, 1);
  return v;
}

// |ov*(tap - sc) + off| into the half-kernel table (standard_grid.py:384-396)
__device__ inline int tab_index(int ov, int tap, int sc, int off) {
  int k = ov * (tap - sc) + off;
  return k < 0 ? -k : k;
}

}  // namespace av
