// degrid.hip -- tiled degridding (replaces dgrid_numba,
// core/imaging/imaging_utils/standard_degrid.py:388-485, zero dphase).
//
// One workgroup per degrid unit (primary records of one bin).  The unit
// copies its tile plus a halo of support/2 + 1 pixels into LDS, cell-major
// and already in visibility-polarisation order (through pol_map), one
// wavefront per row.  Each wavefront then evaluates its staged samples with
// lanes = (sample, v-tap, 16-byte chunk): a lane reads one chunk per u-tap
// (eight lanes cover 128 contiguous bytes, so no bank conflicts), weights it
// by the u and v kernel values, and a DPP reduction over the v-taps gives
// the sample, divided by the kernel-weight sum as in the reference.  Each
// sample is written exactly once.
#include <algorithm>

#include "tile_common.h"

namespace av {

constexpr int kDegridThreads = 256;
constexpr int kDegridXcdGroup = 8;  // matrix-core kernel: workgroups per XCD group (8 measured best of 2-64)
constexpr int kDegridWaves = kDegridThreads / 64;
constexpr int kRegionBatch = 12;  // grid values each thread has in flight per pass

// Flag bytes (0xFF in byte p) of the visibility polarisations whose pol_map
// entry lies outside [0, npol_g): dgrid_numba skips such a polarisation
// (standard_degrid.py:419-420), so it is treated like a flagged one -- left
// alone when incremental, zero otherwise.
template <int NPOL>
__device__ inline uint32_t pol_map_skips(const int32_t *pol_map, int npol_g) {
  uint32_t bad = 0;
  if (pol_map) {
#pragma unroll
    for (int p = 0; p < NPOL; ++p) {
      const int q = pol_map[p];
      if (q < 0 || q >= npol_g) bad |= 0xFFu << (8 * p);
    }
  }
  return bad;
}

// Copy the region [u0 - halo, u0 - halo + rows) x [v0 - halo, v0 - halo +
// cols) of every visibility polarisation's grid plane (through pol_map) into
// LDS, zero outside the grid and beyond (ru_real, rv_real).  The cells are
// numbered polarisation-major, then row, then column, so consecutive threads
// read consecutive cells of a row; each thread issues kRegionBatch loads
// before it waits and writes them with store(p, row, col, value).
template <typename R, int NPOL, typename Store>
__device__ inline void load_region(const R *gplane, size_t pstride, int npol_g,
                                   const int32_t *pol_map, int n_u, int n_v, int u_lo, int v_lo,
                                   int rows, int cols, int ru_real, int rv_real, Store store) {
  constexpr int KB = kRegionBatch / NPOL > 0 ? kRegionBatch / NPOL : 1;  // cells per pol
  const R *src[NPOL];
#pragma unroll
  for (int p = 0; p < NPOL; ++p) {
    int q = pol_map ? pol_map[p] : p % npol_g;
    q = q < 0 ? 0 : (q >= npol_g ? npol_g - 1 : q);
    src[p] = gplane + q * pstride;
  }
  // e / cols by multiply-high, exact for e < 2^16 and 2 <= cols (a region is
  // at least 6 cells wide: tile >= support >= 2 plus a halo of >= 2 per side)
  const int per_pol = rows * cols;
  const uint32_t magic = 0xFFFFFFFFu / (uint32_t)cols + 1u;
  // the whole real region inside the grid: no per-cell bounds test
  const bool inside = u_lo >= 0 && v_lo >= 0 && u_lo + ru_real <= n_u && v_lo + rv_real <= n_v;
  for (int e0 = threadIdx.x; e0 < per_pol; e0 += KB * kDegridThreads) {
    Cplx2<R> val[NPOL][KB];
    int ru[KB], rv[KB];
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      const int e = e0 + k * kDegridThreads;
      ru[k] = (int)__umulhi((uint32_t)e, magic);
      rv[k] = e - ru[k] * cols;
      const int gu = u_lo + ru[k], gv = v_lo + rv[k];
      const bool ok = e < per_pol && ru[k] < ru_real && rv[k] < rv_real &&
                      (inside || (gu >= 0 && gu < n_u && gv >= 0 && gv < n_v));
      const size_t off = ok ? ((size_t)gu * n_v + gv) * 2 : 0;
#pragma unroll
      for (int p = 0; p < NPOL; ++p) {
        val[p][k] = Cplx2<R>{(R)0, (R)0};
        if (ok) val[p][k] = *(const Cplx2<R> *)(src[p] + off);
      }
    }
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      if (e0 + k * kDegridThreads < per_pol) {
#pragma unroll
        for (int p = 0; p < NPOL; ++p) store(p, ru[k], rv[k], val[p][k]);
      }
    }
  }
}

struct DegridLds {
  int tab, reg, st_cu, st_cv, st_pos, st_item, total;
  int ntab, halo, ru, rv;
};

template <typename R, int NPOL, int DVW>
__host__ __device__ inline DegridLds degrid_lds_layout(int ov, int support, int tu, int tv) {
  DegridLds L;
  int sc = support / 2;
  L.ntab = ov * (sc + 1);
  L.halo = sc + 1;
  L.ru = tu + 2 * L.halo;
  L.rv = tv + 2 * L.halo;
  int o = 0;
  auto take = [&](int bytes) {
    int r = o;
    o += (bytes + 15) & ~15;
    return r;
  };
  L.tab = take(L.ntab * (int)sizeof(R));
  L.reg = take(L.ru * L.rv * 2 * NPOL * (int)sizeof(R));
  L.st_cu = take(kDegridThreads * DVW * (int)sizeof(R));
  L.st_cv = take(kDegridThreads * DVW * (int)sizeof(R));
  L.st_pos = take(kDegridThreads * (int)sizeof(int));
  L.st_item = take(kDegridThreads * (int)sizeof(uint32_t));
  L.total = o;
  return L;
}

template <typename R, int NPOL, int DVW>
__global__ __launch_bounds__(kDegridThreads) void k_degrid_tiles(
    PlanDev P, const R *__restrict__ grid, int npol_g, const int32_t *__restrict__ pol_map,
    const uint8_t *__restrict__ flag, const double *__restrict__ cgk, R *__restrict__ vis_out,
    int mode) {
  using CL = CellLayout<R, NPOL>;
  constexpr int C = CL::kC, CC = CL::kCC, H = CL::kH;
  constexpr int SPI = 64 / (H * DVW) > 0 ? 64 / (H * DVW) : 1;  // samples per wave step
  using V = Vec<R, CC>;
  extern __shared__ __align__(16) char smem[];
  const uint32_t u = blockIdx.x;
  if (u >= P.dunit_start[P.n_bins]) return;
  const int S = P.support, sc = S / 2, ov = P.ov;
  const int Tu = P.tile_u, Tv = P.tile_v;
  const DegridLds L = degrid_lds_layout<R, NPOL, DVW>(ov, S, Tu, Tv);
  R *tab = (R *)(smem + L.tab);
  R *reg = (R *)(smem + L.reg);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  R *st_cu = (R *)(smem + L.st_cu) + wave * 64 * DVW;
  R *st_cv = (R *)(smem + L.st_cv) + wave * 64 * DVW;
  int *st_pos = (int *)(smem + L.st_pos) + wave * 64;
  uint32_t *st_item = (uint32_t *)(smem + L.st_item) + wave * 64;
  const bool incremental = (mode & AV_DEGRID_INCREMENTAL) != 0;
  const uint32_t polbad = pol_map_skips<NPOL>(pol_map, npol_g);
  const int Hh = L.halo, RV = L.rv;

  for (int i = tid; i < L.ntab; i += kDegridThreads)
    tab[i] = sizeof(R) == 4 ? (R)P.tab32[i] : (R)cgk[i];

  const uint4 unit = P.dunits[u];
  const int64_t b = unit.x;
  const uint32_t first = unit.y, cnt = unit.z;
  // 32-bit divisions: bins are numbered below 2^32 (64-bit ones cost
  // about a hundred scalar instructions each)
  const uint32_t tile = (uint32_t)b / (uint32_t)P.n_img_chan;
  const int ic = (int)((uint32_t)b - tile * (uint32_t)P.n_img_chan);
  const uint32_t tu_i = tile / (uint32_t)P.ntv;
  const int u0 = (int)tu_i * Tu, v0 = (int)(tile - tu_i * (uint32_t)P.ntv) * Tv;
  const R *gplane = grid + (size_t)ic * npol_g * P.n_u * P.n_v * 2;
  const size_t pstride = (size_t)P.n_u * P.n_v * 2;

  // ---- tile + halo into LDS in visibility-polarisation order (zero outside
  // the grid), cell-major
  load_region<R, NPOL>(gplane, pstride, npol_g, pol_map, P.n_u, P.n_v, u0 - Hh, v0 - Hh, L.ru,
                       RV, L.ru, RV, [&](int p, int ru, int rv, const Cplx2<R> &val) {
                         *(Cplx2<R> *)(reg + (ru * RV + rv) * C + 2 * p) = val;
                       });
  __syncthreads();

  // lanes: (sample slot s, v-tap dv, chunk h)
  const int h = lane % H;
  const int dv = (lane / H) % DVW;
  const int slot = lane / (H * DVW);
  const uint32_t end = first + cnt;
  for (uint32_t r0 = first; r0 < end; r0 += kDegridThreads) {
    const uint32_t r = r0 + wave * 64 + lane;
    int pos = -1;
    if (r < end) {
      const RecWords rw = *(const RecWords *)(P.rec + 3 * (size_t)r);
      const uint32_t g = rw.z;
      if (g != kInvalidGeo) {
        int au, ou, avv, ovv;
        unpack_geo(g, au, ou, avv, ovv);
        pos = (au + kPosBias) | ((avv + kPosBias) << 16);
        st_item[lane] = rw.x;
#pragma unroll
        for (int t = 0; t < DVW; ++t) {
          st_cu[lane * DVW + t] = t < S ? tab[tab_index(ov, t, sc, ou)] : (R)0;
          st_cv[lane * DVW + t] = t < S ? tab[tab_index(ov, t, sc, ovv)] : (R)0;
        }
      }
    }
    st_pos[lane] = pos;
    wave_lds_sync();
    const int nst = (int)min((uint32_t)64, end > r0 + wave * 64 ? end - (r0 + wave * 64) : 0u);
    for (int j0 = 0; j0 < nst; j0 += SPI) {
      const int j = j0 + slot;
      const int pj = (slot < SPI && j < nst) ? st_pos[j] : -1;
      // The kernel-weight sum is formed with exactly the operations applied
      // to the data, so a constant grid degrids to that constant exactly
      // (the reference sums wt*g and wt in the same loop).
      R a[CC];
      R nacc = (R)0;
#pragma unroll
      for (int c = 0; c < CC; ++c) a[c] = (R)0;
      if (pj >= 0 && dv < S) {
        const int au = (pj & 0xFFFF) - kPosBias;
        const int avv = (pj >> 16) - kPosBias;
        const R *base = reg + ((au + Hh - sc) * RV + (avv + Hh - sc + dv)) * C + h * CC;
        for (int du = 0; du < S; ++du) {
          const R cu = st_cu[j * DVW + du];
          const V gv = *(const V *)(base + du * RV * C);
          nacc += cu;
#pragma unroll
          for (int c = 0; c < CC; ++c) a[c] += cu * gv.v[c];
        }
        const R cv = st_cv[j * DVW + dv];
        nacc *= cv;
#pragma unroll
        for (int c = 0; c < CC; ++c) a[c] *= cv;
      }
      // every lane takes part in the reductions over the v-taps
#pragma unroll
      for (int c = 0; c < CC; ++c) a[c] = sum_over_dv<H, DVW>(a[c]);
      nacc = sum_over_dv<H, DVW>(nacc);
      if (pj >= 0 && dv == DVW - 1) {
        const uint32_t item = st_item[j];
        uint32_t fbits = polbad;
        if (flag) {
          if (NPOL == 4)
            fbits |= *(const uint32_t *)(flag + (size_t)item * 4);
          else
            for (int p = 0; p < NPOL; ++p) fbits |= (uint32_t)flag[(size_t)item * NPOL + p] << (8 * p);
        }
        V *dst = (V *)(vis_out + (size_t)item * C + h * CC);
        V o;
#pragma unroll
        for (int c = 0; c < CC; ++c) o.v[c] = incremental ? dst->v[c] : (R)0;
#pragma unroll
        for (int c = 0; c < CC; ++c) {
          const int p = (h * CC + c) >> 1;
          // a zero kernel sum adds nothing (standard_degrid.py:482-485)
          const bool skip = ((fbits >> (8 * p)) & 0xFFu) || nacc == (R)0;
          if (incremental) {
            if (!skip) o.v[c] += a[c] / nacc;
          } else {
            o.v[c] = skip ? (R)0 : a[c] / nacc;
          }
        }
        *dst = o;
      }
    }
    wave_lds_sync();
  }
}

// ---------------------------------------------------------------------------
// Matrix-core degridding.  For a group of 16 records k of one unit,
//   T[k][(v, c)] = sum_u wu_k(u) * G[u][(v, c)]     (product over region rows)
//   V_k[c]       = sum_v wv_k(v) * T[k][(v, c)] / norm_k.
// The first step runs on v_mfma_*_16x16x4 with the records as the 16 result
// rows, four region rows per step (lane l supplies A[l%16][l/16], the
// u-weight of record l%16 at region row 4s + l/16) and the region columns,
// cell-major, in 16-wide blocks read from LDS.  The second step is a
// per-lane weighted sum over the blocks followed by a DPP reduction over the
// lanes holding other cells of the same component.  One more block, of ones,
// gives sum_u wu_k(u); the kernel-weight sum is then formed with exactly the
// operations applied to the data, so a constant grid degrids to that
// constant exactly (the reference sums wt*g and wt side by side).

template <typename R>
struct DegridMfma;
// Result layout (lane l, register r): the FP32 instruction holds row
// 4*(l/16) + r, the FP64 one row l/16 + 4*r; both hold column l%16.
template <>
struct DegridMfma<float> {
  typedef float V __attribute__((ext_vector_type(4)));
  __device__ static V mma(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lk, int r) { return 4 * lk + r; }
};
template <>
struct DegridMfma<double> {
  typedef double V __attribute__((ext_vector_type(4)));
  __device__ static V mma(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lk, int r) { return lk + 4 * r; }
};

struct DegridMfmaLds {
  int tab, reg, sgeo, sitem, scnt, total, ntab, halo, ru, rup, rv, rvp, rs;
};

// The records of a degrid unit are sorted by tile row in LDS, this many at a
// time (one per thread: passes of 512 were slower), so that each 16-record
// group covers a narrow band of region rows.
constexpr int kSortCap = 256;
// Sort keys: tile rows -1 .. tile_u, once per v-half when the records are
// split by v, and one key for records without degrid geometry.  The scan
// takes three keys per lane of one wavefront, and two words after the keys
// hold the start of the second half and the sorted count (with padding).
template <int NB, int NBW>
__host__ __device__ constexpr int degrid_sort_keys(int tu) {
  return (NBW < NB ? 2 : 1) * (tu + 2) + 1;
}
constexpr int kMaxSortKeys = 3 * 64;

// Region rows padded to a multiple of 4 (one matrix step), columns to NB
// blocks of 16 components.  Within a row, component n of the region is
// stored at (n % 16) * NB + n / 16, so the NB B-operands of lane l (column
// l % 16 of every block) are contiguous and load as 16-byte reads; rows are
// padded by 16 bytes so that the four rows of one step fall in different
// banks.
// part / norm for the matrix-core degridder, where norm > 0 is the sum of
// the support's kernel weights.  FP32: hardware reciprocal, quotient, and one
// fused correction step (Markstein's refinement: q + r * (a - b * q) with the
// residual from a fused multiply-add), which is within one unit in the last
// place and exact whenever the quotient is a power-of-two multiple (a = g * b,
// a constant grid of g).  FP64 keeps IEEE division.
__device__ inline float div_by_norm(float a, float b) {
  const float r = __builtin_amdgcn_rcpf(b);
  const float q = a * r;
  return fmaf(fmaf(-b, q, a), r, q);
}
__device__ inline double div_by_norm(double a, double b) { return a / b; }

// x + (x of lane l ^ 16) and x + (x of lane l ^ 32).  FP32 uses the gfx950
// lane swaps: with x in both operands, v_permlane16_swap leaves rows
// (x0, x0, x2, x2) in one result and (x1, x1, x3, x3) in the other (rows of
// 16 lanes), v_permlane32_swap the lower half of x in both halves of one and
// the upper half in both halves of the other.  Their sum has the same two
// addends in every lane as x plus the shuffled x, with no LDS round trip.
__device__ inline float sum_xor16(float x) {
  const auto p =
      __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x), false, false);
  return __uint_as_float(p[0]) + __uint_as_float(p[1]);
}
__device__ inline float sum_xor32(float x) {
  const auto p =
      __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false, false);
  return __uint_as_float(p[0]) + __uint_as_float(p[1]);
}
__device__ inline double sum_xor16(double x) { return x + __shfl_xor(x, 16); }
__device__ inline double sum_xor32(double x) { return x + __shfl_xor(x, 32); }

template <typename R, int NPOL, int NB, int NBW = NB>
__host__ __device__ inline DegridMfmaLds degrid_mfma_layout(int ov, int support, int tu, int tv) {
  DegridMfmaLds L;
  const int sc = support / 2;
  L.ntab = ov * (sc + 1);
  L.halo = sc + 1;
  L.ru = tu + 2 * L.halo;
  L.rup = (L.ru + 3) & ~3;
  L.rv = tv + 2 * L.halo;
  L.rvp = NB * 16 / (2 * NPOL);
  L.rs = NB * 16 + 16 / (int)sizeof(R);  // elements per region row
  L.tab = 0;  // the zero-padded weight table (WTab)
  L.reg = (WTab(ov, support).size() * (int)sizeof(R) + 15) & ~15;
  L.sgeo = (L.reg + L.rup * L.rs * (int)sizeof(R) + 15) & ~15;
  // sorted records: up to 15 padding slots before the second v-half
  L.sitem = L.sgeo + (kSortCap + 16) * 4;
  L.scnt = L.sitem + (kSortCap + 16) * 4;
  L.total = L.scnt + (degrid_sort_keys<NB, NBW>(tu) + 2) * 4;
  return L;
}

template <int NPOL>
__host__ __device__ constexpr int degrid_mfma_blocks_needed(int tv, int support) {
  return ((tv + 2 * (support / 2 + 1)) * 2 * NPOL + 15) / 16;
}

// Records split by v into two halves: a record whose degrid centre is in
// tile columns av < tv/2 goes to the first, the rest to the second, and each
// half's groups compute NBW of the NB column blocks: the first half blocks
// [0, NBW), the second [NB - NBW, NB).  A centre lies in columns -1 .. tv
// (a grid-primary record carries a degrid centre within one pixel of its
// grid centre), and its footprint covers region columns av + 1 .. av + S.
// The split is used when both windows hold every footprint of their half,
// the second window starts on a 16-byte boundary of the region rows, and
// the keys fit the scan.
template <typename R, int NPOL, int NB, int NBW>
__host__ __device__ constexpr bool degrid_halves_fit(int tu, int tv, int support) {
  constexpr int CPB = 16 / (2 * NPOL);
  constexpr int VL = 16 / (int)sizeof(R);
  return NBW < NB && (NB - NBW) % VL == 0 && tv / 2 + support <= NBW * CPB &&
         tv / 2 + 1 >= (NB - NBW) * CPB && tv + support <= NB * CPB - 1 &&
         degrid_sort_keys<NB, NBW>(tu) + 2 <= kMaxSortKeys;
}

// The 16 records of one group (lane l holds record l % 16).
struct DegridGroup {
  uint32_t geo, item, fbits;
};

// One workgroup per degrid unit.  The unit's records (first kSortCap; larger
// units run further passes) are read before the tile region, so both loads
// are in flight together; the records are then sorted by tile row in LDS,
// so that the 16 records of a group span a narrow band of region rows and
// only the row steps of that band are run (the others would add exact
// zeros).  With v-halves (NBW < NB) the sort key puts the records of each
// half together, each half starting on a group boundary, and a group
// computes only its half's NBW column blocks: the blocks outside the window
// hold no tap of the group's records.  A record's flags are read with its group and applied as its
// model is written: a flagged polarisation is set to zero (left alone when
// AV_DEGRID_INCREMENTAL); the finishing pass zeroes the samples no unit
// visits.
// Workgroups per CU the FP32 kernels are compiled for: six with v-halves of
// at most 8 column blocks (80 registers without spills; a 16 x 16 tile's
// workgroup then needs 26,988 bytes of LDS, a sixth of the CU's 160 KB);
// five for up to 12 blocks without halves (96 registers; at six the
// one-polarisation kernels spill).  Other variants keep the compiler's choice.
template <typename R, int NB, int NBW>
constexpr int degrid_mfma_min_blocks() {
  return (sizeof(R) == 4 && NBW < NB && NBW <= 8) ? 6 : (sizeof(R) == 4 && NB <= 12) ? 5 : 1;
}

template <typename R, int NPOL, int NB, int NBW>
__global__ __launch_bounds__(kDegridThreads, (degrid_mfma_min_blocks<R, NB, NBW>()))
void k_degrid_mfma(
    PlanDev P, const R *__restrict__ grid, int npol_g, const int32_t *__restrict__ pol_map,
    const uint8_t *__restrict__ flag, const double *__restrict__ cgk, R *__restrict__ vis_out,
    int mode) {
  using MO = DegridMfma<R>;
  using MV = typename MO::V;
  constexpr int C = 2 * NPOL;
  constexpr int CPB = 16 / C;               // cells per 16-column block
  constexpr int VL = 16 / (int)sizeof(R);   // elements per 16-byte read
  constexpr int HK = kSortCap / kDegridThreads;  // records per thread in a sort pass
  // v-halves: each group computes the NBW column blocks of its half's window
  constexpr bool kHalves = NBW < NB;
  constexpr int KB1 = NB - NBW;  // first block of the second half's window
  static_assert(KB1 % VL == 0, "the second window must start on a 16-byte boundary");
  extern __shared__ __align__(16) char smem[];
  // groups of kDegridXcdGroup consecutive units (neighbouring channels of a
  // tile) run under one L2
  const uint32_t u = xcd_grouped_block(blockIdx.x, kDegridXcdGroup);
  if (u >= P.dunit_start[P.n_bins]) return;
  const int S = P.support, sc = S / 2, ov = P.ov;
  const int Tu = P.tile_u, Tv = P.tile_v;
  const DegridMfmaLds L = degrid_mfma_layout<R, NPOL, NB, NBW>(ov, S, Tu, Tv);
  R *tab = (R *)(smem + L.tab);
  R *reg = (R *)(smem + L.reg);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // scalar group loop
  uint32_t *s_geo = (uint32_t *)(smem + L.sgeo);
  uint32_t *s_item = (uint32_t *)(smem + L.sitem);
  uint32_t *s_cnt = (uint32_t *)(smem + L.scnt);
  const bool incremental = (mode & AV_DEGRID_INCREMENTAL) != 0;
  const uint32_t polbad = pol_map_skips<NPOL>(pol_map, npol_g);
  const int Hh = L.halo, RS = L.rs;
  const int li = lane & 15, lk = lane >> 4;
  const int rkeys = Tu + 2;  // row keys of one half
  const int nkeys = degrid_sort_keys<NB, NBW>(Tu);
  const int half_v = Tv / 2;

  const uint4 unit = P.dunits[u];
  const int64_t b = unit.x;
  const uint32_t first = unit.y, end = unit.y + unit.z;
  // sort key: tile row + 1 of the degrid centre (rows -1 .. Tu), plus rkeys
  // in the second v-half; the last key for a record without degrid geometry
  auto key_of = [&](uint32_t g) {
    const int k = (int)(g & 0xFFu) - kPosBias + 1;
    int key = k < 0 ? 0 : (k > Tu + 1 ? Tu + 1 : k);
    if constexpr (kHalves) {
      const int av = (int)((g >> 16) & 0xFFu) - kPosBias;
      key += av >= half_v ? rkeys : 0;
    }
    return g == kInvalidGeo ? nkeys - 1 : key;
  };
  // records of the first sort pass: in flight during the region load
  // (unconditional loads: a thread past the end rereads the last record and
  // its copy is never used)
  uint32_t pg[HK], pi[HK];
#pragma unroll
  for (int k = 0; k < HK; ++k) {
    const uint32_t r = min(first + tid + k * kDegridThreads, end - 1);
    const RecWords rw = *(const RecWords *)(P.rec + 3 * (size_t)r);
    pg[k] = rw.z;
    pi[k] = rw.x;
  }

  const WTab W(ov, S);
  copy_wtab(tab, W, P);
  // 32-bit divisions: bins are numbered below 2^32 (64-bit ones cost
  // about a hundred scalar instructions each)
  const uint32_t tile = (uint32_t)b / (uint32_t)P.n_img_chan;
  const int ic = (int)((uint32_t)b - tile * (uint32_t)P.n_img_chan);
  const uint32_t tu_i = tile / (uint32_t)P.ntv;
  const int u0 = (int)tu_i * Tu, v0 = (int)(tile - tu_i * (uint32_t)P.ntv) * Tv;
  const R *gplane = grid + (size_t)ic * npol_g * P.n_u * P.n_v * 2;
  const size_t pstride = (size_t)P.n_u * P.n_v * 2;

  // ---- tile + halo into LDS in visibility-polarisation order, zero outside
  // the grid and in the padding
  load_region<R, NPOL>(gplane, pstride, npol_g, pol_map, P.n_u, P.n_v, u0 - Hh, v0 - Hh, L.rup,
                       L.rvp, L.ru, L.rv, [&](int p, int ru, int rv, const Cplx2<R> &val) {
                         const int n = rv * C + 2 * p;  // n, n + 1 share a 16-column block
                         R *dst = reg + ru * RS + (n & 15) * NB + (n >> 4);
                         dst[0] = val.re;
                         dst[NB] = val.im;
                       });

  for (uint32_t c0 = first; c0 < end; c0 += kSortCap) {
    const int n = (int)min((uint32_t)kSortCap, end - c0);
    if (c0 != first) {  // later passes of a large unit
#pragma unroll
      for (int k = 0; k < HK; ++k) {
        const uint32_t r = c0 + min(tid + k * kDegridThreads, n - 1);
        const RecWords rw = *(const RecWords *)(P.rec + 3 * (size_t)r);
        pg[k] = rw.z;
        pi[k] = rw.x;
      }
    }
    // ---- counting sort of the pass's records by key, in LDS
    for (int i = tid; i < nkeys; i += kDegridThreads) s_cnt[i] = 0u;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < HK; ++k)
      if (tid + k * kDegridThreads < n) atomicAdd(&s_cnt[key_of(pg[k])], 1u);
    __syncthreads();
    if (wave == 0) {  // exclusive scan of the counts: three keys per lane
      uint32_t c3[3], t = 0;
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const int kk = 3 * lane + j;
        c3[j] = kk < nkeys ? s_cnt[kk] : 0u;
        t += c3[j];
      }
      uint32_t incl = t;
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t o = __shfl_up(incl, off);
        if (lane >= off) incl += o;
      }
      uint32_t run = incl - t;
      // the second half starts on a group boundary: its keys move up by the
      // padding, and the padding slots hold no record
      uint32_t pad = 0, b1 = 0;
      if constexpr (kHalves) {
        uint32_t h0 = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j) h0 += 3 * lane + j < rkeys ? c3[j] : 0u;
        for (int off = 32; off > 0; off >>= 1) h0 += __shfl_xor(h0, off);
        b1 = (h0 + 15u) & ~15u;
        pad = b1 - h0;
        if ((uint32_t)lane < pad) {
          s_geo[h0 + lane] = kInvalidGeo;
          s_item[h0 + lane] = 0u;
        }
      }
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const int kk = 3 * lane + j;
        if (kk < nkeys) s_cnt[kk] = run + (kk >= rkeys ? pad : 0u);
        run += c3[j];
      }
      if (lane == 0) {
        s_cnt[nkeys] = b1;
        s_cnt[nkeys + 1] = (uint32_t)n + pad;
      }
    }
    __syncthreads();
    const int b1 = __builtin_amdgcn_readfirstlane((int)s_cnt[nkeys]);
    const int np = __builtin_amdgcn_readfirstlane((int)s_cnt[nkeys + 1]);
#pragma unroll
    for (int k = 0; k < HK; ++k)
      if (tid + k * kDegridThreads < n) {
        const uint32_t pos = atomicAdd(&s_cnt[key_of(pg[k])], 1u);
        s_geo[pos] = pg[k];
        s_item[pos] = pi[k];
      }
    __syncthreads();

    // ---- 16-record groups of the sorted pass
    const int ngroups = (np + 15) / 16;
    for (int g = wave; g < ngroups; g += kDegridWaves) {
      const int r = g * 16 + li;
      const uint32_t geo = r < np ? s_geo[r] : kInvalidGeo;
      const uint32_t item = r < np ? s_item[r] : 0u;
      // first column block of the group's window
      const int kb0 = (kHalves && g * 16 >= b1) ? KB1 : 0;
      // the record's flags, used when its model is written (a flagged
      // polarisation is left alone when incremental, set to zero otherwise).
      // The load is unconditional -- a lane without a record reads its slot's
      // sample (0 or a valid one) -- and its result is needed only at the end
      // of the group; neighbouring channels share the flag line in L2
      uint32_t fbits = polbad;
      if (flag) {
        if (NPOL == 4)
          fbits |= *(const uint32_t *)(flag + (size_t)item * 4);
        else if (NPOL == 2)
          fbits |= *(const uint16_t *)(flag + (size_t)item * 2);
        else
          fbits |= flag[item];
      }
      int au, ou, avv, ovv;
      unpack_geo(geo_or_far(geo), au, ou, avv, ovv);
      const int urow0 = au + Hh - sc;   // region row of the first u-tap
      const int vcol0 = avv + Hh - sc;  // region column of the first v-tap
      // the group's band of region rows (records are sorted by row).  Taking
      // the band from lane 0 and the last valid lane with scalar reads was
      // slower: the branch and loop bounds then wait for the record reads.
      const bool valid = geo != kInvalidGeo;
      int rlo = valid ? urow0 : 1 << 20, rhi = valid ? urow0 + S - 1 : -1;
      for (int off = 1; off < 16; off <<= 1) {
        rlo = min(rlo, __shfl_xor(rlo, off));
        rhi = max(rhi, __shfl_xor(rhi, off));
      }
      // every quarter-wave holds the same 16 records, so this is uniform
      if (rlo > rhi) continue;  // no record of the group has a degrid centre

      // Weight-table reads without a multiply or a clamp pair per read:
      // WTab::index(t, off) = (clamp(t, -1, S) + 1) * nw + off + half, and
      // with nw > 0, (clamp(t, -1, S) + 1) * nw = clamp((t + 1) * nw, 0,
      // (S + 1) * nw), where (t + 1) * nw moves by a fixed step from one row
      // step or column block to the next.
      // (The empty asm statements keep a per-group base opaque, so that the
      // compiler does not fold base + step back into one multiply per read.)
      const int nw = W.nw, tcap = (S + 1) * nw;
      auto tap_read = [&](int x, int offh) { return tab[min(max(x, 0), tcap) + offh]; };
      const int uofs = ou + W.half;
      const int ub0 = (lk - urow0 + 1) * nw;

      // row steps outside the band would only add zeros; the first step
      // starts the sums from zero
      MV T[NBW], T1;
      auto row_step = [&](int s4, bool first_step) {
        const int row = 4 * s4 + lk;
        const R a = tap_read(ub0 + s4 * (4 * nw), uofs);  // u-tap row - urow0
        const R *rp = reg + row * RS + li * NB + kb0;
        R bv[NBW];
#pragma unroll
        for (int k = 0; k < NBW; k += VL) {
          const Vec<R, VL> t = *(const Vec<R, VL> *)(rp + k);
#pragma unroll
          for (int e = 0; e < VL; ++e) bv[k + e] = t.v[e];
        }
        const MV zero{0, 0, 0, 0};
#pragma unroll
        for (int k = 0; k < NBW; ++k) T[k] = MO::mma(bv[k], a, first_step ? zero : T[k]);
        T1 = MO::mma((R)1, a, first_step ? zero : T1);
      };
      row_step(rlo >> 2, true);
      for (int s4 = (rlo >> 2) + 1; s4 <= (rhi >> 2); ++s4) row_step(s4, false);
      // ---- weighted sums over v.  The products are transposed (region
      // columns x records), so lane l holds, for its own record li = l % 16,
      // rows MO::row(l/16, r) of every block k: region component
      // 16 k + row, i.e. cell k*CPB + row/C, component row%C.  Every row of
      // the ones block is the record's sum of u-weights.
      constexpr bool kF32 = sizeof(R) == 4;
      // the four rows of a lane lie in one cell (FP32 layout, >= 4 components)
      constexpr bool kOneCell = kF32 && C >= 4;
      R part[4], npart[4];
      // fused multiply-adds: a constant grid g gives T[k] = g * T1 exactly,
      // and each step of part is then exactly g times the same step of npart;
      // the two are combined below by the same adds in the same order
      // The v-weight of this lane's record at region column `cell` comes
      // straight from the weight table; each chain runs over k in order.
      // (window block k is region block kb0 + k)
      const int vcolw = vcol0 - kb0 * CPB;
      const int vofs = ovv + W.half;
      // v-tap of window cell k * CPB + j, scaled by nw: (j - vcolw + 1) * nw + k * CPB * nw
      constexpr int NVB = kOneCell ? 1 : 4;  // one base per cell a lane's rows lie in
      int vb0[NVB];
#pragma unroll
      for (int r = 0; r < NVB; ++r) {
        vb0[r] = (MO::row(lk, r) / C - vcolw + 1) * nw;
        asm("" : "+v"(vb0[r]));
      }
      auto wv_at = [&](int k, int r) { return tap_read(vb0[r] + k * (CPB * nw), vofs); };
#pragma unroll
      for (int r = 0; r < 4; ++r) part[r] = npart[r] = (R)0;
#pragma unroll
      for (int k = 0; k < NBW; ++k) {
        if constexpr (kOneCell) {
          const R wv = wv_at(k, 0);
#pragma unroll
          for (int r = 0; r < 4; ++r) part[r] = fma(T[k][r], wv, part[r]);
          npart[0] = fma(T1[0], wv, npart[0]);
        } else {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const R wv = wv_at(k, r);
            part[r] = fma(T[k][r], wv, part[r]);
            npart[r] = fma(T1[0], wv, npart[r]);
          }
        }
      }
      if constexpr (kOneCell) {
#pragma unroll
        for (int r = 1; r < 4; ++r) npart[r] = npart[0];
      }
      // rows of one component: within the lane, then across quarter-waves
      // (FP32 rows 4 lk + r, FP64 rows lk + 4 r)
      auto combine = [&](R *x) {
        if constexpr ((kF32 && C == 2) || (!kF32 && C == 8)) {
          x[0] += x[2];
          x[1] += x[3];
        }
        if constexpr (!kF32 && C <= 4) x[0] = (x[0] + x[1]) + (x[2] + x[3]);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if constexpr (kF32 && C <= 4) x[r] = sum_xor16(x[r]);
          if constexpr (kF32 || C == 2) x[r] = sum_xor32(x[r]);
        }
      };
      combine(part);
      combine(npart);
      // the lanes holding a record's totals, and the component of total r
      constexpr int kVals = kF32 ? (C == 2 ? 2 : 4) : (C == 8 ? 2 : 1);
      const bool writer = kF32 ? (C == 8 ? lk < 2 : lk == 0) : (C == 2 ? lk < 2 : true);
      auto comp_of = [&](int r) {
        if constexpr (kF32) return C == 8 ? 4 * lk + r : r;
        else return C == 8 ? lk + 4 * r : lk;
      };
      if (writer && geo != kInvalidGeo) {
        R *base = vis_out + (size_t)item * C;
#pragma unroll
        for (int r = 0; r < kVals; ++r) {
          const int c = comp_of(r);
          const R val = div_by_norm(part[r], npart[r]);
          // a zero kernel sum adds nothing (standard_degrid.py:482-485)
          const bool skip = ((fbits >> (8 * (c >> 1))) & 0xFFu) || npart[r] == (R)0;
          if (incremental) {
            if (!skip) base[c] += val;
          } else {
            base[c] = skip ? (R)0 : val;
          }
        }
      }
    }
    __syncthreads();  // the next pass rewrites the sorted records
  }
}

// ---------------------------------------------------------------------------
// Split-FP16 matrix-core degridder (FP32 grids, 16 x 16 tiles, the default
// for them).  For a group of 16 records the products T[(v,c)][k] =
// sum_u G[u][(v,c)] * wu_k(u) run on v_mfma_f32_16x16x32_f16 with K = 32
// region rows at once (the 24 rows of a 16-row tile plus its halo; the other
// 8 read a zero block), each operand split into FP16 hi + lo halves and three
// products per block (hi*hi + hi*lo + lo*hi) with FP32 accumulation.  So no
// row band, no sort by row and no row-step loop: a group is one table read
// per lane per row group, one split of its u-weights, and three products per
// column block of its v-half window.
//
// Scales (powers of two, so exact): the weight table carries 2^14 (u- and
// v-weights at most 1, kept clear of the FP16 subnormals), and the region of
// a unit is scaled by 2^-E with E the exponent of its largest |value|, which
// puts it below 1; both cancel in the division by the kernel-weight sum (the
// ones block and the same v-weights) and the result is multiplied by 2^E.  A
// constant grid g degrids to g exactly: its scaled region is a power-of-two
// multiple m of 1, hi = m and lo = 0, so every product and sum of the data
// is m times the same one of the weight sums.  Grid values of +-inf give NaN
// (the FP32 kernel gives inf).
//
// Region layout in LDS, for the A operands (lane l holds A[l % 16][8 (l/16)
// + j], j = 0..7): hi and lo planes of [row group 0..2][region entry n =
// column * 2npol + component][8 rows] FP16, so the eight rows of a lane are
// one 16-byte read.  Lanes of row group 3 (region rows 24..31) read row group
// 2 again: their u-weights are exactly zero (taps past the support read the
// table's zero row) and the region values are finite, so those products add
// zeros.  Entry n is stored at h3_swz(n), which keeps every 8-entry half of a
// 16-entry block in place, so the 16 lanes of a block read still cover 16
// distinct entries, and sends the stores of consecutive tile columns (8
// entries apart) to distinct banks.  The records of a unit are split by
// v-half only (keys: half 0, half 1, no geometry).
struct DegridH3Lds {
  int tab, ahi, alo, sgeo, sitem, sflag, scnt, red, total, ru, rv, rvp, nbn, halo;
};

__host__ __device__ constexpr int h3_swz(int n) { return n ^ ((n >> 3) & 7); }

// per-(wave, key) counts of a sort pass, then the pass's slot count, window
// starts and padding runs
constexpr int kH3CntWords = 32;
// slots a sort pass may add to its records: up to 15 of padding before each
// window after the first
constexpr int kH3Pad = 48;

template <int NPOL, int NB, int NBW, bool PAIR>
__host__ __device__ inline DegridH3Lds degrid_h3_layout(int ov, int support, int tu, int tv) {
  DegridH3Lds L;
  L.halo = support / 2 + 1;
  L.ru = tu + 2 * L.halo;
  L.rv = (PAIR ? 2 * tv : tv) + 2 * L.halo;
  L.rvp = NB * 16 / (2 * NPOL);
  L.nbn = NB * 16;
  int o = 0;
  auto take = [&](int bytes) {
    const int r = o;
    o += (bytes + 15) & ~15;
    return r;
  };
  L.ahi = take(3 * L.nbn * 16);
  L.alo = take(3 * L.nbn * 16);
  L.tab = take(WTab(ov, support).size() * 4);
  L.sgeo = take((kSortCap + kH3Pad) * 4);
  L.sitem = take((kSortCap + kH3Pad) * 4);
  L.sflag = take((kSortCap + kH3Pad) * 4);
  L.scnt = take(kH3CntWords * 4);
  L.red = take(kDegridWaves * 4);
  L.total = o;
  return L;
}

// The tiles the split-FP16 degridder serves: 16 region rows plus the halo fit
// its 24 rows, and every region value of a pass is loaded by one item of at
// most two per thread.
template <int NPOL, int NB>
__host__ __device__ constexpr bool degrid_h3_fits(int tu, int tv, int support) {
  return tu == 16 && tv == 16 && support >= 2 && tu + 2 * (support / 2 + 1) <= 24 &&
         3 * (NB * 16 / (2 * NPOL)) * NPOL <= 2 * kDegridThreads;
}

// Pair units (four polarisations): the 32 + 2 halo columns of two tiles fit
// the 20 column blocks, and the four 16-column windows starting at columns
// 0, 8, 16 and 24 hold every footprint: a degrid centre lies in pair columns
// -1 .. 32 and a record takes window min(max(av, 0) / 8, 3), whose columns
// cover its footprint av + 1 .. av + S when S <= 7.
template <int NPOL, int NB, int NBW>
__host__ __device__ constexpr bool degrid_h3_pair_fits(int tu, int tv, int support) {
  return NPOL == 4 && NB == 20 && NBW == 8 && degrid_h3_fits<NPOL, NB>(tu, tv, support) &&
         support <= 7 && 2 * tv + 2 * (support / 2 + 1) <= NB * 16 / (2 * NPOL);
}

template <int NPOL, int NB, int NBW, bool PAIR>
__global__ __launch_bounds__(kDegridThreads, PAIR ? 4 : 6)
void k_degrid_h3(PlanDev P, const float *__restrict__ grid, int npol_g,
                 const int32_t *__restrict__ pol_map, const uint8_t *__restrict__ flag,
                 float *__restrict__ vis_out, int mode) {
  using MV = H3Float4;
  constexpr int C = 2 * NPOL;
  constexpr int CPB = 16 / C;  // cells per 16-entry block
  constexpr bool kHalves = NBW < NB;
  // windows of NBW blocks, KSTEP blocks (8 cells) apart: two halves of one
  // tile, or four quarters of a pair of tiles
  constexpr int NWIN = PAIR ? 4 : (kHalves ? 2 : 1);
  constexpr int KSTEP = 4;
  constexpr int HK = kSortCap / kDegridThreads;
  // windows start on multiples of 4 blocks, so a block's swizzle depends only
  // on its index within the window (mod 4)
  static_assert(!kHalves || (NPOL == 4 && (NWIN - 1) * KSTEP + NBW == NB), "windows");
  static_assert(!PAIR || kHalves, "pair units use windows");
  constexpr int nkeys = NWIN + 1;  // the windows, then records without geometry
  constexpr int NVW = HK * kDegridWaves;  // (record slot, wave) pairs of a pass
  // per-(slot, wave, key) bases, then the slot count, the window starts and
  // the padding runs (start, length) before windows 1 .. NWIN - 1
  constexpr int XC = NVW * nkeys;
  static_assert(XC + 1 + 3 * (NWIN - 1) <= kH3CntWords, "sort counts");
  static_assert(16 * (NWIN - 1) <= kH3Pad, "padding slots");
  extern __shared__ __align__(16) char smem[];
  const uint32_t u = xcd_grouped_block(blockIdx.x, kDegridXcdGroup);
  if (u >= (PAIR ? P.punit_start[P.n_pairs] : P.dunit_start[P.n_bins])) return;
  const int S = P.support, sc = S / 2, ov = P.ov;
  const int Tu = P.tile_u, Tv = P.tile_v;
  const DegridH3Lds L = degrid_h3_layout<NPOL, NB, NBW, PAIR>(ov, S, Tu, Tv);
  float *tab = (float *)(smem + L.tab);
  _Float16 *ahi = (_Float16 *)(smem + L.ahi), *alo = (_Float16 *)(smem + L.alo);
  uint32_t *s_geo = (uint32_t *)(smem + L.sgeo);
  uint32_t *s_item = (uint32_t *)(smem + L.sitem);
  uint32_t *s_flag = (uint32_t *)(smem + L.sflag);
  uint32_t *s_cnt = (uint32_t *)(smem + L.scnt);
  float *red = (float *)(smem + L.red);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const bool incremental = (mode & AV_DEGRID_INCREMENTAL) != 0;
  const uint32_t polbad = pol_map_skips<NPOL>(pol_map, npol_g);
  const int Hh = L.halo;
  const int li = lane & 15, lk = lane >> 4;
  const int half_v = Tv / 2;

  const uint4 unit = PAIR ? P.punits[u] : P.dunits[u];
  const int64_t b = unit.x;
  // the unit's records by index: record slots directly, or for a pair unit
  // indices into the primary records of tile 2j (n0 of them, from slot s0)
  // and then those of tile 2j + 1 (from slot s1), whose degrid centres are
  // shifted by Tv columns into the pair's region
  const uint32_t first = unit.y, end = unit.y + unit.z;
  uint32_t s0 = 0, n0 = 0, s1 = 0;
  if constexpr (PAIR) {
    s0 = P.bin_start[b];
    n0 = P.cnt_prim[b];
    s1 = unit.w != kNoBin ? P.bin_start[unit.w] : 0u;
  }
  auto key_of = [&](uint32_t g) {
    if (g == kInvalidGeo) return nkeys - 1;
    if constexpr (NWIN > 1) {
      const int av = (int)((g >> 16) & 0xFFu) - kPosBias;
      return min(max(av, 0) / half_v, NWIN - 1);
    }
    return 0;
  };
  // a record's flags are loaded with it (in flight during the region load)
  auto load_flags = [&](uint32_t item) -> uint32_t {
    if (!flag) return 0u;
    if (NPOL == 4) return *(const uint32_t *)(flag + (size_t)item * 4);
    if (NPOL == 2) return *(const uint16_t *)(flag + (size_t)item * 2);
    return flag[item];
  };
  auto load_record = [&](uint32_t i, uint32_t &geo, uint32_t &item, uint32_t &fb) {
    const uint32_t slot = PAIR ? (i < n0 ? s0 + i : s1 + (i - n0)) : i;
    const RecWords rw = *(const RecWords *)(P.rec + 3 * (size_t)slot);
    geo = rw.z;
    if (PAIR && i >= n0 && geo != kInvalidGeo) geo += (uint32_t)Tv << 16;
    item = rw.x;
    fb = load_flags(rw.x);
  };
  uint32_t pg[HK], pi[HK], pf[HK];
#pragma unroll
  for (int k = 0; k < HK; ++k)
    load_record(min(first + tid + k * kDegridThreads, end - 1), pg[k], pi[k], pf[k]);

  const WTab W(ov, S);
  {
    // the weight table times 2^14
    const int n = W.size();
    for (int i = tid; i < n; i += kDegridThreads) tab[i] = P.wtab32[i] * 16384.0f;
  }
  const uint32_t tile = (uint32_t)b / (uint32_t)P.n_img_chan;
  const int ic = (int)((uint32_t)b - tile * (uint32_t)P.n_img_chan);
  const uint32_t tu_i = tile / (uint32_t)P.ntv;
  const int u0 = (int)tu_i * Tu, v0 = (int)(tile - tu_i * (uint32_t)P.ntv) * Tv;
  const float *gplane = grid + (size_t)ic * npol_g * P.n_u * P.n_v * 2;
  const size_t pstride = (size_t)P.n_u * P.n_v * 2;

  // ---- region: items (row group, column, polarisation), each eight rows of
  // one complex value; at most two items per thread (checked on the host)
  constexpr int kItems = 3 * (NB * 16 / (2 * NPOL)) * NPOL;
  constexpr int kIt = (kItems + kDegridThreads - 1) / kDegridThreads;
  float vre[kIt][8], vim[kIt][8];
  float mx = 0.f;
  {
    const int rvp = L.rvp;
    const int ulo = u0 - Hh, vlo = v0 - Hh;
#pragma unroll
    for (int it = 0; it < kIt; ++it) {
      const int e = tid + it * kDegridThreads;
      const int rv = e % rvp, rest = e / rvp, p = rest % NPOL, kg = rest / NPOL;
      int q = pol_map ? pol_map[p] : p % npol_g;
      q = q < 0 ? 0 : (q >= npol_g ? npol_g - 1 : q);
      const int gv = vlo + rv;
      const bool colok = e < kItems && rv < L.rv && gv >= 0 && gv < P.n_v;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int ru = 8 * kg + j, gu = ulo + ru;
        const bool ok = colok && ru < L.ru && gu >= 0 && gu < P.n_u;
        Cplx2<float> val{0.f, 0.f};
        if (ok) val = *(const Cplx2<float> *)(gplane + q * pstride + ((size_t)gu * P.n_v + gv) * 2);
        vre[it][j] = val.re;
        vim[it][j] = val.im;
        mx = fmaxf(mx, fmaxf(fabsf(val.re), fabsf(val.im)));
      }
    }
  }
  // the unit's scale: the exponent of the largest |value|
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
  if (lane == 0) red[tid >> 6] = mx;
  __syncthreads();
  mx = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
  int E;
  (void)frexpf(mx, &E);  // |value| < 2^E (E = 0 for an all-zero region)
  E = __builtin_amdgcn_readfirstlane(E);
  {
    const int rvp = L.rvp;
#pragma unroll
    for (int it = 0; it < kIt; ++it) {
      const int e = tid + it * kDegridThreads;
      if (e >= kItems) continue;
      const int rv = e % rvp, rest = e / rvp, p = rest % NPOL, kg = rest / NPOL;
      const int n = rv * C + 2 * p;  // region entries n (real), n + 1 (imaginary)
      float sre[8], sim[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        sre[j] = ldexpf(vre[it][j], -E);
        sim[j] = ldexpf(vim[it][j], -E);
      }
      H3U4 hr, lr, hm, lm;
      h3_split8(sre, hr, lr);
      h3_split8(sim, hm, lm);
      const int ore = (kg * L.nbn + h3_swz(n)) * 8, oim = (kg * L.nbn + h3_swz(n + 1)) * 8;
      *(H3U4 *)(ahi + ore) = hr;
      *(H3U4 *)(alo + ore) = lr;
      *(H3U4 *)(ahi + oim) = hm;
      *(H3U4 *)(alo + oim) = lm;
    }
  }

  // this lane's A entries (in halves) for blocks b = 0..3 mod 4: entry 16 b +
  // li of row group min(lk, 2), stored at 16 b + h3_swz(16 b + li) - 16 b
  int aswz[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) aswz[j] = (min(lk, 2) * L.nbn + h3_swz(16 * j + li) - 16 * j) * 8;

  for (uint32_t c0 = first; c0 < end; c0 += kSortCap) {
    const int n = (int)min((uint32_t)kSortCap, end - c0);
    if (c0 != first) {  // later passes of a large unit
#pragma unroll
      for (int k = 0; k < HK; ++k)
        load_record(c0 + min(tid + k * kDegridThreads, n - 1), pg[k], pi[k], pf[k]);
    }
    // ---- stable counting sort of the pass's records by window, in LDS: each
    // wave counts its keys by ballot, one thread turns the (slot, wave, key)
    // counts into bases, and a record's slot is its base plus the number of
    // lower lanes of its wave with the same key (no LDS atomics)
    int key[HK];
    uint64_t kmask[HK][nkeys];
#pragma unroll
    for (int k = 0; k < HK; ++k) {
      const bool valid = tid + k * kDegridThreads < n;
      key[k] = key_of(pg[k]);
#pragma unroll
      for (int kk = 0; kk < nkeys; ++kk) {
        kmask[k][kk] = __ballot(valid && key[k] == kk);
        if (lane == 0)
          s_cnt[(k * kDegridWaves + wave) * nkeys + kk] = (uint32_t)__popcll(kmask[k][kk]);
      }
    }
    __syncthreads();
    if (tid == 0) {
      // the windows' records in window order, each window from a group
      // boundary (the slots of the padding before it hold no record), then
      // the records without geometry; within a key in pass order
      uint32_t cnt[XC];
#pragma unroll
      for (int i = 0; i < XC; ++i) cnt[i] = s_cnt[i];
      uint32_t run = 0;
#pragma unroll
      for (int kk = 0; kk < nkeys; ++kk) {
        if (kk >= 1 && kk < NWIN) {
          const uint32_t r16 = (run + 15u) & ~15u;
          s_cnt[XC + kk] = r16;                        // window kk's first slot
          s_cnt[XC + NWIN + 2 * (kk - 1)] = run;       // its padding
          s_cnt[XC + NWIN + 2 * (kk - 1) + 1] = r16 - run;
          run = r16;
        }
#pragma unroll
        for (int w = 0; w < NVW; ++w) {
          s_cnt[w * nkeys + kk] = run;
          run += cnt[w * nkeys + kk];
        }
      }
      s_cnt[XC] = run;  // slots of the pass, padding included
    }
    __syncthreads();
    const int np = __builtin_amdgcn_readfirstlane((int)s_cnt[XC]);
    int wstart[NWIN];  // first slot of each window (window 0: slot 0)
    wstart[0] = 0;
#pragma unroll
    for (int kk = 1; kk < NWIN; ++kk)
      wstart[kk] = __builtin_amdgcn_readfirstlane((int)s_cnt[XC + kk]);
    if constexpr (NWIN > 1) {
      if (tid < 16 * (NWIN - 1)) {
        const int gi = tid >> 4, j = tid & 15;
        const uint32_t ps = s_cnt[XC + NWIN + 2 * gi], pl = s_cnt[XC + NWIN + 2 * gi + 1];
        if ((uint32_t)j < pl) {
          s_geo[ps + j] = kInvalidGeo;
          s_item[ps + j] = 0u;
          s_flag[ps + j] = 0u;
        }
      }
    }
#pragma unroll
    for (int k = 0; k < HK; ++k)
      if (tid + k * kDegridThreads < n) {
        uint64_t m = kmask[k][0];
#pragma unroll
        for (int kk = 1; kk < nkeys; ++kk) m = key[k] == kk ? kmask[k][kk] : m;
        const uint32_t below = __builtin_amdgcn_mbcnt_hi(
            (uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        const uint32_t pos = s_cnt[(k * kDegridWaves + wave) * nkeys + key[k]] + below;
        s_geo[pos] = pg[k];
        s_item[pos] = pi[k];
        s_flag[pos] = pf[k];
      }
    __syncthreads();

    // ---- 16-record groups of the pass
    const int ngroups = (np + 15) / 16;
    const H3Half8 ones = {1, 1, 1, 1, 1, 1, 1, 1};
    for (int g = wave; g < ngroups; g += kDegridWaves) {
      const int r = g * 16 + li;
      const uint32_t geo = r < np ? s_geo[r] : kInvalidGeo;
      const uint32_t item = r < np ? s_item[r] : 0u;
      int kb0 = 0;  // first block of the group's window
#pragma unroll
      for (int kk = 1; kk < NWIN; ++kk) kb0 += g * 16 >= wstart[kk] ? KSTEP : 0;
      const uint32_t fbits = polbad | (r < np ? s_flag[r] : 0u);
      int au, ou, avv, ovv;
      unpack_geo(geo_or_far(geo), au, ou, avv, ovv);
      const int urow0 = au + Hh - sc;   // region row of the first u-tap
      const int vcol0 = avv + Hh - sc;  // region column of the first v-tap
      const int nw = W.nw, tcap = (S + 1) * nw;
      auto tap_read = [&](int x, int offh) { return tab[min(max(x, 0), tcap) + offh]; };
      // ---- B: this lane's record's u-weights at region rows 8 lk + j
      float wu[8];
      {
        const int uofs = ou + W.half;
        const int ub = (8 * lk - urow0 + 1) * nw;  // (tap + 1) * nw of row 8 lk
#pragma unroll
        for (int j = 0; j < 8; ++j) wu[j] = tap_read(ub + j * nw, uofs);
      }
      H3U4 bhu, blu;
      h3_split8(wu, bhu, blu);
      const H3Half8 bh = __builtin_bit_cast(H3Half8, bhu), bl = __builtin_bit_cast(H3Half8, blu);
      // ---- A: the window's blocks of the region
      const _Float16 *ah_p = ahi + 128 * kb0, *al_p = alo + 128 * kb0;
      MV T[NBW];
#pragma unroll
      for (int k = 0; k < NBW; ++k) {
        const H3Half8 ah = *(const H3Half8 *)(ah_p + aswz[k & 3] + 128 * k);
        const H3Half8 al = *(const H3Half8 *)(al_p + aswz[k & 3] + 128 * k);
        const MV z = {0.f, 0.f, 0.f, 0.f};
        MV t = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, z, 0, 0, 0);
        t = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, t, 0, 0, 0);
        T[k] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, t, 0, 0, 0);
      }
      MV T1;
      {
        const MV z = {0.f, 0.f, 0.f, 0.f};
        T1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ones, bh, z, 0, 0, 0);
        T1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ones, bl, T1, 0, 0, 0);
      }
      // ---- weighted sums over v (lane: record li, entries 4 lk + r of each
      // block: one cell, components 4 lk % C + r, for 2npol >= 4)
      const int vcolw = vcol0 - kb0 * CPB;
      const int vofs = ovv + W.half;
      constexpr bool kOneCell = C >= 4;
      constexpr int NVB = kOneCell ? 1 : 4;
      int vb0[NVB];
#pragma unroll
      for (int r2 = 0; r2 < NVB; ++r2) {
        vb0[r2] = ((4 * lk + r2) / C - vcolw + 1) * nw;
        asm("" : "+v"(vb0[r2]));
      }
      auto wv_at = [&](int k, int r2) { return tap_read(vb0[r2] + k * (CPB * nw), vofs); };
      float part[4], npart[4];
#pragma unroll
      for (int r2 = 0; r2 < 4; ++r2) part[r2] = npart[r2] = 0.f;
#pragma unroll
      for (int k = 0; k < NBW; ++k) {
        if constexpr (kOneCell) {
          const float wv = wv_at(k, 0);
#pragma unroll
          for (int r2 = 0; r2 < 4; ++r2) part[r2] = fmaf(T[k][r2], wv, part[r2]);
          npart[0] = fmaf(T1[0], wv, npart[0]);
        } else {
#pragma unroll
          for (int r2 = 0; r2 < 4; ++r2) {
            const float wv = wv_at(k, r2);
            part[r2] = fmaf(T[k][r2], wv, part[r2]);
            npart[r2] = fmaf(T1[0], wv, npart[r2]);
          }
        }
      }
      if constexpr (kOneCell) {
#pragma unroll
        for (int r2 = 1; r2 < 4; ++r2) npart[r2] = npart[0];
      }
      auto combine = [&](float *x) {
        if constexpr (C == 2) {
          x[0] += x[2];
          x[1] += x[3];
        }
#pragma unroll
        for (int r2 = 0; r2 < 4; ++r2) {
          if constexpr (C <= 4) x[r2] = sum_xor16(x[r2]);
          x[r2] = sum_xor32(x[r2]);
        }
      };
      combine(part);
      combine(npart);
      constexpr int kVals = C == 2 ? 2 : 4;
      const bool writer = C == 8 ? lk < 2 : lk == 0;
      auto comp_of = [&](int r2) { return C == 8 ? 4 * lk + r2 : r2; };
      if (writer && geo != kInvalidGeo) {
        float *base = vis_out + (size_t)item * C;
#pragma unroll
        for (int r2 = 0; r2 < kVals; ++r2) {
          const int c = comp_of(r2);
          const float val = ldexpf(div_by_norm(part[r2], npart[r2]), E);
          // a zero kernel sum adds nothing (standard_degrid.py:482-485)
          const bool skip = ((fbits >> (8 * (c >> 1))) & 0xFFu) || npart[r2] == 0.f;
          if (incremental) {
            if (!skip) base[c] += val;
          } else {
            base[c] = skip ? 0.f : val;
          }
        }
      }
    }
    __syncthreads();  // the next pass rewrites the sorted records
  }
}

// Finishing pass: zero the samples no degrid unit visits (off-grid, NaN uvw,
// channel outside the grid) and, when `flag` is given, the flagged
// polarisations of every sample.  A workgroup covers 64 rows x 64 channels:
// one validity word per channel and contiguous flag reads along channels.
template <typename R, int NPOL>
__global__ __launch_bounds__(256) void k_degrid_finish(PlanDev P, const uint8_t *flag,
                                                      R *vis_out) {
  const int nch = P.row_chan ? 1 : P.n_chan;
  const int64_t ncb = (nch + 63) / 64;
  const int64_t blk = blockIdx.x;
  const int64_t rw = blk / ncb;                 // 64-row block
  const int c = (int)((blk - rw * ncb) * 64) + (threadIdx.x & 63);
  if (c >= nch) return;
  const unsigned long long m = P.dmask[(int64_t)c * P.mask_words_per_chan + rw];
  // this thread's 16 samples: every flag word is loaded (from an in-bounds
  // sample) before any store, then the unvisited and flagged ones are zeroed
  constexpr int K = 64 / 4;
  uint32_t fb[K];
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const int64_t row = min(rw * 64 + (threadIdx.x >> 6) + 4 * j, P.n_rows - 1);
    const int64_t item = P.row_chan ? row : row * P.n_chan + c;
    fb[j] = 0;
    if (flag) {
      if (NPOL == 4)
        fb[j] = *(const uint32_t *)(flag + (size_t)item * 4);
      else if (NPOL == 2)
        fb[j] = *(const uint16_t *)(flag + (size_t)item * 2);
      else
        fb[j] = flag[item];
    }
  }
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const int k = (threadIdx.x >> 6) + 4 * j;
    const int64_t row = rw * 64 + k;
    if (row >= P.n_rows) break;
    const int64_t item = P.row_chan ? row : row * P.n_chan + c;
    const bool visited = (m >> k) & 1ull;
    Cplx2<R> *o = (Cplx2<R> *)vis_out + (size_t)item * NPOL;
#pragma unroll
    for (int p = 0; p < NPOL; ++p)
      if (!visited || ((fb[j] >> (8 * p)) & 0xFFu)) o[p] = Cplx2<R>{(R)0, (R)0};
  }
}

// The same finishing pass for plans with few channels per row (rows that
// carry their own channel included): one thread per sample, in sample
// order, so the writes stay coalesced whatever the channel count.
template <typename R>
__global__ __launch_bounds__(256) void k_degrid_finish_items(PlanDev P, const uint8_t *flag,
                                                            R *vis_out, int npol) {
  const int nch = P.row_chan ? 1 : P.n_chan;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; item < P.n_items;
       item += stride) {
    int64_t row = item;
    int c = 0;
    if (nch > 1) {
      row = item / nch;
      c = (int)(item - row * nch);
    }
    const bool visited = (P.dmask[(int64_t)c * P.mask_words_per_chan + row / 64] >> (row & 63)) & 1ull;
    R *o = vis_out + (size_t)item * npol * 2;
    for (int p = 0; p < npol; ++p)
      if (!visited || (flag && flag[(size_t)item * npol + p])) {
        o[2 * p] = (R)0;
        o[2 * p + 1] = (R)0;
      }
  }
}

// non-incremental mode: zero the samples no unit visits (off-grid, NaN uvw,
// channel outside the grid), read from the plan's degrid-validity bitmask
template <typename R>
__global__ void k_zero_invalid(PlanDev P, R *vis_out, int npol) {
  const int64_t n = P.n_items;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int64_t row = i, chan = 0;
    if (!P.row_chan) {
      row = i / P.n_chan;
      chan = i - row * P.n_chan;
    }
    const unsigned long long m = P.dmask[chan * P.mask_words_per_chan + (row >> 6)];
    if (!((m >> (row & 63)) & 1ull)) {
      R *o = vis_out + (size_t)i * npol * 2;
      for (int k = 0; k < 2 * npol; ++k) o[k] = (R)0;
    }
  }
}

template <typename R, int NPOL, int DVW>
static int launch_degrid(av_plan *plan, const void *grid, int npol_g, const int32_t *pol_map,
                         const uint8_t *flag, const double *cgk, void *vis_out, int mode,
                         hipStream_t s) {
  PlanDev &d = plan->d;
  DegridLds L = degrid_lds_layout<R, NPOL, DVW>(d.ov, d.support, d.tile_u, d.tile_v);
  auto kfn = k_degrid_tiles<R, NPOL, DVW>;
  const DeviceProps &prop = device_props();
  if (prop.error) return hip_fail(prop.error, "av_degrid: device properties");
  if ((size_t)L.total > prop.lds_per_block)
    return fail(AV_ERR_UNSUPPORTED, "av_degrid: tile too large for LDS (" +
                                        std::to_string(L.total) + " bytes)");
  {
    const int rc_lds = ensure_lds_limit<k_degrid_tiles<R, NPOL, DVW>>(L.total, "av_degrid: hipFuncSetAttribute");
    if (rc_lds) return rc_lds;
  }
  int rc = prepare_tables(d, cgk, s);
  if (rc) return rc;
  if (!(mode & AV_DEGRID_INCREMENTAL) && d.n_items > 0) {
    int64_t blocks = std::min<int64_t>((d.n_items + 255) / 256, 8192);
    hipLaunchKernelGGL(k_zero_invalid<R>, dim3((unsigned)blocks), dim3(256), 0, s, d,
                       (R *)vis_out, NPOL);
    AV_LAUNCH_CHECK("k_zero_invalid");
  }
  hipLaunchKernelGGL(kfn, dim3((unsigned)d.max_dunits), dim3(kDegridThreads), L.total, s, d,
                     (const R *)grid, npol_g, pol_map, flag, cgk, (R *)vis_out, mode);
  AV_LAUNCH_CHECK("k_degrid_tiles");
  return AV_OK;
}

template <typename R, int NPOL>
static int launch_degrid_finish(const PlanDev &d, void *vis_out, int mode, hipStream_t s);

template <typename R, int NPOL, int NB, int NBW = NB>
static int launch_degrid_mfma(av_plan *plan, const void *grid, int npol_g, const int32_t *pol_map,
                              const uint8_t *flag, const double *cgk, void *vis_out, int mode,
                              hipStream_t s) {
  PlanDev &d = plan->d;
  const DegridMfmaLds L =
      degrid_mfma_layout<R, NPOL, NB, NBW>(d.ov, d.support, d.tile_u, d.tile_v);
  const DeviceProps &prop = device_props();
  if (prop.error) return hip_fail(prop.error, "av_degrid: device properties");
  if ((size_t)L.total > prop.lds_per_block)
    return fail(AV_ERR_UNSUPPORTED, "av_degrid: tile too large for LDS (" +
                                        std::to_string(L.total) + " bytes)");
  {
    const int rc_lds =
        ensure_lds_limit<k_degrid_mfma<R, NPOL, NB, NBW>>(L.total, "av_degrid: hipFuncSetAttribute");
    if (rc_lds) return rc_lds;
  }
  int rc = prepare_tables(d, cgk, s);
  if (rc) return rc;
  hipLaunchKernelGGL((k_degrid_mfma<R, NPOL, NB, NBW>),
                     dim3((unsigned)xcd_grouped_grid(d.max_dunits, kDegridXcdGroup)),
                     dim3(kDegridThreads), L.total, s, d, (const R *)grid, npol_g, pol_map, flag,
                     cgk, (R *)vis_out, mode);
  AV_LAUNCH_CHECK("k_degrid_mfma");
  return launch_degrid_finish<R, NPOL>(d, vis_out, mode, s);
}

// The finishing pass after a matrix-core degridder: zero the samples no unit
// visited (overwrite mode; the kernels zero the flagged polarisations of the
// samples they visit).
template <typename R, int NPOL>
static int launch_degrid_finish(const PlanDev &d, void *vis_out, int mode, hipStream_t s) {
  if (d.n_rows > 0 && !(mode & AV_DEGRID_INCREMENTAL)) {
    const int64_t nch = d.row_chan ? 1 : d.n_chan;
    if (nch >= 32) {
      const int64_t blocks = d.mask_words_per_chan * ((nch + 63) / 64);
      hipLaunchKernelGGL((k_degrid_finish<R, NPOL>), dim3((unsigned)blocks), dim3(256), 0, s, d,
                         (const uint8_t *)nullptr, (R *)vis_out);
      AV_LAUNCH_CHECK("k_degrid_finish");
    } else {
      const int64_t blocks = std::min<int64_t>((d.n_items + 255) / 256, 16384);
      hipLaunchKernelGGL(k_degrid_finish_items<R>, dim3((unsigned)blocks), dim3(256), 0, s, d,
                         (const uint8_t *)nullptr, (R *)vis_out, NPOL);
      AV_LAUNCH_CHECK("k_degrid_finish_items");
    }
  }
  return AV_OK;
}

// The split-FP16 degridder, when the plan's tiles suit it and its LDS fits;
// with PAIR over the plan's pair units.
template <int NPOL, int NB, int NBW, bool PAIR>
static int try_degrid_h3(av_plan *plan, const void *grid, int npol_g, const int32_t *pol_map,
                         const uint8_t *flag, const double *cgk, void *vis_out, int mode,
                         hipStream_t s, bool &done) {
  PlanDev &d = plan->d;
  if (!degrid_h3_fits<NPOL, NB>(d.tile_u, d.tile_v, d.support)) return AV_OK;
  if (PAIR && (d.n_pairs == 0 || !degrid_h3_pair_fits<NPOL, NB, NBW>(d.tile_u, d.tile_v,
                                                                      d.support)))
    return AV_OK;
  const DegridH3Lds L =
      degrid_h3_layout<NPOL, NB, NBW, PAIR>(d.ov, d.support, d.tile_u, d.tile_v);
  const DeviceProps &prop = device_props();
  if (prop.error) return hip_fail(prop.error, "av_degrid: device properties");
  if ((size_t)L.total > prop.lds_per_block) return AV_OK;
  done = true;
  {
    const int rc_lds = ensure_lds_limit<k_degrid_h3<NPOL, NB, NBW, PAIR>>(
        L.total, "av_degrid: hipFuncSetAttribute");
    if (rc_lds) return rc_lds;
  }
  int rc = prepare_tables(d, cgk, s);
  if (rc) return rc;
  hipLaunchKernelGGL((k_degrid_h3<NPOL, NB, NBW, PAIR>),
                     dim3((unsigned)xcd_grouped_grid(PAIR ? d.max_punits : d.max_dunits,
                                                     kDegridXcdGroup)),
                     dim3(kDegridThreads), L.total, s, d, (const float *)grid, npol_g, pol_map,
                     flag, (float *)vis_out, mode);
  AV_LAUNCH_CHECK("k_degrid_h3");
  return launch_degrid_finish<float, NPOL>(d, vis_out, mode, s);
}

template <typename R, int NPOL>
static int dispatch_degrid(av_plan *plan, const void *grid, int npol_g, const int32_t *pol_map,
                           const uint8_t *flag, const double *cgk, void *vis_out, int mode,
                           hipStream_t s) {
  // regions of at most 16 column blocks run on the matrix cores
  const int nb = degrid_mfma_blocks_needed<NPOL>(plan->d.tile_v, plan->d.support);
  if constexpr (sizeof(R) == 4) {
    // FP32 16 x 16 tiles: split-FP16 products unless FP32 products are asked for
    if (!(mode & AV_DEGRID_F32_PRODUCTS)) {
      bool done = false;
      int rc = AV_OK;
      if (NPOL == 4 && nb == 12 && degrid_halves_fit<R, NPOL, 12, 8>(plan->d.tile_u, plan->d.tile_v,
                                                                     plan->d.support)) {
        // pairs of tiles per unit where the plan has them, else one tile
        rc = try_degrid_h3<4, 20, 8, true>(plan, grid, npol_g, pol_map, flag, cgk, vis_out, mode,
                                           s, done);
        if (!done && !rc)
          rc = try_degrid_h3<4, 12, 8, false>(plan, grid, npol_g, pol_map, flag, cgk, vis_out,
                                              mode, s, done);
      } else if (nb <= 4)
        rc = try_degrid_h3<NPOL, 4, 4, false>(plan, grid, npol_g, pol_map, flag, cgk, vis_out,
                                              mode, s, done);
      else if (nb <= 8)
        rc = try_degrid_h3<NPOL, 8, 8, false>(plan, grid, npol_g, pol_map, flag, cgk, vis_out,
                                              mode, s, done);
      else if (nb <= 12)
        rc = try_degrid_h3<NPOL, 12, 12, false>(plan, grid, npol_g, pol_map, flag, cgk, vis_out,
                                                mode, s, done);
      if (done || rc) return rc;
    }
  }
  if (nb <= 4)
    return launch_degrid_mfma<R, NPOL, 4>(plan, grid, npol_g, pol_map, flag, cgk, vis_out, mode, s);
  if (nb <= 8)
    return launch_degrid_mfma<R, NPOL, 8>(plan, grid, npol_g, pol_map, flag, cgk, vis_out, mode, s);
  if (nb <= 12) {
    // four polarisations: records split by v, eight of the twelve blocks
    if constexpr (NPOL == 4)
      if (nb == 12 && degrid_halves_fit<R, NPOL, 12, 8>(plan->d.tile_u, plan->d.tile_v,
                                                         plan->d.support))
        return launch_degrid_mfma<R, NPOL, 12, 8>(plan, grid, npol_g, pol_map, flag, cgk, vis_out,
                                                  mode, s);
    return launch_degrid_mfma<R, NPOL, 12>(plan, grid, npol_g, pol_map, flag, cgk, vis_out, mode, s);
  }
  if (nb <= 16)
    return launch_degrid_mfma<R, NPOL, 16>(plan, grid, npol_g, pol_map, flag, cgk, vis_out, mode, s);
  if (plan->d.support <= 8)
    return launch_degrid<R, NPOL, 8>(plan, grid, npol_g, pol_map, flag, cgk, vis_out, mode, s);
  return launch_degrid<R, NPOL, 16>(plan, grid, npol_g, pol_map, flag, cgk, vis_out, mode, s);
}

template <typename R>
static int degrid_by_pol(av_plan *plan, const void *grid, int npol_g, const int32_t *pol_map,
                         const uint8_t *flag, int npol, const double *cgk, void *vis_out,
                         int mode, hipStream_t s) {
  switch (npol) {
    case 1: return dispatch_degrid<R, 1>(plan, grid, npol_g, pol_map, flag, cgk, vis_out, mode, s);
    case 2: return dispatch_degrid<R, 2>(plan, grid, npol_g, pol_map, flag, cgk, vis_out, mode, s);
    case 4: return dispatch_degrid<R, 4>(plan, grid, npol_g, pol_map, flag, cgk, vis_out, mode, s);
    default: return fail(AV_ERR_UNSUPPORTED, "av_degrid: npol must be 1, 2 or 4");
  }
}

}  // namespace av

using namespace av;

extern "C" int av_degrid(av_plan *plan, const void *grid, int32_t n_grid_pol,
                         const int32_t *pol_map, const uint8_t *flag, int32_t n_pol,
                         const double *cgk, void *vis_out, int32_t dtype, int32_t mode,
                         void *stream) {
  if (!plan || !grid || !cgk || !vis_out) return fail(AV_ERR_INVALID, "av_degrid: NULL argument");
  if (n_grid_pol < 1 || n_grid_pol > kMaxPol)
    return fail(AV_ERR_UNSUPPORTED, "av_degrid: n_grid_pol must be in [1, 4]");
  if (!plan->d.uvw) return fail(AV_ERR_INVALID, "av_degrid: plan has no uvw (av_plan_set_uvw)");
  if (plan->d.support > kMaxSupport)
    return fail(AV_ERR_UNSUPPORTED, "av_degrid: the plan's support exceeds 16");
  hipStream_t s = (hipStream_t)stream;
  if (dtype == AV_F32)
    return degrid_by_pol<float>(plan, grid, n_grid_pol, pol_map, flag, n_pol, cgk, vis_out, mode, s);
  if (dtype == AV_F64)
    return degrid_by_pol<double>(plan, grid, n_grid_pol, pol_map, flag, n_pol, cgk, vis_out, mode,
                                 s);
  return fail(AV_ERR_INVALID, "av_degrid: dtype must be AV_F32 or AV_F64");
}
