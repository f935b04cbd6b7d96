// grid.hip -- tiled convolutional gridding (replaces standard_grid_jit,
// core/imaging/imaging_utils/standard_grid.py:246-432).
//
// One wavefront per work unit (a bin, or a slice of a large bin); a
// workgroup holds up to four independent wavefronts (fewer when their tiles
// do not fit in LDS together) and the launch covers the plan's unit
// capacity.  A unit owns a tile_u x tile_v tile of every
// polarisation plane of one image channel, held in the wavefront's own LDS
// region cell-major.  Records are staged 64 at a time (one per lane) while
// the next batches' records and sample gathers are already in flight; the
// wavefront then applies the staged samples one after another with plain
// 16-byte read-modify-writes (lanes = (u-tap group, v-tap, chunk)): the lanes
// of one update touch distinct cells and only this wavefront touches the
// tile, so no atomics are needed.  Only taps inside the tile are applied,
// which is why each footprint is recorded in every tile it overlaps.  The
// finished tile is written to HBM once, row by row, coalesced.  A bin split
// over several units adds its partial tiles with global atomics into a tile
// that a pre-pass zeroed.  Sums of weights are FP64: each unit writes its
// partial sums and a reduction pass adds them per plane.
#include <algorithm>

#include "tile_common.h"

namespace av {

constexpr int kGridThreads = 256;
constexpr int kGridMaxWaves = kGridThreads / 64;
constexpr int kCompChunkBatches = 4;  // compensated mode: fold every 4*64 records
constexpr int kGridXcdGroup = 16;     // matrix-core kernel: workgroups per XCD group
constexpr int kGridSubtile = AV_GRID_SUBTILE;

template <typename R>
struct Cplx {
  R re, im;
};

struct GridLds {
  // offsets in bytes inside dynamic LDS; acc/st_d are per wavefront
  int tab, nrm, acc, st_d, wave_bytes, total;
  int ntab, nnrm;
};

template <typename R, int NPOL, int DVW>
__host__ __device__ inline GridLds grid_lds_layout(int ov, int support, int tu, int tv, int nw) {
  GridLds L;
  int sc = support / 2;
  L.ntab = ov * (sc + 1);
  L.nnrm = 2 * (ov / 2 + 2) + 1;  // offsets in [-(ov/2+2), ov/2+2]
  int o = 0;
  auto take = [&](int bytes) {
    int r = o;
    o += (bytes + 15) & ~15;
    return r;
  };
  L.tab = take(L.ntab * (int)sizeof(R));
  L.nrm = take(L.nnrm * (int)sizeof(double));
  const int acc_bytes = (tu * tv * 2 * NPOL * (int)sizeof(R) + 15) & ~15;
  const int std_bytes = 64 * 2 * NPOL * (int)sizeof(R);
  L.acc = o;
  L.st_d = o + acc_bytes;
  L.wave_bytes = acc_bytes + std_bytes;
  o += nw * L.wave_bytes;
  L.total = o;
  return L;
}

// Raw per-record loads of one lane: visibilities, weights, flags.  The loads
// are unconditional -- a lane without a record passes an in-bounds sample
// (its range's last record) and the caller discards the data -- so no branch
// separates a load from its use and the loads stay in flight while the
// previous batch is applied.
template <typename R, int NPOL>
struct RawSample {
  Cplx<R> vv[NPOL];
  R wv[NPOL];
  uint32_t fbits;
};

template <typename R, int NPOL>
__device__ inline void fetch_raw(RawSample<R, NPOL> &s, uint32_t item,
                                 const R *__restrict__ vis, const R *__restrict__ weight,
                                 const uint8_t *__restrict__ flag, bool psf, bool adjoint) {
  s.fbits = 0;
  if (!adjoint) {
    const Vec<R, NPOL> w = *(const Vec<R, NPOL> *)(weight + (size_t)item * NPOL);
#pragma unroll
    for (int p = 0; p < NPOL; ++p) s.wv[p] = w.v[p];
  }
  if (!psf) {
    constexpr int K = CellLayout<R, NPOL>::kCC;  // components per 16-byte piece
#pragma unroll
    for (int k = 0; k < 2 * NPOL; k += K) {
      const Vec<R, K> t = *(const Vec<R, K> *)(vis + (size_t)item * 2 * NPOL + k);
#pragma unroll
      for (int c = 0; c < K; c += 2) s.vv[(k + c) / 2] = Cplx<R>{t.v[c], t.v[c + 1]};
    }
  }
  if (flag && !adjoint) {
    if (NPOL == 4)
      s.fbits = *(const uint32_t *)(flag + (size_t)item * 4);
    else if (NPOL == 2)
      s.fbits = *(const uint16_t *)(flag + (size_t)item * 2);
    else
      s.fbits = flag[item];
  }
}

// One lane's record of the range ending at `end` (end > first): a lane past
// the end reads the range's last record, which keeps the load unconditional;
// the caller knows from base + lane < end whether the record is its own.
__device__ inline uint2 load_record(const PlanDev &P, uint32_t idx, uint32_t end) {
  const RecGridWords v = *(const RecGridWords *)(P.rec + 3 * (size_t)min(idx, end - 1));
  return make_uint2(v.x, v.y);
}

template <typename R, int NPOL, int DVW, bool COMP>
__global__ __launch_bounds__(kGridThreads) void k_grid_tiles(PlanDev P, const R *__restrict__ vis,
                                                             const R *__restrict__ weight,
                                                             const uint8_t *__restrict__ flag,
                                                             const double *__restrict__ cgk,
                                                             R *__restrict__ grid, int mode) {
  using CL = CellLayout<R, NPOL>;
  constexpr int C = CL::kC, CC = CL::kCC, H = CL::kH;
  constexpr int DQ = 64 / (H * DVW) > 0 ? 64 / (H * DVW) : 1;  // u-taps per step
  using V = Vec<R, CC>;
  extern __shared__ __align__(16) char smem[];
  const int S = P.support, sc = S / 2, ov = P.ov;
  const int Tu = P.tile_u, Tv = P.tile_v;
  const int nw = blockDim.x >> 6;  // wavefronts (units) per workgroup
  const GridLds L = grid_lds_layout<R, NPOL, DVW>(ov, S, Tu, Tv, nw);
  R *tab = (R *)(smem + L.tab);
  double *nrm = (double *)(smem + L.nrm);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  R *acc = (R *)(smem + L.acc + wave * L.wave_bytes);
  R *st_d = (R *)(smem + L.st_d + wave * L.wave_bytes);  // [record][C]

  const bool psf = (mode & AV_GRID_PSF) != 0;
  const bool adjoint = (mode & AV_GRID_ADJOINT) != 0;
  const bool accumulate = (mode & AV_GRID_ACCUMULATE) != 0;
  const int half = ov / 2 + 2;

  const uint32_t n_units = P.gunit_start[P.n_bins];
  if ((uint32_t)blockIdx.x * nw >= n_units) return;  // whole workgroup idle
  for (int i = tid; i < L.ntab; i += blockDim.x)
    tab[i] = sizeof(R) == 4 ? (R)P.tab32[i] : (R)cgk[i];
  for (int i = tid; i < L.nnrm; i += blockDim.x) nrm[i] = P.nrm[i];
  __syncthreads();  // the only workgroup barrier: tables ready
  const uint32_t u = blockIdx.x * nw + wave;
  if (u >= n_units) return;
  const int cells = Tu * Tv;

  // update lanes: (u-tap group dq, v-tap dv, chunk h)
  const int h = lane % H;
  const int dv = (lane / H) % DVW;
  const int dq = lane / (H * DVW);
  const bool lane_on = dq < DQ && dv < S;
  // flush layout: a lane owns one column when tile_v divides 64
  const bool rowwise = (64 % Tv) == 0;
  const int fcol = lane % Tv, frow = lane / Tv, fstep = 64 / Tv;

  const uint4 unit = P.gunits[u];
  const int64_t b = unit.x;
  const uint32_t first = unit.y, cnt = unit.z, fl = unit.w;
  // 32-bit divisions: bins are numbered below 2^32 (64-bit ones cost
  // about a hundred scalar instructions each)
  const uint32_t tile = (uint32_t)b / (uint32_t)P.n_img_chan;
  const int ic = (int)((uint32_t)b - tile * (uint32_t)P.n_img_chan);
  const uint32_t tu_i = tile / (uint32_t)P.ntv;
  const int u0 = (int)tu_i * Tu, v0 = (int)(tile - tu_i * (uint32_t)P.ntv) * Tv;
  const uint32_t prim_end = P.bin_start[b] + P.cnt_prim[b];
  R *gplane = grid + (size_t)ic * NPOL * P.n_u * P.n_v * 2;
  const size_t pstride = (size_t)P.n_u * P.n_v * 2;

  if (cnt == 0) {
    if (lane < NPOL) P.upart[(size_t)u * 4 + lane] = 0.0;
    if (!accumulate) {
      for (int idx = lane; idx < NPOL * cells; idx += 64) {
        int p = idx / cells, c = idx - p * cells;
        int gu = u0 + c / Tv, gv = v0 + c % Tv;
        if (gu < P.n_u && gv < P.n_v) {
          Cplx<R> z{(R)0, (R)0};
          *(Cplx<R> *)(gplane + p * pstride + ((size_t)gu * P.n_v + gv) * 2) = z;
        }
      }
    }
    return;
  }

  {
    uint4 *a4 = (uint4 *)acc;
    const int nbytes = cells * C * (int)sizeof(R);
    const int n16 = nbytes / 16;
    for (int i = lane; i < n16; i += 64) a4[i] = make_uint4(0u, 0u, 0u, 0u);
    for (int i = n16 * 16 / (int)sizeof(R) + lane; i < cells * C; i += 64) acc[i] = (R)0;
  }
  // compensated mode: per-lane running sums of the components it flushes
  constexpr int kCpt = COMP ? 32 : 1;
  R csum[kCpt], ccmp[kCpt];
  if (COMP) {
#pragma unroll
    for (int k = 0; k < kCpt; ++k) csum[k] = ccmp[k] = (R)0;
  }
  double sw[NPOL];
#pragma unroll
  for (int p = 0; p < NPOL; ++p) sw[p] = 0.0;

  // software pipeline: records of batch k+2 and sample data of batch k+1
  // are in flight while batch k is applied
  const uint32_t end = first + cnt;
  uint2 r0 = load_record(P, first + lane, end), r1 = load_record(P, first + 64 + lane, end);
  RawSample<R, NPOL> raw0;
  fetch_raw<R, NPOL>(raw0, r0.x, vis, weight, flag, psf, adjoint);
  int batch = 0;
  for (uint32_t base = first; base < end; base += 64, ++batch) {
    // ---- stage this lane's record of the batch
    uint32_t geo = kInvalidGeo;
    const uint32_t g0 = base + lane < end ? r0.y : kInvalidGeo;
    if (g0 != kInvalidGeo) {
      int au, ou, avv, ovv;
      unpack_geo(g0, au, ou, avv, ovv);
      const double norm = nrm[ou + half] * nrm[ovv + half];
      const bool prim = base + lane < prim_end;
      bool any = false;
      R d[C];
#pragma unroll
      for (int p = 0; p < NPOL; ++p) {
        R dre, dim;
        if (adjoint) {
          const R inv = (R)(1.0 / norm);
          dre = raw0.vv[p].re * inv;
          dim = raw0.vv[p].im * inv;
        } else {
          R w = raw0.wv[p];
          if ((raw0.fbits >> (8 * p)) & 0xFFu) w = (R)0;
          if (psf) {
            dre = w;
            dim = (R)0;
          } else {
            dre = raw0.vv[p].re * w;
            dim = raw0.vv[p].im * w;
          }
        }
        // standard_grid.py:378: skip NaN or exactly-zero data
        const bool ok = !(isnan(dre) || isnan(dim)) && !(dre == (R)0 && dim == (R)0);
        if (!ok) {
          dre = (R)0;
          dim = (R)0;
        } else {
          any = true;
          if (prim && !adjoint) sw[p] += (double)raw0.wv[p] * norm;
        }
        d[2 * p] = dre;
        d[2 * p + 1] = dim;
      }
      if (any) {
        geo = g0;
#pragma unroll
        for (int k = 0; k < C; k += CC) {
          V t;
#pragma unroll
          for (int c = 0; c < CC; ++c) t.v[c] = d[k + c];
          *(V *)(st_d + lane * C + k) = t;
        }
      }
    }
    // ---- issue the next loads
    RawSample<R, NPOL> raw1;
    fetch_raw<R, NPOL>(raw1, r1.x, vis, weight, flag, psf, adjoint);
    const uint2 r2 = load_record(P, base + 128 + lane, end);
    wave_lds_sync();
    // ---- apply the staged samples one at a time
    const int nst = (int)min((uint32_t)64, end - base);
    for (int j = 0; j < nst; ++j) {
      const uint32_t g = (uint32_t)__builtin_amdgcn_readlane((int)geo, j);
      if (g == kInvalidGeo) continue;  // wave-uniform
      int au, ou, avv, ovv;
      unpack_geo(g, au, ou, avv, ovv);
      const int rlo = max(0, au - sc), rhi = min(Tu, au - sc + S);
      const int col = avv - sc + dv;
      if (!lane_on || col < 0 || col >= Tv || rlo >= rhi) continue;
      const R cv = tab[tab_index(ov, dv, sc, ovv)];
      const V dd = *(const V *)(st_d + j * C + h * CC);
      R x[CC];
#pragma unroll
      for (int c = 0; c < CC; ++c) x[c] = dd.v[c] * cv;
      for (int row = rlo + dq; row < rhi; row += DQ) {
        const R cu = tab[tab_index(ov, row - au + sc, sc, ou)];
        V *cell = (V *)(acc + (row * Tv + col) * C + h * CC);
        V a = *cell;
#pragma unroll
        for (int c = 0; c < CC; ++c) a.v[c] += x[c] * cu;
        *cell = a;
      }
    }
    wave_lds_sync();
    if (COMP && ((batch + 1) % kCompChunkBatches == 0) && base + 64 < end) {
      // fold the LDS partial sums into the per-lane compensated sums
#pragma unroll
      for (int k = 0; k < kCpt; ++k) {
        const int idx = lane + k * 64;
        if (idx < C * cells) {
          const R x = acc[idx];
          acc[idx] = (R)0;
          const R t = csum[k] + x;  // Neumaier
          if (fabs(csum[k]) >= fabs(x))
            ccmp[k] += (csum[k] - t) + x;
          else
            ccmp[k] += (x - t) + csum[k];
          csum[k] = t;
        }
      }
      wave_lds_sync();
    }
    raw0 = raw1;
    r0 = r1;
    r1 = r2;
  }

  // ---- flush the tile
  const bool multi = (fl & kUnitMulti) != 0;
  if (COMP) {
#pragma unroll
    for (int k = 0; k < kCpt; ++k) {
      const int idx = lane + k * 64;
      if (idx < C * cells) {
        const R x = acc[idx];
        const R t = csum[k] + x;
        if (fabs(csum[k]) >= fabs(x))
          ccmp[k] += (csum[k] - t) + x;
        else
          ccmp[k] += (x - t) + csum[k];
        acc[idx] = t + ccmp[k];
      }
    }
    wave_lds_sync();
  }
  auto flush_cell = [&](int p, int c, int gu, int gv) {
    const Cplx<R> val = *(const Cplx<R> *)(acc + c * C + 2 * p);
    R *dst = gplane + p * pstride + ((size_t)gu * P.n_v + gv) * 2;
    if (multi) {
      if (val.re != (R)0) atomicAdd(dst, val.re);
      if (val.im != (R)0) atomicAdd(dst + 1, val.im);
    } else if (accumulate) {
      Cplx<R> old = *(Cplx<R> *)dst;
      old.re += val.re;
      old.im += val.im;
      *(Cplx<R> *)dst = old;
    } else {
      *(Cplx<R> *)dst = val;
    }
  };
  if (rowwise) {
    for (int p = 0; p < NPOL; ++p)
      for (int uu = frow; uu < Tu; uu += fstep) {
        const int gu = u0 + uu, gv = v0 + fcol;
        if (gu < P.n_u && gv < P.n_v) flush_cell(p, uu * Tv + fcol, gu, gv);
      }
  } else {
    for (int idx = lane; idx < NPOL * cells; idx += 64) {
      const int p = idx / cells, c = idx - p * cells;
      const int gu = u0 + c / Tv, gv = v0 + c % Tv;
      if (gu < P.n_u && gv < P.n_v) flush_cell(p, c, gu, gv);
    }
  }
  // ---- sums of weights (FP64): this unit's partial per polarisation
#pragma unroll
  for (int p = 0; p < NPOL; ++p) {
    double sv = adjoint ? 0.0 : sw[p];
    for (int off = 32; off >= 1; off >>= 1) sv += __shfl_xor(sv, off);
    if (lane == 0) P.upart[(size_t)u * 4 + p] = sv;
  }
}

// ---------------------------------------------------------------------------
// Matrix-core gridder.  The kernel is separable, so the update of one tile by
// a batch of records k is a sum of rank-1 products,
//   tile[u][c][v] += sum_k wu_k(u) * (wv_k(v) * d_k[c]),
// a (tile_u x K) by (K x C*tile_v) matrix product whose K dimension runs over
// records.  v_mfma_*_16x16x4 consumes four records per step: lane l supplies
// A[i = l%16][k = l/16] = wu_k(row i) and, for every 16-column block,
// B[k = l/16][j = l%16] = wv_k(column j) * d_k[c]; the tile accumulates in the
// result registers (lane l holds column l%16 of four rows; MfmaOp::row).
// Rows and columns outside a record's footprint get weight 0, so a record
// costs the same dense blocks wherever it lies in the tile, and no LDS
// read-modify-write is needed at all.  A unit is one wavefront, as above.

template <typename R>
struct MfmaOp;
// Result layout (lane l, register r): the FP32 instruction holds row
// 4*(l/16) + r, the FP64 one row l/16 + 4*r; both hold column l%16.
template <>
struct MfmaOp<float> {
  typedef float V __attribute__((ext_vector_type(4)));
  __device__ static V mma(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lk, int r) { return 4 * lk + r; }
};
template <>
struct MfmaOp<double> {
  typedef double V __attribute__((ext_vector_type(4)));
  __device__ static V mma(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lk, int r) { return lk + 4 * r; }
};

struct MfmaLds {
  int tab, nrm, st_d, st_geo, wave_bytes, total, ntab, nnrm;
  int pad, nw, rows;  // weight table: rows of taps -pad .. S + pad - 1, nw offsets each
};

// The matrix-core kernel's weight table is the call's table (taps -1 .. S)
// with zero rows added on both sides, so that every tap a 16 x 16 block can
// ask for -- tile-relative positions run from -(S - 1 - S/2) to T - 1 + S/2,
// taps from 1 - T to T + S - 2 -- is a row of it and no index needs a clamp.
template <typename R, int NPOL>
__host__ __device__ inline MfmaLds mfma_lds_layout(int ov, int support, int tile) {
  MfmaLds L;
  const int sc = support / 2;
  L.ntab = ov * (sc + 1);
  L.nnrm = 2 * (ov / 2 + 2) + 1;
  L.pad = tile;
  L.nw = WTab(ov, support).nw;
  L.rows = support + 2 * L.pad;
  int o = 0;
  L.tab = o;
  o += (L.rows * L.nw * (int)sizeof(R) + 15) & ~15;
  L.nrm = o;
  o += (L.nnrm * (int)sizeof(double) + 15) & ~15;
  L.st_d = o;                           // per wave: [64][2*NPOL] R
  L.st_geo = 64 * 2 * NPOL * (int)sizeof(R);  // offset inside the wave region: [64] u32
  L.wave_bytes = L.st_geo + 64 * 4;
  L.total = o + kGridMaxWaves * L.wave_bytes;
  return L;
}

// Fill the padded table: rows -1 .. S from the call's table, zeros elsewhere.
// The zero rows are plain stores; only the call's rows are loaded, K per
// thread in flight (one round trip for the common tables).
// `scale` (a power of two, so exact) multiplies every copied weight.
template <typename R>
__device__ inline void copy_wtab_padded(R *w, const MfmaLds &L, const PlanDev &P,
                                        R scale = (R)1) {
  const R *src;
  if constexpr (sizeof(R) == 4)
    src = P.wtab32;
  else
    src = P.wtab64;
  const int S = P.support, n = L.rows * L.nw, nsrc = (S + 2) * L.nw;
  const int lo = (L.pad - 1) * L.nw;  // row pad - 1 of the padded table is tap -1
  for (int i = threadIdx.x; i < lo; i += blockDim.x) w[i] = (R)0;
  for (int i = lo + nsrc + threadIdx.x; i < n; i += blockDim.x) w[i] = (R)0;
  constexpr int K = 4;  // loads in flight per thread
  for (int i0 = threadIdx.x; i0 < nsrc; i0 += K * blockDim.x) {
    R v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = src[min(i0 + k * (int)blockDim.x, nsrc - 1)];
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (i0 + k * (int)blockDim.x < nsrc) w[lo + i0 + k * blockDim.x] = v[k] * scale;
  }
}

// Workgroups per CU the FP32 kernels with up to 8 result blocks are compiled
// for: six (80 registers, no spills; a 16 x 16 tile's workgroup needs 26 KB
// of LDS).  Other variants keep the compiler's choice.
template <typename R, int NPOL, int MB, int VB, bool COMP>
constexpr int grid_mfma_min_blocks() {
  return (sizeof(R) == 4 && !COMP && MB * 2 * NPOL * VB <= 8) ? 6 : 1;
}

template <typename R, int NPOL, int MB, int VB, bool COMP, bool PSF>
__global__ __launch_bounds__(kGridThreads, (grid_mfma_min_blocks<R, NPOL, MB, VB, COMP>()))
void k_grid_mfma(PlanDev P, const R *__restrict__ vis,
                                                            const R *__restrict__ weight,
                                                            const uint8_t *__restrict__ flag,
                                                            const double *__restrict__ cgk,
                                                            R *__restrict__ grid, int mode) {
  using MO = MfmaOp<R>;
  using MV = typename MO::V;
  constexpr int C = 2 * NPOL;
  extern __shared__ __align__(16) char smem[];
  const int S = P.support, sc = S / 2, ov = P.ov;
  const MfmaLds L = mfma_lds_layout<R, NPOL>(ov, S, 16 * (MB > VB ? MB : VB));
  R *tab = (R *)(smem + L.tab);
  double *nrm = (double *)(smem + L.nrm);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  char *wbase = smem + L.st_d + wave * L.wave_bytes;
  R *st_d = (R *)wbase;                           // [record][C]
  // [record] table bases of the u and v taps, 16 bits each (a padded table
  // has at most (16 + 2 * 64) rows x 255 offsets < 2^16 entries)
  uint32_t *st_geo = (uint32_t *)(wbase + L.st_geo);

  constexpr bool psf = PSF;  // PSF data are real: the imaginary blocks are skipped
  const bool adjoint = (mode & AV_GRID_ADJOINT) != 0;
  const bool accumulate = (mode & AV_GRID_ACCUMULATE) != 0;
  const int half = ov / 2 + 2;

  const uint32_t n_units = P.gunit_start[P.n_bins];
  // 16 workgroups of 4 units: the 64 channels of a tile under one L2
  const uint32_t blk = xcd_grouped_block(blockIdx.x, kGridXcdGroup);
  if (blk * kGridMaxWaves >= n_units) return;  // whole workgroup idle
  copy_wtab_padded(tab, L, P);
  for (int i = tid; i < L.nnrm; i += blockDim.x) nrm[i] = P.nrm[i];
  __syncthreads();  // the only workgroup barrier: tables ready
  // the same for every lane: scalar registers for the unit and its bounds
  const uint32_t u = (uint32_t)__builtin_amdgcn_readfirstlane((int)(blk * kGridMaxWaves + wave));
  if (u >= n_units) return;

  const int Tu = 16 * MB, Tv = 16 * VB;
  const uint4 unit = P.gunits[u];
  const int64_t b = unit.x;
  const uint32_t first = unit.y, cnt = unit.z, fl = unit.w;
  // 32-bit divisions: bins are numbered below 2^32 (64-bit ones cost
  // about a hundred scalar instructions each)
  const uint32_t tile = (uint32_t)b / (uint32_t)P.n_img_chan;
  const int ic = (int)((uint32_t)b - tile * (uint32_t)P.n_img_chan);
  const uint32_t tu_i = tile / (uint32_t)P.ntv;
  const int u0 = (int)tu_i * Tu, v0 = (int)(tile - tu_i * (uint32_t)P.ntv) * Tv;
  const uint32_t prim_end = P.bin_start[b] + P.cnt_prim[b];
  R *gplane = grid + (size_t)ic * NPOL * P.n_u * P.n_v * 2;
  const size_t pstride = (size_t)P.n_u * P.n_v * 2;
  const int li = lane & 15, lk = lane >> 4;

  if (cnt == 0) {
    if (lane < NPOL) P.upart[(size_t)u * 4 + lane] = 0.0;
    if (!accumulate) {
      for (int idx = lane; idx < NPOL * Tu * Tv; idx += 64) {
        const int p = idx / (Tu * Tv), c = idx - p * (Tu * Tv);
        const int gu = u0 + c / Tv, gv = v0 + c % Tv;
        if (gu < P.n_u && gv < P.n_v) {
          Cplx<R> z{(R)0, (R)0};
          *(Cplx<R> *)(gplane + p * pstride + ((size_t)gu * P.n_v + gv) * 2) = z;
        }
      }
    }
    return;
  }

  MV acc[MB][C][VB];
#pragma unroll
  for (int m = 0; m < MB; ++m)
#pragma unroll
    for (int c = 0; c < C; ++c)
#pragma unroll
      for (int vb = 0; vb < VB; ++vb) acc[m][c][vb] = MV{0, 0, 0, 0};
  // compensated mode: running (sum, compensation) pairs of the accumulators
  constexpr int kNc = COMP ? MB * C * VB * 4 : 1;
  R csum[kNc], ccmp[kNc];
  if (COMP) {
#pragma unroll
    for (int k = 0; k < kNc; ++k) csum[k] = ccmp[k] = (R)0;
  }
  auto fold = [&](bool reset) {
#pragma unroll
    for (int m = 0; m < MB; ++m)
#pragma unroll
      for (int c = 0; c < C; ++c)
#pragma unroll
        for (int vb = 0; vb < VB; ++vb)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int k = ((m * C + c) * VB + vb) * 4 + r;
            const R x = acc[m][c][vb][r];
            const R t = csum[k] + x;  // Neumaier
            if (fabs(csum[k]) >= fabs(x))
              ccmp[k] += (csum[k] - t) + x;
            else
              ccmp[k] += (x - t) + csum[k];
            csum[k] = t;
            acc[m][c][vb][r] = reset ? (R)0 : t + ccmp[k];
          }
  };
  double sw[NPOL];
#pragma unroll
  for (int p = 0; p < NPOL; ++p) sw[p] = 0.0;

  const uint32_t end = first + cnt;
  uint2 r0 = load_record(P, first + lane, end), r1 = load_record(P, first + 64 + lane, end);
  RawSample<R, NPOL> raw0;
  fetch_raw<R, NPOL>(raw0, r0.x, vis, weight, flag, psf, adjoint);
  int batch = 0;
  for (uint32_t base = first; base < end; base += 64, ++batch) {
    // ---- stage this lane's record: data (zeros when skipped) and geometry
    uint32_t geo = kInvalidGeo;
    R d[C];
#pragma unroll
    for (int k = 0; k < C; ++k) d[k] = (R)0;
    const uint32_t g0 = base + lane < end ? r0.y : kInvalidGeo;
    if (g0 != kInvalidGeo) {
      int au, ou, avv, ovv;
      unpack_geo(g0, au, ou, avv, ovv);
      const double norm = nrm[ou + half] * nrm[ovv + half];
      const bool prim = base + lane < prim_end;
      bool any = false;
#pragma unroll
      for (int p = 0; p < NPOL; ++p) {
        R dre, dim;
        if (adjoint) {
          const R inv = (R)(1.0 / norm);
          dre = raw0.vv[p].re * inv;
          dim = raw0.vv[p].im * inv;
        } else {
          R w = raw0.wv[p];
          if ((raw0.fbits >> (8 * p)) & 0xFFu) w = (R)0;
          if (psf) {
            dre = w;
            dim = (R)0;
          } else {
            dre = raw0.vv[p].re * w;
            dim = raw0.vv[p].im * w;
          }
        }
        // standard_grid.py:378: skip NaN or exactly-zero data
        const bool ok = !(isnan(dre) || isnan(dim)) && !(dre == (R)0 && dim == (R)0);
        if (ok) {
          any = true;
          d[2 * p] = dre;
          d[2 * p + 1] = dim;
          if (prim && !adjoint) sw[p] += (double)raw0.wv[p] * norm;
        }
      }
      if (any) geo = g0;
    }
#pragma unroll
    for (int k = 0; k < C; k += CellLayout<R, NPOL>::kCC) {
      Vec<R, CellLayout<R, NPOL>::kCC> t;
#pragma unroll
      for (int c = 0; c < CellLayout<R, NPOL>::kCC; ++c) t.v[c] = d[k + c];
      *(Vec<R, CellLayout<R, NPOL>::kCC> *)(st_d + lane * C + k) = t;
    }
    {
      // table index of tap (li - position + sc) at this record's offsets is
      // base + li * nw (taps -pad .. are rows 0 ..); a skipped record has
      // zero data, so any in-range base (0) serves it
      uint32_t tb = 0u;
      if (geo != kInvalidGeo) {
        int au, ou, avv, ovv;
        unpack_geo(geo, au, ou, avv, ovv);
        tb = (uint32_t)((sc - au + L.pad) * L.nw + ou + L.nw / 2) |
             ((uint32_t)((sc - avv + L.pad) * L.nw + ovv + L.nw / 2) << 16);
      }
      st_geo[lane] = tb;
    }
    // ---- issue the next loads
    RawSample<R, NPOL> raw1;
    fetch_raw<R, NPOL>(raw1, r1.x, vis, weight, flag, psf, adjoint);
    const uint2 r2 = load_record(P, base + 128 + lane, end);
    wave_lds_sync();
    // ---- four records per matrix-core step
    const int nst = (int)min((uint32_t)64, end - base);
    const int li_nw = li * L.nw;
    for (int s4 = 0; s4 < nst; s4 += 4) {
      const int rec = s4 + lk;
      const uint32_t tb = st_geo[rec];
      const uint32_t tbu = tb & 0xFFFFu, tbv = tb >> 16;
      R a[MB], wv[VB];
#pragma unroll
      for (int m = 0; m < MB; ++m) a[m] = tab[tbu + li_nw + 16 * m * L.nw];
#pragma unroll
      for (int vb = 0; vb < VB; ++vb) wv[vb] = tab[tbv + li_nw + 16 * vb * L.nw];
      R dd[C];
#pragma unroll
      for (int k = 0; k < C; k += CellLayout<R, NPOL>::kCC) {
        const Vec<R, CellLayout<R, NPOL>::kCC> t =
            *(const Vec<R, CellLayout<R, NPOL>::kCC> *)(st_d + rec * C + k);
#pragma unroll
        for (int c = 0; c < CellLayout<R, NPOL>::kCC; ++c) dd[k + c] = t.v[c];
      }
#pragma unroll
      for (int c = 0; c < C; ++c) {
        if constexpr (PSF) {
          if (c & 1) continue;
        }
#pragma unroll
        for (int vb = 0; vb < VB; ++vb) {
          const R bv = wv[vb] * dd[c];
#pragma unroll
          for (int m = 0; m < MB; ++m) acc[m][c][vb] = MO::mma(a[m], bv, acc[m][c][vb]);
        }
      }
    }
    wave_lds_sync();
    if (COMP && ((batch + 1) % kCompChunkBatches == 0) && base + 64 < end) fold(true);
    raw0 = raw1;
    r0 = r1;
    r1 = r2;
  }
  if (COMP) fold(false);

  // ---- flush: lane holds rows MO::row(lane/16, r), column lane%16 of every block
  const bool multi = (fl & kUnitMulti) != 0;
#pragma unroll
  for (int m = 0; m < MB; ++m)
#pragma unroll
    for (int vb = 0; vb < VB; ++vb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gu = u0 + 16 * m + MO::row(lk, r), gv = v0 + 16 * vb + li;
        if (gu >= P.n_u || gv >= P.n_v) continue;
#pragma unroll
        for (int p = 0; p < NPOL; ++p) {
          const R re = acc[m][2 * p][vb][r], im = acc[m][2 * p + 1][vb][r];
          R *dst = gplane + p * pstride + ((size_t)gu * P.n_v + gv) * 2;
          if (multi) {
            if (re != (R)0) atomicAdd(dst, re);
            if (im != (R)0) atomicAdd(dst + 1, im);
          } else if (accumulate) {
            Cplx<R> old = *(Cplx<R> *)dst;
            old.re += re;
            old.im += im;
            *(Cplx<R> *)dst = old;
          } else {
            *(Cplx<R> *)dst = Cplx<R>{re, im};
          }
        }
      }
  // ---- sums of weights (FP64): this unit's partial per polarisation
#pragma unroll
  for (int p = 0; p < NPOL; ++p) {
    double sv = adjoint ? 0.0 : sw[p];
    for (int off = 32; off >= 1; off >>= 1) sv += __shfl_xor(sv, off);
    if (lane == 0) P.upart[(size_t)u * 4 + p] = sv;
  }
}

// sums of weights per (image channel, polarisation) from the unit partials:
// one workgroup per image channel walks the channel's bins in tile order and
// reduces in a fixed order, so repeated calls give bit-identical sums
__global__ __launch_bounds__(256) void k_sumwt_reduce(PlanDev P, int npol, double *sumwt) {
  __shared__ double red[4][kMaxPol];
  const int ic = blockIdx.x;
  const int64_t ntiles = (int64_t)P.ntu * P.ntv;
  double acc[kMaxPol] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t t = threadIdx.x; t < ntiles; t += blockDim.x) {
    const int64_t b = t * P.n_img_chan + ic;
    for (uint32_t u = P.gunit_start[b]; u < P.gunit_start[b + 1]; ++u)
      for (int p = 0; p < npol; ++p) acc[p] += P.upart[(size_t)u * 4 + p];
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int p = 0; p < npol; ++p) {
    double v = acc[p];
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    if (lane == 0) red[wave][p] = v;
  }
  __syncthreads();
  if (threadIdx.x < npol) {
    const int p = threadIdx.x;
    double v = red[0][p];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) v += red[w][p];
    sumwt[ic * npol + p] += v;
  }
}

// zero the tiles of bins that several units add into (overwrite mode): one
// thread per unit reads the unit list in one coalesced pass; the thread of a
// split bin's first unit clears that tile (such bins are few: those with more
// records than a unit holds)
template <typename R>
__global__ __launch_bounds__(256) void k_zero_multi(PlanDev P, R *grid, int npol) {
  const uint32_t n_units = P.gunit_start[P.n_bins];
  const int Tu = P.tile_u, Tv = P.tile_v;
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n_units;
       u += gridDim.x * blockDim.x) {
    const uint4 unit = P.gunits[u];
    if (!((unit.w & kUnitMulti) && (unit.w & kUnitFirst))) continue;
    const uint32_t tile = unit.x / (uint32_t)P.n_img_chan;
    const int ic = (int)(unit.x - tile * (uint32_t)P.n_img_chan);
    const uint32_t tu_i = tile / (uint32_t)P.ntv;
    const int u0 = (int)tu_i * Tu, v0 = (int)(tile - tu_i * (uint32_t)P.ntv) * Tv;
    const int u1 = min(u0 + Tu, P.n_u), v1 = min(v0 + Tv, P.n_v);
    for (int p = 0; p < npol; ++p)
      for (int gu = u0; gu < u1; ++gu) {
        Cplx<R> *row = (Cplx<R> *)(grid + (((size_t)ic * npol + p) * P.n_u + gu) * P.n_v * 2);
        for (int gv = v0; gv < v1; ++gv) row[gv] = Cplx<R>{(R)0, (R)0};
      }
  }
}

__global__ void k_zero_f64(double *p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

template <typename R, int NPOL, int DVW, bool COMP>
static int launch_grid(av_plan *plan, const void *vis, const void *weight, const uint8_t *flag,
                       const double *cgk, void *grid, double *sumwt, int mode, hipStream_t s) {
  PlanDev &d = plan->d;
  auto kfn = k_grid_tiles<R, NPOL, DVW, COMP>;
  const DeviceProps &prop = device_props();
  if (prop.error) return hip_fail(prop.error, "av_grid: device properties");
  // as many wavefronts per workgroup as LDS holds tiles for, up to four
  int nw = kGridMaxWaves;
  GridLds L = grid_lds_layout<R, NPOL, DVW>(d.ov, d.support, d.tile_u, d.tile_v, nw);
  while (nw > 1 && (size_t)L.total > prop.lds_per_block)
    L = grid_lds_layout<R, NPOL, DVW>(d.ov, d.support, d.tile_u, d.tile_v, --nw);
  if ((size_t)L.total > prop.lds_per_block)
    return fail(AV_ERR_UNSUPPORTED, "av_grid: tile too large for LDS (" +
                                        std::to_string(L.total) + " bytes)");
  {
    const int rc_lds = ensure_lds_limit<k_grid_tiles<R, NPOL, DVW, COMP>>(L.total, "av_grid: hipFuncSetAttribute");
    if (rc_lds) return rc_lds;
  }
  int rc = prepare_tables(d, cgk, s);
  if (rc) return rc;
  // one wavefront per unit; the plan's capacity bounds the unit count
  const unsigned nwg = (unsigned)((d.max_gunits + nw - 1) / nw);
  hipLaunchKernelGGL(kfn, dim3(nwg), dim3(64 * nw), L.total, s, d,
                     (const R *)vis, (const R *)weight, flag, cgk, (R *)grid, mode);
  AV_LAUNCH_CHECK("k_grid_tiles");
  if (!(mode & AV_GRID_ADJOINT)) {
    hipLaunchKernelGGL(k_sumwt_reduce, dim3((unsigned)d.n_img_chan), dim3(256), 0, s, d, NPOL,
                       sumwt);
    AV_LAUNCH_CHECK("k_sumwt_reduce");
  }
  return AV_OK;
}

template <typename R, int NPOL, int MB, int VB, bool COMP, bool PSF>
static int launch_grid_mfma(av_plan *plan, const void *vis, const void *weight,
                            const uint8_t *flag, const double *cgk, void *grid, double *sumwt,
                            int mode, hipStream_t s) {
  PlanDev &d = plan->d;
  auto kfn = k_grid_mfma<R, NPOL, MB, VB, COMP, PSF>;
  const DeviceProps &prop = device_props();
  if (prop.error) return hip_fail(prop.error, "av_grid: device properties");
  const MfmaLds L = mfma_lds_layout<R, NPOL>(d.ov, d.support, 16 * (MB > VB ? MB : VB));
  {
    const int rc_lds = ensure_lds_limit<k_grid_mfma<R, NPOL, MB, VB, COMP, PSF>>(L.total, "av_grid: hipFuncSetAttribute");
    if (rc_lds) return rc_lds;
  }
  int rc = prepare_tables(d, cgk, s);
  if (rc) return rc;
  const unsigned nwg = (unsigned)xcd_grouped_grid(
      (d.max_gunits + kGridMaxWaves - 1) / kGridMaxWaves, kGridXcdGroup);
  hipLaunchKernelGGL(kfn, dim3(nwg), dim3(kGridThreads), L.total, s, d, (const R *)vis,
                     (const R *)weight, flag, cgk, (R *)grid, mode);
  AV_LAUNCH_CHECK("k_grid_mfma");
  if (!(mode & AV_GRID_ADJOINT)) {
    hipLaunchKernelGGL(k_sumwt_reduce, dim3((unsigned)d.n_img_chan), dim3(256), 0, s, d, NPOL,
                       sumwt);
    AV_LAUNCH_CHECK("k_sumwt_reduce");
  }
  return AV_OK;
}

// ---------------------------------------------------------------------------
// Split-FP16 matrix-core gridder (FP32 grids, 16 x 16 tiles, the default for
// them).  The FP16 product v_mfma_f32_16x16x32_f16 takes 32 records per
// instruction in about the time the FP32 16x16x4 product takes 4, so the
// tile update out[u][(c,v)] += sum_k wu_k(u) * (wv_k(v) * d_k[c]) runs on it
// with each FP32 operand x split into two FP16 halves, x = hi + lo with
// hi = fp16(x) and lo = fp16(x - hi), and three products per block:
// hi*hi + hi*lo + lo*hi, accumulated in FP32.  The split keeps 22 bits of
// each operand (the dropped lo*lo term is below 2^-22 relative), against 24
// for an FP32 product, and each product is exact in the FP32 accumulator.
//
// FP16 has a narrow range, so every operand is scaled by a power of two
// (exactly): the weight table by 2^14 (weights are at most 1, so the table
// stays below 2^14 and its values above the FP16 subnormals down to 4e-9),
// and the data of a unit by 2^-E with E the exponent of the largest |d| the
// unit has seen so far, which puts them below 1.  The products wu * wv * d
// are then below 2^28 (the u- and v-weights both carry 2^14) and the FP32
// accumulators hold the tile times 2^(28 - E).  When a batch of 64 records
// raises E, the accumulators are rescaled by the power of two first, so
// records of any magnitude mix; the flush multiplies by 2^(E - 28).  Errors are therefore relative to the
// largest data of the unit, as for FP32 sums.  Data of +-inf make the cells
// they touch NaN (the FP32 kernel gives inf there).
//
// Lane l of the 16x16x32 product holds A[l % 16][8 * (l / 16) + j] and
// B[8 * (l / 16) + j][l % 16], j = 0..7, and result rows 4 * (l / 16) + r of
// column l % 16 (checked on the GPU).  So lane l forms, for records
// 8 * (l / 16) + j of the step, the u-weight of tile row l % 16 and the
// v-weight of tile column l % 16 times each component.  Staged data are
// component-major ([c][64 records]), so the eight records of a lane are two
// 16-byte LDS reads per component.
constexpr int kH3TabExp = 14;  // weight table scale 2^14
constexpr int kH3NoScale = -100000;


struct H3Lds {
  int tab, nrm, st_d, st_gu, st_gv, wave_bytes, total, nw, pad, rows, nnrm;
};

// The padded FP32 weight table (as the 16 x 16 FP32 matrix-core kernel's,
// times 2^14), the FP64 kernel-sum table, and per wavefront the staged data
// [c][64 records] and the records' two table offsets in bytes.
template <int NPOL>
__host__ __device__ inline H3Lds h3_lds_layout(int ov, int support) {
  const MfmaLds M = mfma_lds_layout<float, NPOL>(ov, support, 16);
  H3Lds L;
  L.tab = M.tab;
  L.nrm = M.nrm;
  L.nw = M.nw;
  L.pad = M.pad;
  L.rows = M.rows;
  L.nnrm = M.nnrm;
  L.st_d = M.st_d;
  L.st_gu = 64 * 2 * NPOL * 4;
  L.st_gv = L.st_gu + 64 * 4;
  L.wave_bytes = L.st_gv + 64 * 4;
  L.total = L.st_d + kGridMaxWaves * L.wave_bytes;
  return L;
}

template <int NPOL, bool PSF>
__global__ __launch_bounds__(kGridThreads, (NPOL == 4 ? 4 : 5))
void k_grid_h3(PlanDev P, const float *__restrict__ vis, const float *__restrict__ weight,
               const uint8_t *__restrict__ flag, float *__restrict__ grid, int mode) {
  constexpr int C = 2 * NPOL;
  extern __shared__ __align__(16) char smem[];
  const int S = P.support, sc = S / 2, ov = P.ov;
  const H3Lds L = h3_lds_layout<NPOL>(ov, S);
  float *tab = (float *)(smem + L.tab);
  const char *tabc = smem + L.tab;
  double *nrm = (double *)(smem + L.nrm);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  char *wbase = smem + L.st_d + wave * L.wave_bytes;
  float *st_d = (float *)wbase;                      // [c][64 records]
  uint32_t *st_gu = (uint32_t *)(wbase + L.st_gu);  // [64 records] u-tap table offsets, bytes
  uint32_t *st_gv = (uint32_t *)(wbase + L.st_gv);  // [64 records] v-tap table offsets, bytes

  const bool adjoint = (mode & AV_GRID_ADJOINT) != 0;
  const bool accumulate = (mode & AV_GRID_ACCUMULATE) != 0;
  const int half = ov / 2 + 2;

  const uint32_t n_units = P.gunit_start[P.n_bins];
  const uint32_t blk = xcd_grouped_block(blockIdx.x, kGridXcdGroup);
  if (blk * kGridMaxWaves >= n_units) return;
  // The tables (the padded weight table times 2^14, as copy_wtab_padded
  // builds it, and the kernel-sum table) are copied into LDS while the unit
  // and its first two batches of records load: the table rows and the unit
  // are requested first, the records as soon as the unit arrives, and the
  // LDS stores wait only for the table rows.
  constexpr int kTabK = 4;  // table loads in flight per thread
  const MfmaLds M = mfma_lds_layout<float, NPOL>(ov, S, 16);
  const int nsrc = (S + 2) * M.nw, tlo = (M.pad - 1) * M.nw, ntab = M.rows * M.nw;
  float tv[kTabK];
#pragma unroll
  for (int k = 0; k < kTabK; ++k) tv[k] = P.wtab32[min(tid + k * kGridThreads, nsrc - 1)];
  const double nv = P.nrm[min(tid, L.nnrm - 1)];
  const uint32_t u = (uint32_t)__builtin_amdgcn_readfirstlane((int)(blk * kGridMaxWaves + wave));
  const bool live = u < n_units;
  uint4 unit = make_uint4(0u, 0u, 0u, 0u);
  if (live) unit = P.gunits[u];
  const uint32_t first = unit.y, cnt = unit.z, fl = unit.w;
  const uint32_t end = first + cnt;
  // unconditional loads (a branch here makes the compiler wait for them, and
  // for the table rows before them, at its join): an empty unit reads record
  // 0, which every plan holds
  const uint32_t lfirst = cnt ? first : 0u, lend = cnt ? end : 1u;
  uint2 r0 = load_record(P, lfirst + lane, lend), r1 = load_record(P, lfirst + 64 + lane, lend);
  {
    const float scale = (float)(1 << kH3TabExp);
    for (int i = tid; i < tlo; i += kGridThreads) tab[i] = 0.f;
    for (int i = tlo + nsrc + tid; i < ntab; i += kGridThreads) tab[i] = 0.f;
#pragma unroll
    for (int k = 0; k < kTabK; ++k)
      if (tid + k * kGridThreads < nsrc) tab[tlo + tid + k * kGridThreads] = tv[k] * scale;
    for (int i = tid + kTabK * kGridThreads; i < nsrc; i += kGridThreads)
      tab[tlo + i] = P.wtab32[i] * scale;  // tables of more than kTabK rows of threads
    if (tid < L.nnrm) nrm[tid] = nv;
    for (int i = tid + kGridThreads; i < L.nnrm; i += kGridThreads) nrm[i] = P.nrm[i];
  }
  __syncthreads();
  if (!live) return;

  const int64_t b = unit.x;
  const uint32_t tile = (uint32_t)b / (uint32_t)P.n_img_chan;
  const int ic = (int)((uint32_t)b - tile * (uint32_t)P.n_img_chan);
  const uint32_t tu_i = tile / (uint32_t)P.ntv;
  const int u0 = (int)tu_i * 16, v0 = (int)(tile - tu_i * (uint32_t)P.ntv) * 16;
  const uint32_t prim_end = P.bin_start[b] + P.cnt_prim[b];
  float *gplane = grid + (size_t)ic * NPOL * P.n_u * P.n_v * 2;
  const size_t pstride = (size_t)P.n_u * P.n_v * 2;
  const int li = lane & 15, lk = lane >> 4;

  if (cnt == 0) {
    if (lane < NPOL) P.upart[(size_t)u * 4 + lane] = 0.0;
    if (!accumulate) {
      for (int idx = lane; idx < NPOL * 256; idx += 64) {
        const int p = idx >> 8, c = idx & 255;
        const int gu = u0 + (c >> 4), gv = v0 + (c & 15);
        if (gu < P.n_u && gv < P.n_v)
          *(Cplx<float> *)(gplane + p * pstride + ((size_t)gu * P.n_v + gv) * 2) =
              Cplx<float>{0.f, 0.f};
      }
    }
    return;
  }

  H3Float4 acc[C];
#pragma unroll
  for (int c = 0; c < C; ++c) acc[c] = H3Float4{0.f, 0.f, 0.f, 0.f};
  int E = kH3NoScale;  // the unit's data exponent (uniform)
  double sw[NPOL];
#pragma unroll
  for (int p = 0; p < NPOL; ++p) sw[p] = 0.0;

  RawSample<float, NPOL> raw0;
  fetch_raw<float, NPOL>(raw0, r0.x, vis, weight, flag, PSF, adjoint);
  for (uint32_t base = first; base < end; base += 64) {
    // ---- this lane's record: data (zeros when skipped) and geometry
    uint32_t geo = kInvalidGeo;
    float d[C];
#pragma unroll
    for (int k = 0; k < C; ++k) d[k] = 0.f;
    const uint32_t g0 = base + lane < end ? r0.y : kInvalidGeo;
    if (g0 != kInvalidGeo) {
      int au, ou, avv, ovv;
      unpack_geo(g0, au, ou, avv, ovv);
      const double norm = nrm[ou + half] * nrm[ovv + half];
      const bool prim = base + lane < prim_end;
      bool any = false;
#pragma unroll
      for (int p = 0; p < NPOL; ++p) {
        float dre, dim;
        if (adjoint) {
          const float inv = (float)(1.0 / norm);
          dre = raw0.vv[p].re * inv;
          dim = raw0.vv[p].im * inv;
        } else {
          float w = raw0.wv[p];
          if ((raw0.fbits >> (8 * p)) & 0xFFu) w = 0.f;
          if (PSF) {
            dre = w;
            dim = 0.f;
          } else {
            dre = raw0.vv[p].re * w;
            dim = raw0.vv[p].im * w;
          }
        }
        // standard_grid.py:378: skip NaN or exactly-zero data
        const bool ok = !(isnan(dre) || isnan(dim)) && !(dre == 0.f && dim == 0.f);
        if (ok) {
          any = true;
          d[2 * p] = dre;
          d[2 * p + 1] = dim;
          if (prim && !adjoint) sw[p] += (double)raw0.wv[p] * norm;
        }
      }
      if (any) geo = g0;
    }
    // ---- the unit's scale: raise E to this batch's largest |d| (a power of
    // two, so rescaling the accumulators is exact)
    float m = 0.f;
#pragma unroll
    for (int k = 0; k < C; ++k) m = fmaxf(m, fabsf(d[k]));
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    m = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, m)));
    if (m > 0.f) {
      int e;
      (void)frexpf(m, &e);  // m < 2^e
      e = __builtin_amdgcn_readfirstlane(e);
      if (e > E) {
        if (E != kH3NoScale) {
#pragma unroll
          for (int c = 0; c < C; ++c)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[c][r] = ldexpf(acc[c][r], E - e);
        }
        E = e;
      }
    }
#pragma unroll
    for (int k = 0; k < C; ++k) st_d[k * 64 + lane] = E == kH3NoScale ? 0.f : ldexpf(d[k], -E);
    {
      // byte offsets of tap (row) 0 of the record's u and v weights; a
      // skipped record has zero data, so any in-range offset (0) serves it
      uint32_t ou_b = 0u, ov_b = 0u;
      if (geo != kInvalidGeo) {
        int au, ou, avv, ovv;
        unpack_geo(geo, au, ou, avv, ovv);
        ou_b = 4u * (uint32_t)((sc - au + L.pad) * L.nw + ou + L.nw / 2);
        ov_b = 4u * (uint32_t)((sc - avv + L.pad) * L.nw + ovv + L.nw / 2);
      }
      st_gu[lane] = ou_b;
      st_gv[lane] = ov_b;
    }
    // ---- issue the next loads
    RawSample<float, NPOL> raw1;
    fetch_raw<float, NPOL>(raw1, r1.x, vis, weight, flag, PSF, adjoint);
    const uint2 r2 = load_record(P, base + 128 + lane, end);
    wave_lds_sync();
    // ---- 32 records per step: records s + 8 * lk + j of the batch
    const int nst = (int)min((uint32_t)64, end - base);
    const uint32_t li_b = 4u * (uint32_t)(li * L.nw);  // this lane's tile row / column
    for (int s = 0; s < nst; s += 32) {
      const int rb = s + 8 * lk;
      const uint4 ua = *(const uint4 *)(st_gu + rb), ub = *(const uint4 *)(st_gu + rb + 4);
      const uint4 va = *(const uint4 *)(st_gv + rb), vb = *(const uint4 *)(st_gv + rb + 4);
      const uint32_t uo[8] = {ua.x, ua.y, ua.z, ua.w, ub.x, ub.y, ub.z, ub.w};
      const uint32_t vo[8] = {va.x, va.y, va.z, va.w, vb.x, vb.y, vb.z, vb.w};
      float a[8], wv[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        a[j] = *(const float *)(tabc + uo[j] + li_b);
        wv[j] = *(const float *)(tabc + vo[j] + li_b);
      }
      H3U4 ahu, alu;
      h3_split8(a, ahu, alu);
      const H3Half8 ah = __builtin_bit_cast(H3Half8, ahu), al = __builtin_bit_cast(H3Half8, alu);
#pragma unroll
      for (int c = 0; c < C; ++c) {
        if constexpr (PSF) {
          if (c & 1) continue;
        }
        const H3Float4 da = *(const H3Float4 *)(st_d + c * 64 + rb);
        const H3Float4 db = *(const H3Float4 *)(st_d + c * 64 + rb + 4);
        const float dv[8] = {da.x, da.y, da.z, da.w, db.x, db.y, db.z, db.w};
        H3U4 bhu, blu;
        h3_split_products8(wv, dv, bhu, blu);
        const H3Half8 bh = __builtin_bit_cast(H3Half8, bhu), bl = __builtin_bit_cast(H3Half8, blu);
        acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, acc[c], 0, 0, 0);
        acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, acc[c], 0, 0, 0);
        acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, acc[c], 0, 0, 0);
      }
    }
    wave_lds_sync();
    raw0 = raw1;
    r0 = r1;
    r1 = r2;
  }

  // ---- flush: undo the scales; lane holds rows 4 * lk + r, column li
  const int fexp = E - 2 * kH3TabExp;  // both operands carry the table scale
  const bool multi = (fl & kUnitMulti) != 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int gu = u0 + 4 * lk + r, gv = v0 + li;
    if (gu >= P.n_u || gv >= P.n_v) continue;
#pragma unroll
    for (int p = 0; p < NPOL; ++p) {
      float re = 0.f, im = 0.f;
      if (E != kH3NoScale) {
        re = ldexpf(acc[2 * p][r], fexp);
        im = ldexpf(acc[2 * p + 1][r], fexp);
      }
      float *dst = gplane + p * pstride + ((size_t)gu * P.n_v + gv) * 2;
      if (multi) {
        if (re != 0.f) atomicAdd(dst, re);
        if (im != 0.f) atomicAdd(dst + 1, im);
      } else if (accumulate) {
        Cplx<float> old = *(Cplx<float> *)dst;
        old.re += re;
        old.im += im;
        *(Cplx<float> *)dst = old;
      } else {
        *(Cplx<float> *)dst = Cplx<float>{re, im};
      }
    }
  }
#pragma unroll
  for (int p = 0; p < NPOL; ++p) {
    double sv = adjoint ? 0.0 : sw[p];
    for (int off = 32; off >= 1; off >>= 1) sv += __shfl_xor(sv, off);
    if (lane == 0) P.upart[(size_t)u * 4 + p] = sv;
  }
}

template <int NPOL, bool PSF>
static int launch_grid_h3(av_plan *plan, const void *vis, const void *weight,
                          const uint8_t *flag, void *grid, double *sumwt, int mode,
                          hipStream_t s) {
  PlanDev &d = plan->d;
  const H3Lds L = h3_lds_layout<NPOL>(d.ov, d.support);
  {
    const int rc_lds =
        ensure_lds_limit<k_grid_h3<NPOL, PSF>>(L.total, "av_grid: hipFuncSetAttribute");
    if (rc_lds) return rc_lds;
  }
  const unsigned nwg = (unsigned)xcd_grouped_grid(
      (d.max_gunits + kGridMaxWaves - 1) / kGridMaxWaves, kGridXcdGroup);
  hipLaunchKernelGGL((k_grid_h3<NPOL, PSF>), dim3(nwg), dim3(kGridThreads), L.total, s, d,
                     (const float *)vis, (const float *)weight, flag, (float *)grid, mode);
  AV_LAUNCH_CHECK("k_grid_h3");
  if (!(mode & AV_GRID_ADJOINT)) {
    hipLaunchKernelGGL(k_sumwt_reduce, dim3((unsigned)d.n_img_chan), dim3(256), 0, s, d, NPOL,
                       sumwt);
    AV_LAUNCH_CHECK("k_sumwt_reduce");
  }
  return AV_OK;
}

// The split-FP16 kernel serves FP32 16 x 16 tiles without compensated sums
// whenever its padded weight table fits in a workgroup's LDS.
template <int NPOL>
static int try_grid_h3(av_plan *plan, const void *vis, const void *weight, const uint8_t *flag,
                       const double *cgk, void *grid, double *sumwt, int mode, hipStream_t s,
                       bool &done) {
  const PlanDev &d = plan->d;
  if (d.tile_u != 16 || d.tile_v != 16 || (mode & AV_GRID_COMPENSATED)) return AV_OK;
  const DeviceProps &prop = device_props();
  if (prop.error) return hip_fail(prop.error, "av_grid: device properties");
  if ((size_t)h3_lds_layout<NPOL>(d.ov, d.support).total > prop.lds_per_block) return AV_OK;
  int rc = prepare_tables(d, cgk, s);
  if (rc) return rc;
  done = true;
  if (mode & AV_GRID_PSF)
    return launch_grid_h3<NPOL, true>(plan, vis, weight, flag, grid, sumwt, mode, s);
  return launch_grid_h3<NPOL, false>(plan, vis, weight, flag, grid, sumwt, mode, s);
}

// ---------------------------------------------------------------------------
// Sub-tile scheduled matrix-core gridder (FP32, 16 x 16 tiles).  A record's
// footprint (at most S x S cells) covers only a few of the tile's sixteen
// 4 x 4 sub-tiles, yet a 16 x 16 x 4 product spends its full 16 x 16 block
// on every record.  v_mfma_f32_4x4x1_16b_f32 instead performs sixteen
// independent rank-1 4 x 4 updates, one per block of four lanes, and each
// block may take a different record: block b accumulates sub-tile
// (b / 4, b % 4), and in every round it applies the next record of its own
// list of the batch's records that touch that sub-tile.  A batch of 64
// records then takes max_b |list_b| rounds (about a quarter of the records
// for a dense tile, never more than the records) of 2 * npol instructions of
// 8 cycles, where the 16 x 16 x 4 form takes 2 * npol instructions of 32
// cycles per four records.
//
// Lane l of a 4 x 4 x 1 product supplies A[block l/4][row l%4] and
// B[block l/4][column l%4], and holds rows 0..3 (registers) of column l%4 of
// block l/4 of the result.  So lane l evaluates, for its block's record of
// the round, the u-weight at tile row 4*(b/4) + l%4 and the v-weight at tile
// column 4*(b%4) + l%4 times each data component.  A list entry beyond a
// block's records is 64, a staging slot that holds zero data.
//
// Weight table: a block is only ever given records whose footprint reaches
// it, so every tap it reads lies within 3 rows of the footprint (taps -3 ..
// S + 3 cover odd and even supports).  The LDS table is the call's table
// (taps -1 .. S) padded to taps -4 .. S + 3 with zero rows, and to at least
// 16 rows, so that the null record -- table base row 0 -- reads in-bounds
// finite weights for every row of every block (its data are zero).  A
// record's stored table base is offset by 16 rows so that it is never
// negative; the lane's row offset carries the -16.
constexpr int kSubNull = 64;  // staging slot of the null record
constexpr int kSubPad = 4;    // zero rows below tap 0 (and above tap S - 1, at least)
constexpr int kSubShift = 16;

struct SubLds {
  int tab, nrm, wave, st_d, st_geo, list, wave_bytes, total, pad, nw, rows, nnrm;
};

template <int NPOL>
__host__ __device__ inline SubLds sub_lds_layout(int ov, int support) {
  SubLds L;
  L.nnrm = 2 * (ov / 2 + 2) + 1;
  L.pad = kSubPad;
  L.nw = WTab(ov, support).nw;
  L.rows = support + 2 * kSubPad > 16 ? support + 2 * kSubPad : 16;
  int o = 0;
  L.tab = o;
  o += (L.rows * L.nw * 4 + 15) & ~15;
  L.nrm = o;
  o += (L.nnrm * 8 + 15) & ~15;
  L.wave = o;
  // per wave: 65 staged records' data and table bases, then the block lists
  // (16 x 64 slot numbers, plus 16 bytes that a read one round ahead may touch)
  L.st_d = 0;
  L.st_geo = (kSubNull + 1) * 2 * NPOL * 4;
  L.list = (L.st_geo + (kSubNull + 1) * 4 + 15) & ~15;
  L.wave_bytes = L.list + 16 * 64 + 16;
  L.total = o + kGridMaxWaves * L.wave_bytes;
  return L;
}

// Sub-tiles (bit 4 * bu + bv) touched by a footprint centred at tile-local
// (au, av) with half-width sc, clipped to the 16 x 16 tile.
__device__ inline uint32_t subtile_mask(int au, int av, int sc) {
  const int rlo = max(au - sc, 0) >> 2, rhi = min(au + sc, 15) >> 2;
  const int clo = max(av - sc, 0) >> 2, chi = min(av + sc, 15) >> 2;
  if (rlo > rhi || clo > chi) return 0u;
  const uint32_t rows = ((2u << rhi) - 1u) & ~((1u << rlo) - 1u);
  const uint32_t cols = ((2u << chi) - 1u) & ~((1u << clo) - 1u);
  // rows bit bu -> bit 4 * bu, times the column bits (no carries)
  const uint32_t spread = (rows & 1u) | ((rows & 2u) << 3) | ((rows & 4u) << 6) | ((rows & 8u) << 9);
  return spread * cols;
}

// MODE: 0 visibilities x weights, 1 weights only (PSF), 2 visibilities / norm
// (the adjoint of the degridder).  A template parameter, so that no mode's
// arithmetic is evaluated speculatively in another's.
template <int NPOL, int MODE>
__global__ __launch_bounds__(kGridThreads, 5)
void k_grid_sub(PlanDev P, const float *__restrict__ vis, const float *__restrict__ weight,
                const uint8_t *__restrict__ flag, float *__restrict__ grid, int mode) {
  typedef float V4 __attribute__((ext_vector_type(4)));
  constexpr int C = 2 * NPOL;
  constexpr bool psf = MODE == 1, adjoint = MODE == 2;
  using CL = CellLayout<float, NPOL>;
  extern __shared__ __align__(16) char smem[];
  const int S = P.support, sc = S / 2, ov = P.ov;
  const SubLds L = sub_lds_layout<NPOL>(ov, S);
  float *tab = (float *)(smem + L.tab);
  double *nrm = (double *)(smem + L.nrm);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  char *wbase = smem + L.wave + wave * L.wave_bytes;
  float *st_d = (float *)(wbase + L.st_d);          // [65][C]
  uint32_t *st_geo = (uint32_t *)(wbase + L.st_geo);  // [65] table bases, 16 + 16 bits
  uint8_t *list = (uint8_t *)(wbase + L.list);       // [16][64] slot numbers
  const bool accumulate = (mode & AV_GRID_ACCUMULATE) != 0;
  const int half = ov / 2 + 2;

  const uint32_t n_units = P.gunit_start[P.n_bins];
  const uint32_t blk = xcd_grouped_block(blockIdx.x, kGridXcdGroup);
  if (blk * kGridMaxWaves >= n_units) return;  // whole workgroup idle
  {
    MfmaLds ML;
    ML.pad = L.pad;
    ML.nw = L.nw;
    ML.rows = L.rows;
    copy_wtab_padded(tab, ML, P);
  }
  for (int i = tid; i < L.nnrm; i += blockDim.x) nrm[i] = P.nrm[i];
  // the null record: zero data; its table bases point at row 0
  if (lane < C) st_d[kSubNull * C + lane] = 0.0f;
  if (lane == 0) st_geo[kSubNull] = (uint32_t)(kSubShift * L.nw) * 0x10001u;
  __syncthreads();  // the only workgroup barrier: tables ready
  const uint32_t u = (uint32_t)__builtin_amdgcn_readfirstlane((int)(blk * kGridMaxWaves + wave));
  if (u >= n_units) return;

  const uint4 unit = P.gunits[u];
  const int64_t b = unit.x;
  const uint32_t first = unit.y, cnt = unit.z, fl = unit.w;
  const uint32_t tile = (uint32_t)b / (uint32_t)P.n_img_chan;
  const int ic = (int)((uint32_t)b - tile * (uint32_t)P.n_img_chan);
  const uint32_t tu_i = tile / (uint32_t)P.ntv;
  const int u0 = (int)tu_i * 16, v0 = (int)(tile - tu_i * (uint32_t)P.ntv) * 16;
  const uint32_t prim_end = P.bin_start[b] + P.cnt_prim[b];
  float *gplane = grid + (size_t)ic * NPOL * P.n_u * P.n_v * 2;
  const size_t pstride = (size_t)P.n_u * P.n_v * 2;

  if (cnt == 0) {
    if (lane < NPOL) P.upart[(size_t)u * 4 + lane] = 0.0;
    if (!accumulate) {
      for (int idx = lane; idx < NPOL * 256; idx += 64) {
        const int p = idx >> 8, c = idx & 255;
        const int gu = u0 + (c >> 4), gv = v0 + (c & 15);
        if (gu < P.n_u && gv < P.n_v)
          *(Cplx<float> *)(gplane + p * pstride + ((size_t)gu * P.n_v + gv) * 2) =
              Cplx<float>{0.0f, 0.0f};
      }
    }
    return;
  }

  // this lane's block and its row / column inside the block
  const int blk_id = lane >> 2, li = lane & 3;
  const int bu = blk_id >> 2, bv = blk_id & 3;
  // table index of tile row r for a record with stored base tb: tb + (r - 16) nw
  const int rowoff = (4 * bu + li - kSubShift) * L.nw;
  const int coloff = (4 * bv + li - kSubShift) * L.nw;
  const int lbase = blk_id * 64;

  V4 acc[C];
#pragma unroll
  for (int c = 0; c < C; ++c) acc[c] = V4{0.0f, 0.0f, 0.0f, 0.0f};
  double sw[NPOL];
#pragma unroll
  for (int p = 0; p < NPOL; ++p) sw[p] = 0.0;

  const uint32_t end = first + cnt;
  uint2 r0 = load_record(P, first + lane, end), r1 = load_record(P, first + 64 + lane, end);
  RawSample<float, NPOL> raw0;
  fetch_raw<float, NPOL>(raw0, r0.x, vis, weight, flag, psf, adjoint);
  for (uint32_t base = first; base < end; base += 64) {
    // ---- stage this lane's record: data (zeros when skipped) and geometry
    uint32_t geo = kInvalidGeo;
    float d[C];
#pragma unroll
    for (int k = 0; k < C; ++k) d[k] = 0.0f;
    const uint32_t g0 = base + lane < end ? r0.y : kInvalidGeo;
    if (g0 != kInvalidGeo) {
      int au, ou, avv, ovv;
      unpack_geo(g0, au, ou, avv, ovv);
      const double norm = nrm[ou + half] * nrm[ovv + half];
      const bool prim = base + lane < prim_end;
      float inv = 0.0f;
      if constexpr (adjoint) inv = (float)(1.0 / norm);
      bool any = false;
#pragma unroll
      for (int p = 0; p < NPOL; ++p) {
        float dre, dim;
        if constexpr (adjoint) {
          dre = raw0.vv[p].re * inv;
          dim = raw0.vv[p].im * inv;
        } else {
          float w = raw0.wv[p];
          if ((raw0.fbits >> (8 * p)) & 0xFFu) w = 0.0f;
          if constexpr (psf) {
            dre = w;
            dim = 0.0f;
          } else {
            dre = raw0.vv[p].re * w;
            dim = raw0.vv[p].im * w;
          }
        }
        // standard_grid.py:378: skip NaN or exactly-zero data
        const bool ok = !(isnan(dre) || isnan(dim)) && !(dre == 0.0f && dim == 0.0f);
        if (ok) {
          any = true;
          d[2 * p] = dre;
          d[2 * p + 1] = dim;
          if constexpr (!adjoint)
            if (prim) sw[p] += (double)raw0.wv[p] * norm;
        }
      }
      if (any) geo = g0;
    }
#pragma unroll
    for (int k = 0; k < C; k += CL::kCC) {
      Vec<float, CL::kCC> t;
#pragma unroll
      for (int c = 0; c < CL::kCC; ++c) t.v[c] = d[k + c];
      *(Vec<float, CL::kCC> *)(st_d + lane * C + k) = t;
    }
    uint32_t touched = 0u;
    {
      uint32_t tb = 0u;
      if (geo != kInvalidGeo) {
        int au, ou, avv, ovv;
        unpack_geo(geo, au, ou, avv, ovv);
        // tile row r is tap r - au + sc, table row r - au + sc + pad
        tb = (uint32_t)((sc - au + kSubPad + kSubShift) * L.nw + ou + L.nw / 2) |
             ((uint32_t)((sc - avv + kSubPad + kSubShift) * L.nw + ovv + L.nw / 2) << 16);
        touched = subtile_mask(au, avv, sc);
      }
      st_geo[lane] = tb;
    }
    // ---- the blocks' lists: every entry starts as the null record, then
    // each record is appended to the list of every sub-tile it touches
    // (the 16 bytes past the lists are read one round ahead by the last block)
    const uint4 nulls = make_uint4(0x40404040u, 0x40404040u, 0x40404040u, 0x40404040u);
    *(uint4 *)(list + lane * 16) = nulls;
    if (lane == 0) *(uint4 *)(list + 1024) = nulls;
    wave_lds_sync();
    uint32_t rounds = 0;
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const bool t = (touched >> s) & 1u;
      const uint64_t bal = __ballot(t);
      const uint32_t pos = __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
      if (t) list[s * 64 + pos] = (uint8_t)lane;
      rounds = max(rounds, (uint32_t)__popcll(bal));
    }
    // ---- issue the next loads
    RawSample<float, NPOL> raw1;
    fetch_raw<float, NPOL>(raw1, r1.x, vis, weight, flag, psf, adjoint);
    const uint2 r2 = load_record(P, base + 128 + lane, end);
    wave_lds_sync();
    // ---- rounds: every block applies the next record of its list.  The
    // next round's slot, table bases and data are read while this round's
    // products run (a list entry past a block's records is the null slot).
    const int nrounds = __builtin_amdgcn_readfirstlane((int)rounds);
    auto load_slot = [&](int k, uint32_t &tb, float *dd) {
      tb = st_geo[k];
#pragma unroll
      for (int q = 0; q < C; q += CL::kCC) {
        const Vec<float, CL::kCC> t = *(const Vec<float, CL::kCC> *)(st_d + k * C + q);
#pragma unroll
        for (int c = 0; c < CL::kCC; ++c) dd[q + c] = t.v[c];
      }
    };
    uint32_t tb_n;
    float dd_n[C];
    load_slot(list[lbase], tb_n, dd_n);
    for (int i = 0; i < nrounds; ++i) {
      const uint32_t tb = tb_n;
      float dd[C];
#pragma unroll
      for (int c = 0; c < C; ++c) dd[c] = dd_n[c];
      const float a = tab[(tb & 0xFFFFu) + rowoff];
      const float wv = tab[(tb >> 16) + coloff];
      // LDS reads return in order: keep this round's table reads ahead of
      // the next round's, so the products wait for them alone
      __builtin_amdgcn_sched_barrier(0);
      load_slot(list[lbase + i + 1], tb_n, dd_n);
#pragma unroll
      for (int c = 0; c < C; ++c) {
        if constexpr (psf) {
          if (c & 1) continue;
        }
        acc[c] = __builtin_amdgcn_mfma_f32_4x4x1f32(a, wv * dd[c], acc[c], 0, 0, 0);
      }
    }
    wave_lds_sync();
    raw0 = raw1;
    r0 = r1;
    r1 = r2;
  }

  // ---- flush: lane holds rows 0..3 of column li of sub-tile (bu, bv)
  const bool multi = (fl & kUnitMulti) != 0;
  const int gv = v0 + 4 * bv + li;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int gu = u0 + 4 * bu + r;
    if (gu >= P.n_u || gv >= P.n_v) continue;
#pragma unroll
    for (int p = 0; p < NPOL; ++p) {
      const float re = acc[2 * p][r], im = acc[2 * p + 1][r];
      float *dst = gplane + p * pstride + ((size_t)gu * P.n_v + gv) * 2;
      if (multi) {
        if (re != 0.0f) atomicAdd(dst, re);
        if (im != 0.0f) atomicAdd(dst + 1, im);
      } else if (accumulate) {
        Cplx<float> old = *(Cplx<float> *)dst;
        old.re += re;
        old.im += im;
        *(Cplx<float> *)dst = old;
      } else {
        *(Cplx<float> *)dst = Cplx<float>{re, im};
      }
    }
  }
  // ---- sums of weights (FP64): this unit's partial per polarisation
#pragma unroll
  for (int p = 0; p < NPOL; ++p) {
    double sv = adjoint ? 0.0 : sw[p];
    for (int off = 32; off >= 1; off >>= 1) sv += __shfl_xor(sv, off);
    if (lane == 0) P.upart[(size_t)u * 4 + p] = sv;
  }
}

template <int NPOL, int MODE>
static int launch_grid_sub(av_plan *plan, const void *vis, const void *weight,
                           const uint8_t *flag, void *grid, double *sumwt, int mode,
                           hipStream_t s) {
  PlanDev &d = plan->d;
  const SubLds L = sub_lds_layout<NPOL>(d.ov, d.support);
  {
    const int rc_lds =
        ensure_lds_limit<k_grid_sub<NPOL, MODE>>(L.total, "av_grid: hipFuncSetAttribute");
    if (rc_lds) return rc_lds;
  }
  const unsigned nwg = (unsigned)xcd_grouped_grid(
      (d.max_gunits + kGridMaxWaves - 1) / kGridMaxWaves, kGridXcdGroup);
  hipLaunchKernelGGL((k_grid_sub<NPOL, MODE>), dim3(nwg), dim3(kGridThreads), L.total, s, d,
                     (const float *)vis, (const float *)weight, flag, (float *)grid, mode);
  AV_LAUNCH_CHECK("k_grid_sub");
  if (!(mode & AV_GRID_ADJOINT)) {
    hipLaunchKernelGGL(k_sumwt_reduce, dim3((unsigned)d.n_img_chan), dim3(256), 0, s, d, NPOL,
                       sumwt);
    AV_LAUNCH_CHECK("k_sumwt_reduce");
  }
  return AV_OK;
}

// The sub-tile kernel serves FP32 16 x 16 tiles without compensated sums
// whenever its padded weight table fits in a workgroup's LDS.
template <int NPOL>
static int try_grid_sub(av_plan *plan, const void *vis, const void *weight, const uint8_t *flag,
                        const double *cgk, void *grid, double *sumwt, int mode, hipStream_t s,
                        bool &done) {
  const PlanDev &d = plan->d;
  if (d.tile_u != 16 || d.tile_v != 16 || (mode & AV_GRID_COMPENSATED)) return AV_OK;
  const DeviceProps &prop = device_props();
  if (prop.error) return hip_fail(prop.error, "av_grid: device properties");
  if ((size_t)sub_lds_layout<NPOL>(d.ov, d.support).total > prop.lds_per_block) return AV_OK;
  int rc = prepare_tables(d, cgk, s);
  if (rc) return rc;
  done = true;
  if (mode & AV_GRID_PSF)
    return launch_grid_sub<NPOL, 1>(plan, vis, weight, flag, grid, sumwt, mode, s);
  if (mode & AV_GRID_ADJOINT)
    return launch_grid_sub<NPOL, 2>(plan, vis, weight, flag, grid, sumwt, mode, s);
  return launch_grid_sub<NPOL, 0>(plan, vis, weight, flag, grid, sumwt, mode, s);
}

// Result blocks a matrix-core unit may hold in registers: 16 for FP32 (64
// registers), 8 for FP64 and for the compensated mode.
template <typename R, bool COMP>
constexpr int mfma_max_blocks() {
  return (sizeof(R) == 8 || COMP) ? 8 : 16;
}

template <typename R, int NPOL, int MB, int VB, bool COMP>
static int try_mfma(av_plan *plan, const void *vis, const void *weight, const uint8_t *flag,
                    const double *cgk, void *grid, double *sumwt, int mode, hipStream_t s,
                    bool &done) {
  if constexpr (MB * 2 * NPOL * VB <= mfma_max_blocks<R, COMP>()) {
    if (plan->d.tile_u == 16 * MB && plan->d.tile_v == 16 * VB) {
      // the padded weight table grows with the tile, the support and the
      // oversampling: when it does not fit in LDS the LDS tile kernel runs
      const DeviceProps &prop = device_props();
      if (prop.error) return hip_fail(prop.error, "av_grid: device properties");
      const MfmaLds L = mfma_lds_layout<R, NPOL>(plan->d.ov, plan->d.support,
                                                 16 * (MB > VB ? MB : VB));
      if ((size_t)L.total > prop.lds_per_block) return AV_OK;
      done = true;
      if (mode & AV_GRID_PSF)
        return launch_grid_mfma<R, NPOL, MB, VB, COMP, true>(plan, vis, weight, flag, cgk, grid,
                                                             sumwt, mode, s);
      return launch_grid_mfma<R, NPOL, MB, VB, COMP, false>(plan, vis, weight, flag, cgk, grid,
                                                            sumwt, mode, s);
    }
  }
  return AV_OK;
}

template <typename R, int NPOL, bool COMP>
static int grid_mfma_shapes(av_plan *plan, const void *vis, const void *weight,
                            const uint8_t *flag, const double *cgk, void *grid, double *sumwt,
                            int mode, hipStream_t s, bool &done) {
  int rc = AV_OK;
#define AV_TRY(MB, VB)                                                                       \
  if (!done) rc = try_mfma<R, NPOL, MB, VB, COMP>(plan, vis, weight, flag, cgk, grid, sumwt, \
                                                  mode, s, done);
  AV_TRY(1, 1) AV_TRY(1, 2) AV_TRY(2, 1) AV_TRY(2, 2) AV_TRY(1, 4) AV_TRY(2, 4)
#undef AV_TRY
  return rc;
}

template <typename R, int NPOL>
static int dispatch_grid(av_plan *plan, const void *vis, const void *weight, const uint8_t *flag,
                         const double *cgk, void *grid, double *sumwt, int mode, hipStream_t s) {
  const bool comp = (mode & AV_GRID_COMPENSATED) != 0;
  const PlanDev &d = plan->d;
  if (comp && sizeof(R) != 4)
    return fail(AV_ERR_UNSUPPORTED, "av_grid: compensated mode is for FP32 grids");
  // tiles of 16 x 16 multiples whose result blocks fit in registers run on
  // the matrix cores; other tile shapes use the LDS tile kernel
  bool done = false;
  int rc;
  if constexpr (sizeof(R) == 4) {
    if (!comp && (mode & kGridSubtile)) {
      rc = try_grid_sub<NPOL>(plan, vis, weight, flag, cgk, grid, sumwt, mode, s, done);
      if (done || rc) return rc;
    }
    if (!comp && !(mode & AV_GRID_F32_PRODUCTS)) {
      rc = try_grid_h3<NPOL>(plan, vis, weight, flag, cgk, grid, sumwt, mode, s, done);
      if (done || rc) return rc;
    }
    if (comp)
      rc = grid_mfma_shapes<R, NPOL, true>(plan, vis, weight, flag, cgk, grid, sumwt, mode, s,
                                           done);
    else
      rc = grid_mfma_shapes<R, NPOL, false>(plan, vis, weight, flag, cgk, grid, sumwt, mode, s,
                                            done);
  } else {
    rc = grid_mfma_shapes<R, NPOL, false>(plan, vis, weight, flag, cgk, grid, sumwt, mode, s, done);
  }
  if (done) return rc;
  if (comp && 2 * NPOL * d.tile_u * d.tile_v > 32 * 64)
    return fail(AV_ERR_UNSUPPORTED, "av_grid: compensated mode needs tile area <= 1024/npol");
  if constexpr (sizeof(R) == 4) {
    if (comp) {
      if (d.support <= 8)
        return launch_grid<R, NPOL, 8, true>(plan, vis, weight, flag, cgk, grid, sumwt, mode, s);
      return launch_grid<R, NPOL, 16, true>(plan, vis, weight, flag, cgk, grid, sumwt, mode, s);
    }
  }
  if (d.support <= 8)
    return launch_grid<R, NPOL, 8, false>(plan, vis, weight, flag, cgk, grid, sumwt, mode, s);
  return launch_grid<R, NPOL, 16, false>(plan, vis, weight, flag, cgk, grid, sumwt, mode, s);
}

// Overwrite mode, before a tile kernel: zero the sums of weights and the
// tiles of split bins (the units of a split bin add into them).
int zero_before_grid(const PlanDev &d, void *grid, int npol, bool fp64, double *sumwt,
                     hipStream_t s) {
  int64_t n = (int64_t)d.n_img_chan * npol;
  hipLaunchKernelGGL(k_zero_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, sumwt, n);
  AV_LAUNCH_CHECK("k_zero_f64");
  const unsigned zb = (unsigned)std::min<int64_t>((d.max_gunits + 255) / 256, 8192);
  if (fp64)
    hipLaunchKernelGGL(k_zero_multi<double>, dim3(zb), dim3(256), 0, s, d, (double *)grid, npol);
  else
    hipLaunchKernelGGL(k_zero_multi<float>, dim3(zb), dim3(256), 0, s, d, (float *)grid, npol);
  AV_LAUNCH_CHECK("k_zero_multi");
  return AV_OK;
}

// After a tile kernel: the per-unit partial sums of weights into sumwt.
int reduce_sumwt(const PlanDev &d, int npol, double *sumwt, hipStream_t s) {
  hipLaunchKernelGGL(k_sumwt_reduce, dim3((unsigned)d.n_img_chan), dim3(256), 0, s, d, npol,
                     sumwt);
  AV_LAUNCH_CHECK("k_sumwt_reduce");
  return AV_OK;
}

template <typename R>
static int grid_by_pol(av_plan *plan, const void *vis, const void *weight, const uint8_t *flag,
                       int npol, const double *cgk, void *grid, double *sumwt, int mode,
                       hipStream_t s) {
  PlanDev &d = plan->d;
  if (!(mode & AV_GRID_ACCUMULATE)) {
    const int rc = zero_before_grid(d, grid, npol, sizeof(R) == 8, sumwt, s);
    if (rc) return rc;
  }
  switch (npol) {
    case 1: return dispatch_grid<R, 1>(plan, vis, weight, flag, cgk, grid, sumwt, mode, s);
    case 2: return dispatch_grid<R, 2>(plan, vis, weight, flag, cgk, grid, sumwt, mode, s);
    case 4: return dispatch_grid<R, 4>(plan, vis, weight, flag, cgk, grid, sumwt, mode, s);
    default: return fail(AV_ERR_UNSUPPORTED, "av_grid: npol must be 1, 2 or 4");
  }
}

}  // namespace av

using namespace av;

extern "C" int av_grid(av_plan *plan, const void *vis, const void *weight, const uint8_t *flag,
                       int32_t n_pol, const double *cgk, void *grid, double *sumwt, int32_t dtype,
                       int32_t mode, void *stream) {
  if (!plan || !cgk || !grid || !sumwt) return fail(AV_ERR_INVALID, "av_grid: NULL argument");
  const bool psf = (mode & AV_GRID_PSF) != 0, adjoint = (mode & AV_GRID_ADJOINT) != 0;
  if (!psf && !vis) return fail(AV_ERR_INVALID, "av_grid: vis is NULL");
  if (!adjoint && !weight) return fail(AV_ERR_INVALID, "av_grid: weight is NULL");
  if (psf && adjoint) return fail(AV_ERR_INVALID, "av_grid: PSF and ADJOINT are exclusive");
  if (!plan->d.uvw) return fail(AV_ERR_INVALID, "av_grid: plan has no uvw (av_plan_set_uvw)");
  if (plan->d.support > kMaxSupport)
    return fail(AV_ERR_UNSUPPORTED, "av_grid: the plan's support exceeds 16");
  hipStream_t s = (hipStream_t)stream;
  if (dtype == AV_F32)
    return grid_by_pol<float>(plan, vis, weight, flag, n_pol, cgk, grid, sumwt, mode, s);
  if (dtype == AV_F64)
    return grid_by_pol<double>(plan, vis, weight, flag, n_pol, cgk, grid, sumwt, mode, s);
  return fail(AV_ERR_INVALID, "av_grid: dtype must be AV_F32 or AV_F64");
}
