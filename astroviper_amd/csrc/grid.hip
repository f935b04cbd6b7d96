// grid.hip -- tiled convolutional gridding (replaces standard_grid_jit,
// core/imaging/imaging_utils/standard_grid.py:246-432).
//
// One workgroup per work unit (a bin, or a slice of a large bin); the launch
// covers the plan's unit capacity and surplus workgroups exit at once.  A
// unit owns a tile_u x tile_v tile of every polarisation plane of one image
// channel, held in LDS cell-major.  Records are staged 256 at a time; each of
// the four wavefronts owns a band of tile rows and walks all staged samples,
// applying the taps that fall in its band with plain 16-byte read-modify-
// writes (lanes = (u-tap group, v-tap, chunk); one sample at a time, so the
// lanes of an update touch distinct cells and no atomics are needed).  Only
// taps inside the tile are applied, which is why each footprint is recorded
// in every tile it overlaps.  The finished tile is written to HBM once, row
// by row, coalesced.  A bin split over several units adds its partial tiles
// with global atomics into a tile that a pre-pass zeroed.  Sums of weights
// are FP64: each unit writes its partial sums and a reduction pass adds them
// per plane.
#include <algorithm>

#include "tile_common.h"

namespace av {

constexpr int kGridThreads = 256;
constexpr int kGridWaves = kGridThreads / 64;
constexpr int kCompChunkRounds = 4;  // compensated mode: fold every 4*256 records

template <typename R>
struct Cplx {
  R re, im;
};

struct GridLds {
  // offsets in bytes inside dynamic LDS
  int tab, nrm, acc, st_d, st_cu, st_cv, st_pos, red, total;
  int ntab, nnrm;
};

template <typename R, int NPOL, int DVW>
__host__ __device__ inline GridLds grid_lds_layout(int ov, int support, int tu, int tv) {
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
  L.acc = take(tu * tv * 2 * NPOL * (int)sizeof(R));
  L.st_d = take(kGridThreads * 2 * NPOL * (int)sizeof(R));
  L.st_cu = take(kGridThreads * DVW * (int)sizeof(R));
  L.st_cv = take(kGridThreads * DVW * (int)sizeof(R));
  L.st_pos = take(kGridThreads * (int)sizeof(int));
  L.red = take(kGridWaves * NPOL * (int)sizeof(double) + 16);
  L.total = o;
  return L;
}

template <typename R, int NPOL, int DVW, bool COMP>
__global__ __launch_bounds__(kGridThreads) void k_grid_tiles(PlanDev P, const R *__restrict__ vis,
                                                             const R *__restrict__ weight,
                                                             const uint8_t *__restrict__ flag,
                                                             const double *__restrict__ cgk,
                                                             R *__restrict__ grid, int mode) {
  using CL = CellLayout<R, NPOL>;
  constexpr int C = CL::kC, CC = CL::kCC, H = CL::kH;
  constexpr int DQ = 64 / (H * DVW) > 0 ? 64 / (H * DVW) : 1;  // u-taps per wave step
  using V = Vec<R, CC>;
  extern __shared__ __align__(16) char smem[];
  const uint32_t u = blockIdx.x;
  if (u >= P.gunit_start[P.n_bins]) return;
  const int S = P.support, sc = S / 2, ov = P.ov;
  const int Tu = P.tile_u, Tv = P.tile_v;
  const GridLds L = grid_lds_layout<R, NPOL, DVW>(ov, S, Tu, Tv);
  R *tab = (R *)(smem + L.tab);
  double *nrm = (double *)(smem + L.nrm);
  R *acc = (R *)(smem + L.acc);
  R *st_d = (R *)(smem + L.st_d);    // [record][C]
  R *st_cu = (R *)(smem + L.st_cu);  // [record][DVW]
  R *st_cv = (R *)(smem + L.st_cv);  // [record][DVW]
  int *st_pos = (int *)(smem + L.st_pos);
  double *red = (double *)(smem + L.red);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

  const bool psf = (mode & AV_GRID_PSF) != 0;
  const bool adjoint = (mode & AV_GRID_ADJOINT) != 0;
  const bool accumulate = (mode & AV_GRID_ACCUMULATE) != 0;
  const int half = ov / 2 + 2;

  for (int i = tid; i < L.ntab; i += kGridThreads)
    tab[i] = sizeof(R) == 4 ? (R)P.tab32[i] : (R)cgk[i];
  for (int i = tid; i < L.nnrm; i += kGridThreads) nrm[i] = P.nrm[i];
  const int cells = Tu * Tv;

  // update lanes: (u-tap group dq, v-tap dv, chunk h)
  const int h = lane % H;
  const int dv = (lane / H) % DVW;
  const int dq = lane / (H * DVW);
  const bool lane_on = dq < DQ && dv < S;
  // this wavefront's band of tile rows
  const int band_lo = (wave * Tu) / kGridWaves, band_hi = ((wave + 1) * Tu) / kGridWaves;
  // flush layout: a thread owns one column of the tile when tile_v divides
  // the workgroup, so the flush loops need no division
  const bool rowwise = (kGridThreads % Tv) == 0;
  const int fcol = tid % Tv, frow = tid / Tv, fstep = kGridThreads / Tv;

  const uint4 unit = P.gunits[u];
  const int64_t b = unit.x;
  const uint32_t first = unit.y, cnt = unit.z, fl = unit.w;
  const int ic = (int)(b % P.n_img_chan);
  const int64_t tile = b / P.n_img_chan;
  const int u0 = (int)(tile / P.ntv) * Tu, v0 = (int)(tile % P.ntv) * Tv;
  const uint32_t prim_end = P.bin_start[b] + P.cnt_prim[b];
  R *gplane = grid + (size_t)ic * NPOL * P.n_u * P.n_v * 2;
  const size_t pstride = (size_t)P.n_u * P.n_v * 2;

  if (cnt == 0) {
    if (tid < NPOL) P.upart[(size_t)u * 4 + tid] = 0.0;
    if (!accumulate) {
      for (int idx = tid; idx < NPOL * cells; idx += kGridThreads) {
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
    for (int i = tid; i < n16; i += kGridThreads) a4[i] = make_uint4(0u, 0u, 0u, 0u);
    for (int i = n16 * 16 / (int)sizeof(R) + tid; i < cells * C; i += kGridThreads) acc[i] = (R)0;
  }
  // compensated mode: per-thread running sums of the components it flushes
  constexpr int kCpt = COMP ? 32 : 1;
  R csum[kCpt], ccmp[kCpt];
  if (COMP) {
#pragma unroll
    for (int k = 0; k < kCpt; ++k) csum[k] = ccmp[k] = (R)0;
  }
  double sw[NPOL];
#pragma unroll
  for (int p = 0; p < NPOL; ++p) sw[p] = 0.0;

  const uint32_t end = first + cnt;
  int round = 0;
  for (uint32_t r0 = first; r0 < end; r0 += kGridThreads, ++round) {
    __syncthreads();  // previous batch fully applied (and tile zeroed)
    // ---- stage: one record per thread
    const uint32_t r = r0 + tid;
    int pos = -1;
    if (r < end) {
      const uint32_t g = P.rec_g[r];
      const uint32_t item = P.rec_idx[r];
      if (g != kInvalidGeo) {
        int au, ou, avv, ovv;
        unpack_geo(g, au, ou, avv, ovv);
        const double norm = nrm[ou + half] * nrm[ovv + half];
        const bool prim = r < prim_end;
        R wv[NPOL];
        Cplx<R> vv[NPOL];
#pragma unroll
        for (int p = 0; p < NPOL; ++p) {
          if (!adjoint) wv[p] = weight[(size_t)item * NPOL + p];
          if (!psf) vv[p] = ((const Cplx<R> *)vis)[(size_t)item * NPOL + p];
        }
        uint32_t fbits = 0;
        if (flag && !adjoint) {
          if (NPOL == 4)
            fbits = *(const uint32_t *)(flag + (size_t)item * 4);
          else
            for (int p = 0; p < NPOL; ++p) fbits |= (uint32_t)flag[(size_t)item * NPOL + p] << (8 * p);
        }
        bool any = false;
        R d[C];
#pragma unroll
        for (int p = 0; p < NPOL; ++p) {
          R dre, dim;
          if (adjoint) {
            R inv = (R)(1.0 / norm);
            dre = vv[p].re * inv;
            dim = vv[p].im * inv;
          } else {
            R w = wv[p];
            if ((fbits >> (8 * p)) & 0xFFu) w = (R)0;
            if (psf) {
              dre = w;
              dim = (R)0;
            } else {
              dre = vv[p].re * w;
              dim = vv[p].im * w;
            }
          }
          // standard_grid.py:378: skip NaN or exactly-zero data
          bool ok = !(isnan(dre) || isnan(dim)) && !(dre == (R)0 && dim == (R)0);
          if (!ok) {
            dre = (R)0;
            dim = (R)0;
          } else {
            any = true;
            if (prim && !adjoint) sw[p] += (double)wv[p] * norm;
          }
          d[2 * p] = dre;
          d[2 * p + 1] = dim;
        }
        if (any) {
          pos = (au + kPosBias) | ((avv + kPosBias) << 16);
#pragma unroll
          for (int k = 0; k < C; k += CC) {
            V t;
#pragma unroll
            for (int c = 0; c < CC; ++c) t.v[c] = d[k + c];
            *(V *)(st_d + tid * C + k) = t;
          }
#pragma unroll
          for (int t = 0; t < DVW; ++t) {
            st_cu[tid * DVW + t] = t < S ? tab[tab_index(ov, t, sc, ou)] : (R)0;
            st_cv[tid * DVW + t] = t < S ? tab[tab_index(ov, t, sc, ovv)] : (R)0;
          }
        }
      }
    }
    st_pos[tid] = pos;
    __syncthreads();
    // ---- apply: every wavefront walks the staged samples, own band only
    const int nst = (int)min((uint32_t)kGridThreads, end - r0);
    for (int j = 0; j < nst; ++j) {
      const int pj = st_pos[j];
      if (pj < 0) continue;
      const int au = (pj & 0xFFFF) - kPosBias;
      const int avv = (pj >> 16) - kPosBias;
      const int rlo = max(band_lo, au - sc), rhi = min(band_hi, au - sc + S);
      if (rlo >= rhi) continue;  // nothing in this band (wave-uniform)
      const int col = avv - sc + dv;
      if (!lane_on || col < 0 || col >= Tv) continue;
      const R cv = st_cv[j * DVW + dv];
      const V dd = *(const V *)(st_d + j * C + h * CC);
      R x[CC];
#pragma unroll
      for (int c = 0; c < CC; ++c) x[c] = dd.v[c] * cv;
      for (int row = rlo + dq; row < rhi; row += DQ) {
        const R cu = st_cu[j * DVW + (row - (au - sc))];
        V *cell = (V *)(acc + (row * Tv + col) * C + h * CC);
        V a = *cell;
#pragma unroll
        for (int c = 0; c < CC; ++c) a.v[c] += x[c] * cu;
        *cell = a;
      }
    }
    if (COMP && ((round + 1) % kCompChunkRounds == 0) && r0 + kGridThreads < end) {
      // fold the LDS partial sums into the per-thread compensated sums
      __syncthreads();
#pragma unroll
      for (int k = 0; k < kCpt; ++k) {
        int idx = tid + k * kGridThreads;
        if (idx < C * cells) {
          R x = acc[idx];
          acc[idx] = (R)0;
          R t = csum[k] + x;  // Neumaier
          if (fabs(csum[k]) >= fabs(x))
            ccmp[k] += (csum[k] - t) + x;
          else
            ccmp[k] += (x - t) + csum[k];
          csum[k] = t;
        }
      }
    }
  }
  __syncthreads();

  // ---- flush the tile
  const bool multi = (fl & kUnitMulti) != 0;
  if (COMP) {
#pragma unroll
    for (int k = 0; k < kCpt; ++k) {
      int idx = tid + k * kGridThreads;
      if (idx < C * cells) {
        R x = acc[idx];
        R t = csum[k] + x;
        if (fabs(csum[k]) >= fabs(x))
          ccmp[k] += (csum[k] - t) + x;
        else
          ccmp[k] += (x - t) + csum[k];
        acc[idx] = t + ccmp[k];
      }
    }
    __syncthreads();
  }
  auto flush_cell = [&](int p, int c, int gu, int gv) {
    Cplx<R> val = *(const Cplx<R> *)(acc + c * C + 2 * p);
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
    for (int idx = tid; idx < NPOL * cells; idx += kGridThreads) {
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
    if (lane == 0) red[wave * NPOL + p] = sv;
  }
  __syncthreads();
  if (tid < NPOL) {
    double sv = 0.0;
    for (int w = 0; w < kGridWaves; ++w) sv += red[w * NPOL + tid];
    P.upart[(size_t)u * 4 + tid] = sv;
  }
}

// sums of weights per (image channel, polarisation) from the unit partials,
// gathered in LDS (FP64 LDS atomics) and added to HBM once per workgroup
__global__ __launch_bounds__(256) void k_sumwt_reduce(PlanDev P, int npol, double *sumwt,
                                                     int use_lds) {
  extern __shared__ double sacc[];
  const int nplanes = P.n_img_chan * npol;
  const uint32_t n_units = P.gunit_start[P.n_bins];
  if (use_lds) {
    for (int i = threadIdx.x; i < nplanes; i += blockDim.x) sacc[i] = 0.0;
    __syncthreads();
  }
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n_units;
       u += gridDim.x * blockDim.x) {
    const int ic = (int)(P.gunits[u].x % P.n_img_chan);
    for (int p = 0; p < npol; ++p) {
      const double v = P.upart[(size_t)u * 4 + p];
      if (v != 0.0) {
        if (use_lds)
          atomicAdd(&sacc[ic * npol + p], v);
        else
          atomicAdd(&sumwt[ic * npol + p], v);
      }
    }
  }
  if (use_lds) {
    __syncthreads();
    for (int i = threadIdx.x; i < nplanes; i += blockDim.x)
      if (sacc[i] != 0.0) atomicAdd(&sumwt[i], sacc[i]);
  }
}

// zero the tiles of bins that several units add into (overwrite mode)
template <typename R>
__global__ void k_zero_multi(PlanDev P, R *grid, int npol) {
  const uint32_t n_units = P.gunit_start[P.n_bins];
  const int Tu = P.tile_u, Tv = P.tile_v, cells = Tu * Tv;
  for (uint32_t u = blockIdx.x; u < n_units; u += gridDim.x) {
    const uint4 unit = P.gunits[u];
    if (!((unit.w & kUnitMulti) && (unit.w & kUnitFirst))) continue;
    const int64_t b = unit.x;
    const int ic = (int)(b % P.n_img_chan);
    const int64_t tile = b / P.n_img_chan;
    const int u0 = (int)(tile / P.ntv) * Tu, v0 = (int)(tile % P.ntv) * Tv;
    for (int idx = threadIdx.x; idx < npol * cells; idx += blockDim.x) {
      int p = idx / cells, c = idx - p * cells;
      int gu = u0 + c / Tv, gv = v0 + c % Tv;
      if (gu < P.n_u && gv < P.n_v) {
        R *dst = grid + (((size_t)ic * npol + p) * P.n_u + gu) * P.n_v * 2 + (size_t)gv * 2;
        dst[0] = (R)0;
        dst[1] = (R)0;
      }
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
  GridLds L = grid_lds_layout<R, NPOL, DVW>(d.ov, d.support, d.tile_u, d.tile_v);
  auto kfn = k_grid_tiles<R, NPOL, DVW, COMP>;
  const DeviceProps &prop = device_props();
  if (prop.error) return hip_fail(prop.error, "av_grid: device properties");
  if (d.tile_u < kGridWaves) return fail(AV_ERR_UNSUPPORTED, "av_grid: tile_u must be at least 4");
  if ((size_t)L.total > prop.lds_per_block)
    return fail(AV_ERR_UNSUPPORTED, "av_grid: tile too large for LDS (" +
                                        std::to_string(L.total) + " bytes)");
  AV_HIP(hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize,
                             L.total));
  int rc = prepare_tables(d, cgk, s);
  if (rc) return rc;
  // one workgroup per unit; the plan's capacity bounds the unit count
  hipLaunchKernelGGL(kfn, dim3((unsigned)d.max_gunits), dim3(kGridThreads), L.total, s, d,
                     (const R *)vis, (const R *)weight, flag, cgk, (R *)grid, mode);
  AV_LAUNCH_CHECK("k_grid_tiles");
  if (!(mode & AV_GRID_ADJOINT)) {
    const int nplanes = d.n_img_chan * NPOL;
    const int use_lds = nplanes * 8 <= 65536;
    hipLaunchKernelGGL(k_sumwt_reduce, dim3(prop.num_cus), dim3(256),
                       use_lds ? nplanes * 8 : 0, s, d, NPOL, sumwt, use_lds);
    AV_LAUNCH_CHECK("k_sumwt_reduce");
  }
  return AV_OK;
}

template <typename R, int NPOL>
static int dispatch_grid(av_plan *plan, const void *vis, const void *weight, const uint8_t *flag,
                         const double *cgk, void *grid, double *sumwt, int mode, hipStream_t s) {
  const bool comp = (mode & AV_GRID_COMPENSATED) != 0;
  const PlanDev &d = plan->d;
  if (comp) {
    if (sizeof(R) != 4)
      return fail(AV_ERR_UNSUPPORTED, "av_grid: compensated mode is for FP32 grids");
    if (2 * NPOL * d.tile_u * d.tile_v > 32 * kGridThreads)
      return fail(AV_ERR_UNSUPPORTED, "av_grid: compensated mode needs tile area <= 4096/npol");
    if (d.tile_u < kGridWaves)
      return fail(AV_ERR_UNSUPPORTED, "av_grid: tile_u must be at least 4");
  }
  if (d.support <= 8) {
    if (comp) return launch_grid<R, NPOL, 8, true>(plan, vis, weight, flag, cgk, grid, sumwt, mode, s);
    return launch_grid<R, NPOL, 8, false>(plan, vis, weight, flag, cgk, grid, sumwt, mode, s);
  }
  if (comp) return launch_grid<R, NPOL, 16, true>(plan, vis, weight, flag, cgk, grid, sumwt, mode, s);
  return launch_grid<R, NPOL, 16, false>(plan, vis, weight, flag, cgk, grid, sumwt, mode, s);
}

template <typename R>
static int grid_by_pol(av_plan *plan, const void *vis, const void *weight, const uint8_t *flag,
                       int npol, const double *cgk, void *grid, double *sumwt, int mode,
                       hipStream_t s) {
  PlanDev &d = plan->d;
  if (!(mode & AV_GRID_ACCUMULATE)) {
    int64_t n = (int64_t)d.n_img_chan * npol;
    hipLaunchKernelGGL(k_zero_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, sumwt, n);
    AV_LAUNCH_CHECK("k_zero_f64");
    hipLaunchKernelGGL(k_zero_multi<R>, dim3(1024), dim3(256), 0, s, d, (R *)grid, npol);
    AV_LAUNCH_CHECK("k_zero_multi");
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
  hipStream_t s = (hipStream_t)stream;
  if (dtype == AV_F32)
    return grid_by_pol<float>(plan, vis, weight, flag, n_pol, cgk, grid, sumwt, mode, s);
  if (dtype == AV_F64)
    return grid_by_pol<double>(plan, vis, weight, flag, n_pol, cgk, grid, sumwt, mode, s);
  return fail(AV_ERR_INVALID, "av_grid: dtype must be AV_F32 or AV_F64");
}
