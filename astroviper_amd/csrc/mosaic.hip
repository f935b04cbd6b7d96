// mosaic.hip -- the mosaic / A-projection gridder (replaces mosaic_grid_jit,
// core/imaging/imaging_utils/mosaic_grid.py:306-486).
//
// Each sample is convolved with a non-separable, oversampled complex
// convolution function (already multiplied by the field's phase gradient)
// chosen by its baseline, channel and polarisation, over a support that the
// same three indices choose.  The samples come binned by a plan whose
// support is at least max_support = max(weight_support): every footprint
// then lies inside the tiles the plan records the sample in, and the plan's
// own bounds test keeps every sample that the reference's test (u_centre +-
// max_support inside the grid, mosaic_grid.py:379-383) keeps.  That test is
// applied here as records are staged, with max_support reduced from the
// table on the device.
//
// One workgroup per work unit, one wavefront per polarisation.  The unit's
// tile, all its polarisations (polarisation-major, so the lanes of a
// wavefront, which run along v, touch consecutive cells), is accumulated in
// LDS.  Records are staged 64 at a time, one per lane, each wavefront
// staging its own polarisation (FP64 position, kernel offsets,
// convolution-function base and supports, weighted data); every wavefront
// then takes the staged records
// one after another, its lanes over the taps of its own polarisation.  A
// tap reads its complex kernel value from the convolution function in
// global memory (read-only, shared by every unit of the same baseline
// group, so it stays in L2) and adds kernel * data into its tile cell: the
// taps of one record are distinct cells and only this wavefront touches the
// tile, so no atomics.  The tile then goes to HBM once, split bins adding
// theirs with atomics.
//
// The taps of a record that count (inside the tile and inside the
// convolution function) form a rectangle, which staging works out once per
// record and polarisation, one record per lane, with its first kernel-value
// index and cell.  The wavefronts then walk only that rectangle: a tap is a
// quotient by the rectangle's width and two 24-bit multiply-adds from
// wave-uniform bases, with no bounds tests, and taps outside the tile
// (most of a record near a tile edge) cost nothing.
//
// The kernel sum of a record (the reference's `norm`, over all its taps)
// depends only on its convolution function and the kernel offsets of its
// first tap, so it is tabulated per call, summed in the reference's order
// (v outer, u inner, in FP64), and each wavefront adds its primary
// records' share of the sum of weights as it stages them: weight * Re(norm)
// for the PSF, weight * Re(norm^2) for data (mosaic_grid.py:470-484).
#include "tile_common.h"

namespace av {

constexpr int kMosaicStage = 64;
// units per XCD group: a tile's channels are consecutive units, so a group
// of 64 keeps a tile's units (and the samples they share) on one XCD's L2
constexpr uint32_t kMosaicGroup = 64;

template <typename R>
struct MosaicArgs {
  const R *vis, *weight;
  const int32_t *cf_baseline_map, *cf_chan_map, *cf_pol_map;
  const R *cfpg;  // complex [ncb][ncc][ncp][cu][cv]
  const int32_t *weight_support;  // [ncb][ncc][ncp][2]
  int n_baseline, ncb, ncc, ncp, cu_n, cv_n, ov_u, ov_v;
  R *grid;
  // kernel sums [ncb * ncc * ncp][ndu][ndv] (re, im), for first-tap kernel
  // offsets cu_n / 2 + du_lo + du and cv_n / 2 + dv_lo + dv
  const double *norms;
  int du_lo, ndu, dv_lo, ndv;
  const int *max_support = nullptr;  // max(weight_support), one value on the device
};

// max(weight_support): the reference's max_support_center
// (mosaic_grid.py:345), by one workgroup of 256 threads
__global__ __launch_bounds__(256) void k_mosaic_max_support(const int32_t *ws, int n, int *out) {
  __shared__ int red[4];
  int m = INT_MIN;
  for (int i = threadIdx.x; i < n; i += 256) m = max(m, ws[i]);
  for (int off = 32; off >= 1; off >>= 1) m = max(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) *out = max(max(red[0], red[1]), max(red[2], red[3]));
}

// per staged record and polarisation: the h x w rectangle of its taps that
// count (n = h * w, 0 for a record with nothing to add here).  Tap t of it
// is row q = t / w, column p = t - q * w; its kernel value is element
// cfoff + q * ov_u * cv_n + p * ov_v of the convolution functions and its
// cell is cell0 + q * tile_v + p.
template <typename R, int NPOL>
struct MosaicStaged {
  int n[NPOL], w[NPOL], cfoff[NPOL], cell0[NPOL];
  float rw[NPOL];  // 1 / w rounded up: (int)(t * rw) = t / w for t < 2^16
  R dre[NPOL], dim[NPOL];
};

// floor(x / d) for d > 0
__device__ inline int floor_div(int x, int d) { return x >= 0 ? x / d : -((-x + d - 1) / d); }

// The kernel-sum table: one thread per (convolution function, first-tap
// offsets), summing the in-range taps as the reference does.
template <typename R>
__global__ __launch_bounds__(256) void k_mosaic_norms(MosaicArgs<R> A, int ncf) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = (int64_t)A.ndu * A.ndv;
  if (i >= per * ncf) return;
  const int b = (int)(i / per), rem = (int)(i - (int64_t)b * per);
  const int du = rem / A.ndv, dv = rem - du * A.ndv;
  const int cfu0 = A.cu_n / 2 + A.du_lo + du, cfv0 = A.cv_n / 2 + A.dv_lo + dv;
  const int su = A.weight_support[2 * b], sv = A.weight_support[2 * b + 1];
  const int scu = su / 2, scv = sv / 2;
  const Cplx2<R> *cf = (const Cplx2<R> *)A.cfpg + (size_t)b * A.cu_n * A.cv_n;
  double nre = 0.0, nim = 0.0;
  for (int iv = -scv; iv < sv - scv; ++iv) {
    const int cfv = A.ov_v * iv + cfv0;
    for (int iu = -scu; iu < su - scu; ++iu) {
      const int cfu = A.ov_u * iu + cfu0;
      if (cfu < 0 || cfu >= A.cu_n || cfv < 0 || cfv >= A.cv_n) continue;
      const Cplx2<R> c = cf[(size_t)cfu * A.cv_n + cfv];
      nre = nre + (double)c.re;
      nim = nim + (double)c.im;
    }
  }
  ((double2 *)A.norms)[i] = make_double2(nre, nim);
}

template <typename R, int NPOL>
__global__ __launch_bounds__(64 * NPOL) void k_mosaic_grid(PlanDev P, MosaicArgs<R> A, int psf,
                                                           int accumulate) {
  extern __shared__ __align__(16) char smem[];
  using St = MosaicStaged<R, NPOL>;
  const int T = P.tile_u, TV = P.tile_v;
  Cplx2<R> *tile = (Cplx2<R> *)smem;  // [pol][u][v]
  St *st = (St *)(smem + (size_t)T * TV * NPOL * sizeof(Cplx2<R>));
  const int tid = threadIdx.x, lane = tid & 63;
  const int wp = __builtin_amdgcn_readfirstlane(tid >> 6);  // this wavefront's polarisation
  const uint32_t u = xcd_grouped_block(blockIdx.x, kMosaicGroup);
  const uint32_t n_units = P.gunit_start[P.n_bins];
  if (u >= n_units) return;
  const uint4 unit = P.gunits[u];
  const uint32_t b = unit.x, first = unit.y, cnt = unit.z, fl = unit.w;
  const uint32_t tile_i = b / (uint32_t)P.n_img_chan;
  const int ic = (int)(b - tile_i * (uint32_t)P.n_img_chan);
  const uint32_t tu_i = tile_i / (uint32_t)P.ntv;
  const int u0 = (int)tu_i * T, v0 = (int)(tile_i - tu_i * (uint32_t)P.ntv) * TV;
  const uint32_t prim_end = P.bin_start[b] + P.cnt_prim[b];
  R *gplane = A.grid + (size_t)ic * NPOL * P.n_u * P.n_v * 2;
  const size_t pstride = (size_t)P.n_u * P.n_v * 2;
  const int ncf = A.ncb * A.ncc * A.ncp;

  for (int i = tid; i < T * TV * NPOL; i += 64 * NPOL) tile[i] = Cplx2<R>{(R)0, (R)0};
  double swl = 0.0;  // this lane's records' sum of weights, this wavefront's polarisation
  const int ccu = A.cu_n / 2, ccv = A.cv_n / 2;
  const int cfsize = A.cu_n * A.cv_n;  // the host bounds ncf * cfsize below 2^31
  const int cfrow = A.ov_u * A.cv_n;   // elements between taps of one column
  const Cplx2<R> *cfall = (const Cplx2<R> *)A.cfpg;
  const int msup = *A.max_support;

  for (uint32_t base = first; base < first + cnt; base += kMosaicStage) {
    uint64_t live;  // the staged records with taps for this wavefront's polarisation
    __syncthreads();  // the previous batch's staged records are consumed
    // ---- stage this lane's record: every wavefront forms the record's
    // geometry and its own polarisation's rectangle and data, so the four
    // chains of loads run side by side
    {
      St &s = st[lane];
      int pn = 0, pw = 1, pcfoff = 0, pcell0 = 0;
      float prw = 1.0f;
      R pdre = (R)0, pdim = (R)0;
      const uint32_t r = base + lane;
      if (r < first + cnt) {
        const uint32_t item = P.rec[3 * (size_t)r];
        const uint32_t geo = P.rec[3 * (size_t)r + 1];
        int au, ou, avv, ovv;
        unpack_geo(geo, au, ou, avv, ovv);
        const int64_t row = (int64_t)(item / (uint32_t)P.n_chan);
        const int ch = (int)(item - (uint32_t)row * (uint32_t)P.n_chan);
        const R w = A.weight[(size_t)item * NPOL + wp];
        R vre = (R)0, vim = (R)0;
        if (!psf) {
          vre = A.vis[((size_t)item * NPOL + wp) * 2];
          vim = A.vis[((size_t)item * NPOL + wp) * 2 + 1];
        }
        // a 32-bit remainder: row < 2^32, as item = row * n_chan + ch is 32-bit
        const int cfb = A.cf_baseline_map[(uint32_t)row % (uint32_t)A.n_baseline];
        const int cfc = A.cf_chan_map[ch];
        const int idx = (cfb * A.ncc + cfc) * A.ncp + A.cf_pol_map[wp];
        // the reference's offsets at this call's oversampling, from the
        // same FP64 positions the plan binned (mosaic_grid.py:376-386)
        const double uu = P.uvw[row * 3] * P.gscale_u[ch] + (double)(P.n_u / 2);
        const double vv = P.uvw[row * 3 + 1] * P.gscale_v[ch] + (double)(P.n_v / 2);
        const int gcu = u0 + au, gcv = v0 + avv;
        const int cu = au, cv = avv;
        // the reference's bounds test (mosaic_grid.py:379-383), written
        // without sums that could overflow for any table value
        const bool inside = gcu >= msup && P.n_u - gcu > msup && gcv >= msup && P.n_v - gcv > msup;
        const int cfu0 = (int)floor(((double)gcu - uu) * (double)A.ov_u + 0.5) + ccu;
        const int cfv0 = (int)floor(((double)gcv - vv) * (double)A.ov_v + 0.5) + ccv;
        const bool prim = r < prim_end;
        const int du = min(max(cfu0 - ccu - A.du_lo, 0), A.ndu - 1);
        const int dv = min(max(cfv0 - ccv - A.dv_lo, 0), A.ndv - 1);
        const bool cf_ok = idx >= 0 && idx < ncf;
        const int idc = cf_ok ? idx : 0;  // in-bounds reads whatever the maps hold
        const int wsu = A.weight_support[2 * idc], wsv = A.weight_support[2 * idc + 1];
        const double2 nm = ((const double2 *)A.norms)[((size_t)idc * A.ndu + du) * A.ndv + dv];
        R dre, dim;
        if (psf) {
          dre = w;
          dim = (R)0;
        } else {
          dre = vre * w;
          dim = vim * w;
        }
        // mosaic_grid.py:412: skip NaN or exactly-zero data (and, here,
        // convolution functions outside the table)
        if (inside && cf_ok && !(isnan(dre) || isnan(dim)) && !(dre == (R)0 && dim == (R)0)) {
          // taps iu in [-wsu / 2, wsu - wsu / 2) whose kernel row
          // ov_u * iu + cfu0 lies in [0, cu_n) and whose cell cu + iu lies
          // in the tile; the same along v
          const int scu = wsu / 2, scv = wsv / 2;
          const int iu_lo = max(max(-scu, -floor_div(cfu0, A.ov_u)), -cu);
          const int iu_hi = min(min(wsu - scu - 1, floor_div(A.cu_n - 1 - cfu0, A.ov_u)), T - 1 - cu);
          const int iv_lo = max(max(-scv, -floor_div(cfv0, A.ov_v)), -cv);
          const int iv_hi = min(min(wsv - scv - 1, floor_div(A.cv_n - 1 - cfv0, A.ov_v)), TV - 1 - cv);
          if (iu_hi >= iu_lo && iv_hi >= iv_lo) {
            pw = iv_hi - iv_lo + 1;
            pn = (iu_hi - iu_lo + 1) * pw;
            pcfoff = idx * cfsize + (A.ov_u * iu_lo + cfu0) * A.cv_n + A.ov_v * iv_lo + cfv0;
            pcell0 = (cu + iu_lo) * TV + cv + iv_lo;
            prw = (1.0f / (float)pw) * (1.0f + 0x1p-20f);
          }
          pdre = dre;
          pdim = dim;
          if (prim) {
            const double wd = (double)w;
            swl += psf ? wd * nm.x : wd * (nm.x * nm.x - nm.y * nm.y);
          }
        }
      }
      live = __ballot(pn > 0);
      s.n[wp] = pn;
      s.w[wp] = pw;
      s.cfoff[wp] = pcfoff;
      s.cell0[wp] = pcell0;
      s.rw[wp] = prw;
      s.dre[wp] = pdre;
      s.dim[wp] = pdim;
    }
    __syncthreads();
    // ---- the staged records, one after another; lanes over the taps of
    // this wavefront's polarisation, v fastest (t = iu' * sv + iv')
    Cplx2<R> *ptile = tile + (size_t)wp * T * TV;
    for (; live; live &= live - 1) {
      const St &s = st[__builtin_ctzll(live)];
      // every lane reads the same record: its values are wave-uniform
      const int n = __builtin_amdgcn_readfirstlane(s.n[wp]);
      const int w = __builtin_amdgcn_readfirstlane(s.w[wp]);
      const int cell0 = __builtin_amdgcn_readfirstlane(s.cell0[wp]);
      const float rw = __builtin_bit_cast(
          float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, s.rw[wp])));
      const Cplx2<R> *cf = cfall + __builtin_amdgcn_readfirstlane(s.cfoff[wp]);
      const R dre = s.dre[wp], dim = s.dim[wp];
      for (int t = lane; t < n; t += 64) {
        const int q = (int)((float)t * rw);
        const int pc = t - (int)__umul24(q, w);
        // an unsigned offset: no sign extension before the 64-bit address
        const uint32_t off = (uint32_t)__umul24(q, cfrow) + (uint32_t)__umul24(pc, A.ov_v);
        const Cplx2<R> c = cf[off];
        Cplx2<R> &cell = ptile[cell0 + (int)__umul24(q, TV) + pc];
        // the product first, then the cell, as the reference adds them
        cell.re += fma(c.re, dre, -(c.im * dim));
        cell.im += fma(c.re, dim, c.im * dre);
      }
    }
  }
  __syncthreads();
  // ---- flush the tile (cells outside the grid dropped)
  const bool multi = (fl & kUnitMulti) != 0;
  for (int i = tid; i < T * TV * NPOL; i += 64 * NPOL) {
    const int p = i / (T * TV), cell = i - p * (T * TV);
    const int gu = u0 + cell / TV, gv = v0 + cell % TV;
    if (gu >= P.n_u || gv >= P.n_v) continue;
    const Cplx2<R> val = tile[i];
    R *dst = gplane + p * pstride + ((size_t)gu * P.n_v + gv) * 2;
    if (multi) {
      if (val.re != (R)0) atomicAdd(dst, val.re);
      if (val.im != (R)0) atomicAdd(dst + 1, val.im);
    } else if (accumulate) {
      dst[0] += val.re;
      dst[1] += val.im;
    } else {
      dst[0] = val.re;
      dst[1] = val.im;
    }
  }
  for (int off = 32; off >= 1; off >>= 1) swl += __shfl_xor(swl, off);
  if (lane == 0) P.upart[(size_t)u * 4 + wp] = swl;
}

template <typename R, int NPOL>
static int launch_mosaic(av_plan *plan, MosaicArgs<R> A, double *sumwt, int mode,
                         hipStream_t s) {
  PlanDev &d = plan->d;
  const size_t lds = (size_t)d.tile_u * d.tile_v * NPOL * sizeof(Cplx2<R>) +
                     (size_t)kMosaicStage * sizeof(MosaicStaged<R, NPOL>);
  const DeviceProps &prop = device_props();
  if (prop.error) return hip_fail(prop.error, "av_mosaic_grid: device properties");
  if (lds > prop.lds_per_block)
    return fail(AV_ERR_UNSUPPORTED, "av_mosaic_grid: tile too large for LDS (" +
                                        std::to_string(lds) + " bytes)");
  {
    const int rc = ensure_lds_limit<k_mosaic_grid<R, NPOL>>((int)lds,
                                                            "av_mosaic_grid: hipFuncSetAttribute");
    if (rc) return rc;
  }
  const bool accumulate = (mode & AV_GRID_ACCUMULATE) != 0;
  if (!accumulate) {
    const int rc = zero_before_grid(d, A.grid, NPOL, sizeof(R) == 8, sumwt, s);
    if (rc) return rc;
  }
  // the kernel-sum table, in stream-ordered memory freed after the gridder
  // (first-tap offsets: floor(x * ov + 0.5) for x in (-1/2, 1/2], within
  // [-ov / 2 - 1, ov / 2 + 1])
  const int64_t ncf = (int64_t)A.ncb * A.ncc * A.ncp;
  if (ncf * A.cu_n * A.cv_n >= (int64_t)1 << 31)
    return fail(AV_ERR_UNSUPPORTED, "av_mosaic_grid: convolution functions of 2^31 elements or more");
  if ((int64_t)A.ov_u * A.cv_n >= 1 << 24 || A.ov_v >= 1 << 24)
    return fail(AV_ERR_UNSUPPORTED, "av_mosaic_grid: ov_u * cv_n must be below 2^24");
  if (d.tile_u * d.tile_v > 1 << 16)
    return fail(AV_ERR_UNSUPPORTED, "av_mosaic_grid: tile of more than 2^16 cells");
  A.du_lo = -(A.ov_u / 2) - 1;
  A.ndu = A.ov_u + 3;
  A.dv_lo = -(A.ov_v / 2) - 1;
  A.ndv = A.ov_v + 3;
  const int64_t entries = ncf * A.ndu * A.ndv;
  // one stream-ordered block: the kernel sums, then max(weight_support)
  void *norms = nullptr;
  AV_HIP(hipMallocAsync(&norms, (size_t)entries * 2 * sizeof(double) + 16, s));
  A.norms = (const double *)norms;
  int *msup = (int *)((char *)norms + (size_t)entries * 2 * sizeof(double));
  A.max_support = msup;
  hipLaunchKernelGGL((k_mosaic_norms<R>), dim3((unsigned)((entries + 255) / 256)), dim3(256), 0, s,
                     A, (int)ncf);
  int rc = hipGetLastError() == hipSuccess ? AV_OK : fail(AV_ERR_HIP, "k_mosaic_norms launch");
  if (rc == AV_OK) {
    hipLaunchKernelGGL(k_mosaic_max_support, dim3(1), dim3(256), 0, s, A.weight_support,
                       (int)(2 * ncf), msup);
    if (hipGetLastError() != hipSuccess) rc = fail(AV_ERR_HIP, "k_mosaic_max_support launch");
  }
  if (rc == AV_OK) {
    hipLaunchKernelGGL((k_mosaic_grid<R, NPOL>),
                       dim3((unsigned)xcd_grouped_grid(d.max_gunits, kMosaicGroup)),
                       dim3(64 * NPOL), lds, s, d, A, (mode & AV_GRID_PSF) ? 1 : 0,
                       accumulate ? 1 : 0);
    if (hipGetLastError() != hipSuccess) rc = fail(AV_ERR_HIP, "k_mosaic_grid launch");
  }
  const hipError_t fe = hipFreeAsync(norms, s);
  if (rc) return rc;
  if (fe != hipSuccess) return hip_fail(fe, "av_mosaic_grid: hipFreeAsync");
  return reduce_sumwt(d, NPOL, sumwt, s);
}

template <typename R>
static int mosaic_by_pol(av_plan *plan, MosaicArgs<R> A, int npol, double *sumwt, int mode,
                         hipStream_t s) {
  switch (npol) {
    case 1: return launch_mosaic<R, 1>(plan, A, sumwt, mode, s);
    case 2: return launch_mosaic<R, 2>(plan, A, sumwt, mode, s);
    case 4: return launch_mosaic<R, 4>(plan, A, sumwt, mode, s);
    default: return fail(AV_ERR_UNSUPPORTED, "av_mosaic_grid: npol must be 1, 2 or 4");
  }
}

}  // namespace av

using namespace av;

extern "C" int av_mosaic_grid(av_plan *plan, const void *vis, const void *weight, int32_t n_pol,
                              int32_t n_baseline, const int32_t *cf_baseline_map,
                              const int32_t *cf_chan_map, const int32_t *cf_pol_map,
                              const void *cfpg, int32_t ncb, int32_t ncc, int32_t ncp,
                              int32_t cu_n,
                              int32_t cv_n, const int32_t *weight_support, int32_t ov_u,
                              int32_t ov_v, void *grid, double *sumwt, int32_t dtype,
                              int32_t mode, void *stream) {
  if (!plan || !weight || !cf_baseline_map || !cf_chan_map || !cf_pol_map || !cfpg ||
      !weight_support || !grid || !sumwt)
    return fail(AV_ERR_INVALID, "av_mosaic_grid: NULL argument");
  const bool psf = (mode & AV_GRID_PSF) != 0;
  if (!psf && !vis) return fail(AV_ERR_INVALID, "av_mosaic_grid: vis is NULL");
  if (mode & ~(AV_GRID_PSF | AV_GRID_ACCUMULATE))
    return fail(AV_ERR_UNSUPPORTED, "av_mosaic_grid: mode takes AV_GRID_PSF and "
                                    "AV_GRID_ACCUMULATE only");
  const PlanDev &d = plan->d;
  if (!d.uvw) return fail(AV_ERR_INVALID, "av_mosaic_grid: plan has no uvw (av_plan_set_uvw)");
  if (d.row_channels || d.geometry != AV_GEOM_REFERENCE || d.conjugate)
    return fail(AV_ERR_UNSUPPORTED, "av_mosaic_grid: needs an AV_GEOM_REFERENCE plan "
                                    "without row channels");
  if (n_baseline < 1 || d.n_rows % n_baseline != 0)
    return fail(AV_ERR_INVALID, "av_mosaic_grid: n_baseline must divide the plan's rows");
  if (ncb < 1 || ncc < 1 || ncp < 1 || cu_n < 1 || cv_n < 1 || ov_u < 1 || ov_v < 1)
    return fail(AV_ERR_INVALID, "av_mosaic_grid: bad convolution-function shape");
  hipStream_t s = (hipStream_t)stream;
  if (dtype == AV_F32) {
    MosaicArgs<float> A{(const float *)vis, (const float *)weight, cf_baseline_map, cf_chan_map,
                        cf_pol_map, (const float *)cfpg, weight_support, n_baseline, ncb, ncc,
                        ncp, cu_n, cv_n, ov_u, ov_v, (float *)grid, nullptr, 0, 0, 0, 0};
    return mosaic_by_pol<float>(plan, A, n_pol, sumwt, mode, s);
  }
  if (dtype == AV_F64) {
    MosaicArgs<double> A{(const double *)vis, (const double *)weight, cf_baseline_map,
                         cf_chan_map, cf_pol_map, (const double *)cfpg, weight_support,
                         n_baseline, ncb, ncc, ncp, cu_n, cv_n, ov_u, ov_v, (double *)grid,
                         nullptr, 0, 0, 0, 0};
    return mosaic_by_pol<double>(plan, A, n_pol, sumwt, mode, s);
  }
  return fail(AV_ERR_INVALID, "av_mosaic_grid: dtype must be AV_F32 or AV_F64");
}
