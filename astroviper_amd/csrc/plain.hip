// plain.hip -- the straightforward GPU gridder/degridder used as the
// comparison baseline: one thread per sample of one visibility channel, no
// binning, no LDS; the grid is updated with global atomics (gridding) or
// gathered straight from HBM (degridding).  Same FP64 geometry rules as the
// tiled path (standard_grid.py:322-352, standard_degrid.py:435-462).
#include "tile_common.h"

namespace av {

// Sum of x over the lanes of a wavefront that share lane % npol (npol a
// power of two dividing 64): every such lane ends with the total.
__device__ inline double wave_sum_per_pol(double x, int npol) {
  for (int off = 32; off >= npol; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

template <typename R>
__global__ __launch_bounds__(256) void k_grid_plain(const double *uvw, int64_t n_rows, int n_chan,
                                                    int chan, double scale_u, double scale_v,
                                                    int img_chan, const R *vis, const R *weight,
                                                    const uint8_t *flag, int npol,
                                                    const double *cgk, int support, int ov,
                                                    int n_u, int n_v, R *grid, double *sumwt,
                                                    int mode) {
  // i = row * npol + p and 256 is a multiple of npol, so p = lane % npol.
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t row = i / npol;
  const int p = (int)(i - row * npol);
  const int sc = support / 2;
  double wnorm = 0.0;  // this sample's w * norm; summed over the wavefront below
  int iu, iv, ou, ovv;
  if (i < n_rows * npol &&
      grid_geometry(uvw[row * 3], uvw[row * 3 + 1], scale_u, scale_v, n_u, n_v, sc, ov, iu, iv,
                    ou, ovv)) {
    const size_t s = ((size_t)row * n_chan + chan) * npol + p;
    R w = weight[s];
    const R wsel = w;
    if (flag && flag[s]) w = (R)0;
    R dre, dim;
    if (mode & AV_GRID_PSF) {
      dre = w;
      dim = (R)0;
    } else {
      dre = vis[2 * s] * w;
      dim = vis[2 * s + 1] * w;
    }
    if (!(isnan(dre) || isnan(dim) || (dre == (R)0 && dim == (R)0))) {
      R *g = grid + ((size_t)img_chan * npol + p) * n_u * n_v * 2;
      double norm = 0.0;
      for (int tv = 0; tv < support; ++tv) {
        const double cv = cgk[tab_index(ov, tv, sc, ovv)];
        for (int tu = 0; tu < support; ++tu) {
          const double cu = cgk[tab_index(ov, tu, sc, ou)];
          const R conv = (R)(cu * cv);
          const size_t cell = (size_t)(iu + tu - sc) * n_v + (iv + tv - sc);
          atomicAdd(&g[2 * cell], conv * dre);
          atomicAdd(&g[2 * cell + 1], conv * dim);
          norm += cu * cv;
        }
      }
      wnorm = (double)wsel * norm;
    }
  }
  // one sum-of-weights atomic per polarisation per wavefront (every lane of
  // a launch adds to the same img_chan)
  wnorm = wave_sum_per_pol(wnorm, npol);
  if ((threadIdx.x & 63) < npol && wnorm != 0.0)
    atomicAdd(&sumwt[(size_t)img_chan * npol + p], wnorm);
}

template <typename R>
__global__ __launch_bounds__(256) void k_degrid_plain(const double *uvw, int64_t n_rows,
                                                      int n_chan, int chan, double freq,
                                                      double dscale_u, double dscale_v,
                                                      int img_chan, const R *grid, int npol_g,
                                                      const uint8_t *flag, int npol,
                                                      const double *cgk, int support, int ov,
                                                      int n_u, int n_v, R *vis_out, int mode) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_rows * npol) return;
  const int64_t row = i / npol;
  const int p = (int)(i - row * npol);
  const size_t s = ((size_t)row * n_chan + chan) * npol + p;
  const bool incremental = (mode & AV_DEGRID_INCREMENTAL) != 0;
  const int sc = support / 2;
  int lu, lv, fu, fv;
  bool ok = !(flag && flag[s]) &&
            degrid_geometry(uvw[row * 3], uvw[row * 3 + 1], freq, dscale_u, dscale_v,
                            (double)n_u / 2.0, (double)n_v / 2.0, n_u, n_v, support, ov, lu, lv,
                            fu, fv);
  if (!ok) {
    if (!incremental) {
      vis_out[2 * s] = (R)0;
      vis_out[2 * s + 1] = (R)0;
    }
    return;
  }
  const R *g = grid + ((size_t)img_chan * npol_g + (p % npol_g)) * n_u * n_v * 2;
  R are = (R)0, aim = (R)0;
  double norm = 0.0;
  for (int tu = 0; tu < support; ++tu) {
    const double cu = cgk[tab_index(ov, tu, sc, fu)];
    for (int tv = 0; tv < support; ++tv) {
      const double cv = cgk[tab_index(ov, tv, sc, fv)];
      const R wt = (R)(cu * cv);
      const size_t cell = (size_t)(lu + tu - sc) * n_v + (lv + tv - sc);
      are += wt * g[2 * cell];
      aim += wt * g[2 * cell + 1];
      norm += cu * cv;
    }
  }
  // the reference adds nvalue / norm only when norm != 0
  // (standard_degrid.py:482-485): a zero norm leaves the value as it was
  const R nv = (R)norm;
  if (incremental) {
    if (nv != (R)0) {
      vis_out[2 * s] += are / nv;
      vis_out[2 * s + 1] += aim / nv;
    }
  } else {
    vis_out[2 * s] = nv != (R)0 ? are / nv : (R)0;
    vis_out[2 * s + 1] = nv != (R)0 ? aim / nv : (R)0;
  }
}

}  // namespace av

using namespace av;

extern "C" {

int av_grid_plain(const double *uvw, int64_t n_rows, int32_t n_chan, int32_t chan, double freq,
                  int32_t img_chan, const void *vis, const void *weight, const uint8_t *flag,
                  int32_t n_pol, const double *cgk, int32_t support, int32_t oversampling,
                  double cell_u, double cell_v, int32_t n_u, int32_t n_v, void *grid,
                  double *sumwt, int32_t dtype, int32_t mode, void *stream) {
  if (!uvw || !weight || !cgk || !grid || !sumwt || (!(mode & AV_GRID_PSF) && !vis))
    return fail(AV_ERR_INVALID, "av_grid_plain: NULL argument");
  if (chan < 0 || chan >= n_chan || support < 2 || support > kMaxSupport)
    return fail(AV_ERR_INVALID, "av_grid_plain: bad channel or support");
  if (n_pol != 1 && n_pol != 2 && n_pol != 4)
    return fail(AV_ERR_UNSUPPORTED, "av_grid_plain: npol must be 1, 2 or 4");
  if (n_rows == 0) return AV_OK;
  // numpy: -(freq_chan * delta_lm[0] * n_uv[0]) / c   (standard_grid.py:298-299)
  const double su = -(freq * cell_u * (double)n_u) / kLightSpeed;
  const double sv = -(freq * cell_v * (double)n_v) / kLightSpeed;
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = n_rows * n_pol;
  dim3 g((unsigned)((n + 255) / 256)), b(256);
  if (dtype == AV_F32)
    hipLaunchKernelGGL(k_grid_plain<float>, g, b, 0, s, uvw, n_rows, n_chan, chan, su, sv,
                       img_chan, (const float *)vis, (const float *)weight, flag, n_pol, cgk,
                       support, oversampling, n_u, n_v, (float *)grid, sumwt, mode);
  else if (dtype == AV_F64)
    hipLaunchKernelGGL(k_grid_plain<double>, g, b, 0, s, uvw, n_rows, n_chan, chan, su, sv,
                       img_chan, (const double *)vis, (const double *)weight, flag, n_pol, cgk,
                       support, oversampling, n_u, n_v, (double *)grid, sumwt, mode);
  else
    return fail(AV_ERR_INVALID, "av_grid_plain: bad dtype");
  AV_LAUNCH_CHECK("k_grid_plain");
  return AV_OK;
}

int av_degrid_plain(const double *uvw, int64_t n_rows, int32_t n_chan, int32_t chan, double freq,
                    int32_t img_chan, const void *grid, int32_t n_grid_pol, const uint8_t *flag,
                    int32_t n_pol, const double *cgk, int32_t support, int32_t oversampling,
                    double cell_u, double cell_v, int32_t n_u, int32_t n_v, void *vis_out,
                    int32_t dtype, int32_t mode, void *stream) {
  if (!uvw || !grid || !cgk || !vis_out)
    return fail(AV_ERR_INVALID, "av_degrid_plain: NULL argument");
  if (chan < 0 || chan >= n_chan || n_pol < 1 || n_grid_pol < 1 || support < 2 ||
      support > kMaxSupport)
    return fail(AV_ERR_INVALID, "av_degrid_plain: bad channel, npol or support");
  if (n_rows == 0) return AV_OK;
  const double du = (-cell_u) * (double)n_u, dv = (-cell_v) * (double)n_v;
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = n_rows * n_pol;
  dim3 g((unsigned)((n + 255) / 256)), b(256);
  if (dtype == AV_F32)
    hipLaunchKernelGGL(k_degrid_plain<float>, g, b, 0, s, uvw, n_rows, n_chan, chan, freq, du, dv,
                       img_chan, (const float *)grid, n_grid_pol, flag, n_pol, cgk, support,
                       oversampling, n_u, n_v, (float *)vis_out, mode);
  else if (dtype == AV_F64)
    hipLaunchKernelGGL(k_degrid_plain<double>, g, b, 0, s, uvw, n_rows, n_chan, chan, freq, du,
                       dv, img_chan, (const double *)grid, n_grid_pol, flag, n_pol, cgk,
                       support, oversampling, n_u, n_v, (double *)vis_out, mode);
  else
    return fail(AV_ERR_INVALID, "av_degrid_plain: bad dtype");
  AV_LAUNCH_CHECK("k_degrid_plain");
  return AV_OK;
}

}  // extern "C"
