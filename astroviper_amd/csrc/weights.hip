// weights.hip -- imaging weights on the GPU (replaces the numba kernels of
// core/imaging/imaging_weighting/grid_imaging_weights.py: grid_imaging_weights_jit
// :105-206 and degrid_imaging_weights_jit :274-443).
//
// Nearest-pixel gridding of the data weights (and their conjugate pixels)
// into a float64 density grid per channel, and the per-sample Briggs
// reweighting from that grid.  One thread per sample, in blocks of time
// steps and channel by channel within a block, so that a wavefront's
// samples share a channel: the sum of weights of a wavefront is reduced
// first and added with one atomic.  Grid cells take FP64 atomics
// (one sample adds to two cells); the sums therefore depend on the order of
// arrival in their last bits only.
#include "avgrid_internal.h"

namespace av {

// grid_imaging_weights_jit: the pixel of position x is int(x + 0.5),
// truncated toward zero (:176-181); positions beyond +-1e9 are rejected
// before the conversion (they fail the bounds test anyway)
__device__ inline bool nearest_pixel(double x, int n, int &i) {
  const double y = x + 0.5;
  if (!(y > -1.0e9 && y < 1.0e9)) return false;
  i = (int)y;
  return i >= 0 && i < n;
}

// Sample i of a launch (one thread per (row, channel); rows are time-major,
// n_baseline per time step): blocks of kTimeBlock time steps, within a block
// channel by channel, within a channel baseline by baseline with time
// fastest.  A wavefront then holds consecutive time samples of one baseline,
// which usually fall in the same pixel, and a block's uvw rows (and the
// lines of neighbouring channels' weights) are read by all its channels
// while they are still in cache: taking whole channels in turn read them
// from memory once per channel (grid pass 12.9 -> 9.3 ms, reweighting
// 7.6 -> 3.5 ms on 64 channels x 2016 baselines x 1000 times).
constexpr int64_t kTimeBlock = 16;
__device__ inline void sample_of(int64_t i, int64_t n_rows, int64_t nb, int nch, int &ch,
                                 int64_t &row) {
  const int64_t nt = n_rows / nb;
  const int64_t full = kTimeBlock * nb * nch;  // samples of a full block
  const int64_t k = i / full;
  const int64_t j = i - k * full;
  const int64_t bt = min(kTimeBlock, nt - k * kTimeBlock);  // time steps of this block (the last may be short)
  ch = (int)(j / (bt * nb));
  const int64_t rem = j - (int64_t)ch * bt * nb;
  const int64_t b = rem / bt;
  row = (k * kTimeBlock + (rem - b * bt)) * nb + b;
}

// Sum v over the runs of equal keys in lane order (the wavefront's lanes
// all active); returns true in the last lane of each run, which then holds
// the run's total.  Runs are summed by doubling steps, in an order fixed by
// the lanes.
__device__ inline bool run_sum(int64_t key, double &v) {
  const int lane = (int)__lane_id();
  // the first lane of this lane's run
  const int64_t kp = __shfl_up(key, 1, 64);
  int start = (lane == 0 || kp != key) ? lane : 0;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) start = max(start, __shfl_up(start, d, 64) * (lane >= d));
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const double o = __shfl_up(v, d, 64);
    if (lane - d >= start) v += o;
  }
  const int64_t kn = __shfl_down(key, 1, 64);
  return lane == 63 || kn != key;
}

__global__ __launch_bounds__(256) void k_grid_imaging_weights(
    const double *__restrict__ uvw, int64_t n_rows, int64_t n_baseline,
    const double *__restrict__ freq, int n_chan, const double *__restrict__ weight, int wstride,
    int n_u, int n_v, double cell_u, double cell_v, double *__restrict__ grid,
    double *__restrict__ sum_weight) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = n_rows * n_chan;
  int ch = n_chan - 1;
  int64_t row = 0;
  if (i < total) sample_of(i, n_rows, n_baseline, n_chan, ch, row);
  double add = 0.0;
  // this sample's pixel and conjugate pixel (-1: none), and its weight
  int64_t pix = -1, cpix = -1;
  double w = 0.0;
  if (i < total) {
    // uv_scale = -(freq * delta_lm * n_uv) / c  (:150-152)
    const double su = -(freq[ch] * cell_u * (double)n_u) / kLightSpeed;
    const double sv = -(freq[ch] * cell_v * (double)n_v) / kLightSpeed;
    const double u = uvw[row * 3] * su, v = uvw[row * 3 + 1] * sv;
    int iu, iv;
    if (!isnan(u) && !isnan(v) && nearest_pixel(u + (double)(n_u / 2), n_u, iu) &&
        nearest_pixel(v + (double)(n_v / 2), n_v, iv)) {
      w = weight[(row * n_chan + ch) * wstride];
      if (!isnan(w)) {
        pix = ((int64_t)ch * n_u + iu) * n_v + iv;
        // the conjugate pixel; the reference does not bound it (on row or
        // column 0 of an even grid it is one past the end): skipped there
        int iuc, ivc;
        if (nearest_pixel(-u + (double)(n_u / 2), n_u, iuc) &&
            nearest_pixel(-v + (double)(n_v / 2), n_v, ivc))
          cpix = ((int64_t)ch * n_u + iuc) * n_v + ivc;
        add = 2.0 * w;
      }
    }
  }
  // one atomic per run of equal pixels in the wavefront (skipped samples
  // take keys of their own below -1, so they form runs that add nothing)
  const int lane = (int)__lane_id();
  {
    double v = pix >= 0 ? w : 0.0;
    if (run_sum(pix >= 0 ? pix : -2 - lane, v) && pix >= 0) atomicAdd(&grid[pix], v);
  }
  {
    double v = cpix >= 0 ? w : 0.0;
    if (run_sum(cpix >= 0 ? cpix : -2 - lane, v) && cpix >= 0) atomicAdd(&grid[cpix], v);
  }
  // one atomic per wavefront when all its samples share the channel
  const int ch0 = __shfl(ch, 0, 64);
  const bool uniform = __all(ch == ch0);
  if (uniform) {
    for (int off = 32; off >= 1; off >>= 1) add += __shfl_xor(add, off, 64);
    if (lane == 0 && add != 0.0) atomicAdd(&sum_weight[ch0], add);
  } else if (add != 0.0) {
    atomicAdd(&sum_weight[ch], add);
  }
}

__global__ __launch_bounds__(256) void k_degrid_imaging_weights(
    const double *__restrict__ grid, const double *__restrict__ briggs,
    const double *__restrict__ uvw, int64_t n_rows, int64_t n_baseline,
    const double *__restrict__ freq, int n_chan, const double *__restrict__ weight, int n_u,
    int n_v, double cell_u, double cell_v, double *__restrict__ out,
    int32_t *__restrict__ nan_flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = n_rows * n_chan;
  if (i >= total) return;
  int ch;
  int64_t row;
  sample_of(i, n_rows, n_baseline, n_chan, ch, row);
  const int64_t k = row * n_chan + ch;
  const double su = -(freq[ch] * cell_u * (double)n_u) / kLightSpeed;
  const double sv = -(freq[ch] * cell_v * (double)n_v) / kLightSpeed;
  const double u = uvw[row * 3] * su, v = uvw[row * 3 + 1] * sv;
  int iu, iv;
  // outside the grid the imaging weight stays 0 (the output starts at zero, :260)
  if (isnan(u) || isnan(v) || !nearest_pixel(u + (double)(n_u / 2), n_u, iu) ||
      !nearest_pixel(v + (double)(n_v / 2), n_v, iv)) {
    out[k] = 0.0;
    return;
  }
  const double w = weight[k];
  double r = w;
  if (!isnan(w) && w != 0.0) {
    const double g = grid[((size_t)ch * n_u + iu) * n_v + iv];
    if (!isnan(g) && g != 0.0) {
      const double b0 = briggs[ch], b1 = briggs[n_chan + ch];
      if (isnan(b0) || isnan(b1)) atomicOr(nan_flag, 1);
      const double denom = b0 * g + b1;  // Briggs denominator a * G + b (:430-440)
      r = r / denom;
    }
  }
  out[k] = r;
}

}  // namespace av

using namespace av;

extern "C" {

int av_grid_imaging_weights(const double *uvw, int64_t n_rows, int32_t n_baseline,
                            const double *freq,
                            int32_t n_chan, const double *weight, int32_t weight_stride,
                            int32_t n_u, int32_t n_v, double cell_u, double cell_v, double *grid,
                            double *sum_weight, void *stream) {
  if (!freq || !grid || !sum_weight || (n_rows > 0 && (!uvw || !weight)))
    return fail(AV_ERR_INVALID, "av_grid_imaging_weights: NULL argument");
  if (n_rows < 0 || n_chan < 1 || n_u < 1 || n_v < 1 || weight_stride < 1 || n_baseline < 1 ||
      n_rows % n_baseline != 0)
    return fail(AV_ERR_INVALID, "av_grid_imaging_weights: bad sizes");
  const int64_t total = n_rows * n_chan;
  if (total == 0) return AV_OK;
  if ((total + 255) / 256 > 0x7FFFFFFF)
    return fail(AV_ERR_UNSUPPORTED, "av_grid_imaging_weights: too many samples");
  hipLaunchKernelGGL(k_grid_imaging_weights, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, uvw, n_rows, (int64_t)n_baseline, freq, n_chan, weight,
                     weight_stride, n_u,
                     n_v, cell_u, cell_v, grid, sum_weight);
  AV_LAUNCH_CHECK("k_grid_imaging_weights");
  return AV_OK;
}

int av_degrid_imaging_weights(const double *grid, const double *briggs, const double *uvw,
                              int64_t n_rows, int32_t n_baseline, const double *freq,
                              int32_t n_chan,
                              const double *weight, int32_t n_u, int32_t n_v, double cell_u,
                              double cell_v, double *imaging_weight, int32_t *nan_flag,
                              void *stream) {
  if (!grid || !briggs || !freq || !nan_flag || (n_rows > 0 && (!uvw || !weight || !imaging_weight)))
    return fail(AV_ERR_INVALID, "av_degrid_imaging_weights: NULL argument");
  if (n_rows < 0 || n_chan < 1 || n_u < 1 || n_v < 1 || n_baseline < 1 ||
      n_rows % n_baseline != 0)
    return fail(AV_ERR_INVALID, "av_degrid_imaging_weights: bad sizes");
  const int64_t total = n_rows * n_chan;
  if (total == 0) return AV_OK;
  if ((total + 255) / 256 > 0x7FFFFFFF)
    return fail(AV_ERR_UNSUPPORTED, "av_degrid_imaging_weights: too many samples");
  hipLaunchKernelGGL(k_degrid_imaging_weights, dim3((unsigned)((total + 255) / 256)), dim3(256),
                     0, (hipStream_t)stream, grid, briggs, uvw, n_rows, (int64_t)n_baseline, freq,
                     n_chan, weight,
                     n_u, n_v, cell_u, cell_v, imaging_weight, nan_flag);
  AV_LAUNCH_CHECK("k_degrid_imaging_weights");
  return AV_OK;
}

}  // extern "C"
