// gcf.hip -- the pieces of the mosaic gridder's convolution functions that
// are evaluated per pixel: the airy-disk aperture patterns of the antenna
// types (airy_disk_rorder and casa_airy_disk_rorder,
// core/imaging/imaging_utils/make_pb_symmetric.py:148-280, on the grid that
// make_baseline_patterns gives them, make_gridding_convolution_function.py:
// 196-229) and the phase gradient of a field (make_phase_gradient, :426-485).
//
// One thread per output value.  Each value is formed with the reference's
// operations in the reference's order (NumPy evaluates left to right, one
// rounding per operation; this file is compiled without contraction), with
// the device J1 Bessel function in place of scipy.special.jn(1, .).
#include "avgrid_internal.h"

namespace av {

// Pattern value of one antenna type at scaled radius r (r > 0):
//   airy, unblocked:  2 J1(r) / r
//   airy, blocked:    (2 J1(r) / r - 2 e J1(r e) / r) / (1 - e^2),  e = b / d
//   casa, blocked:    (A 2 J1(r) / r - 2 J1(r L) / (r L)) / (A - 1),
//                     L = d / b, A = L^2
__device__ inline double aperture_value(double r, double dish, double blockage, int casa) {
  if (blockage == 0.0) return 2.0 * j1(r) / r;
  if (!casa) {
    const double e = blockage / dish;
    return (2.0 * j1(r) / r - 2.0 * e * j1(r * e) / r) / (1.0 - e * e);
  }
  const double lr = dish / blockage;
  const double ar = lr * lr;
  return (ar * 2.0 * j1(r) / r - 2.0 * j1(r * lr) / (r * lr)) / (ar - 1.0);
}

__global__ __launch_bounds__(256) void k_aperture_patterns(
    const double *__restrict__ kf, int nf, const double *__restrict__ dish,
    const double *__restrict__ blockage, int nant, int n0, int n1, int c0, int c1, double cell0,
    double cell1, int ipower, int casa, double *__restrict__ out) {
  const int64_t npix = (int64_t)n0 * n1;
  const int64_t total = npix * nf * nant;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t af = i / npix, pix = i - af * npix;
    const int a = (int)(af / nf), f = (int)(af - (int64_t)a * nf);
    const int ix = (int)(pix / n1), iy = (int)(pix - (int64_t)ix * n1);
    double v = 1.0;  // the centre value (make_pb_symmetric.py:205-206)
    if (ix != c0 || iy != c1) {
      const double x = (double)(ix - c0) * cell0, y = (double)(iy - c1) * cell1;
      const double r = sqrt(x * x + y * y) * kf[f] * (dish[a] / 2.0);
      v = aperture_value(r, dish[a], blockage[a], casa);
      if (ipower == 2) v = v * v;
    }
    out[i] = v;
  }
}

// exp(1j (x pix0 + y pix1)) over x in [-cx, nx - cx), y in [-cy, ny - cy),
// one [nx][ny] complex128 plane per field
__global__ __launch_bounds__(256) void k_phase_gradient(const double *__restrict__ pix,
                                                        int nfield, int nx, int ny, int cx,
                                                        int cy, double *__restrict__ out) {
  const int64_t n = (int64_t)nfield * nx * ny;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t fld = i / ((int64_t)nx * ny), rem = i - fld * nx * ny;
    const int ix = (int)(rem / ny), iy = (int)(rem - (int64_t)ix * ny);
    const double t = (double)(ix - cx) * pix[2 * fld] + (double)(iy - cy) * pix[2 * fld + 1];
    out[2 * i] = cos(t);
    out[2 * i + 1] = sin(t);
  }
}

static unsigned grid_for(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536));
}

}  // namespace av

using namespace av;

extern "C" {

int av_gcf_aperture_patterns(const double *kf, int32_t n_freq, const double *dish,
                             const double *blockage, int32_t n_ant, int32_t n0, int32_t n1,
                             int32_t c0, int32_t c1, double cell0, double cell1, int32_t ipower,
                             int32_t casa, double *out, void *stream) {
  if (!kf || !dish || !blockage || !out)
    return fail(AV_ERR_INVALID, "av_gcf_aperture_patterns: NULL argument");
  if (n_freq < 1 || n_ant < 1 || n0 < 1 || n1 < 1 || (ipower != 1 && ipower != 2))
    return fail(AV_ERR_INVALID, "av_gcf_aperture_patterns: bad sizes or ipower");
  const int64_t total = (int64_t)n0 * n1 * n_freq * n_ant;
  hipLaunchKernelGGL(k_aperture_patterns, dim3(grid_for(total)), dim3(256), 0,
                     (hipStream_t)stream, kf, n_freq, dish, blockage, n_ant, n0, n1, c0, c1,
                     cell0, cell1, ipower, casa ? 1 : 0, out);
  AV_LAUNCH_CHECK("k_aperture_patterns");
  return AV_OK;
}

int av_gcf_phase_gradient(const double *pix, int32_t n_field, int32_t nx, int32_t ny,
                          int32_t cx, int32_t cy, double *out, void *stream) {
  if (!pix || !out) return fail(AV_ERR_INVALID, "av_gcf_phase_gradient: NULL argument");
  if (n_field < 1 || nx < 1 || ny < 1)
    return fail(AV_ERR_INVALID, "av_gcf_phase_gradient: bad sizes");
  const int64_t total = (int64_t)n_field * nx * ny;
  hipLaunchKernelGGL(k_phase_gradient, dim3(grid_for(total)), dim3(256), 0, (hipStream_t)stream,
                     pix, n_field, nx, ny, cx, cy, out);
  AV_LAUNCH_CHECK("k_phase_gradient");
  return AV_OK;
}

}  // extern "C"
