// post.hip -- the passes around the FFT of grid2image_spheroid_ms4
// (core/imaging/imaging_utils/standard_grid.py:693-708): ifftshift of the
// grid before the inverse FFT, and fftshift + real part + taper correction
// + N^2/sumwt normalisation after it, fused into one pass.
#include "avgrid_internal.h"

namespace av {

template <typename R, typename O>
__global__ void k_ifftshift(const R *in, O *out, int64_t n_planes, int n_u, int n_v) {
  // np.fft.ifftshift: out[k] = in[(k + n//2) % n] on each of the two axes
  const int64_t plane = (int64_t)n_u * n_v;
  const int64_t total = n_planes * plane;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = i / plane;
    const int64_t c = i - p * plane;
    const int ku = (int)(c / n_v), kv = (int)(c - (int64_t)ku * n_v);
    int su = ku + n_u / 2;
    if (su >= n_u) su -= n_u;
    int sv = kv + n_v / 2;
    if (sv >= n_v) sv -= n_v;
    const R *src = in + (p * plane + (int64_t)su * n_v + sv) * 2;
    out[i * 2] = (O)src[0];
    out[i * 2 + 1] = (O)src[1];
  }
}

template <typename R, typename O>
__global__ void k_normalise(const R *fft, const double *sumwt, const double *tu,
                            const double *tv, int64_t n_planes, int n_u, int n_v, O *out) {
  // np.fft.fftshift: out[k] = y[(k + n - n//2) % n]; then
  // real / corrTerm * nx * ny / sumwt, left to right, in FP64.
  const int64_t plane = (int64_t)n_u * n_v;
  const int64_t total = n_planes * plane;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = i / plane;
    const int64_t c = i - p * plane;
    const int ku = (int)(c / n_v), kv = (int)(c - (int64_t)ku * n_v);
    int su = ku + n_u - n_u / 2;
    if (su >= n_u) su -= n_u;
    int sv = kv + n_v - n_v / 2;
    if (sv >= n_v) sv -= n_v;
    const double re = (double)fft[(p * plane + (int64_t)su * n_v + sv) * 2];
    const double corr = tu[ku] * tv[kv];
    out[i] = (O)(re / corr * (double)n_u * (double)n_v / sumwt[p]);
  }
}

// The half spectrum whose complex-to-real inverse FFT is the shifted real
// image: for even n_u and n_v, with X = ifftshift(in),
//   real(fftshift(ifft2(X))) = irfft2(Y),
//   Y[ku][kv] = (-1)^(ku + kv) * (X[ku][kv] + conj(X[-ku][-kv])) / 2,
// kv in [0, n_v / 2], indices mod n.  The bracket is X's Hermitian part
// (whose inverse transform is the real part of X's), and the sign moves the
// result by half a period on each axis, which is fftshift for even n.  Y is
// exactly Hermitian on its self-conjugate columns, as the complex-to-real
// transform assumes.  out is complex FP64 [n_planes][n_u][n_v / 2 + 1].
template <typename R>
__global__ void k_ifftshift_hermitian(const R *in, double *out, int64_t n_planes, int n_u,
                                      int n_v) {
  const int nh = n_v / 2 + 1;
  const int64_t hplane = (int64_t)n_u * nh, plane = (int64_t)n_u * n_v;
  const int64_t total = n_planes * hplane;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = i / hplane;
    const int64_t c = i - p * hplane;
    const int ku = (int)(c / nh), kv = (int)(c - (int64_t)ku * nh);
    // X[k] = in[(k + n / 2) mod n]; X[-k] = in[(n / 2 - k) mod n]
    int su = ku + n_u / 2;
    su -= su >= n_u ? n_u : 0;
    int sv = kv + n_v / 2;
    sv -= sv >= n_v ? n_v : 0;
    int mu = n_u / 2 - ku;
    mu += mu < 0 ? n_u : 0;
    int mv = n_v / 2 - kv;
    mv += mv < 0 ? n_v : 0;
    const R *a = in + (p * plane + (int64_t)su * n_v + sv) * 2;
    const R *b = in + (p * plane + (int64_t)mu * n_v + mv) * 2;
    const double sgn = ((ku + kv) & 1) ? -0.5 : 0.5;
    out[i * 2] = sgn * ((double)a[0] + (double)b[0]);
    out[i * 2 + 1] = sgn * ((double)a[1] - (double)b[1]);
  }
}

// The real image of k_ifftshift_hermitian's transform (already shifted):
// image / corrTerm * nx * ny / sumwt, left to right, in FP64, as k_normalise.
template <typename O>
__global__ void k_normalise_real(const double *img, const double *sumwt, const double *tu,
                                 const double *tv, int64_t n_planes, int n_u, int n_v, O *out) {
  const int64_t plane = (int64_t)n_u * n_v;
  const int64_t total = n_planes * plane;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = i / plane;
    const int64_t c = i - p * plane;
    const int ku = (int)(c / n_v), kv = (int)(c - (int64_t)ku * n_v);
    const double corr = tu[ku] * tv[kv];
    out[i] = (O)(img[i] / corr * (double)n_u * (double)n_v / sumwt[p]);
  }
}

// The whole spectrum of a real image's FFT from its half [n_u][n_v / 2 + 1]
// (a real-to-complex FFT's output): F[ku][kv] = conj(F[-ku][n_v - kv]) for
// kv > n_v / 2, indices mod n.  With `sign`, each value is multiplied by
// (-1)^(ku + kv), the output side of fftshift(fft2(ifftshift(x))) =
// S * fft2(S * x) for axes that are multiples of 4.
template <typename R>
__global__ void k_hermitian_expand(const R *half, R *out, int64_t n_planes, int n_u, int n_v,
                                   int sign) {
  const int nh = n_v / 2 + 1;
  const int64_t plane = (int64_t)n_u * n_v;
  const int64_t total = n_planes * plane;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = i / plane;
    const int64_t c = i - p * plane;
    const int ku = (int)(c / n_v), kv = (int)(c - (int64_t)ku * n_v);
    const bool direct = kv < nh;
    const int su = direct ? ku : (ku == 0 ? 0 : n_u - ku);
    const int sv = direct ? kv : n_v - kv;
    const R *src = half + ((p * n_u + su) * nh + sv) * 2;
    R re = src[0], im = direct ? src[1] : -src[1];
    if (sign && ((ku + kv) & 1)) {
      re = -re;
      im = -im;
    }
    out[i * 2] = re;
    out[i * 2 + 1] = im;
  }
}

// out[k] = scale * in[(k + shift) % n] (per axis), optionally divided by the
// taper at the source or the destination index, real or complex in/out.
template <typename R, typename O, bool IN_REAL, bool OUT_REAL>
__global__ void k_shift_taper(const R *in, O *out, int64_t n_planes, int n_u, int n_v,
                              int su, int sv, const double *tu, const double *tv, int taper,
                              double scale) {
  const int64_t plane = (int64_t)n_u * n_v;
  const int64_t total = n_planes * plane;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = i / plane;
    const int64_t c = i - p * plane;
    const int ku = (int)(c / n_v), kv = (int)(c - (int64_t)ku * n_v);
    int iu = ku + su;
    iu -= (iu >= n_u) ? n_u : 0;
    int iv = kv + sv;
    iv -= (iv >= n_v) ? n_v : 0;
    const int64_t src = p * plane + (int64_t)iu * n_v + iv;
    double re = IN_REAL ? (double)in[src] : (double)in[2 * src];
    double im = IN_REAL ? 0.0 : (double)in[2 * src + 1];
    double f = scale;
    if (taper == 1) f /= tu[iu] * tv[iv];
    if (taper == 2) f /= tu[ku] * tv[kv];
    if (OUT_REAL) {
      out[i] = (O)(re * f);
    } else {
      out[2 * i] = (O)(re * f);
      out[2 * i + 1] = (O)(im * f);
    }
  }
}

static unsigned grid_blocks(int64_t total) {
  int64_t b = (total + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (unsigned)b;
}

}  // namespace av

using namespace av;

extern "C" {

int av_ifftshift(const void *in, void *out, int64_t n_planes, int32_t n_u, int32_t n_v,
                 int32_t dtype, int32_t out_dtype, void *stream) {
  if (!in || !out || n_planes < 0 || n_u < 1 || n_v < 1)
    return fail(AV_ERR_INVALID, "av_ifftshift: bad arguments");
  if (in == out) return fail(AV_ERR_INVALID, "av_ifftshift: in and out must differ");
  const int64_t total = n_planes * n_u * n_v;
  if (total == 0) return AV_OK;
  hipStream_t s = (hipStream_t)stream;
  dim3 g(grid_blocks(total)), b(256);
  if (dtype == AV_F32 && out_dtype == AV_F32)
    hipLaunchKernelGGL((k_ifftshift<float, float>), g, b, 0, s, (const float *)in, (float *)out,
                       n_planes, n_u, n_v);
  else if (dtype == AV_F32 && out_dtype == AV_F64)
    hipLaunchKernelGGL((k_ifftshift<float, double>), g, b, 0, s, (const float *)in,
                       (double *)out, n_planes, n_u, n_v);
  else if (dtype == AV_F64 && out_dtype == AV_F64)
    hipLaunchKernelGGL((k_ifftshift<double, double>), g, b, 0, s, (const double *)in,
                       (double *)out, n_planes, n_u, n_v);
  else if (dtype == AV_F64 && out_dtype == AV_F32)
    hipLaunchKernelGGL((k_ifftshift<double, float>), g, b, 0, s, (const double *)in,
                       (float *)out, n_planes, n_u, n_v);
  else
    return fail(AV_ERR_INVALID, "av_ifftshift: bad dtype");
  AV_LAUNCH_CHECK("k_ifftshift");
  return AV_OK;
}

int av_shift_taper(const void *in, void *out, int64_t n_planes, int32_t n_u, int32_t n_v,
                   int32_t shift_u, int32_t shift_v, const double *taper_u,
                   const double *taper_v, int32_t taper, double scale, int32_t in_dtype,
                   int32_t in_real, int32_t out_dtype, int32_t out_real, void *stream) {
  if (!in || !out || in == out || n_planes < 0 || n_u < 1 || n_v < 1 || shift_u < 0 ||
      shift_u >= n_u || shift_v < 0 || shift_v >= n_v || taper < 0 || taper > 2 ||
      (taper && (!taper_u || !taper_v)))
    return fail(AV_ERR_INVALID, "av_shift_taper: bad arguments");
  const int64_t total = n_planes * n_u * n_v;
  if (total == 0) return AV_OK;
  hipStream_t s = (hipStream_t)stream;
  dim3 g(grid_blocks(total)), b(256);
#define AV_ST(R, O, IR, OR_)                                                                  \
  hipLaunchKernelGGL((k_shift_taper<R, O, IR, OR_>), g, b, 0, s, (const R *)in, (O *)out,      \
                     n_planes, n_u, n_v, shift_u, shift_v, taper_u, taper_v, taper, scale)
  const bool f32i = in_dtype == AV_F32, f32o = out_dtype == AV_F32;
  if ((in_dtype != AV_F32 && in_dtype != AV_F64) || (out_dtype != AV_F32 && out_dtype != AV_F64))
    return fail(AV_ERR_INVALID, "av_shift_taper: bad dtype");
  if (f32i && f32o) {
    if (in_real && out_real) AV_ST(float, float, true, true);
    else if (in_real) AV_ST(float, float, true, false);
    else if (out_real) AV_ST(float, float, false, true);
    else AV_ST(float, float, false, false);
  } else if (f32i) {
    if (in_real && out_real) AV_ST(float, double, true, true);
    else if (in_real) AV_ST(float, double, true, false);
    else if (out_real) AV_ST(float, double, false, true);
    else AV_ST(float, double, false, false);
  } else if (f32o) {
    if (in_real && out_real) AV_ST(double, float, true, true);
    else if (in_real) AV_ST(double, float, true, false);
    else if (out_real) AV_ST(double, float, false, true);
    else AV_ST(double, float, false, false);
  } else {
    if (in_real && out_real) AV_ST(double, double, true, true);
    else if (in_real) AV_ST(double, double, true, false);
    else if (out_real) AV_ST(double, double, false, true);
    else AV_ST(double, double, false, false);
  }
#undef AV_ST
  AV_LAUNCH_CHECK("k_shift_taper");
  return AV_OK;
}

int av_ifftshift_hermitian(const void *in, void *out, int64_t n_planes, int32_t n_u,
                           int32_t n_v, int32_t dtype, void *stream) {
  if (!in || !out || n_planes < 0 || n_u < 2 || n_v < 2)
    return fail(AV_ERR_INVALID, "av_ifftshift_hermitian: bad arguments");
  if ((n_u | n_v) & 1)
    return fail(AV_ERR_UNSUPPORTED, "av_ifftshift_hermitian: n_u and n_v must be even");
  const int64_t total = n_planes * n_u * (int64_t)(n_v / 2 + 1);
  if (total == 0) return AV_OK;
  hipStream_t s = (hipStream_t)stream;
  dim3 g(grid_blocks(total)), b(256);
  if (dtype == AV_F32)
    hipLaunchKernelGGL(k_ifftshift_hermitian<float>, g, b, 0, s, (const float *)in,
                       (double *)out, n_planes, n_u, n_v);
  else if (dtype == AV_F64)
    hipLaunchKernelGGL(k_ifftshift_hermitian<double>, g, b, 0, s, (const double *)in,
                       (double *)out, n_planes, n_u, n_v);
  else
    return fail(AV_ERR_INVALID, "av_ifftshift_hermitian: bad dtype");
  AV_LAUNCH_CHECK("k_ifftshift_hermitian");
  return AV_OK;
}

int av_normalise_real(const double *image, const double *sumwt, const double *taper_u,
                      const double *taper_v, int64_t n_planes, int32_t n_u, int32_t n_v,
                      void *out, int32_t out_dtype, void *stream) {
  if (!image || !sumwt || !taper_u || !taper_v || !out || n_planes < 0 || n_u < 1 || n_v < 1)
    return fail(AV_ERR_INVALID, "av_normalise_real: bad arguments");
  const int64_t total = n_planes * n_u * n_v;
  if (total == 0) return AV_OK;
  hipStream_t s = (hipStream_t)stream;
  dim3 g(grid_blocks(total)), b(256);
  if (out_dtype == AV_F64)
    hipLaunchKernelGGL(k_normalise_real<double>, g, b, 0, s, image, sumwt, taper_u, taper_v,
                       n_planes, n_u, n_v, (double *)out);
  else if (out_dtype == AV_F32)
    hipLaunchKernelGGL(k_normalise_real<float>, g, b, 0, s, image, sumwt, taper_u, taper_v,
                       n_planes, n_u, n_v, (float *)out);
  else
    return fail(AV_ERR_INVALID, "av_normalise_real: bad dtype");
  AV_LAUNCH_CHECK("k_normalise_real");
  return AV_OK;
}

int av_hermitian_expand(const void *half, void *out, int64_t n_planes, int32_t n_u, int32_t n_v,
                        int32_t sign, int32_t dtype, void *stream) {
  if (!half || !out || n_planes < 0 || n_u < 1 || n_v < 1)
    return fail(AV_ERR_INVALID, "av_hermitian_expand: bad arguments");
  if (sign && ((n_u | n_v) & 3))
    return fail(AV_ERR_INVALID, "av_hermitian_expand: sign needs axes that are multiples of 4");
  const int64_t total = n_planes * n_u * (int64_t)n_v;
  if (total == 0) return AV_OK;
  hipStream_t s = (hipStream_t)stream;
  dim3 g(grid_blocks(total)), b(256);
  if (dtype == AV_F32)
    hipLaunchKernelGGL(k_hermitian_expand<float>, g, b, 0, s, (const float *)half, (float *)out,
                       n_planes, n_u, n_v, sign);
  else if (dtype == AV_F64)
    hipLaunchKernelGGL(k_hermitian_expand<double>, g, b, 0, s, (const double *)half,
                       (double *)out, n_planes, n_u, n_v, sign);
  else
    return fail(AV_ERR_INVALID, "av_hermitian_expand: bad dtype");
  AV_LAUNCH_CHECK("k_hermitian_expand");
  return AV_OK;
}

int av_normalise(const void *fft_planes, const double *sumwt, const double *taper_u,
                 const double *taper_v, int64_t n_planes, int32_t n_u, int32_t n_v,
                 int32_t dtype, void *out, int32_t out_dtype, void *stream) {
  if (!fft_planes || !sumwt || !taper_u || !taper_v || !out || n_planes < 0 || n_u < 1 || n_v < 1)
    return fail(AV_ERR_INVALID, "av_normalise: bad arguments");
  const int64_t total = n_planes * n_u * n_v;
  if (total == 0) return AV_OK;
  hipStream_t s = (hipStream_t)stream;
  dim3 g(grid_blocks(total)), b(256);
  if (dtype == AV_F32 && out_dtype == AV_F64)
    hipLaunchKernelGGL((k_normalise<float, double>), g, b, 0, s, (const float *)fft_planes, sumwt,
                       taper_u, taper_v, n_planes, n_u, n_v, (double *)out);
  else if (dtype == AV_F32 && out_dtype == AV_F32)
    hipLaunchKernelGGL((k_normalise<float, float>), g, b, 0, s, (const float *)fft_planes, sumwt,
                       taper_u, taper_v, n_planes, n_u, n_v, (float *)out);
  else if (dtype == AV_F64 && out_dtype == AV_F64)
    hipLaunchKernelGGL((k_normalise<double, double>), g, b, 0, s, (const double *)fft_planes,
                       sumwt, taper_u, taper_v, n_planes, n_u, n_v, (double *)out);
  else if (dtype == AV_F64 && out_dtype == AV_F32)
    hipLaunchKernelGGL((k_normalise<double, float>), g, b, 0, s, (const double *)fft_planes,
                       sumwt, taper_u, taper_v, n_planes, n_u, n_v, (float *)out);
  else
    return fail(AV_ERR_INVALID, "av_normalise: bad dtype");
  AV_LAUNCH_CHECK("k_normalise");
  return AV_OK;
}

}  // extern "C"
