// hogbom.hip -- Hogbom CLEAN on the GPU (replaces the _hogbom_ext extension
// of core/imaging/deconvolvers/hogbom: src/hclean.cpp maximg :51-81 and
// clean :106-210).
//
// CLEAN is a sequence of dependent iterations -- find the residual's peak,
// subtract gain * peak * PSF shifted to it -- so it runs as ONE cooperative
// launch that keeps every workgroup resident and separates iterations with
// a grid-wide barrier, instead of two launches per iteration.  Workgroup g
// owns a band of image rows.  In every iteration it
//   1. reduces the peak candidate of its band (largest |residual| inside the
//      clean box and the mask; ties go to the lowest row-major index, as the
//      reference's strict > scan does) and publishes it;
//   2. passes the grid barrier, then reduces all bands' candidates itself,
//      so every workgroup reaches the same peak and the same stop decision;
//   3. subtracts gain * peak * PSF over its band and, in the same pass,
//      finds its band's candidate for the next iteration.
// Candidates are double-buffered by iteration parity, so one barrier per
// iteration is enough.  All arithmetic is in the image's type, with separate
// roundings (no contraction), as the reference's C++ does.
#include <algorithm>
#include <type_traits>

#include "avgrid_internal.h"

namespace av {

constexpr int kHogThreads = 1024;
constexpr int kHogMaxGroups = 1024;
constexpr uint32_t kHogSpinLimit = 1u << 25;  // s_sleep rounds before a barrier gives up

// work buffer layout (bytes): header, then two candidate arrays
struct HogHeader {
  uint32_t barrier;  // arrivals, monotonically increasing within a launch
  uint32_t error;    // set when a barrier wait exceeded its spin limit
  int32_t iter, px, py, reason;  // final state; reason 0 = stopped / done, 1 = report
  double maxval;                 // signed peak of the last iteration examined
};
constexpr size_t kHogHeaderBytes = 64;
template <typename T>
struct HogCand {
  T absval, value;
  int32_t idx, pad;
};
constexpr size_t kHogWorkBytes = kHogHeaderBytes + 2 * kHogMaxGroups * sizeof(HogCand<double>);

template <typename T>
__device__ inline bool hog_better(T a, int ia, T b, int ib) {
  return a > b || (a == b && a > (T)0 && ia < ib);
}

// Workgroup reduction of (absval, value, idx); the result is valid in every
// thread.  sv/sx/si are LDS slots for one entry per wave.
template <typename T>
__device__ inline void hog_reduce_wg(T &a, T &v, int &i, T *sa, T *sv, int *si) {
  for (int off = 32; off >= 1; off >>= 1) {
    const T oa = __shfl_xor(a, off, 64);
    const T ov = __shfl_xor(v, off, 64);
    const int oi = __shfl_xor(i, off, 64);
    if (hog_better(oa, oi, a, i)) {
      a = oa;
      v = ov;
      i = oi;
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    sa[wave] = a;
    sv[wave] = v;
    si[wave] = i;
  }
  __syncthreads();
  a = sa[0];
  v = sv[0];
  i = si[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
    if (hog_better(sa[w], si[w], a, i)) {
      a = sa[w];
      v = sv[w];
      i = si[w];
    }
  __syncthreads();
}

template <typename T>
__device__ inline T hog_pow2(T e) {
  if constexpr (sizeof(T) == 4)
    return powf(2.0f, e);
  else
    return pow(2.0, e);
}

// Grid-wide barrier number k of this launch (every workgroup arrives once
// per barrier).  Returns false when the wait exceeded its spin limit (some
// workgroup never arrived); the caller then stops and the host reports it.
__device__ inline bool hog_grid_barrier(HogHeader *h, uint32_t k, uint32_t groups, bool *lds_ok) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(&h->barrier, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t target = (k + 1) * groups;
    uint32_t spins = 0;
    bool ok = true;
    while (__hip_atomic_load(&h->barrier, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(2);
      if (++spins > kHogSpinLimit ||
          __hip_atomic_load(&h->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        ok = false;
        __hip_atomic_store(&h->error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
    }
    *lds_ok = ok;
  }
  __syncthreads();
  return *lds_ok;
}

template <typename T>
__global__ __launch_bounds__(kHogThreads) void k_hogbom(
    T *__restrict__ model, T *__restrict__ resid, const T *__restrict__ psf,
    const T *__restrict__ mask, int nx, int ny, int xbeg, int xend, int ybeg, int yend,
    int niter, int siter, int start, int report_cycle, int skip_report, T gain, T thres,
    T cspeedup, int px0, int py0, char *work) {
  HogHeader *h = (HogHeader *)work;
  HogCand<T> *cand = (HogCand<T> *)(work + kHogHeaderBytes);
  __shared__ T s_a[kHogThreads / 64], s_v[kHogThreads / 64];
  __shared__ int s_i[kHogThreads / 64];
  __shared__ bool s_ok;
  const uint32_t G = gridDim.x, g = blockIdx.x;
  const int rpw = (ny + (int)G - 1) / (int)G;
  const int r0 = min(ny, (int)g * rpw), r1 = min(ny, r0 + rpw);
  const int npix = (r1 - r0) * nx;
  const int tid = threadIdx.x;
  const bool domask = mask != nullptr;

  // The band's pixels q = tid + j * B, j = 0, 1, ... (row r0 + q / nx,
  // column q % nx) are walked U at a time (a 2048-row image on 256 bands of
  // 1024 threads gives each thread 16 pixels), the row and column
  // advanced incrementally (one division per thread, not per pixel), and the
  // loads of each group of U pixels issued before their use.
  constexpr int U = sizeof(T) == 4 ? 16 : 8;
  const int B = (int)blockDim.x;
  const int drow = B / nx, dcol = B - drow * nx;
  const int iy_t = r0 + tid / nx, ix_t = tid - (tid / nx) * nx;
  auto in_box = [&](int ix, int iy) {
    return ix >= xbeg && ix < xend && iy >= ybeg && iy < yend &&
           (!domask || mask[iy * nx + ix] > (T)0.5);
  };
  // one pass over the band: subtract pv * PSF shifted to (px, py) inside the
  // window [x1, x2] x [y1, y2] (when SUB), then this thread's candidate for
  // the peak of the clean box
  auto band_pass = [&](auto sub_tag, T pv, int px, int py, int x1, int x2, int y1, int y2,
                       T &ba, T &bv, int &bi) {
    constexpr bool SUB = decltype(sub_tag)::value;
    ba = (T)0;
    bv = (T)0;
    bi = 0x7FFFFFFF;
    int ix = ix_t, iy = iy_t;
    for (int q0 = tid; q0 < npix; q0 += U * B) {
      int cx[U], cy[U];
      bool ok[U], win[U];
      T r[U], pp[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        cx[u] = ix;
        cy[u] = iy;
        ok[u] = q0 + u * B < npix;
        win[u] = false;
        pp[u] = (T)0;
        r[u] = ok[u] ? resid[iy * nx + ix] : (T)0;
        if constexpr (SUB) {
          const int qx = nx / 2 + ix - px, qy = ny / 2 + iy - py;
          win[u] = ok[u] && iy >= y1 && iy <= y2 && ix >= x1 && ix <= x2 && qx >= 0 &&
                   qx < nx && qy >= 0 && qy < ny;
          pp[u] = win[u] ? psf[qy * nx + qx] : (T)0;
        }
        ix += dcol;
        iy += drow;
        if (ix >= nx) {
          ix -= nx;
          ++iy;
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int idx = cy[u] * nx + cx[u];
        if constexpr (SUB) {
          if (win[u]) {
            const T prod = pv * pp[u];
            r[u] = r[u] - prod;
            resid[idx] = r[u];
          }
        }
        if (ok[u] && in_box(cx[u], cy[u])) {
          const T a = fabs(r[u]);
          if (hog_better(a, idx, ba, bi)) {
            ba = a;
            bv = r[u];
            bi = idx;
          }
        }
      }
    }
  };
  T ba, bv;
  int bi;
  band_pass(std::false_type{}, (T)0, 0, 0, 0, -1, 0, -1, ba, bv, bi);

  int px = px0, py = py0;
  T maxval = (T)0;
  int iter = start, reason = 0;
  bool ok = true;
  for (uint32_t k = 0;; ++k, ++iter) {
    if (iter >= niter) break;
    // ---- publish this band's candidate, then reduce all bands
    hog_reduce_wg(ba, bv, bi, s_a, s_v, s_i);
    HogCand<T> *cur = cand + (size_t)(iter & 1) * kHogMaxGroups;
    if (tid == 0) cur[g] = HogCand<T>{ba, bv, bi, 0};
    if (!(ok = hog_grid_barrier(h, k, G, &s_ok))) break;
    T ga = (T)0, gv = (T)0;
    int gi = 0x7FFFFFFF;
    for (uint32_t j = tid; j < G; j += blockDim.x) {
      const T a = __hip_atomic_load(&cur[j].absval, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const T v = __hip_atomic_load(&cur[j].value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const int i = __hip_atomic_load(&cur[j].idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (hog_better(a, i, ga, gi)) {
        ga = a;
        gv = v;
        gi = i;
      }
    }
    hog_reduce_wg(ga, gv, gi, s_a, s_v, s_i);
    // ---- the peak (hclean.cpp:127-145): a box with no |r| > 0 keeps the
    // previous position, and the value there is read back
    if (ga > (T)0) {
      py = gi / nx;
      px = gi - py * nx;
      maxval = gv;
    } else {
      maxval = __hip_atomic_load(&resid[py * nx + px], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const T cthres = cspeedup > (T)0 ? thres * hog_pow2((T)(iter - siter) / cspeedup) : thres;
    if (ga < cthres) break;
    if (report_cycle > 0 && iter != skip_report &&
        (iter == siter || (iter % report_cycle) == 1)) {
      reason = 1;
      break;
    }
    // ---- add the component and subtract the shifted PSF (hclean.cpp:166-196)
    const T pv = gain * maxval;
    if (tid == 0 && py >= r0 && py < r1) model[py * nx + px] += pv;
    const int x1 = max(0, px - nx / 2), y1 = max(0, py - ny / 2);
    const int x2 = min(nx - 1, px + nx / 2 - 1), y2 = min(ny - 1, py + ny / 2 - 1);
    band_pass(std::true_type{}, pv, px, py, x1, x2, y1, y2, ba, bv, bi);
  }
  if (g == 0 && tid == 0) {
    h->iter = iter > niter ? niter : iter;
    h->px = px;
    h->py = py;
    h->reason = ok ? reason : -1;
    h->maxval = (double)maxval;
  }
}

// maxval per workgroup slice, then one reduction: min and max of the
// (masked) image starting from +-1e20 (hclean.cpp:51-81).
template <typename T>
__global__ __launch_bounds__(256) void k_maximg(const T *__restrict__ img,
                                                const T *__restrict__ mask, int64_t n, T *out) {
  __shared__ T smin[4], smax[4];
  T lo = (T)1e20, hi = (T)-1e20;
  for (int64_t k = threadIdx.x; k < n; k += blockDim.x) {
    if (!mask || mask[k] > (T)0.5) {
      const T w = img[k];
      if (w > hi) hi = w;
      if (w < lo) lo = w;
    }
  }
  for (int off = 32; off >= 1; off >>= 1) {
    const T ol = __shfl_xor(lo, off, 64), oh = __shfl_xor(hi, off, 64);
    lo = ol < lo ? ol : lo;
    hi = oh > hi ? oh : hi;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    smin[wave] = lo;
    smax[wave] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w) {
      lo = smin[w] < lo ? smin[w] : lo;
      hi = smax[w] > hi ? smax[w] : hi;
    }
    out[0] = lo;
    out[1] = hi;
  }
}

__global__ void k_hog_clear(uint32_t *w, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) w[i] = 0u;
}

template <typename T>
static int launch_hogbom(void *model, void *resid, const void *psf, const void *mask, int nx,
                         int ny, int xbeg, int xend, int ybeg, int yend, int niter, int siter,
                         int start, int report_cycle, int skip_report, double gain, double thres,
                         double cspeedup, int px0, int py0, void *work, hipStream_t s) {
  const DeviceProps &prop = device_props();
  if (prop.error) return hip_fail(prop.error, "av_hogbom_clean: device properties");
  int per_cu = 0;
  AV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_hogbom<T>, kHogThreads, 0));
  int groups = std::min(std::min(prop.num_cus * std::max(per_cu, 0), kHogMaxGroups), ny);
  if (groups < 1) return fail(AV_ERR_HIP, "av_hogbom_clean: no resident workgroup possible");
  groups = std::min(groups, prop.num_cus);  // one band per CU
  hipLaunchKernelGGL(k_hog_clear, dim3(1), dim3(64), 0, s, (uint32_t *)work,
                     (int)(kHogHeaderBytes / 4));
  AV_LAUNCH_CHECK("k_hog_clear");
  T *m = (T *)model, *r = (T *)resid;
  const T *p = (const T *)psf, *k = (const T *)mask;
  T tg = (T)gain, tt = (T)thres, tc = (T)cspeedup;
  char *w = (char *)work;
  void *args[] = {&m,     &r,           &p,          &k,  &nx, &ny, &xbeg, &xend,
                  &ybeg,  &yend,        &niter,      &siter, &start, &report_cycle,
                  &skip_report, &tg, &tt, &tc, &px0, &py0, &w};
  AV_HIP(hipLaunchCooperativeKernel((const void *)k_hogbom<T>, dim3((unsigned)groups),
                                    dim3(kHogThreads), args, 0, s));
  return AV_OK;
}

}  // namespace av

using namespace av;

extern "C" {

int av_hogbom_work_bytes(int64_t *bytes) {
  if (!bytes) return fail(AV_ERR_INVALID, "av_hogbom_work_bytes: NULL argument");
  *bytes = (int64_t)kHogWorkBytes;
  return AV_OK;
}

int av_hogbom_maximg(const void *image, const void *mask, int32_t nx, int32_t ny, int32_t dtype,
                     void *out, void *stream) {
  if (!image || !out || nx < 1 || ny < 1)
    return fail(AV_ERR_INVALID, "av_hogbom_maximg: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = (int64_t)nx * ny;
  if (dtype == AV_F32)
    hipLaunchKernelGGL(k_maximg<float>, dim3(1), dim3(256), 0, s, (const float *)image,
                       (const float *)mask, n, (float *)out);
  else if (dtype == AV_F64)
    hipLaunchKernelGGL(k_maximg<double>, dim3(1), dim3(256), 0, s, (const double *)image,
                       (const double *)mask, n, (double *)out);
  else
    return fail(AV_ERR_INVALID, "av_hogbom_maximg: dtype must be AV_F32 or AV_F64");
  AV_LAUNCH_CHECK("k_maximg");
  return AV_OK;
}

int av_hogbom_clean(void *model, void *residual, const void *psf, const void *mask, int32_t nx,
                    int32_t ny, int32_t xbeg, int32_t xend, int32_t ybeg, int32_t yend,
                    int32_t niter, int32_t siter, int32_t start, int32_t report_cycle,
                    int32_t skip_report, double gain, double thres, double cspeedup,
                    int32_t px0, int32_t py0, int32_t dtype, void *work, void *stream) {
  if (!model || !residual || !psf || !work)
    return fail(AV_ERR_INVALID, "av_hogbom_clean: NULL argument");
  if (nx < 1 || ny < 1 || (int64_t)nx * ny >= ((int64_t)1 << 31))
    return fail(AV_ERR_INVALID, "av_hogbom_clean: image size out of range");
  if (!(0 <= xbeg && xbeg < xend && xend <= nx && 0 <= ybeg && ybeg < yend && yend <= ny))
    return fail(AV_ERR_INVALID, "av_hogbom_clean: clean box outside the image");
  if (px0 < 0 || px0 >= nx || py0 < 0 || py0 >= ny)
    return fail(AV_ERR_INVALID, "av_hogbom_clean: previous peak outside the image");
  hipStream_t s = (hipStream_t)stream;
  if (dtype == AV_F32)
    return launch_hogbom<float>(model, residual, psf, mask, nx, ny, xbeg, xend, ybeg, yend,
                                niter, siter, start, report_cycle, skip_report, gain, thres,
                                cspeedup, px0, py0, work, s);
  if (dtype == AV_F64)
    return launch_hogbom<double>(model, residual, psf, mask, nx, ny, xbeg, xend, ybeg, yend,
                                 niter, siter, start, report_cycle, skip_report, gain, thres,
                                 cspeedup, px0, py0, work, s);
  return fail(AV_ERR_INVALID, "av_hogbom_clean: dtype must be AV_F32 or AV_F64");
}

}  // extern "C"
