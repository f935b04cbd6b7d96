// common.hip -- error reporting, device info, fills, scans and the
// prolate-spheroidal table kernels of libavgrid.so.
#include <cstdio>
#include <mutex>

#include "avgrid_internal.h"

namespace av {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg) {
  set_error(msg);
  return code;
}

int hip_fail(hipError_t e, const char *what) {
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return e == hipErrorOutOfMemory ? AV_ERR_NOMEM : AV_ERR_HIP;
}

const DeviceProps &device_props() {
  static std::mutex mu;
  static DeviceProps cache[64];
  static bool have[64] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  static DeviceProps bad;
  if (e != hipSuccess || dev < 0 || dev >= 64) {
    bad.error = e != hipSuccess ? e : hipErrorInvalidDevice;
    return bad;
  }
  std::lock_guard<std::mutex> lock(mu);
  if (!have[dev]) {
    hipDeviceProp_t prop;
    DeviceProps p;
    p.error = hipGetDeviceProperties(&prop, dev);
    p.num_cus = prop.multiProcessorCount;
    p.lds_per_block = prop.sharedMemPerBlock;
    p.lds_per_cu = prop.maxSharedMemoryPerMultiProcessor;
    cache[dev] = p;
    have[dev] = p.error == hipSuccess;
    if (!have[dev]) return cache[dev];
  }
  return cache[dev];
}

// ------------------------------------------------------------- fill ----
__global__ void k_fill_u32(uint32_t *p, uint32_t v, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

int launch_fill_u32(uint32_t *p, uint32_t v, int64_t n, hipStream_t s) {
  if (n <= 0) return AV_OK;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s, p, v, n);
  AV_LAUNCH_CHECK("k_fill_u32");
  return AV_OK;
}

// ------------------------------------------------------------- scan ----
// Exclusive prefix sum of n uint32 values into out[0..n], out[n] = total.
// Three passes: per-block local scan (4096 values per block), scan of the
// block totals in one block, and a final add.
constexpr int kScanThreads = 256;
constexpr int kScanPerThread = 16;
constexpr int kScanBlock = kScanThreads * kScanPerThread;

__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t *sh, uint32_t &total) {
  // Hillis-Steele over blockDim.x (<= 1024) values in LDS.
  int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < (int)blockDim.x; off <<= 1) {
    uint32_t add = t >= off ? sh[t - off] : 0u;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  uint32_t incl = sh[t];
  total = sh[blockDim.x - 1];
  __syncthreads();
  return incl - v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_local(const uint32_t *in, uint32_t *out,
                                                             int64_t n, uint32_t *block_sums) {
  __shared__ uint32_t sh[kScanThreads];
  int64_t base = (int64_t)blockIdx.x * kScanBlock + (int64_t)threadIdx.x * kScanPerThread;
  uint32_t vals[kScanPerThread];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanPerThread; ++k) {
    int64_t i = base + k;
    vals[k] = i < n ? in[i] : 0u;
    s += vals[k];
  }
  uint32_t total;
  uint32_t excl = block_exclusive_scan(s, sh, total);
#pragma unroll
  for (int k = 0; k < kScanPerThread; ++k) {
    int64_t i = base + k;
    if (i < n) out[i] = excl;
    excl += vals[k];
  }
  if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void k_scan_sums(uint32_t *block_sums, int64_t nblk,
                                                    uint32_t *grand_total) {
  __shared__ uint32_t sh[1024];
  // each thread handles a contiguous run of ceil(nblk/1024) sums
  int64_t per = (nblk + 1023) / 1024;
  int64_t b0 = (int64_t)threadIdx.x * per;
  uint32_t s = 0;
  for (int64_t k = 0; k < per; ++k)
    if (b0 + k < nblk) s += block_sums[b0 + k];
  uint32_t total;
  uint32_t excl = block_exclusive_scan(s, sh, total);
  for (int64_t k = 0; k < per; ++k) {
    if (b0 + k < nblk) {
      uint32_t v = block_sums[b0 + k];
      block_sums[b0 + k] = excl;
      excl += v;
    }
  }
  if (threadIdx.x == 0) *grand_total = total;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_add(uint32_t *out, int64_t n,
                                                           const uint32_t *block_sums) {
  int64_t base = (int64_t)blockIdx.x * kScanBlock;
  uint32_t add = block_sums[blockIdx.x];
  for (int k = threadIdx.x; k < kScanBlock; k += kScanThreads) {
    int64_t i = base + k;
    if (i < n) out[i] += add;
  }
}

int scan_exclusive_u32(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *tmp,
                       hipStream_t s) {
  int64_t nblk = (n + kScanBlock - 1) / kScanBlock;
  if (nblk < 1) nblk = 1;
  hipLaunchKernelGGL(k_scan_local, dim3((unsigned)nblk), dim3(kScanThreads), 0, s, in, out, n,
                     tmp);
  AV_LAUNCH_CHECK("k_scan_local");
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, tmp, nblk, out + n);
  AV_LAUNCH_CHECK("k_scan_sums");
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nblk), dim3(kScanThreads), 0, s, out, n, tmp);
  AV_LAUNCH_CHECK("k_scan_add");
  return AV_OK;
}

// -------------------------------------------------- prolate spheroid ----
// Schwab's rational approximation (M=6, alpha=1) as in
// core/imaging/imaging_utils/gcf_prolate_spheroidal.py:102-158: the two
// coefficient tables are the published ones and must be identical.
__constant__ double c_ps_p[2][5] = {
    {8.203343e-2, -3.644705e-1, 6.278660e-1, -5.335581e-1, 2.312756e-1},
    {4.028559e-3, -3.697768e-2, 1.021332e-1, -1.201436e-1, 6.412774e-2}};
__constant__ double c_ps_q[2][3] = {{1.0000000e0, 8.212018e-1, 2.078043e-1},
                                    {1.0000000e0, 9.599102e-1, 2.918724e-1}};

__device__ void ps_eval(double uin, double &grdsf, double &kern) {
  double u = fabs(uin);
  int part = 0;
  double uend = 0.0;  // |u| > 1 (or NaN) keeps the zero defaults
  if (u >= 0.0 && u < 0.75) {
    part = 0;
    uend = 0.75;
  }
  if (u >= 0.75 && u <= 1.0) {
    part = 1;
    uend = 1.0;
  }
  double d = u * u - uend * uend;
  double d2 = d * d, d3 = d2 * d, d4 = d2 * d2;
  double top = c_ps_p[part][0];
  top += c_ps_p[part][1] * d;
  top += c_ps_p[part][2] * d2;
  top += c_ps_p[part][3] * d3;
  top += c_ps_p[part][4] * d4;
  double bot = c_ps_q[part][0];
  bot += c_ps_q[part][1] * d;
  bot += c_ps_q[part][2] * d2;
  double g = 0.0;
  if (bot > 0.0) g = top / bot;
  if (u > 1.0) g = 0.0;
  grdsf = g;
  kern = (1.0 - u * u) * g;
}

__global__ void k_ps_function(const double *u, int64_t n, double *grdsf, double *kern) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double g, k;
  ps_eval(u[i], g, k);
  if (grdsf) grdsf[i] = g;
  if (kern) kern[i] = k;
}

// create_prolate_spheroidal_kernel_1D (gcf_prolate_spheroidal.py:87-99)
__global__ void k_ps_kernel_1d(int ov, int support, double *out) {
  int sc = support / 2;
  int n = ov * sc, ntot = ov * (sc + 1);
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ntot) return;
  double v = 0.0;
  if (i < n) {
    double g;
    ps_eval((double)i / (double)(sc * ov), g, v);
  }
  out[i] = v;
}

// grdsf(|2 * coordinates(n)|) (gcf_prolate_spheroidal.py:73-77, :161-163)
__global__ void k_ps_correction_1d(int n, double *out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double c = (double)(i - n / 2) / (double)n;
  double g, k;
  ps_eval(fabs(2.0 * c), g, k);
  out[i] = g;
}

// kernel_image of create_prolate_spheroidal_kernel (gcf_prolate_spheroidal.py:72-79)
__global__ void k_ps_taper(int nu, int nv, double *out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)nu * nv) return;
  int a = (int)(i / nv), b = (int)(i - (int64_t)a * nv);
  double gu, gv, k;
  ps_eval(fabs(2.0 * ((double)(a - nu / 2) / (double)nu)), gu, k);
  ps_eval(fabs(2.0 * ((double)(b - nv / 2) / (double)nv)), gv, k);
  out[i] = gu * gv;
}

// gcf_prolate_spheroidal.py:65-67: np.outer of two kernel rows
__global__ void k_ps_outer(const double *k1, int nrow, int nfill, int support, double *k4) {
  int64_t total = (int64_t)nrow * nrow * support * support;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  int j = (int)(i % support);
  int64_t r = i / support;
  int ii = (int)(r % support);
  r /= support;
  int y = (int)(r % nrow);
  int x = (int)(r / nrow);
  double v = 0.0;
  if (x < nfill && y < nfill) v = k1[(int64_t)x * support + ii] * k1[(int64_t)y * support + j];
  k4[i] = v;
}

}  // namespace av

using namespace av;

extern "C" {

const char *av_last_error(void) { return g_last_error.c_str(); }

int av_version(void) { return AV_ABI_VERSION; }

int av_device_info(int32_t *num_cus, int32_t *lds_bytes_per_cu, int32_t *xcds) {
  int dev = 0;
  AV_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  AV_HIP(hipGetDeviceProperties(&prop, dev));
  if (num_cus) *num_cus = prop.multiProcessorCount;
  if (lds_bytes_per_cu) *lds_bytes_per_cu = (int32_t)prop.maxSharedMemoryPerMultiProcessor;
  if (xcds) *xcds = 8;
  return AV_OK;
}

int av_ps_function(const double *u, int64_t n, double *grdsf, double *kern, void *stream) {
  if (n < 0 || (n > 0 && !u)) return fail(AV_ERR_INVALID, "av_ps_function: bad arguments");
  if (n == 0) return AV_OK;
  hipLaunchKernelGGL(k_ps_function, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, u, n, grdsf, kern);
  AV_LAUNCH_CHECK("k_ps_function");
  return AV_OK;
}

int av_ps_kernel_1d(int32_t oversampling, int32_t support, double *out, void *stream) {
  if (oversampling < 1 || support < 2 || !out)
    return fail(AV_ERR_INVALID, "av_ps_kernel_1d: need oversampling >= 1, support >= 2");
  int ntot = oversampling * (support / 2 + 1);
  hipLaunchKernelGGL(k_ps_kernel_1d, dim3((ntot + 255) / 256), dim3(256), 0,
                     (hipStream_t)stream, oversampling, support, out);
  AV_LAUNCH_CHECK("k_ps_kernel_1d");
  return AV_OK;
}

int av_ps_correction_1d(int32_t n, double *out, void *stream) {
  if (n < 1 || !out) return fail(AV_ERR_INVALID, "av_ps_correction_1d: bad arguments");
  hipLaunchKernelGGL(k_ps_correction_1d, dim3((n + 255) / 256), dim3(256), 0,
                     (hipStream_t)stream, n, out);
  AV_LAUNCH_CHECK("k_ps_correction_1d");
  return AV_OK;
}

int av_ps_taper(int32_t n_u, int32_t n_v, double *out, void *stream) {
  if (n_u < 1 || n_v < 1 || !out) return fail(AV_ERR_INVALID, "av_ps_taper: bad arguments");
  int64_t n = (int64_t)n_u * n_v;
  hipLaunchKernelGGL(k_ps_taper, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, n_u, n_v, out);
  AV_LAUNCH_CHECK("k_ps_taper");
  return AV_OK;
}

int av_ps_outer(const double *k1, int32_t nrow, int32_t nfill, int32_t support, double *k4,
                void *stream) {
  if (nrow < 1 || support < 1 || nfill < 0 || nfill > nrow || !k1 || !k4)
    return fail(AV_ERR_INVALID, "av_ps_outer: bad arguments");
  int64_t total = (int64_t)nrow * nrow * support * support;
  hipLaunchKernelGGL(k_ps_outer, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, k1, nrow, nfill, support, k4);
  AV_LAUNCH_CHECK("k_ps_outer");
  return AV_OK;
}

}  // extern "C"
