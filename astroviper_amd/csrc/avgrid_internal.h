// avgrid_internal.h -- shared device helpers and the plan layout of
// libavgrid.so.  Compiled for gfx950 only (wave64, 160 KB LDS per CU).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <cmath>
#include <string>

#include "../../include/avgrid.h"

namespace av {

// ------------------------------------------------------------ errors ----
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);

#define AV_HIP(call)                                            \
  do {                                                          \
    hipError_t _e = (call);                                     \
    if (_e != hipSuccess) return ::av::hip_fail(_e, #call);     \
  } while (0)

#define AV_LAUNCH_CHECK(name)                                   \
  do {                                                          \
    hipError_t _e = hipGetLastError();                          \
    if (_e != hipSuccess) return ::av::hip_fail(_e, name);      \
  } while (0)

constexpr double kLightSpeed = 299792458.0;  // standard_grid.py:296
constexpr uint32_t kInvalidGeo = 0xFFFFFFFFu;
constexpr int kPosBias = 16;   // tile-relative positions are stored + 16
constexpr int kOffBias = 128;  // kernel offsets are stored + 128
constexpr int kMaxSupport = 16;      // av_grid / av_degrid kernels
constexpr int kMaxPlanSupport = 63;  // binning only (the mosaic gridder)
constexpr int kMaxPol = 4;
constexpr int kMaxDevices = 64;
constexpr int kChunkThreads = 1024;       // threads of a chunked-binning workgroup
constexpr int kMaxChunkTiles = 16384;     // tiles whose two counters fit in 128 KB of LDS
constexpr int kMaxRowsPerChunk = 124 * kChunkThreads;  // rows of one chunked-binning workgroup
constexpr int kRowBlockRows = 8 * kChunkThreads;  // rows of a row-block binning workgroup

// Packed geometry of one sample relative to a tile origin:
// bits 0-7 u position + 16, 8-15 u offset + 128, 16-23 v position + 16,
// 24-31 v offset + 128.
__host__ __device__ inline uint32_t pack_geo(int au, int ou, int av_, int ov) {
  return (uint32_t)(au + kPosBias) | ((uint32_t)(ou + kOffBias) << 8) |
         ((uint32_t)(av_ + kPosBias) << 16) | ((uint32_t)(ov + kOffBias) << 24);
}
// A valid geometry with both centres at +239 (far outside any tile and
// region) and zero kernel offsets: substituting it for kInvalidGeo gives
// every tap weight zero through in-bounds table reads, with no branch.
constexpr uint32_t kFarGeo = 0x80FF80FFu;
__device__ inline uint32_t geo_or_far(uint32_t g) { return g == kInvalidGeo ? kFarGeo : g; }

// A record's words as vectors of 4-byte alignment: one 12-byte store per
// primary record (one 8-byte store per duplicate), and the grid kernel reads
// the first two words only.
typedef uint32_t RecWords __attribute__((ext_vector_type(3), aligned(4)));
typedef uint32_t RecGridWords __attribute__((ext_vector_type(2), aligned(4)));

__device__ inline void unpack_geo(uint32_t g, int &au, int &ou, int &av_, int &ov) {
  au = (int)(g & 0xFFu) - kPosBias;
  ou = (int)((g >> 8) & 0xFFu) - kOffBias;
  av_ = (int)((g >> 16) & 0xFFu) - kPosBias;
  ov = (int)((g >> 24) & 0xFFu) - kOffBias;
}

// --------------------------------------------------- FP64 geometry ----
// Library-wide -ffp-contract=off keeps every product and sum below a
// separately rounded IEEE operation, as in the reference's numba code.

// standard_grid.py:322-352.  u = uvw_u * scale_u (scale precomputed per
// channel as -(f*dl*N)/c), centre int(u + N//2 + 0.5) truncated toward zero,
// bounds iu-sc >= 0 and iu+sc < N, offset floor((iu - pos)*ov + 0.5).
__device__ inline bool grid_geometry(double uvw_u, double uvw_v, double scale_u,
                                     double scale_v, int n_u, int n_v, int sc,
                                     int ov, int &iu, int &iv, int &ou, int &ovv) {
  double u = uvw_u * scale_u;
  double v = uvw_v * scale_v;
  if (isnan(u) || isnan(v)) return false;
  double u_pos = u + (double)(n_u / 2);
  double v_pos = v + (double)(n_v / 2);
  double xu = u_pos + 0.5, xv = v_pos + 0.5;
  // Anything outside +-1e9 fails the bounds test; reject before the cast.
  if (!(xu > -1.0e9 && xu < 1.0e9 && xv > -1.0e9 && xv < 1.0e9)) return false;
  int cu = (int)xu;  // int(): truncation toward zero
  int cv = (int)xv;
  if (!((cu + sc < n_u) && (cv + sc < n_v) && (cu - sc >= 0) && (cv - sc >= 0)))
    return false;
  double uo = (double)cu - u_pos;
  double vo = (double)cv - v_pos;
  iu = cu;
  iv = cv;
  ou = (int)floor(uo * (double)ov + 0.5);
  ovv = (int)floor(vo * (double)ov + 0.5);
  return true;
}

// x / c, correctly rounded, for the speed of light c (exactly representable)
// by Markstein's correction step (P. Markstein, "Computation of elementary
// functions on the IBM RISC System/6000 processor", IBM J. Res. Dev. 34(1),
// 1990; J.-M. Muller et al., Handbook of Floating-Point Arithmetic, 2nd ed.,
// section 4.6): with y = RN(1/c) and q0 = RN(x y), within one ulp of x/c,
// the remainder r = x - q0 c is exact in one fma and RN(q0 + r y) = RN(x/c)
// whenever x/c is a normal number.  Three operations instead of a division
// sequence; NaN and infinite x give NaN, which the callers reject as they
// would the division's NaN or infinity.  Checked bit for bit against x / c
// on 5e8 inputs over the normal range.
__device__ inline double div_light_speed(double x) {
  constexpr double y = 1.0 / kLightSpeed;
  const double q0 = x * y;
  const double r = fma(-q0, kLightSpeed, x);
  return fma(r, y, q0);
}

// standard_degrid.py:435-462 with scale = -dl*N (standard_degrid.py:572),
// offset N/2 (float), pos = scale*uvw*freq/c + offset evaluated left to right,
// loc = round-half-even(pos), off = round-half-even((loc - pos)*ov).
// The reference bounds the u index by grid.shape[3] and v by grid.shape[2]
// (standard_degrid.py:407,457-462); both those and the true axis bounds
// are applied (they differ only for non-square grids).
__device__ inline bool degrid_geometry(double uvw_u, double uvw_v, double freq,
                                       double dscale_u, double dscale_v,
                                       double doff_u, double doff_v, int n_u,
                                       int n_v, int support, int ov, int &lu,
                                       int &lv, int &fu, int &fv) {
  double pu = div_light_speed(dscale_u * uvw_u * freq) + doff_u;
  double pv = div_light_speed(dscale_v * uvw_v * freq) + doff_v;
  if (isnan(pu) || isnan(pv)) return false;
  double ru = rint(pu), rv = rint(pv);
  if (!(ru > -1.0e9 && ru < 1.0e9 && rv > -1.0e9 && rv < 1.0e9)) return false;
  int x0 = (int)ru, y0 = (int)rv;
  int sb = -(support / 2);
  int se = sb + support;
  int bu = n_u < n_v ? n_u : n_v;  // min(shape[3] bound, true bound)
  int bv = bu;
  if (!((x0 + sb >= 0) && (x0 + se - 1 < bu) && (y0 + sb >= 0) && (y0 + se - 1 < bv)))
    return false;
  lu = x0;
  lv = y0;
  fu = (int)rint(((double)x0 - pu) * (double)ov);
  fv = (int)rint(((double)y0 - pv) * (double)ov);
  return true;
}

// ------------------------------------------------------------- plan ----
struct PlanDev {
  // sizes
  int64_t n_rows;
  int32_t n_chan, n_img_chan, n_u, n_v;
  double cell_u, cell_v;
  double dscale_u, dscale_v, doff_u, doff_v;  // degrid pixel transform
  int32_t support, ov, tile_u, tile_v, geometry, unit_cap;
  int32_t conjugate;  // AV_GEOM_CONJUGATE: samples at their conjugate centres
  // x / tile as __umulhi(x, tile_*_mul), exact for 0 <= x < 2^24 (tiles <= 128)
  uint32_t tile_u_mul, tile_v_mul;
  int32_t row_channels;  // every row is one sample at its own channel
  int32_t ntu, ntv;      // tiles per axis
  int64_t n_bins;        // ntu * ntv * n_img_chan
  int64_t n_items;       // samples (row, chan)
  int64_t max_records;   // capacity of the record arrays
  int64_t max_gunits, max_dunits;
  // per-channel data
  double *gscale_u, *gscale_v;  // [n_chan] grid scale -(f*dl*N)/c
  double *freq;                 // [n_chan] copy of the frequencies
  int32_t *chan_map;            // [n_chan]
  int32_t *row_chan;            // [n_rows] or nullptr (not owned)
  const double *uvw;            // [n_rows][3] (not owned; valid until next set)
  // bins
  uint32_t *cnt_all, *cnt_prim;     // [n_bins]
  uint32_t *bin_start;              // [n_bins + 1]
  uint32_t *cur_prim, *cur_dup;     // [n_bins]
  uint32_t *gunit_start, *dunit_start;  // [n_bins + 1]
  uint32_t *scan_tmp;               // block sums
  // records, three words each: {sample, packed grid geometry, packed
  // degrid geometry}; the third word is written for primary records only
  uint32_t *rec;                    // [max_records][3]
  // work units: x = bin, y = first record, z = count, w = flags
  uint4 *gunits, *dunits;
  // pair units of the split-FP16 degridder (16 x 16 tiles, support <= 7):
  // tiles 2j and 2j + 1 along v of one image channel share a unit, pair p =
  // (tile row, j, image channel) channel-minor; a unit is {bin of tile 2j,
  // first index into the pair's primary records (those of tile 2j, then
  // those of tile 2j + 1), count, bin of tile 2j + 1 or kNoBin}
  int64_t n_pairs, max_punits;      // 0 when the plan has no pair units
  uint32_t *punit_start;            // [n_pairs + 1]
  uint4 *punits;                    // [max_punits]
  // validity bitmasks, one bit per sample: grid-valid, degrid-valid; bit
  // (chan, row) is bit row % 64 of word chan * mask_words_per_chan + row / 64
  // (chan 0 when rows carry their own channel)
  unsigned long long *gmask, *dmask;  // [n_words]
  int64_t n_words, mask_words_per_chan;
  // chunked binning: a chunk is rows_per_chunk rows of one channel
  int32_t chunked;                  // chunked binning possible for this plan
  int32_t rows_per_chunk, chunks_per_chan;
  // per chunk and tile, the chunk's primary (chunk_hist) and duplicate
  // (chunk_dup) counts, each then replaced by the first slot of that run
  uint32_t *chunk_hist;             // [n_chan * chunks_per_chan][ntu * ntv]
  uint32_t *chunk_dup;              // [n_chan * chunks_per_chan][ntu * ntv]
  int32_t chunk_scatter;            // records written through per-chunk LDS cursors
  int32_t order_nb;                 // baselines per time step of time-major rows (1: row order)
  // row-block binning (rows carry their own channel): kRowBlockRows rows per
  // workgroup, per-tile counters and cursors in LDS; the counting pass keeps
  // each block's own-channel counts (prim | dup << 16) for the scatter pass
  int32_t row_blocks;
  uint32_t *row_hist;               // [row blocks][ntu * ntv], or NULL
  // per-call tables: FP32 kernel table, FP64 per-axis kernel sums by offset
  float *tab32;                     // [ov * (support/2 + 1)]
  double *nrm;                      // [2 * (ov/2 + 2) + 1]
  float *wtab32;                    // zero-padded weight table [(support + 2) * nrm size]
  double *wtab64;                   // the same in FP64
  // per grid unit, per polarisation: sum of weight * norm (FP64)
  double *upart;                    // [max_gunits * 4]
  uint32_t *counters;               // scratch for statistics
};

constexpr uint32_t kNoBin = 0xFFFFFFFFu;  // a pair unit's missing second tile
constexpr uint32_t kUnitMulti = 1u;  // bin split over several units
constexpr uint32_t kUnitFirst = 2u;  // first unit of its bin

struct av_plan_host;  // defined in plan.hip

// launch helpers shared by the translation units
int launch_fill_u32(uint32_t *p, uint32_t v, int64_t n, hipStream_t s);
int scan_exclusive_u32(const uint32_t *in, uint32_t *out, int64_t n,
                       uint32_t *tmp, hipStream_t s);
int prepare_tables(const PlanDev &d, const double *cgk, hipStream_t s);
// grid.hip: overwrite-mode zeroing before a tile kernel (sums of weights and
// split-bin tiles), and the per-unit sums of weights into sumwt after it
int zero_before_grid(const PlanDev &d, void *grid, int npol, bool fp64, double *sumwt,
                     hipStream_t s);
int reduce_sumwt(const PlanDev &d, int npol, double *sumwt, hipStream_t s);

// properties of the current device, queried once per device
struct DeviceProps {
  hipError_t error;
  int num_cus;
  size_t lds_per_block, lds_per_cu;
};
const DeviceProps &device_props();

// Raise a kernel's dynamic-LDS limit to `bytes` the first time a launch needs
// it on a device.  Later launches make no runtime call, so they can also be
// recorded under stream capture once the kernel has run eagerly.
template <auto Kernel>
int ensure_lds_limit(int bytes, const char *what) {
  static std::atomic<int> limit[kMaxDevices];
  int dev = 0;
  AV_HIP(hipGetDevice(&dev));
  if (dev >= 0 && dev < kMaxDevices && limit[dev].load(std::memory_order_relaxed) >= bytes)
    return AV_OK;
  hipError_t e = hipFuncSetAttribute((const void *)Kernel,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e != hipSuccess) return hip_fail(e, what);
  if (dev >= 0 && dev < kMaxDevices) {
    int cur = limit[dev].load(std::memory_order_relaxed);
    while (cur < bytes && !limit[dev].compare_exchange_weak(cur, bytes)) {
    }
  }
  return AV_OK;
}

}  // namespace av

struct av_plan {
  av::PlanDev d;
  void *alloc_base;   // single device allocation holding every plan array
  size_t alloc_bytes;
  int device;
};

// Validate a plan's parameters, fill plan->d's sizes and return in *bytes
// the device memory its arrays need; with base non-NULL also place the
// arrays in [base, base + *bytes).  Defined in plan.hip.
int plan_layout(int64_t n_rows, int32_t n_chan, int32_t n_img_chan, int32_t n_u, int32_t n_v,
                double cell_u, double cell_v, int32_t support, int32_t oversampling,
                int32_t tile_u, int32_t tile_v, int32_t geometry, int32_t unit_cap, av_plan *p,
                char *base, size_t *bytes);
