// noalloc.hip -- av_grid_noalloc: binning and gridding in one call that
// allocates no device memory.  The plan's arrays (records, bins, units,
// masks, tables, partial sums) are carved from `scratch`, a file-scope
// device array of kScratchBytes that every GPU running the library holds
// until the process ends.  Host code never passes `scratch` itself to a
// kernel or a copy: that name is a host-side stand-in, not a device address.
// Its address on each device comes from hipGetSymbolAddress, taken once per
// device (with that device current) on the first call there.
//
// Every call on one GPU shares that GPU's array and keeps the last call's
// contents; the binning passes initialise every array they read, so nothing
// needs clearing between calls.  Calls that overlap on one GPU (different
// streams or threads) race on the array.
#include <mutex>

#include "avgrid_internal.h"

namespace av {

constexpr size_t kScratchBytes = (size_t)32 << 20;  // 32 MiB
constexpr int kMaxScratchDevices = 64;

static __device__ __align__(256) char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static char *g_scratch_ptr[kMaxScratchDevices];  // device address of `scratch`, per device

// Address of the current device's scratch array (resolved on first use).
static int scratch_on_current_device(char **out, int *device) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "av_grid_noalloc: hipGetDevice");
  if (dev < 0 || dev >= kMaxScratchDevices)
    return fail(AV_ERR_UNSUPPORTED, "av_grid_noalloc: device index beyond 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "av_grid_noalloc: hipGetSymbolAddress");
    g_scratch_ptr[dev] = (char *)p;
  }
  *out = g_scratch_ptr[dev];
  *device = dev;
  return AV_OK;
}

}  // namespace av

using namespace av;

extern "C" int av_grid_noalloc(const double *uvw, int64_t n_rows, const double *freq,
                               int32_t n_chan, const int32_t *chan_map, int32_t n_img_chan,
                               int32_t n_u, int32_t n_v, double cell_u, double cell_v,
                               int32_t support, int32_t oversampling, int32_t tile_u,
                               int32_t tile_v, const void *vis, const void *weight,
                               const uint8_t *flag, int32_t n_pol, const double *cgk,
                               void *grid, double *sumwt, int32_t dtype, int32_t mode,
                               void *stream) {
  av_plan plan{};
  size_t bytes = 0;
  int rc = plan_layout(n_rows, n_chan, n_img_chan, n_u, n_v, cell_u, cell_v, support,
                       oversampling, tile_u, tile_v, AV_GEOM_REFERENCE, 0, &plan, nullptr,
                       &bytes);
  if (rc) return rc;
  if (bytes > kScratchBytes)
    return fail(AV_ERR_INVALID, "av_grid_noalloc: this call's plan needs " +
                                    std::to_string(bytes) + " bytes, beyond the " +
                                    std::to_string(kScratchBytes) +
                                    "-byte static scratch (use av_plan_create)");
  char *base = nullptr;
  int device = 0;
  if ((rc = scratch_on_current_device(&base, &device))) return rc;
  if ((rc = plan_layout(n_rows, n_chan, n_img_chan, n_u, n_v, cell_u, cell_v, support,
                        oversampling, tile_u, tile_v, AV_GEOM_REFERENCE, 0, &plan, base,
                        &bytes)))
    return rc;
  plan.device = device;
  if ((rc = av_plan_set_uvw(&plan, uvw, freq, chan_map, nullptr, stream))) return rc;
  return av_grid(&plan, vis, weight, flag, n_pol, cgk, grid, sumwt, dtype, mode, stream);
}
