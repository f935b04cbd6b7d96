// plan.hip -- binning of samples into (uv tile, image channel) bins.
//
// A sample is one (row, channel) pair with all its polarisations.  Its grid
// centre (standard_grid.py:322-352) and degrid centre
// (standard_degrid.py:435-462) are computed once, in FP64, here.  The sample
// is recorded in the bin of its anchor tile ("primary" record, carrying both
// geometries) and, for gridding, in every other tile its support footprint
// overlaps ("duplicate" records), so that each tile of the grid is produced
// by exactly one owner with no halo exchange.  Within a bin, primary records
// come first.  Bins are ordered tile-major, channel-minor, so neighbouring
// channels of the same rows are processed close together in time.
//
// Binning is a counting sort.  The counts come from chunks (one channel, a
// range of rows) gathered in LDS when a channel's tiles fit there, else from
// global atomics (very many tiles).  The records are then written with one
// cursor per bin, the samples taken in order so that the records a
// wavefront writes land in neighbouring bins.  When rows carry their own
// channel, blocks of consecutive rows count and place their records through
// per-tile counters and cursors in LDS instead (row-block binning).  Validity bitmasks are channel-major: bit (chan, row) is
// bit row % 64 of word chan * ceil(n_rows / 64) + row / 64 (chan 0 when
// every row carries its own channel).
#include <algorithm>

#include "avgrid_internal.h"

namespace av {

struct ItemGeo {
  int64_t row;
  int chan, ic;
  bool gvalid, dvalid;
  int iu, iv, ou, ov;  // grid centre and kernel offsets
  int lu, lv, fu, fv;  // degrid centre and kernel offsets
};

// The per-channel constants of the geometry: loaded once per channel where
// a loop visits many rows of one channel (the chunked binning passes),
// instead of once per sample behind the sample's own loads.
struct ChanGeo {
  int chan, ic, ch;  // channel, image channel, channel the tables are read at
  bool ok;           // both in range
  double freq, gsu, gsv;
};

__device__ inline ChanGeo chan_geo(const PlanDev &P, int chan) {
  ChanGeo c;
  c.chan = chan;
  c.ok = chan >= 0 && chan < P.n_chan;
  c.ch = c.ok ? chan : 0;
  c.ic = P.chan_map[c.ch];
  c.ok = c.ok && c.ic >= 0 && c.ic < P.n_img_chan;
  c.freq = P.freq[c.ch];
  c.gsu = P.gscale_u[c.ch];
  c.gsv = P.gscale_v[c.ch];
  return c;
}

// Geometry of the sample of row `row` at channel c, whose u and v are uu
// and vv.
__device__ inline void sample_geometry_at(const PlanDev &P, int64_t row, const ChanGeo &c,
                                          double uu, double vv, ItemGeo &g) {
  g.row = row;
  g.chan = c.chan;
  g.ic = c.ic;
  const bool chan_ok = c.ok;
  const int sc = P.support / 2;
  g.dvalid = degrid_geometry(uu, vv, c.freq, P.dscale_u, P.dscale_v, P.doff_u, P.doff_v,
                             P.n_u, P.n_v, P.support, P.ov, g.lu, g.lv, g.fu, g.fv) &&
             chan_ok;
  if (P.geometry == AV_GEOM_DEGRID) {
    g.gvalid = g.dvalid;
    g.iu = g.lu;
    g.iv = g.lv;
    g.ou = g.fu;
    g.ov = g.fv;
  } else {
    g.gvalid = grid_geometry(uu, vv, c.gsu, c.gsv, P.n_u, P.n_v, sc, P.ov, g.iu, g.iv, g.ou,
                             g.ov) &&
               chan_ok;
    if (P.conjugate) {
      // the second scatter of do_imaging_weight (standard_grid.py:330-340,
      // 389-425): the conjugate centre int(-u + N/2 + 0.5), the sample's own
      // kernel offsets and bounds test, no degrid record
      g.dvalid = false;
      if (g.gvalid) {
        const double u = uu * c.gsu, v = vv * c.gsv;
        g.iu = (int)(-u + (double)(P.n_u / 2) + 0.5);
        g.iv = (int)(-v + (double)(P.n_v / 2) + 0.5);
      }
    }
  }
}

// Geometry of the sample of row `row` at channel `chan` (the row's own
// channel when rows carry one).
__device__ inline void sample_geometry(const PlanDev &P, int64_t row, int chan, ItemGeo &g) {
  const ChanGeo c = chan_geo(P, P.row_chan ? P.row_chan[row] : chan);
  sample_geometry_at(P, row, c, P.uvw[row * 3 + 0], P.uvw[row * 3 + 1], g);
}

__device__ inline int64_t bin_of(const PlanDev &P, int tu, int tv, int ic) {
  return ((int64_t)tu * P.ntv + tv) * P.n_img_chan + ic;
}

__global__ void k_chan_setup(PlanDev P, const double *freq, const int32_t *chan_map) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= P.n_chan) return;
  double f = freq[c];
  P.freq[c] = f;
  // numpy: -(freq_chan * delta_lm[0] * n_uv[0]) / c   (standard_grid.py:298-299)
  P.gscale_u[c] = -(f * P.cell_u * (double)P.n_u) / kLightSpeed;
  P.gscale_v[c] = -(f * P.cell_v * (double)P.n_v) / kLightSpeed;
  int m;
  if (chan_map)
    m = chan_map[c];
  else
    m = (P.n_img_chan == P.n_chan) ? c : 0;
  P.chan_map[c] = m;
}

// Records of one sample.  A grid-valid sample gets a primary record in the
// tile of its grid centre (carrying the degrid geometry too when the two
// centres are within one pixel, which is always the case for the reference
// transforms) and a duplicate in every other tile its footprint overlaps.
// A degrid-valid sample whose degrid centre is elsewhere (a caller-supplied
// degrid transform) gets its own primary record in the tile of that centre.
struct RecPlan {
  bool gprim, dsep;        // grid-primary record, separate degrid record
  bool share;              // grid-primary record also carries degrid geometry
  int tu, tv, du_t, dv_t;  // grid anchor tile, degrid anchor tile
  int tu0, tu1, tv0, tv1;  // tiles overlapped by the grid footprint
};

__device__ inline int tile_u_of(const PlanDev &P, int x) {
  return (int)__umulhi((uint32_t)x, P.tile_u_mul);
}
__device__ inline int tile_v_of(const PlanDev &P, int x) {
  return (int)__umulhi((uint32_t)x, P.tile_v_mul);
}

__device__ inline RecPlan record_plan(const PlanDev &P, const ItemGeo &g) {
  RecPlan r;
  const int sc = P.support / 2;
  r.gprim = g.gvalid;
  r.share = g.gvalid && g.dvalid && abs(g.lu - g.iu) <= 1 && abs(g.lv - g.iv) <= 1;
  r.dsep = g.dvalid && !r.share;
  r.tu = r.tv = r.du_t = r.dv_t = 0;
  r.tu0 = r.tu1 = r.tv0 = r.tv1 = 0;
  if (g.gvalid) {
    // tile numbers by multiply-high of positions clamped into [0, n) (n <
    // 2^24): a valid footprint lies inside the grid, a conjugate one may
    // reach past an edge, whose cells are dropped
    r.tu = tile_u_of(P, min(max(g.iu, 0), P.n_u - 1));
    r.tv = tile_v_of(P, min(max(g.iv, 0), P.n_v - 1));
    r.tu0 = tile_u_of(P, max(g.iu - sc, 0));
    r.tu1 = tile_u_of(P, min(g.iu + P.support - sc - 1, P.n_u - 1));
    r.tv0 = tile_v_of(P, max(g.iv - sc, 0));
    r.tv1 = tile_v_of(P, min(g.iv + P.support - sc - 1, P.n_v - 1));
  }
  if (r.dsep) {
    r.du_t = tile_u_of(P, g.lu);
    r.dv_t = tile_v_of(P, g.lv);
  }
  return r;
}

__device__ inline int64_t item_index(const PlanDev &P, int64_t row, int chan) {
  return P.row_chan ? row : row * P.n_chan + chan;
}

// The validity bits of a wavefront's 64 consecutive rows of one channel
// (rows from a multiple of 64): one mask word, written by lane 0.
__device__ inline void write_masks(const PlanDev &P, int64_t row, int chan, bool gvalid,
                                   bool dvalid) {
  const unsigned long long gm = __ballot(gvalid), dm = __ballot(dvalid);
  const int64_t first = row - __lane_id();
  if (__lane_id() == 0 && first < P.n_rows) {
    const int64_t w = (P.row_chan ? 0 : (int64_t)chan * P.mask_words_per_chan) + first / 64;
    P.gmask[w] = gm;
    P.dmask[w] = dm;
  }
}

// The records of one sample counted straight into the global bins.
__device__ inline void count_global(const PlanDev &P, const ItemGeo &g, const RecPlan &rp) {
  if (rp.gprim) {
    const int64_t pb = bin_of(P, rp.tu, rp.tv, g.ic);
    atomicAdd(&P.cnt_prim[pb], 1u);
    atomicAdd(&P.cnt_all[pb], 1u);
    for (int a = rp.tu0; a <= rp.tu1; ++a)
      for (int b = rp.tv0; b <= rp.tv1; ++b)
        if (a != rp.tu || b != rp.tv) atomicAdd(&P.cnt_all[bin_of(P, a, b, g.ic)], 1u);
  }
  if (rp.dsep) {
    const int64_t db = bin_of(P, rp.du_t, rp.dv_t, g.ic);
    atomicAdd(&P.cnt_prim[db], 1u);
    atomicAdd(&P.cnt_all[db], 1u);
  }
}

// Global-atomic binning, pass 1: count records per bin.  Threads walk the
// (channel, row) space with rows padded to a multiple of 64 per channel.
__global__ __launch_bounds__(256) void k_count(PlanDev P) {
  const int64_t rows_pad = P.mask_words_per_chan * 64;
  const int nch = P.row_chan ? 1 : P.n_chan;
  const int64_t total = rows_pad * nch;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;  // a multiple of 64
  for (int64_t it = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; it < total; it += stride) {
    const int chan = (int)(it / rows_pad);
    const int64_t row = it - (int64_t)chan * rows_pad;
    ItemGeo g;
    g.gvalid = g.dvalid = false;
    if (row < P.n_rows) sample_geometry(P, row, chan, g);
    write_masks(P, row, chan, g.gvalid, g.dvalid);
    if (!(g.gvalid || g.dvalid)) continue;
    count_global(P, g, record_plan(P, g));
  }
}

__global__ void k_cursors(PlanDev P) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= P.n_bins) return;
  uint32_t s = P.bin_start[b];
  P.cur_prim[b] = s;
  P.cur_dup[b] = s + P.cnt_prim[b];
}

// One store per record: 12 bytes for a primary record, 8 for a duplicate
// (degrid units cover the primary records of a bin only, so a duplicate's
// degrid word is never read).
__device__ inline void write_records(const PlanDev &P, const ItemGeo &g, const RecPlan &rp,
                                     int64_t item, uint32_t slot, int a, int b, bool primary) {
  const int u0 = a * P.tile_u, v0 = b * P.tile_v;
  uint32_t *dst = P.rec + 3 * (size_t)slot;
  const uint32_t gg = pack_geo(g.iu - u0, g.ou, g.iv - v0, g.ov);
  if (primary) {
    const uint32_t dg = rp.share ? pack_geo(g.lu - u0, g.fu, g.lv - v0, g.fv) : kInvalidGeo;
    *(RecWords *)dst = RecWords{(uint32_t)item, gg, dg};
  } else {
    *(RecGridWords *)dst = RecGridWords{(uint32_t)item, gg};
  }
}

__device__ inline void write_degrid_record(const PlanDev &P, const ItemGeo &g, const RecPlan &rp,
                                           int64_t item, uint32_t slot) {
  const int u0 = rp.du_t * P.tile_u, v0 = rp.dv_t * P.tile_v;
  *(RecWords *)(P.rec + 3 * (size_t)slot) =
      RecWords{(uint32_t)item, kInvalidGeo, pack_geo(g.lu - u0, g.fu, g.lv - v0, g.fv)};
}

// The records of one sample, their slots taken from the global bin cursors.
__device__ inline void scatter_global(const PlanDev &P, const ItemGeo &g, const RecPlan &rp,
                                      int64_t item) {
  if (rp.gprim) {
    const uint32_t slot = atomicAdd(&P.cur_prim[bin_of(P, rp.tu, rp.tv, g.ic)], 1u);
    write_records(P, g, rp, item, slot, rp.tu, rp.tv, true);
    for (int a = rp.tu0; a <= rp.tu1; ++a)
      for (int b = rp.tv0; b <= rp.tv1; ++b)
        if (a != rp.tu || b != rp.tv) {
          const uint32_t s2 = atomicAdd(&P.cur_dup[bin_of(P, a, b, g.ic)], 1u);
          write_records(P, g, rp, item, s2, a, b, false);
        }
  }
  if (rp.dsep) {
    const uint32_t slot = atomicAdd(&P.cur_prim[bin_of(P, rp.du_t, rp.dv_t, g.ic)], 1u);
    write_degrid_record(P, g, rp, item, slot);
  }
}

// Binning pass 2: write the records, taking slots from the bins' cursors.
// Threads walk the samples in order, so a wavefront covers the channels of
// one or a few rows: their bins are neighbours (bins are channel-minor) and
// so are the records they write.
__global__ __launch_bounds__(256) void k_scatter(PlanDev P) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t item = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; item < P.n_items;
       item += stride) {
    int64_t row = item;
    int chan = 0;
    if (!P.row_chan) {  // samples number below 2^32: a 32-bit division
      const uint32_t r32 = (uint32_t)item / (uint32_t)P.n_chan;
      row = r32;
      chan = (int)((uint32_t)item - r32 * (uint32_t)P.n_chan);
    }
    ItemGeo g;
    sample_geometry(P, row, chan, g);
    if (!(g.gvalid || g.dvalid)) continue;
    scatter_global(P, g, record_plan(P, g), item);
  }
}

// ---- chunked binning ------------------------------------------------------
// A chunk is rows [r0, r0 + rows_per_chunk) of one channel.  Per tile it
// counts primaries (chunk_hist) and duplicates (chunk_dup), 32 bits each.

// Chunked pass 1: per-tile counts of one chunk, gathered in LDS.
__global__ __launch_bounds__(kChunkThreads) void k_count_chunks(PlanDev P) {
  extern __shared__ uint32_t h[];  // [n_tiles] primaries, then [n_tiles] duplicates
  const int64_t w = blockIdx.x;
  const int chan = (int)(w / P.chunks_per_chan);
  const int64_t r0 = (w - (int64_t)chan * P.chunks_per_chan) * P.rows_per_chunk;
  const int ntiles = P.ntu * P.ntv;
  uint32_t *hd = h + ntiles;
  for (int i = threadIdx.x; i < 2 * ntiles; i += blockDim.x) h[i] = 0u;
  const ChanGeo cg = chan_geo(P, chan);
  __syncthreads();
  for (int64_t row = r0 + threadIdx.x; row < r0 + P.rows_per_chunk; row += blockDim.x) {
    ItemGeo g;
    g.gvalid = g.dvalid = false;
    if (row < P.n_rows) sample_geometry_at(P, row, cg, P.uvw[row * 3 + 0], P.uvw[row * 3 + 1], g);
    write_masks(P, row, chan, g.gvalid, g.dvalid);
    if (!(g.gvalid || g.dvalid)) continue;
    const RecPlan rp = record_plan(P, g);
    if (rp.gprim) {
      atomicAdd(&h[rp.tu * P.ntv + rp.tv], 1u);
      for (int a = rp.tu0; a <= rp.tu1; ++a)
        for (int b = rp.tv0; b <= rp.tv1; ++b)
          if (a != rp.tu || b != rp.tv) atomicAdd(&hd[a * P.ntv + b], 1u);
    }
    if (rp.dsep) atomicAdd(&h[rp.du_t * P.ntv + rp.dv_t], 1u);
  }
  __syncthreads();
  uint32_t *dst = P.chunk_hist + w * ntiles, *dstd = P.chunk_dup + w * ntiles;
  for (int i = threadIdx.x; i < ntiles; i += blockDim.x) {
    dst[i] = h[i];
    dstd[i] = hd[i];
  }
}

// Chunked pass 2: the counts of every bin, summed over the chunks of the
// channels that map to its image channel.
__global__ void k_chunk_totals(PlanDev P) {
  const int ntiles = P.ntu * P.ntv;
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)P.n_img_chan * ntiles) return;
  const int ic = (int)(id / ntiles);
  const int t = (int)(id - (int64_t)ic * ntiles);
  uint32_t sp = 0, sd = 0;
  for (int c = 0; c < P.n_chan; ++c) {
    if (P.chan_map[c] != ic) continue;
    for (int k = 0; k < P.chunks_per_chan; ++k) {
      const int64_t i = ((int64_t)c * P.chunks_per_chan + k) * ntiles + t;
      sp += P.chunk_hist[i];
      sd += P.chunk_dup[i];
    }
  }
  const int64_t b = (int64_t)t * P.n_img_chan + ic;
  P.cnt_prim[b] = sp;
  P.cnt_all[b] = sp + sd;
}

// Chunked pass 3 (after the bin offsets): the first slot of every chunk's
// run in every bin.  The first primary slot overwrites the primary counts
// (chunk_hist), the first duplicate slot the duplicate counts (chunk_dup).
// When image channels gather several channels (continuum), a bin takes its
// runs row-chunk by row-chunk, channel by channel within one: the channels
// of a row, which share cache lines, then lie a run apart in the bin
// instead of a whole channel's records apart.  Otherwise (one channel per
// image channel) a bin takes its channel's runs in row order.
__global__ void k_chunk_bases(PlanDev P) {
  const int ntiles = P.ntu * P.ntv;
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)P.n_img_chan * ntiles) return;
  const int ic = (int)(id / ntiles);
  const int t = (int)(id - (int64_t)ic * ntiles);
  const int64_t b = (int64_t)t * P.n_img_chan + ic;
  uint32_t sp = P.bin_start[b];
  uint32_t sd = sp + P.cnt_prim[b];
  auto take = [&](int c, int k) {
    const int64_t i = ((int64_t)c * P.chunks_per_chan + k) * ntiles + t;
    const uint32_t np = P.chunk_hist[i], nd = P.chunk_dup[i];
    P.chunk_hist[i] = sp;
    P.chunk_dup[i] = sd;
    sp += np;
    sd += nd;
  };
  if (P.n_img_chan < P.n_chan) {
    for (int k = 0; k < P.chunks_per_chan; ++k)
      for (int c = 0; c < P.n_chan; ++c)
        if (P.chan_map[c] == ic) take(c, k);
  } else {
    for (int c = 0; c < P.n_chan; ++c) {
      if (P.chan_map[c] != ic) continue;
      for (int k = 0; k < P.chunks_per_chan; ++k) take(c, k);
    }
  }
}

// Slots for one record per lane from the LDS cursors `cur` (key: the tile,
// or -1 for no record).  A lane whose key equals its left neighbour's joins
// that lane's run; the first lane of a run takes the whole run's slots with
// one LDS atomic, and the run's lanes get consecutive slots in lane order,
// so their record stores are contiguous.  Called by all 64 lanes together.
__device__ inline uint32_t run_slots(uint32_t *cur, int key) {
  if (__ballot(key >= 0) == 0ull) return 0u;
  const int lane = (int)__lane_id();
  const int left = __shfl_up(key, 1u);
  const bool start = lane == 0 || key != left;
  const unsigned long long starts = __ballot(start);
  const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  const int first = 63 - __clzll((long long)(starts & upto));
  const unsigned long long after = starts & ~upto;
  const int end = after ? __ffsll((long long)after) - 1 : 64;
  uint32_t base = 0u;
  if (start && key >= 0) base = atomicAdd(&cur[key], (uint32_t)(end - lane));
  return (uint32_t)__shfl((int)base, first) + (uint32_t)(lane - first);
}

// Chunked pass 4: a workgroup per chunk writes the chunk's records.  Its
// runs' first slots are loaded into LDS cursors, and the records take their
// slots from LDS atomics, one per run of equal tiles in a wavefront: no
// global atomics.  When the rows are time-major with order_nb baselines per
// time step, the chunk's rows are visited baseline by baseline: a
// wavefront then holds consecutive times of one baseline, which mostly fall
// in the same tile, so its stores to a bin are contiguous.  (Numbering the
// chunks so that 16, 32 or 64 consecutive chunks of a channel share an
// XCD's L2 measured the same as plain numbering.)
__global__ __launch_bounds__(kChunkThreads) void k_scatter_chunks(PlanDev P) {
  extern __shared__ uint32_t lds[];
  const int ntiles = P.ntu * P.ntv;
  uint32_t *cp = lds, *cd = lds + ntiles;  // [n_tiles] each
  const int64_t w = blockIdx.x;
  const int chan = (int)(w / P.chunks_per_chan);
  const int64_t r0 = (w - (int64_t)chan * P.chunks_per_chan) * P.rows_per_chunk;
  const int64_t r1 = min(r0 + (int64_t)P.rows_per_chunk, P.n_rows);
  const uint32_t *bp = P.chunk_hist + w * ntiles, *bd = P.chunk_dup + w * ntiles;
  for (int i = threadIdx.x; i < ntiles; i += blockDim.x) {
    cp[i] = bp[i];
    cd[i] = bd[i];
  }
  __syncthreads();
  // positions p = b * nt + t over the nt time steps the chunk touches, row
  // (t_lo + t) * nb + b; rows outside [r0, r1) are skipped (nb = 1: p = row - r0)
  const int64_t nb = P.order_nb;
  const int64_t t_lo = r0 / nb;
  const uint32_t nt = (uint32_t)((r1 - 1) / nb - t_lo + 1);
  const uint32_t n_pos = nt * (uint32_t)nb;
  const uint64_t nt_mul = ((1ull << 32) + nt - 1) / nt;  // floor(p * nt_mul / 2^32) - p / nt in {0, 1}
  const ChanGeo cg = chan_geo(P, chan);
  const uint32_t lane = __lane_id();
  for (uint32_t p0 = threadIdx.x - lane; p0 < n_pos; p0 += blockDim.x) {
    const uint32_t p = p0 + lane;
    int64_t row = -1;
    if (p < n_pos) {
      uint32_t b = (uint32_t)(((uint64_t)p * nt_mul) >> 32);
      if (b * nt > p) --b;
      const int64_t r = (t_lo + (int64_t)(p - b * nt)) * nb + b;
      if (r >= r0 && r < r1) row = r;
    }
    ItemGeo g;
    g.gvalid = g.dvalid = false;
    if (row >= 0) sample_geometry_at(P, row, cg, P.uvw[row * 3 + 0], P.uvw[row * 3 + 1], g);
    const RecPlan rp = record_plan(P, g);
    const int64_t item = row * P.n_chan + chan;
    const uint32_t slot = run_slots(cp, rp.gprim ? rp.tu * P.ntv + rp.tv : -1);
    if (rp.gprim) write_records(P, g, rp, item, slot, rp.tu, rp.tv, true);
    // duplicates: the other tiles of the footprint's (at most) 2 x 2 (a
    // tile is at least `support` wide), in a fixed order so that
    // neighbouring lanes' runs line up
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int a = (k & 2) ? rp.tu1 : rp.tu0, b = (k & 1) ? rp.tv1 : rp.tv0;
      const bool dup = rp.gprim && rp.tu0 <= rp.tu1 && rp.tv0 <= rp.tv1 &&
                       (!(k & 2) || rp.tu1 != rp.tu0) && (!(k & 1) || rp.tv1 != rp.tv0) &&
                       (a != rp.tu || b != rp.tv);
      const uint32_t s2 = run_slots(cd, dup ? a * P.ntv + b : -1);
      if (dup) write_records(P, g, rp, item, s2, a, b, false);
    }
    if (rp.dsep) {
      const uint32_t s3 = atomicAdd(&cp[rp.du_t * P.ntv + rp.dv_t], 1u);
      write_degrid_record(P, g, rp, item, s3);
    }
  }
}

// ---- row-block binning (rows carry their own channel) ---------------------
// A workgroup takes kRowBlockRows consecutive rows.  Rows of one plane are
// normally stored together, so the records of the block's "own" image
// channel (that of its first row) are gathered per tile in LDS and reach the
// global bins with one atomic per tile the block touches; the records of any
// other image channel take the global path.  This keeps the hot central
// tiles of a plane from being hit by every wavefront on the chip at once.

// Image channel of a row, or -1 when the row's channel is out of range.
__device__ inline int row_image_channel(const PlanDev &P, int64_t row) {
  if (row >= P.n_rows) return -1;
  const int c = P.row_chan[row];
  if (c < 0 || c >= P.n_chan) return -1;
  const int ic = P.chan_map[c];
  return (ic >= 0 && ic < P.n_img_chan) ? ic : -1;
}

// Row-block pass 1: per-tile counts, (primaries) | (duplicates << 16); a row
// adds at most two primaries and one duplicate to a tile and a block has
// kRowBlockRows < 32768 rows, so neither half overflows.
__global__ __launch_bounds__(kChunkThreads) void k_count_rows(PlanDev P) {
  extern __shared__ uint32_t h[];  // [n_tiles]
  const int64_t r0 = (int64_t)blockIdx.x * kRowBlockRows;
  const int ntiles = P.ntu * P.ntv;
  const int own = row_image_channel(P, r0);
  for (int i = threadIdx.x; i < ntiles; i += blockDim.x) h[i] = 0u;
  __syncthreads();
  for (int64_t row = r0 + threadIdx.x; row < r0 + kRowBlockRows; row += blockDim.x) {
    ItemGeo g;
    g.gvalid = g.dvalid = false;
    if (row < P.n_rows) sample_geometry(P, row, 0, g);
    write_masks(P, row, 0, g.gvalid, g.dvalid);
    if (!(g.gvalid || g.dvalid)) continue;
    const RecPlan rp = record_plan(P, g);
    if (g.ic != own) {
      count_global(P, g, rp);
      continue;
    }
    if (rp.gprim) {
      atomicAdd(&h[rp.tu * P.ntv + rp.tv], 1u);
      for (int a = rp.tu0; a <= rp.tu1; ++a)
        for (int b = rp.tv0; b <= rp.tv1; ++b)
          if (a != rp.tu || b != rp.tv) atomicAdd(&h[a * P.ntv + b], 1u << 16);
    }
    if (rp.dsep) atomicAdd(&h[rp.du_t * P.ntv + rp.dv_t], 1u);
  }
  __syncthreads();
  if (own < 0) return;
  uint32_t *keep = P.row_hist + (int64_t)blockIdx.x * ntiles;
  for (int i = threadIdx.x; i < ntiles; i += blockDim.x) {
    const uint32_t hv = h[i];
    keep[i] = hv;
    if (!hv) continue;
    const int64_t b = (int64_t)i * P.n_img_chan + own;
    const uint32_t np = hv & 0xFFFFu;
    if (np) atomicAdd(&P.cnt_prim[b], np);
    atomicAdd(&P.cnt_all[b], np + (hv >> 16));
  }
}

// Row-block pass 2: the block loads its own-channel counts per tile (kept by
// pass 1), reserves one run of slots per tile from each bin cursor
// (primaries and duplicates), and hands out the slots of each run from LDS
// cursors.
__global__ __launch_bounds__(kChunkThreads) void k_scatter_rows(PlanDev P) {
  extern __shared__ uint32_t lds[];
  const int ntiles = P.ntu * P.ntv;
  uint32_t *cp = lds, *cd = lds + ntiles;  // [n_tiles] each
  const int64_t r0 = (int64_t)blockIdx.x * kRowBlockRows;
  const int64_t r1 = min(r0 + (int64_t)kRowBlockRows, P.n_rows);
  const int own = row_image_channel(P, r0);
  if (own >= 0) {
    const uint32_t *kept = P.row_hist + (int64_t)blockIdx.x * ntiles;
    for (int i = threadIdx.x; i < ntiles; i += blockDim.x) {
      const uint32_t hv = kept[i];
      cp[i] = hv & 0xFFFFu;
      cd[i] = hv >> 16;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < ntiles; i += blockDim.x) {
      const int64_t b = (int64_t)i * P.n_img_chan + own;
      if (cp[i]) cp[i] = atomicAdd(&P.cur_prim[b], cp[i]);
      if (cd[i]) cd[i] = atomicAdd(&P.cur_dup[b], cd[i]);
    }
    __syncthreads();
  }
  for (int64_t row = r0 + threadIdx.x; row < r1; row += blockDim.x) {
    ItemGeo g;
    sample_geometry(P, row, 0, g);
    if (!(g.gvalid || g.dvalid)) continue;
    const RecPlan rp = record_plan(P, g);
    if (g.ic != own) {
      scatter_global(P, g, rp, row);
      continue;
    }
    if (rp.gprim) {
      const uint32_t slot = atomicAdd(&cp[rp.tu * P.ntv + rp.tv], 1u);
      write_records(P, g, rp, row, slot, rp.tu, rp.tv, true);
      for (int a = rp.tu0; a <= rp.tu1; ++a)
        for (int b = rp.tv0; b <= rp.tv1; ++b)
          if (a != rp.tu || b != rp.tv) {
            const uint32_t s2 = atomicAdd(&cd[a * P.ntv + b], 1u);
            write_records(P, g, rp, row, s2, a, b, false);
          }
    }
    if (rp.dsep) {
      const uint32_t slot = atomicAdd(&cp[rp.du_t * P.ntv + rp.dv_t], 1u);
      write_degrid_record(P, g, rp, row, slot);
    }
  }
}

// units per bin: grid units cover every bin (an empty bin still owns its
// tile), degrid units only the primary records.
__global__ void k_unit_counts(PlanDev P, uint32_t *nug, uint32_t *nud) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= P.n_bins) return;
  uint32_t cap = (uint32_t)P.unit_cap;
  uint32_t na = P.cnt_all[b], np = P.cnt_prim[b];
  nug[b] = na == 0 ? 1u : (na + cap - 1) / cap;
  nud[b] = (np + cap - 1) / cap;
}

__global__ void k_units(PlanDev P) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= P.n_bins) return;
  uint32_t cap = (uint32_t)P.unit_cap;
  uint32_t s = P.bin_start[b];
  uint32_t na = P.cnt_all[b], np = P.cnt_prim[b];
  uint32_t g0 = P.gunit_start[b], g1 = P.gunit_start[b + 1];
  uint32_t ng = g1 - g0;
  for (uint32_t k = 0; k < ng; ++k) {
    uint32_t first = s + k * cap;
    uint32_t cnt = na > k * cap ? min(cap, na - k * cap) : 0u;
    uint32_t fl = (ng > 1 ? kUnitMulti : 0u) | (k == 0 ? kUnitFirst : 0u);
    P.gunits[g0 + k] = make_uint4((uint32_t)b, first, cnt, fl);
  }
  uint32_t d0 = P.dunit_start[b], d1 = P.dunit_start[b + 1];
  for (uint32_t k = 0; k < d1 - d0; ++k) {
    uint32_t first = s + k * cap;
    uint32_t cnt = min(cap, np - k * cap);
    P.dunits[d0 + k] = make_uint4((uint32_t)b, first, cnt, 0u);
  }
}

// Pair units: the bins of tiles 2j and 2j + 1 (along v) of pair p, and the
// primary records of both.
__device__ inline void pair_bins(const PlanDev &P, int64_t p, int64_t &b0, bool &has1,
                                 uint32_t &n) {
  const int ntv2 = (P.ntv + 1) / 2;
  const int64_t tp = p / P.n_img_chan;
  const int ic = (int)(p - tp * P.n_img_chan);
  const int tu = (int)(tp / ntv2), tv = 2 * (int)(tp - (int64_t)tu * ntv2);
  b0 = ((int64_t)tu * P.ntv + tv) * P.n_img_chan + ic;
  has1 = tv + 1 < P.ntv;
  n = P.cnt_prim[b0] + (has1 ? P.cnt_prim[b0 + P.n_img_chan] : 0u);
}

__global__ void k_pair_counts(PlanDev P, uint32_t *npu) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P.n_pairs) return;
  int64_t b0;
  bool has1;
  uint32_t n;
  pair_bins(P, p, b0, has1, n);
  const uint32_t cap = (uint32_t)P.unit_cap;
  npu[p] = (n + cap - 1) / cap;
}

__global__ void k_pair_units(PlanDev P) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P.n_pairs) return;
  int64_t b0;
  bool has1;
  uint32_t n;
  pair_bins(P, p, b0, has1, n);
  const uint32_t cap = (uint32_t)P.unit_cap;
  const uint32_t k0 = P.punit_start[p], nu = P.punit_start[p + 1] - k0;
  const uint32_t b1 = has1 ? (uint32_t)(b0 + P.n_img_chan) : kNoBin;
  for (uint32_t k = 0; k < nu; ++k)
    P.punits[k0 + k] = make_uint4((uint32_t)b0, k * cap, min(cap, n - k * cap), b1);
}

// per-call tables: the FP32 kernel table; for every kernel offset, the FP64
// sum of the support taps of one axis (norm = nrm[ou] * nrm[ov]); and the
// zero-padded weight table by (tap, offset) in FP32 and FP64.  Offsets past
// ov/2 (never produced by a sample) read 0 beyond the end of the table.
__global__ void k_prepare_tables(PlanDev P, const double *cgk) {
  const int sc = P.support / 2, ov = P.ov;
  const int ntab = ov * (sc + 1), half = ov / 2 + 2, nnrm = 2 * half + 1;
  const int nw = (P.support + 2) * nnrm;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ntab + nnrm + nw;
       i += gridDim.x * blockDim.x) {
    if (i < ntab) {
      P.tab32[i] = (float)cgk[i];
    } else if (i < ntab + nnrm) {
      const int off = i - ntab - half;
      double s = 0.0;
      for (int t = 0; t < P.support; ++t) {
        int k = ov * (t - sc) + off;
        k = k < 0 ? -k : k;
        s += k < ntab ? cgk[k] : 0.0;
      }
      P.nrm[i - ntab] = s;
    } else {
      const int j = i - ntab - nnrm;
      const int tc = j / nnrm - 1, off = j - (j / nnrm) * nnrm - half;
      double v = 0.0;
      if (tc >= 0 && tc < P.support) {
        int k = ov * (tc - sc) + off;
        k = k < 0 ? -k : k;
        v = k < ntab ? cgk[k] : 0.0;
      }
      P.wtab32[j] = (float)v;
      P.wtab64[j] = v;
    }
  }
}

int prepare_tables(const PlanDev &d, const double *cgk, hipStream_t s) {
  hipLaunchKernelGGL(k_prepare_tables, dim3(4), dim3(256), 0, s, d, cgk);
  AV_LAUNCH_CHECK("k_prepare_tables");
  return AV_OK;
}

__global__ void k_popcount(const unsigned long long *m, int64_t n, uint32_t *out) {
  uint32_t c = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    c += (uint32_t)__popcll(m[i]);
  for (int off = 32; off >= 1; off >>= 1) c += __shfl_xor(c, off);
  if (__lane_id() == 0 && c) atomicAdd(out, c);
}

}  // namespace av

using namespace av;

static size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

// Validate the plan parameters, fill plan->d's sizes and compute the bytes
// of device memory the plan's arrays need; with base non-NULL, also point
// the arrays into [base, base + bytes).  Shared by av_plan_create (one
// hipMalloc) and av_grid_noalloc (the library's static scratch array).
int plan_layout(int64_t n_rows, int32_t n_chan, int32_t n_img_chan, int32_t n_u, int32_t n_v,
                double cell_u, double cell_v, int32_t support, int32_t oversampling,
                int32_t tile_u, int32_t tile_v, int32_t geometry, int32_t unit_cap, av_plan *p,
                char *base, size_t *bytes) {
  if (n_rows < 0 || n_chan < 1 || n_img_chan < 1 || n_u < 1 || n_v < 1)
    return fail(AV_ERR_INVALID, "av_plan_create: sizes must be positive");
  // plans bin footprints up to kMaxPlanSupport wide (the mosaic gridder's
  // max_support); av_grid and av_degrid take supports up to
  // kMaxSupport
  if (support < 2 || support > kMaxPlanSupport)
    return fail(AV_ERR_UNSUPPORTED, "av_plan_create: support must be in [2, 63]");
  if (oversampling < 1 || oversampling > 250)
    return fail(AV_ERR_UNSUPPORTED, "av_plan_create: oversampling must be in [1, 250]");
  if (tile_u == 0) tile_u = 32;
  if (tile_v == 0) tile_v = 32;
  if (tile_u < support || tile_v < support || tile_u > 128 || tile_v > 128)
    return fail(AV_ERR_INVALID, "av_plan_create: tiles must be in [support, 128]");
  if (n_u >= (1 << 24) || n_v >= (1 << 24))
    return fail(AV_ERR_UNSUPPORTED, "av_plan_create: grid sides must be below 2^24");
  const bool row_channels = (geometry & AV_PLAN_ROW_CHANNELS) != 0;
  const bool conjugate = (geometry & AV_GEOM_CONJUGATE) != 0;
  geometry &= ~(AV_PLAN_ROW_CHANNELS | AV_GEOM_CONJUGATE);
  if (geometry != AV_GEOM_REFERENCE && geometry != AV_GEOM_DEGRID)
    return fail(AV_ERR_INVALID, "av_plan_create: unknown geometry");
  if (conjugate && geometry != AV_GEOM_REFERENCE)
    return fail(AV_ERR_INVALID, "av_plan_create: AV_GEOM_CONJUGATE needs AV_GEOM_REFERENCE");
  if (unit_cap == 0) unit_cap = 8192;
  if (unit_cap < 64) return fail(AV_ERR_INVALID, "av_plan_create: unit_cap < 64");
  int64_t n_items = row_channels ? n_rows : n_rows * (int64_t)n_chan;
  if (n_items > ((int64_t)1 << 32) / 5 - 1)
    return fail(AV_ERR_UNSUPPORTED, "av_plan_create: more than 858993458 samples per plan");

  PlanDev &d = p->d;
  d.n_rows = n_rows;
  d.n_chan = n_chan;
  d.n_img_chan = n_img_chan;
  d.n_u = n_u;
  d.n_v = n_v;
  d.cell_u = cell_u;
  d.cell_v = cell_v;
  // degrid_spheroid_ms4: scale = -pixelincr * n, offset = n / 2 (float)
  d.dscale_u = (-cell_u) * (double)n_u;
  d.dscale_v = (-cell_v) * (double)n_v;
  d.doff_u = (double)n_u / 2.0;
  d.doff_v = (double)n_v / 2.0;
  d.support = support;
  d.ov = oversampling;
  d.tile_u = tile_u;
  d.tile_v = tile_v;
  d.tile_u_mul = 0xFFFFFFFFu / (uint32_t)tile_u + 1u;
  d.tile_v_mul = 0xFFFFFFFFu / (uint32_t)tile_v + 1u;
  d.geometry = geometry;
  d.conjugate = conjugate ? 1 : 0;
  d.row_channels = row_channels;
  d.unit_cap = unit_cap;
  d.ntu = (n_u + tile_u - 1) / tile_u;
  d.ntv = (n_v + tile_v - 1) / tile_v;
  d.n_bins = (int64_t)d.ntu * d.ntv * n_img_chan;
  d.n_items = n_items;
  // a footprint overlaps at most 2x2 tiles, plus one separate degrid record
  d.max_records = 5 * n_items;
  d.max_gunits = d.n_bins + d.max_records / unit_cap + 1;
  // a sample has at most two primary records (grid + separate degrid)
  d.max_dunits = d.n_bins + 2 * n_items / unit_cap + 1;
  const bool pairs = tile_u == 16 && tile_v == 16 && support <= 7;
  d.n_pairs = pairs ? (int64_t)d.ntu * ((d.ntv + 1) / 2) * n_img_chan : 0;
  d.max_punits = pairs ? d.n_pairs + 2 * n_items / unit_cap + 1 : 0;
  d.row_chan = nullptr;
  d.order_nb = 1;
  d.uvw = nullptr;

  const int64_t nb = d.n_bins;
  int64_t scan_blocks = (std::max<int64_t>(nb, n_items) + 4095) / 4096 + 1;
  d.mask_words_per_chan = (n_rows + 63) / 64;
  const int64_t mask_words = d.mask_words_per_chan * (row_channels ? 1 : n_chan) + 1;
  // chunked binning when a channel's tiles fit in LDS
  const int64_t ntiles = (int64_t)d.ntu * d.ntv;
  d.chunked = !row_channels && ntiles <= kMaxChunkTiles && n_rows > 0;
  d.rows_per_chunk = (int32_t)std::min<int64_t>(
      kMaxRowsPerChunk, std::max<int64_t>(kChunkThreads,
                                          (n_rows + kChunkThreads - 1) / kChunkThreads * kChunkThreads));
  d.chunks_per_chan = (int32_t)((n_rows + d.rows_per_chunk - 1) / d.rows_per_chunk);
  const int64_t chunk_cells = d.chunked ? (int64_t)n_chan * d.chunks_per_chan * ntiles : 0;
  // row-block counts, kept for plans whose rows carry their channel when a
  // block's counters and cursors fit in 160 KB of LDS
  const int64_t row_cells = (row_channels && n_rows > 0 && 8 * ntiles <= 160 * 1024)
                                ? (n_rows + kRowBlockRows - 1) / kRowBlockRows * ntiles
                                : 0;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off += align_up(bytes);
    return o;
  };
  size_t o_gsu = take(sizeof(double) * n_chan), o_gsv = take(sizeof(double) * n_chan),
         o_freq = take(sizeof(double) * n_chan), o_cmap = take(sizeof(int32_t) * n_chan),
         o_call = take(4 * nb), o_cprim = take(4 * nb), o_bstart = take(4 * (nb + 1)),
         o_curp = take(4 * nb), o_curd = take(4 * nb), o_gus = take(4 * (nb + 1)),
         o_dus = take(4 * (nb + 1)), o_scan = take(4 * scan_blocks),
         o_rec = take(12 * std::max<int64_t>(d.max_records, 1)), o_gun = take(sizeof(uint4) * d.max_gunits),
         o_dun = take(sizeof(uint4) * d.max_dunits), o_pus = take(4 * (d.n_pairs + 1)),
         o_pun = take(sizeof(uint4) * d.max_punits),
         o_gm = take(8 * mask_words), o_dm = take(8 * mask_words),
         o_chh = take(4 * chunk_cells), o_chd = take(4 * chunk_cells), o_rh = take(4 * row_cells),
         o_tab = take(sizeof(float) * oversampling * (support / 2 + 1)),
         o_nrm = take(sizeof(double) * (2 * (oversampling / 2 + 2) + 1)),
         o_wt32 = take(sizeof(float) * (support + 2) * (2 * (oversampling / 2 + 2) + 1)),
         o_wt64 = take(sizeof(double) * (support + 2) * (2 * (oversampling / 2 + 2) + 1)),
         o_upart = take(sizeof(double) * 4 * d.max_gunits), o_ctr = take(4 * 16);
  *bytes = off;
  if (!base) return AV_OK;
  char *c = base;
  d.gscale_u = (double *)(c + o_gsu);
  d.gscale_v = (double *)(c + o_gsv);
  d.freq = (double *)(c + o_freq);
  d.chan_map = (int32_t *)(c + o_cmap);
  d.cnt_all = (uint32_t *)(c + o_call);
  d.cnt_prim = (uint32_t *)(c + o_cprim);
  d.bin_start = (uint32_t *)(c + o_bstart);
  d.cur_prim = (uint32_t *)(c + o_curp);
  d.cur_dup = (uint32_t *)(c + o_curd);
  d.gunit_start = (uint32_t *)(c + o_gus);
  d.dunit_start = (uint32_t *)(c + o_dus);
  d.scan_tmp = (uint32_t *)(c + o_scan);
  d.rec = (uint32_t *)(c + o_rec);
  d.gunits = (uint4 *)(c + o_gun);
  d.dunits = (uint4 *)(c + o_dun);
  d.punit_start = d.n_pairs ? (uint32_t *)(c + o_pus) : nullptr;
  d.punits = d.n_pairs ? (uint4 *)(c + o_pun) : nullptr;
  d.gmask = (unsigned long long *)(c + o_gm);
  d.dmask = (unsigned long long *)(c + o_dm);
  d.n_words = mask_words - 1;
  d.chunk_hist = chunk_cells ? (uint32_t *)(c + o_chh) : nullptr;
  d.chunk_dup = chunk_cells ? (uint32_t *)(c + o_chd) : nullptr;
  d.row_hist = row_cells ? (uint32_t *)(c + o_rh) : nullptr;
  d.tab32 = (float *)(c + o_tab);
  d.nrm = (double *)(c + o_nrm);
  d.wtab32 = (float *)(c + o_wt32);
  d.wtab64 = (double *)(c + o_wt64);
  d.upart = (double *)(c + o_upart);
  d.counters = (uint32_t *)(c + o_ctr);
  p->alloc_base = base;
  p->alloc_bytes = off;
  return AV_OK;
}

extern "C" {

int av_plan_create(int64_t n_rows, int32_t n_chan, int32_t n_img_chan, int32_t n_u, int32_t n_v,
                   double cell_u, double cell_v, int32_t support, int32_t oversampling,
                   int32_t tile_u, int32_t tile_v, int32_t geometry, int32_t unit_cap,
                   av_plan **out) {
  if (!out) return fail(AV_ERR_INVALID, "av_plan_create: out is NULL");
  *out = nullptr;
  av_plan probe{};
  size_t bytes = 0;
  int rc = plan_layout(n_rows, n_chan, n_img_chan, n_u, n_v, cell_u, cell_v, support,
                       oversampling, tile_u, tile_v, geometry, unit_cap, &probe, nullptr, &bytes);
  if (rc) return rc;
  void *base = nullptr;
  hipError_t e = hipMalloc(&base, bytes);
  if (e != hipSuccess) return hip_fail(e, "av_plan_create: hipMalloc");
  av_plan *p = new av_plan();
  rc = plan_layout(n_rows, n_chan, n_img_chan, n_u, n_v, cell_u, cell_v, support, oversampling,
                   tile_u, tile_v, geometry, unit_cap, p, (char *)base, &bytes);
  if (rc == AV_OK) {
    e = hipGetDevice(&p->device);
    if (e != hipSuccess) rc = hip_fail(e, "av_plan_create: hipGetDevice");
  }
  if (rc) {
    (void)hipFree(base);
    delete p;
    return rc;
  }
  *out = p;
  return AV_OK;
}

int av_plan_destroy(av_plan *plan) {
  if (!plan) return AV_OK;
  const hipError_t e = plan->alloc_base ? hipFree(plan->alloc_base) : hipSuccess;
  delete plan;
  return e == hipSuccess ? AV_OK : hip_fail(e, "av_plan_destroy: hipFree");
}

int av_plan_set_uvw(av_plan *plan, const double *uvw, const double *freq,
                    const int32_t *chan_map, const int32_t *row_chan, void *stream) {
  if (!plan || !freq || (plan->d.n_rows > 0 && !uvw))
    return fail(AV_ERR_INVALID, "av_plan_set_uvw: NULL argument");
  if (row_chan && plan->d.n_chan < 1)
    return fail(AV_ERR_INVALID, "av_plan_set_uvw: row_chan needs n_chan >= 1");
  if (plan->d.row_channels && !row_chan)
    return fail(AV_ERR_INVALID,
                "av_plan_set_uvw: the plan was created for per-row channels; row_chan is NULL");
  hipStream_t s = (hipStream_t)stream;
  PlanDev &d = plan->d;
  d.uvw = uvw;
  d.row_chan = (int32_t *)row_chan;
  if (row_chan) d.n_items = d.n_rows;  // one sample per row
  else d.n_items = d.n_rows * (int64_t)d.n_chan;
  d.n_words = d.mask_words_per_chan * (row_chan ? 1 : d.n_chan);
  hipLaunchKernelGGL(k_chan_setup, dim3((d.n_chan + 255) / 256), dim3(256), 0, s, d, freq,
                     chan_map);
  AV_LAUNCH_CHECK("k_chan_setup");
  int rc;
  const unsigned bb = (unsigned)((d.n_bins + 255) / 256);
  // row-block binning needs a counter and a cursor per tile in LDS
  const int64_t ntiles = (int64_t)d.ntu * d.ntv;
  d.row_blocks = 0;
  d.chunk_scatter = 0;
  if (d.n_rows > 0 && (row_chan || d.chunked)) {
    const DeviceProps &prop = device_props();
    if (prop.error) return hip_fail(prop.error, "av_plan_set_uvw: device properties");
    const bool fits = (size_t)(2 * sizeof(uint32_t) * ntiles) <= prop.lds_per_block;
    if (row_chan) d.row_blocks = fits && d.row_hist != nullptr;
    else d.chunk_scatter = fits;
  }
  const int64_t nchunks = (int64_t)d.n_chan * d.chunks_per_chan;
  const unsigned row_grid = (unsigned)((d.n_rows + kRowBlockRows - 1) / kRowBlockRows);
  // pass 1: counts per bin
  if (d.row_blocks) {
    if ((rc = launch_fill_u32(d.cnt_all, 0u, d.n_bins, s))) return rc;
    if ((rc = launch_fill_u32(d.cnt_prim, 0u, d.n_bins, s))) return rc;
    const int lds = (int)(sizeof(uint32_t) * ntiles);
    if ((rc = ensure_lds_limit<k_count_rows>(lds, "av_plan_set_uvw: hipFuncSetAttribute")))
      return rc;
    hipLaunchKernelGGL(k_count_rows, dim3(row_grid), dim3(kChunkThreads), lds, s, d);
    AV_LAUNCH_CHECK("k_count_rows");
  } else if (d.chunked && !row_chan) {
    const int lds = (int)(2 * sizeof(uint32_t) * d.ntu * d.ntv);
    if ((rc = ensure_lds_limit<k_count_chunks>(lds, "av_plan_set_uvw: hipFuncSetAttribute")))
      return rc;
    hipLaunchKernelGGL(k_count_chunks, dim3((unsigned)nchunks), dim3(kChunkThreads), lds, s, d);
    AV_LAUNCH_CHECK("k_count_chunks");
    const int64_t pairs = (int64_t)d.n_img_chan * d.ntu * d.ntv;
    hipLaunchKernelGGL(k_chunk_totals, dim3((unsigned)((pairs + 255) / 256)), dim3(256), 0, s,
                       d);
    AV_LAUNCH_CHECK("k_chunk_totals");
  } else {
    if ((rc = launch_fill_u32(d.cnt_all, 0u, d.n_bins, s))) return rc;
    if ((rc = launch_fill_u32(d.cnt_prim, 0u, d.n_bins, s))) return rc;
    const int64_t span = d.mask_words_per_chan * 64 * (row_chan ? 1 : d.n_chan);
    int64_t blocks = (span + 255) / 256;
    blocks = std::max<int64_t>(1, std::min<int64_t>(blocks, 16384));
    hipLaunchKernelGGL(k_count, dim3((unsigned)blocks), dim3(256), 0, s, d);
    AV_LAUNCH_CHECK("k_count");
  }
  // pass 2: bin offsets, cursors, records
  if ((rc = scan_exclusive_u32(d.cnt_all, d.bin_start, d.n_bins, d.scan_tmp, s))) return rc;
  if (d.chunk_scatter) {
    const int64_t pairs = (int64_t)d.n_img_chan * d.ntu * d.ntv;
    hipLaunchKernelGGL(k_chunk_bases, dim3((unsigned)((pairs + 255) / 256)), dim3(256), 0, s, d);
    AV_LAUNCH_CHECK("k_chunk_bases");
    const int lds = (int)(2 * sizeof(uint32_t) * ntiles);
    if ((rc = ensure_lds_limit<k_scatter_chunks>(lds, "av_plan_set_uvw: hipFuncSetAttribute")))
      return rc;
    hipLaunchKernelGGL(k_scatter_chunks, dim3((unsigned)nchunks), dim3(kChunkThreads), lds, s, d);
    AV_LAUNCH_CHECK("k_scatter_chunks");
  } else if (d.row_blocks) {
    hipLaunchKernelGGL(k_cursors, dim3(bb), dim3(256), 0, s, d);
    AV_LAUNCH_CHECK("k_cursors");
    const int lds = (int)(2 * sizeof(uint32_t) * ntiles);
    if ((rc = ensure_lds_limit<k_scatter_rows>(lds, "av_plan_set_uvw: hipFuncSetAttribute")))
      return rc;
    hipLaunchKernelGGL(k_scatter_rows, dim3(row_grid), dim3(kChunkThreads), lds, s, d);
    AV_LAUNCH_CHECK("k_scatter_rows");
  } else {
    hipLaunchKernelGGL(k_cursors, dim3(bb), dim3(256), 0, s, d);
    AV_LAUNCH_CHECK("k_cursors");
    int64_t blocks = (d.n_items + 255) / 256;
    blocks = std::max<int64_t>(1, std::min<int64_t>(blocks, 16384));
    hipLaunchKernelGGL(k_scatter, dim3((unsigned)blocks), dim3(256), 0, s, d);
    AV_LAUNCH_CHECK("k_scatter");
  }
  // cursors are free again: reuse them for the per-bin unit counts
  hipLaunchKernelGGL(k_unit_counts, dim3(bb), dim3(256), 0, s, d, d.cur_prim, d.cur_dup);
  AV_LAUNCH_CHECK("k_unit_counts");
  if ((rc = scan_exclusive_u32(d.cur_prim, d.gunit_start, d.n_bins, d.scan_tmp, s))) return rc;
  if ((rc = scan_exclusive_u32(d.cur_dup, d.dunit_start, d.n_bins, d.scan_tmp, s))) return rc;
  hipLaunchKernelGGL(k_units, dim3(bb), dim3(256), 0, s, d);
  AV_LAUNCH_CHECK("k_units");
  if (d.n_pairs > 0) {
    // the bins' cursors are free again: the pairs' unit counts
    const unsigned pb = (unsigned)((d.n_pairs + 255) / 256);
    hipLaunchKernelGGL(k_pair_counts, dim3(pb), dim3(256), 0, s, d, d.cur_prim);
    AV_LAUNCH_CHECK("k_pair_counts");
    if ((rc = scan_exclusive_u32(d.cur_prim, d.punit_start, d.n_pairs, d.scan_tmp, s))) return rc;
    hipLaunchKernelGGL(k_pair_units, dim3(pb), dim3(256), 0, s, d);
    AV_LAUNCH_CHECK("k_pair_units");
  }
  return AV_OK;
}

int av_plan_set_degrid_transform(av_plan *plan, double scale_u, double scale_v,
                                 double offset_u, double offset_v) {
  if (!plan) return fail(AV_ERR_INVALID, "av_plan_set_degrid_transform: NULL plan");
  plan->d.dscale_u = scale_u;
  plan->d.dscale_v = scale_v;
  plan->d.doff_u = offset_u;
  plan->d.doff_v = offset_v;
  return AV_OK;
}

int av_plan_set_row_order(av_plan *plan, int32_t n_baseline) {
  if (!plan) return fail(AV_ERR_INVALID, "av_plan_set_row_order: NULL plan");
  PlanDev &d = plan->d;
  if (n_baseline < 1 || (d.n_rows > 0 && d.n_rows % n_baseline != 0))
    return fail(AV_ERR_INVALID, "av_plan_set_row_order: n_baseline must be >= 1 and divide n_rows");
  // a chunk is visited baseline by baseline only when it spans at least 16
  // time steps; otherwise the runs per bin stay short and row order is kept
  d.order_nb = (n_baseline > 1 && (int64_t)n_baseline * 16 <= d.rows_per_chunk) ? n_baseline : 1;
  return AV_OK;
}

int av_plan_stats(av_plan *plan, int64_t out[8], void *stream) {
  if (!plan || !out) return fail(AV_ERR_INVALID, "av_plan_stats: NULL argument");
  hipStream_t s = (hipStream_t)stream;
  const PlanDev &d = plan->d;
  int rc = launch_fill_u32(d.counters, 0u, 2, s);
  if (rc) return rc;
  if (d.n_words > 0) {
    hipLaunchKernelGGL(k_popcount, dim3(256), dim3(256), 0, s, d.gmask, d.n_words, d.counters);
    AV_LAUNCH_CHECK("k_popcount");
    hipLaunchKernelGGL(k_popcount, dim3(256), dim3(256), 0, s, d.dmask, d.n_words,
                       d.counters + 1);
    AV_LAUNCH_CHECK("k_popcount");
  }
  uint32_t ctr[2], tot[3];
  AV_HIP(hipMemcpyAsync(ctr, d.counters, sizeof(ctr), hipMemcpyDeviceToHost, s));
  AV_HIP(hipMemcpyAsync(&tot[0], d.bin_start + d.n_bins, 4, hipMemcpyDeviceToHost, s));
  AV_HIP(hipMemcpyAsync(&tot[1], d.gunit_start + d.n_bins, 4, hipMemcpyDeviceToHost, s));
  AV_HIP(hipMemcpyAsync(&tot[2], d.dunit_start + d.n_bins, 4, hipMemcpyDeviceToHost, s));
  AV_HIP(hipStreamSynchronize(s));
  out[0] = d.n_items;
  out[1] = ctr[0];
  out[2] = ctr[1];
  out[3] = tot[0];
  out[4] = d.n_bins;
  out[5] = tot[1];
  out[6] = tot[2];
  out[7] = d.n_items - ctr[1];  // samples no degrid unit visits
  return AV_OK;
}

}  // extern "C"
