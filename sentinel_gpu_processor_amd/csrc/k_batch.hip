// Batched burst pairs: every stage of the pair chain runs once for all pairs
// of the batch, as a persistent kernel walking a flattened work-item space
// (rows, tiles or patches of every pair, pairs of uneven size back to back).
// A work item is mapped to its pair by a binary search over the stage's
// prefix array.  The per-item device code is the one the single-pair entry
// points run (geom_dev.h, gold_dev.h, gold_band_dev.h, ifg_dev.h), so each
// pair's products are bit-identical to those entry points.
//
// The launch sequence can be captured once into a HIP graph (on the batch's
// own stream) and replayed: every kernel writes all of its outputs, the
// work-queue counters are zeroed by a kernel (no memset), and the raw bursts
// are never written.
#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

#include "geom_dev.h"
#include "gold_dev.h"
#include "gold_band_dev.h"
#include "ifg_dev.h"
#include "pair_lds_dev.h"

namespace s1ifg {

int goldstein_grid(int nrow, int ncol, int *pi_n, int *pj_n);

namespace {

// Descriptors and prefix arrays are read through the constant address space
// (scalar loads into SGPRs, not vector registers).
#define S1C __attribute__((address_space(4)))

struct PairDev {
  const S1C GeomDev *gr, *gs;
  OffPoly op;
  GridDev grid;
  DemDev dem;
  const float *dh;
  const float2 *ref_d, *sec_d;
  RowTab *rows;
  ColTab *cols;
  float2 *ifg;
  float *pr, *ps;
  float2 *patches, *filt;
  float *coh;
  int nr_ml, nc_ml, pi_n, pj_n;
};

// One burst to deramp (bursts shared by several pairs of a stack appear once).
struct ImgDev {
  const S1C GeomDev *g;
  const float2 *in;
  float2 *out;
};

enum { kStTables, kStDeramp, kStPair, kStGold, kStGather, kStCoh, kNumStages };

__device__ __forceinline__ const GeomDev &geo(const S1C GeomDev *g) { return *(const GeomDev *)g; }

// Largest p in [0, n) with off[p] <= idx, for 0 <= idx < off[n]; hint is
// checked first (the previous item's entry: consecutive items of a block
// usually stay in one pair).
__device__ __forceinline__ int find_item(const S1C long *__restrict__ off, int n, long idx,
                                         int hint = -1) {
  if (hint >= 0 && off[hint] <= idx && idx < off[hint + 1]) return hint;
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (off[mid] <= idx)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Grid tables of every pair; one thread per table entry.
__global__ void kb_tables(const S1C PairDev *__restrict__ P, const S1C long *__restrict__ off, int n) {
  const long total = off[n];
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long)gridDim.x * blockDim.x) {
    int p = find_item(off, n, t);
    const PairDev &d = *(const PairDev *)&P[p];
    grid_table_entry(d.grid, d.dem, (int)(t - off[p]), d.rows, d.cols);
  }
}

// Deramp of every distinct burst of the batch (deramp_tile).  Work item:
// kDerampRows burst lines x kDerampChunk sample pairs (one pair per thread).
constexpr int kDerampChunk = 256;

__global__ __launch_bounds__(kDerampChunk) void kb_deramp(const S1C ImgDev *__restrict__ I,
                                                          const S1C long *__restrict__ off,
                                                          int nimg) {
  const long total = off[nimg];
  int k = -1;
  for (long t = blockIdx.x; t < total; t += gridDim.x) {
    k = find_item(off, nimg, t, k);
    const ImgDev &im = *(const ImgDev *)&I[k];
    const GeomDev &g = geo(im.g);
    const int N = g.samples;
    const int nchunk = (N + 2 * kDerampChunk - 1) / (2 * kDerampChunk);
    const long lt = t - off[k];
    const int rg = (int)(lt / nchunk);
    const bool vec = (N % 2 == 0) && (((uintptr_t)im.in | (uintptr_t)im.out) % 16 == 0);
    deramp_tile(g, im.in, im.out, rg * kDerampRows, (int)(lt - (long)rg * nchunk) * kDerampChunk,
                vec);
  }
}

// Fused pair kernel over tiles of blockDim.x x blockDim.y multilooked pixels.
template <int INTERP>
__global__ __launch_bounds__(256) void kb_pair(const S1C PairDev *__restrict__ P,
                                               const S1C long *__restrict__ off, int n, int rl,
                                               int cl) {
  const long total = off[n];
  int ph = -1;
  for (long t = blockIdx.x; t < total; t += gridDim.x) {
    const int p = ph = find_item(off, n, t, ph);
    const PairDev &d = *(const PairDev *)&P[p];
    const long lt = t - off[p];
    const int tiles_x = (d.nc_ml + blockDim.x - 1) / blockDim.x;
    const int ty = (int)(lt / tiles_x), tx = (int)(lt - (long)ty * tiles_x);
    const int cm = tx * blockDim.x + threadIdx.x, rm = ty * blockDim.y + threadIdx.y;
    if (rm < d.nr_ml && cm < d.nc_ml)
      pair_ml_pixel<INTERP>(geo(d.gr), geo(d.gs), d.op, d.grid, d.dem, d.rows, d.cols, d.dh, d.ref_d,
                            d.sec_d, rl, cl, rm, cm, d.nc_ml, d.ifg, d.pr, d.ps);
  }
}

// Goldstein FP16 split: (band, chunk) items of every pair on persistent
// wavefronts, overlap-add fused (gold_band_dev.h); after its first item a
// wavefront takes the next from the stage's queue counter.
__global__ __launch_bounds__(64 * kBandWaves) __attribute__((amdgpu_waves_per_eu(2))) void kb_gold_band(
    const S1C PairDev *__restrict__ P, const S1C long *__restrict__ off, int n, float alpha,
    unsigned long long *ctr) {
  __shared__ BandLds L;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long total = off[n];
  const long nw = (long)gridDim.x * kBandWaves;
  long t = (long)blockIdx.x * kBandWaves + wave;
  if (t >= total) return;
  mf::Consts K;
  mf::build_consts(lane >> 5, lane & 31, K, L.u[wave]);
  wave_sync_lds();  // L.u[wave] written
  for (; t < total; t = band_next(ctr, nw)) {
    const int p = find_item(off, n, t);
    const PairDev &d = *(const PairDev *)&P[p];
    const long lt = t - off[p];
    const int nchunk = band_nchunk(d.pj_n);
    band_item(K, L.u[wave], d.ifg, d.nr_ml, d.nc_ml, d.pi_n, d.pj_n, (int)(lt / nchunk), (int)(lt % nchunk),
              alpha, d.patches, L.buf[wave]);
  }
}

// Work-queue counters (one per stage), zeroed at the start of every run.
__global__ void kb_reset(unsigned long long *ctr, int n) {
  if ((int)threadIdx.x < n) ctr[threadIdx.x] = 0ull;
}

// Sinc pair kernel with LDS-staged windows over 16 x 16-cell tiles of every
// pair (rl * cl <= kPairMaxPix), one block per tile in tile order.  RAW:
// ref_d / sec_d are the raw bursts, deramped as they are staged.
template <int NP, int CL, bool RAW>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void kb_pair_tiles(
    const S1C PairDev *__restrict__ P, const S1C long *__restrict__ off, int n, int rl, int cl) {
  __shared__ PairLds L;
  const long t = blockIdx.x;
  if (t >= off[n]) return;
  const int p = find_item(off, n, t);
  const PairDev &d = *(const PairDev *)&P[p];
  const long lt = t - off[p];
  const int tiles_x = (d.nc_ml + kPairTile - 1) / kPairTile;
  const int ty = (int)(lt / tiles_x), tx = (int)(lt - (long)ty * tiles_x);
  pair_tile_sinc<NP, CL, RAW>(geo(d.gr), geo(d.gs), d.op, d.grid, d.dem, d.rows, d.cols, d.dh, d.ref_d,
                 d.sec_d, rl, cl, ty, tx, d.nr_ml, d.nc_ml, d.ifg, d.pr, d.ps, L);
}

// The same with persistent blocks that take tiles from a work-queue
// counter, so the tiles in flight on the chip stay together in the bursts
// (cache reuse) whatever the grid size; the next tile is fetched while the
// current one runs.  (A separate kernel: one inlined copy of the tile code
// per kernel keeps either within the register budget.)
template <int NP, int CL, bool RAW>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void kb_pair_tiles_q(
    const S1C PairDev *__restrict__ P, const S1C long *__restrict__ off, int n, int rl, int cl,
    unsigned long long *ctr) {
  __shared__ PairLds L;
  __shared__ long slot[2];
  const long total = off[n];
  if (threadIdx.x == 0) slot[0] = (long)atomicAdd(ctr, 1ull);
  __syncthreads();
  int cur = 0, p = -1;
  for (;;) {
    const long t = slot[cur];
    if (t >= total) break;
    // the next item is fetched while this one runs
    long nxt = 0;
    if (threadIdx.x == 0) nxt = (long)atomicAdd(ctr, 1ull);
    p = find_item(off, n, t, p);
    const PairDev &d = *(const PairDev *)&P[p];
    const long lt = t - off[p];
    const int tiles_x = (d.nc_ml + kPairTile - 1) / kPairTile;
    const int ty = (int)(lt / tiles_x), tx = (int)(lt - (long)ty * tiles_x);
    pair_tile_sinc<NP, CL, RAW>(geo(d.gr), geo(d.gs), d.op, d.grid, d.dem, d.rows, d.cols, d.dh, d.ref_d,
                   d.sec_d, rl, cl, ty, tx, d.nr_ml, d.nc_ml, d.ifg, d.pr, d.ps, L);
    if (threadIdx.x == 0) slot[cur ^ 1] = nxt;
    cur ^= 1;
    __syncthreads();
  }
}

constexpr int kGoldGroups = 4;

// FP32 Goldstein patch pass: 4 groups of 32 lanes per block, one patch each.
__global__ __launch_bounds__(128) void kb_gold_fp32(const S1C PairDev *__restrict__ P,
                                                    const S1C long *__restrict__ off, int n,
                                                    float alpha) {
  __shared__ float2 T[kGoldGroups][kGoldN * kGoldPad];
  __shared__ float M[kGoldGroups][kGoldN * kGoldPad];
  const int lane = threadIdx.x & 31;
  const int sub = threadIdx.x >> 5;
  const long total = off[n];
  for (long base = (long)blockIdx.x * kGoldGroups; base < total;
       base += (long)gridDim.x * kGoldGroups) {
    const long patch = base + sub;
    const bool active = patch < total;
    const int p = find_item(off, n, active ? patch : base);
    const PairDev &d = *(const PairDev *)&P[p];
    const long lp = active ? patch - off[p] : 0;
    fp32_patch_group(d.ifg, d.nr_ml, d.nc_ml, d.pj_n, lp, active, alpha,
                     d.patches + (size_t)lp * (kGoldN * kGoldN), T[sub], M[sub], lane);
    __syncthreads();
  }
}

// Overlap-add gather over 64 x 4 tiles of every filtered image.
__global__ __launch_bounds__(256) void kb_gather(const S1C PairDev *__restrict__ P,
                                                 const S1C long *__restrict__ off, int n,
                                                 int band) {
  __shared__ BandSumLds L;
  const long total = off[n];
  int ph = -1;
  for (long t = blockIdx.x; t < total; t += gridDim.x) {
    const int p = ph = find_item(off, n, t, ph);
    const PairDev &d = *(const PairDev *)&P[p];
    const long lt = t - off[p];
    const int tiles_x = (d.nc_ml + kSumTX - 1) / kSumTX;
    const int ty = (int)(lt / tiles_x), tx = (int)(lt - (long)ty * tiles_x);
    if (band) {
      band_sum_tile(d.patches, d.nr_ml, d.nc_ml, d.pi_n, d.pj_n, tx, ty, d.filt, L);
      continue;
    }
    const int x = tx * kSumTX + threadIdx.x, y = ty * kSumTY + threadIdx.y;
    if (x < d.nc_ml && y < d.nr_ml)
      d.filt[(size_t)y * d.nc_ml + x] =
          gold_gather_pixel(d.patches, d.nr_ml, d.nc_ml, d.pi_n, d.pj_n, y, x);
  }
}

// Coherence over strips of 64 x kStripRows: separable window sums
// (coherence_strip), the strip's 64 x 4 LDS tiles for larger windows, or
// the look-cell estimator (win = 1).
__global__ __launch_bounds__(256) void kb_coherence(const S1C PairDev *__restrict__ P,
                                                    const S1C long *__restrict__ off, int n,
                                                    int win, int fill, uint32_t seed) {
  __shared__ union {
    CohTile t;
    CohStrip s;
  } L;
  const long total = off[n];
  int ph = -1;
  for (long t = blockIdx.x; t < total; t += gridDim.x) {
    const int p = ph = find_item(off, n, t, ph);
    const PairDev &d = *(const PairDev *)&P[p];
    const long lt = t - off[p];
    const int strips_x = (d.nc_ml + kCohTX - 1) / kCohTX;
    const int sy = (int)(lt / strips_x), sx = (int)(lt - (long)sy * strips_x);
    if (win > 1 && win / 2 <= kStripMaxHalo) {
      coherence_strip(d.ifg, d.pr, d.ps, d.nr_ml, d.nc_ml, win, fill, seed, d.coh, sx, sy, L.s);
    } else {
      // the strip's 64 x 4 tiles
      for (int ty = sy * (kStripRows / kCohTY); ty < (sy + 1) * (kStripRows / kCohTY); ++ty) {
        if (ty * kCohTY >= d.nr_ml) break;
        if (win == 1) {
          const int c = sx * kCohTX + threadIdx.x, r = ty * kCohTY + threadIdx.y;
          if (r < d.nr_ml && c < d.nc_ml) {
            size_t o = (size_t)r * d.nc_ml + c;
            d.coh[o] = coh_fill(coherence_cell(d.ifg[o], d.pr[o], d.ps[o]), fill, seed, o);
          }
        } else {
          coherence_tile(d.ifg, d.pr, d.ps, d.nr_ml, d.nc_ml, win, fill, seed, d.coh, sx, ty,
                         L.t);
        }
      }
    }
  }
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace
}  // namespace s1ifg

using namespace s1ifg;

struct s1ifg_batch {
  s1ifg_ctx *ctx;
  int n;
  s1ifg_batch_opts opts;
  std::vector<long> off[kNumStages];  // host copies of the prefix arrays
  std::vector<PairDev> pairs;
  std::vector<ImgDev> imgs;
  void *d_buf = nullptr;
  size_t bytes = 0;
  PairDev *d_pairs = nullptr;
  ImgDev *d_imgs = nullptr;
  long *d_off[kNumStages] = {};
  unsigned long long *d_ctr = nullptr;  // work-queue counters, one per stage
  hipStream_t stream = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

int fill_defaults(s1ifg_batch_opts *o) {
  if (o->pair_block_x == 0) o->pair_block_x = 64;
  if (o->pair_blocks_per_cu == 0) o->pair_blocks_per_cu = 8;
  if (o->gold_blocks_per_cu == 0) o->gold_blocks_per_cu = 2;
  if (o->stream_blocks_per_cu == 0) o->stream_blocks_per_cu = 8;
  const int bx = o->pair_block_x;
  if (bx != 16 && bx != 32 && bx != 64 && bx != 128 && bx != 256)
    return set_error(S1IFG_EINVAL, "pair_block_x %d (16, 32, 64, 128 or 256)", bx);
  for (int v : {o->pair_blocks_per_cu, o->gold_blocks_per_cu, o->stream_blocks_per_cu})
    if (v != S1IFG_BLOCK_PER_ITEM && (v < 1 || v > 64))
      return set_error(S1IFG_EINVAL, "blocks per CU must be in [1, 64] or S1IFG_BLOCK_PER_ITEM");
  return S1IFG_OK;
}

bool use_pair_tiles(const s1ifg_batch_opts &o) {
  return o.interp == S1IFG_INTERP_SINC && o.rl * o.cl <= kPairMaxPix;
}

// Pair-kernel tile prefix for the current tile width.
void pair_prefix(s1ifg_batch *b) {
  const bool tiles = use_pair_tiles(b->opts);
  const int bx = tiles ? kPairTile : b->opts.pair_block_x;
  const int by = tiles ? kPairTile : 256 / bx;
  std::vector<long> &o = b->off[kStPair];
  o.assign(b->n + 1, 0);
  for (int p = 0; p < b->n; ++p) {
    const PairDev &d = b->pairs[p];
    long tiles = (long)((d.nc_ml + bx - 1) / bx) * ((d.nr_ml + by - 1) / by);
    o[p + 1] = o[p] + tiles;
  }
}

void drop_graph(s1ifg_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// Grid for `items` work items, `per_block` per block: capped at
// blocks_per_cu blocks per CU (persistent blocks), or one block per
// per_block items for S1IFG_BLOCK_PER_ITEM.
unsigned grid_of(long items, long per_block, long ncu, int blocks_per_cu) {
  const long want = (items + per_block - 1) / per_block;
  const long cap = blocks_per_cu == S1IFG_BLOCK_PER_ITEM ? want : ncu * blocks_per_cu;
  return (unsigned)std::min(std::min(want, cap), (long)INT_MAX);
}

// The launch sequence, enqueued on stream s.
int enqueue(s1ifg_batch *b, hipStream_t s) {
  const s1ifg_batch_opts &o = b->opts;
  const long ncu = device_cu_count(b->ctx->device);
  const int n = b->n;
  const S1C long *off[kNumStages];
  for (int st = 0; st < kNumStages; ++st) off[st] = (const S1C long *)b->d_off[st];
  const S1C PairDev *pairs = (const S1C PairDev *)b->d_pairs;
  auto total = [&](int st) { return b->off[st].back(); };
  hipLaunchKernelGGL(kb_reset, dim3(1), dim3(64), 0, s, b->d_ctr, (int)kNumStages);
  S1_LAUNCH_CHECK("kb_reset");
  if (long t = total(kStTables)) {
    hipLaunchKernelGGL(kb_tables, dim3(grid_of(t, 256, ncu, o.stream_blocks_per_cu)), dim3(256),
                       0, s, pairs, off[kStTables], n);
    S1_LAUNCH_CHECK("kb_tables");
  }
  if (long t = total(kStDeramp)) {
    hipLaunchKernelGGL(kb_deramp, dim3(grid_of(t, 1, ncu, o.stream_blocks_per_cu)), dim3(256), 0,
                       s, (const S1C ImgDev *)b->d_imgs, off[kStDeramp], (int)b->imgs.size());
    S1_LAUNCH_CHECK("kb_deramp");
  }
  if (long t = total(kStPair)) {
    dim3 blk(o.pair_block_x, 256 / o.pair_block_x);
    dim3 grd(grid_of(t, 1, ncu, o.pair_blocks_per_cu));
    // the 1 x 4 looks of the workloads with the look shape compiled in
    const bool l14 = o.rl == 1 && o.cl == 4;
    auto tiles = o.raw ? (l14 ? kb_pair_tiles<4, 4, true> : kb_pair_tiles<0, 0, true>)
                       : (l14 ? kb_pair_tiles<4, 4, false> : kb_pair_tiles<0, 0, false>);
    auto tiles_q = o.raw ? (l14 ? kb_pair_tiles_q<4, 4, true> : kb_pair_tiles_q<0, 0, true>)
                         : (l14 ? kb_pair_tiles_q<4, 4, false> : kb_pair_tiles_q<0, 0, false>);
    if (use_pair_tiles(o) && (long)grd.x >= t)
      hipLaunchKernelGGL(tiles, grd, dim3(256), 0, s, pairs, off[kStPair], n, o.rl, o.cl);
    else if (use_pair_tiles(o))
      hipLaunchKernelGGL(tiles_q, grd, dim3(256), 0, s, pairs, off[kStPair], n, o.rl, o.cl,
                         b->d_ctr + kStPair);
    else if (o.interp == S1IFG_INTERP_BILINEAR)
      hipLaunchKernelGGL(kb_pair<S1IFG_INTERP_BILINEAR>, grd, blk, 0, s, pairs, off[kStPair],
                         n, o.rl, o.cl);
    else
      hipLaunchKernelGGL(kb_pair<S1IFG_INTERP_SINC>, grd, blk, 0, s, pairs, off[kStPair], n,
                         o.rl, o.cl);
    S1_LAUNCH_CHECK("kb_pair");
  }
  if (long t = total(kStGold)) {
    if (o.gold_mode == S1IFG_GOLD_FP16_SPLIT)
      hipLaunchKernelGGL(kb_gold_band, dim3(grid_of(t, kBandWaves, ncu, o.gold_blocks_per_cu)),
                         dim3(64 * kBandWaves), 0, s, pairs, off[kStGold], n, o.alpha,
                         b->d_ctr + kStGold);
    else
      hipLaunchKernelGGL(kb_gold_fp32, dim3(grid_of(t, kGoldGroups, ncu, o.gold_blocks_per_cu == S1IFG_BLOCK_PER_ITEM ? S1IFG_BLOCK_PER_ITEM : 8 * o.gold_blocks_per_cu)),
                         dim3(32 * kGoldGroups), 0, s, pairs, off[kStGold], n, o.alpha);
    S1_LAUNCH_CHECK("kb_gold");
  }
  if (long t = total(kStGather)) {
    hipLaunchKernelGGL(kb_gather, dim3(grid_of(t, 1, ncu, o.stream_blocks_per_cu)), dim3(64, 4),
                       0, s, pairs, off[kStGather], n, o.gold_mode == S1IFG_GOLD_FP16_SPLIT);
    S1_LAUNCH_CHECK("kb_gather");
  }
  if (long t = total(kStCoh)) {
    hipLaunchKernelGGL(kb_coherence, dim3(grid_of(t, 1, ncu, o.stream_blocks_per_cu)),
                       dim3(kCohTX, kCohTY), 0, s, pairs, off[kStCoh], n, o.coh_win,
                       o.coh_fill, o.coh_seed);
    S1_LAUNCH_CHECK("kb_coherence");
  }
  return S1IFG_OK;
}

int upload_prefix(s1ifg_batch *b, int st) {
  S1_HIP(hipMemcpy(b->d_off[st], b->off[st].data(), b->off[st].size() * sizeof(long),
                   hipMemcpyHostToDevice));
  return S1IFG_OK;
}

}  // namespace

extern "C" {

int s1ifg_batch_create(s1ifg_ctx *ctx, const s1ifg_pair_task *tasks, int n,
                       const s1ifg_batch_opts *opts, s1ifg_batch **out) {
  if (!out) return set_error(S1IFG_EINVAL, "null output pointer");
  *out = nullptr;
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!tasks || n < 1 || !opts) return set_error(S1IFG_EINVAL, "empty batch");
  s1ifg_batch_opts o = *opts;
  if ((rc = fill_defaults(&o))) return rc;
  if (o.interp != S1IFG_INTERP_BILINEAR && o.interp != S1IFG_INTERP_SINC)
    return set_error(S1IFG_EINVAL, "interp %d", o.interp);
  if (o.rl < 1 || o.cl < 1) return set_error(S1IFG_EINVAL, "looks %d x %d", o.rl, o.cl);
  if (o.gold_mode != S1IFG_GOLD_FP32 && o.gold_mode != S1IFG_GOLD_FP16_SPLIT)
    return set_error(S1IFG_EINVAL, "gold_mode %d", o.gold_mode);
  if (!(o.alpha >= 0.f) || !std::isfinite(o.alpha)) return set_error(S1IFG_EINVAL, "alpha %g", o.alpha);
  if (o.coh_win < 1 || o.coh_win % 2 == 0)
    return set_error(S1IFG_EINVAL, "coherence window %d (odd, >= 1)", o.coh_win);
  if (o.coh_win / 2 > kCohMaxHalo)
    return set_error(S1IFG_ERANGE, "batched coherence window at most %d", 2 * kCohMaxHalo + 1);
  if (o.coh_fill < S1IFG_COH_FILL_CLAMP || o.coh_fill > S1IFG_COH_FILL_HASH)
    return set_error(S1IFG_EINVAL, "coherence fill %d", o.coh_fill);
  o.raw = o.raw != 0;
  if (o.raw && !use_pair_tiles(o))
    return set_error(S1IFG_EINVAL, "raw bursts need sinc interpolation and rl * cl <= %d",
                     kPairMaxPix);

  std::vector<PairDev> pairs(n);
  std::vector<ImgDev> imgs;
  std::vector<const s1ifg_geom *> img_geom;
  size_t tab_bytes = 0, patch_bytes = 0;
  std::vector<size_t> tab_off(n), patch_off(n);
  for (int p = 0; p < n; ++p) {
    const s1ifg_pair_task &t = tasks[p];
    PairDev &d = pairs[p];
    std::memset(&d, 0, sizeof(d));
    if ((rc = check_geom(ctx, t.g_ref)) || (rc = check_geom(ctx, t.g_sec))) return rc;
    if ((rc = make_grid(&t.grid, &d.grid)) || (rc = make_dem(&t.dem, &d.dem))) return rc;
    if ((rc = make_offpoly(t.use_offpoly ? t.offpoly_sec : nullptr, &d.op))) return rc;
    if (o.raw ? (!t.ref_slc || !t.sec_slc || !t.ifg_ml)
              : (!t.ref_deramped || !t.sec_deramped || !t.ifg_ml))
      return set_error(S1IFG_EINVAL, "pair %d: null data pointer", p);
    if (!o.raw && (t.ref_deramped == t.ref_slc || t.sec_deramped == t.sec_slc ||
                   t.ref_deramped == t.sec_deramped))
      return set_error(S1IFG_EINVAL, "pair %d: deramped bursts must not alias the raw bursts", p);
    // distinct bursts to deramp (raw NULL: the deramped buffer is an input);
    // none with raw bursts
    for (int side = 0; side < 2 && !o.raw; ++side) {
      const s1ifg_geom *g = side ? t.g_sec : t.g_ref;
      const void *raw = side ? (const void *)t.sec_slc : (const void *)t.ref_slc;
      void *der = side ? (void *)t.sec_deramped : (void *)t.ref_deramped;
      bool seen = false;
      for (size_t q = 0; q < imgs.size(); ++q) {
        if (imgs[q].out == (float2 *)der || (raw && imgs[q].in == (const float2 *)raw)) {
          if (imgs[q].out != (float2 *)der || imgs[q].in != (const float2 *)raw ||
              img_geom[q] != g)
            return set_error(S1IFG_EINVAL,
                             "pair %d: a burst buffer is shared with a different burst", p);
          seen = true;
        }
      }
      if (!seen && raw) {
        ImgDev im;
        im.g = (const S1C GeomDev *)g->d_dev;
        im.in = (const float2 *)raw;
        im.out = (float2 *)der;
        imgs.push_back(im);
        img_geom.push_back(g);
      }
    }
    d.gr = (const S1C GeomDev *)t.g_ref->d_dev;
    d.gs = (const S1C GeomDev *)t.g_sec->d_dev;
    d.dh = t.dh;
    d.ref_d = (const float2 *)(o.raw ? t.ref_slc : t.ref_deramped);
    d.sec_d = (const float2 *)(o.raw ? t.sec_slc : t.sec_deramped);
    d.ifg = (float2 *)t.ifg_ml;
    d.pr = t.pow_ref;
    d.ps = t.pow_sec;
    d.filt = (float2 *)t.filt;
    d.coh = t.coh;
    d.nr_ml = d.grid.nrow / o.rl;
    d.nc_ml = d.grid.ncol / o.cl;
    if (d.nr_ml < 1 || d.nc_ml < 1)
      return set_error(S1IFG_EINVAL, "pair %d: grid smaller than one look cell", p);
    if (d.coh && (!d.pr || !d.ps))
      return set_error(S1IFG_EINVAL, "pair %d: coherence needs both power images", p);
    if (d.filt) {
      if (d.nr_ml < kGoldN || d.nc_ml < kGoldN)
        return set_error(S1IFG_EINVAL, "pair %d: multilooked image %d x %d smaller than the %d-pixel window",
                         p, d.nr_ml, d.nc_ml, kGoldN);
      if (d.filt == d.ifg) return set_error(S1IFG_EINVAL, "pair %d: in-place filtering is not supported", p);
      goldstein_grid(d.nr_ml, d.nc_ml, &d.pi_n, &d.pj_n);
      if (o.gold_mode == S1IFG_GOLD_FP16_SPLIT && !band_offsets_fit(d.pi_n, d.pj_n))
        return set_error(S1IFG_ERANGE, "pair %d: multilooked image %d x %d too large for the fp16_split band sums",
                         p, d.nr_ml, d.nc_ml);
      patch_off[p] = patch_bytes;
      patch_bytes += align256(o.gold_mode == S1IFG_GOLD_FP16_SPLIT
                                  ? band_buffer_bytes(d.pi_n, d.pj_n)
                                  : (size_t)d.pi_n * d.pj_n * kGoldN * kGoldN * sizeof(float2));
    }
    tab_off[p] = tab_bytes;
    tab_bytes += align256(sizeof(RowTab) * (size_t)d.grid.nrow) +
                 align256(sizeof(ColTab) * (size_t)d.grid.ncol);
  }

  s1ifg_batch *b = new s1ifg_batch();
  b->ctx = ctx;
  b->n = n;
  b->opts = o;
  b->pairs = pairs;
  b->imgs = imgs;
  const int nimg = (int)imgs.size();
  // prefix arrays
  for (int st = 0; st < kNumStages; ++st) b->off[st].assign(st == kStDeramp ? nimg + 1 : n + 1, 0);
  for (int q = 0; q < nimg; ++q)
    b->off[kStDeramp][q + 1] =
        b->off[kStDeramp][q] +
        (long)((img_geom[q]->dev.lines + kDerampRows - 1) / kDerampRows) *
            ((img_geom[q]->dev.samples + 2 * kDerampChunk - 1) / (2 * kDerampChunk));
  for (int p = 0; p < n; ++p) {
    const PairDev &d = b->pairs[p];
    long tab = d.grid.nrow > d.grid.ncol ? d.grid.nrow : d.grid.ncol;
    b->off[kStTables][p + 1] = b->off[kStTables][p] + tab;
    const long gitems = o.gold_mode == S1IFG_GOLD_FP16_SPLIT
                            ? band_items(d.pi_n, d.pj_n)
                            : (long)d.pi_n * d.pj_n;
    b->off[kStGold][p + 1] = b->off[kStGold][p] + (d.filt ? gitems : 0);
    long gtiles = (long)((d.nc_ml + 63) / 64) * ((d.nr_ml + 3) / 4);
    b->off[kStGather][p + 1] = b->off[kStGather][p] + (d.filt ? gtiles : 0);
    long ctiles =
        (long)((d.nc_ml + kCohTX - 1) / kCohTX) * ((d.nr_ml + kStripRows - 1) / kStripRows);
    b->off[kStCoh][p + 1] = b->off[kStCoh][p] + (d.coh ? ctiles : 0);
  }
  pair_prefix(b);

  // one device buffer: descriptors | prefix arrays | grid tables | patches
  size_t off_pairs = 0;
  size_t sz = align256(sizeof(PairDev) * n);
  size_t off_imgs = sz;
  sz += align256(sizeof(ImgDev) * (nimg > 0 ? nimg : 1));
  size_t off_pref[kNumStages];
  for (int st = 0; st < kNumStages; ++st) {
    off_pref[st] = sz;
    sz += align256(sizeof(long) * (2 * n + 1));
  }
  size_t off_ctr = sz;
  sz += align256(sizeof(unsigned long long) * kNumStages);
  size_t off_tab = sz;
  sz += tab_bytes;
  size_t off_patch = sz;
  sz += patch_bytes;
  (void)hipSetDevice(ctx->device);
  if (hipMalloc(&b->d_buf, sz) != hipSuccess) {
    (void)hipGetLastError();
    delete b;
    return set_error(S1IFG_ENOMEM, "batch scratch of %zu bytes", sz);
  }
  b->bytes = sz;
  char *base = (char *)b->d_buf;
  b->d_pairs = (PairDev *)(base + off_pairs);
  b->d_imgs = (ImgDev *)(base + off_imgs);
  b->d_ctr = (unsigned long long *)(base + off_ctr);
  for (int st = 0; st < kNumStages; ++st) b->d_off[st] = (long *)(base + off_pref[st]);
  for (int p = 0; p < n; ++p) {
    PairDev &d = b->pairs[p];
    d.rows = (RowTab *)(base + off_tab + tab_off[p]);
    d.cols = (ColTab *)(base + off_tab + tab_off[p] + align256(sizeof(RowTab) * (size_t)d.grid.nrow));
    d.patches = d.filt ? (float2 *)(base + off_patch + patch_off[p]) : nullptr;
  }
  int err = S1IFG_OK;
  if (hipMemcpy(b->d_pairs, b->pairs.data(), sizeof(PairDev) * n, hipMemcpyHostToDevice) != hipSuccess ||
      (nimg > 0 && hipMemcpy(b->d_imgs, b->imgs.data(), sizeof(ImgDev) * nimg,
                             hipMemcpyHostToDevice) != hipSuccess))
    err = set_error(S1IFG_EHIP, "descriptor upload failed");
  for (int st = 0; st < kNumStages && !err; ++st) err = upload_prefix(b, st);
  if (!err && (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
               hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming) != hipSuccess ||
               hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming) != hipSuccess))
    err = set_error(S1IFG_EHIP, "batch stream / event creation failed");
  if (err) {
    s1ifg_batch_destroy(b);
    return err;
  }
  *out = b;
  return S1IFG_OK;
}

void s1ifg_batch_destroy(s1ifg_batch *b) {
  if (!b) return;
  (void)hipSetDevice(b->ctx->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  if (b->ev_in) (void)hipEventDestroy(b->ev_in);
  if (b->ev_out) (void)hipEventDestroy(b->ev_out);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  if (b->d_buf) (void)hipFree(b->d_buf);
  delete b;
}

int s1ifg_batch_run(s1ifg_batch *b, int use_graph) {
  if (!b) return set_error(S1IFG_EINVAL, "null batch");
  int rc = check_ctx(b->ctx);
  if (rc) return rc;
  if (!use_graph) return enqueue(b, b->ctx->stream);
  if (!b->exec) {
    S1_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    rc = enqueue(b, b->stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return hip_error(e, "hipStreamEndCapture");
    b->graph = g;
    e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      drop_graph(b);
      return hip_error(e, "hipGraphInstantiate");
    }
  }
  // the graph runs on the batch stream, ordered after and before the
  // context stream's work
  S1_HIP(hipEventRecord(b->ev_in, b->ctx->stream));
  S1_HIP(hipStreamWaitEvent(b->stream, b->ev_in, 0));
  S1_HIP(hipGraphLaunch(b->exec, b->stream));
  S1_HIP(hipEventRecord(b->ev_out, b->stream));
  S1_HIP(hipStreamWaitEvent(b->ctx->stream, b->ev_out, 0));
  return S1IFG_OK;
}

int s1ifg_batch_bytes(const s1ifg_batch *b, int64_t *bytes) {
  if (!b || !bytes) return set_error(S1IFG_EINVAL, "null argument");
  *bytes = (int64_t)b->bytes;
  return S1IFG_OK;
}

int s1ifg_batch_get_opts(const s1ifg_batch *b, s1ifg_batch_opts *opts) {
  if (!b || !opts) return set_error(S1IFG_EINVAL, "null argument");
  *opts = b->opts;
  return S1IFG_OK;
}

int s1ifg_batch_set_launch(s1ifg_batch *b, const s1ifg_batch_opts *opts) {
  if (!b || !opts) return set_error(S1IFG_EINVAL, "null argument");
  s1ifg_batch_opts o = b->opts;
  o.pair_block_x = opts->pair_block_x;
  o.pair_blocks_per_cu = opts->pair_blocks_per_cu;
  o.gold_blocks_per_cu = opts->gold_blocks_per_cu;
  o.stream_blocks_per_cu = opts->stream_blocks_per_cu;
  int rc = fill_defaults(&o);
  if (rc) return rc;
  if ((rc = check_ctx(b->ctx))) return rc;
  // the prefix array is read by kernels already enqueued
  S1_HIP(hipStreamSynchronize(b->ctx->stream));
  S1_HIP(hipStreamSynchronize(b->stream));
  drop_graph(b);
  b->opts = o;
  pair_prefix(b);
  return upload_prefix(b, kStPair);
}

}  // extern "C"
