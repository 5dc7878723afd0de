// PS phase similarity (the reference's phase_similarity executable,
// phase_similarity.cu:116-548) on one resident stack.
//
// Layout: the interferogram stack arrives image after image
// ([nifg][nrow][ncol] complex).  The similarity kernels gather whole
// per-pixel series of neighbours, so a preparation pass rewrites the stack
// pixel-major (each pixel's nifg phasors contiguous, row stride rounded up
// to an even count so that 16-byte loads fetch two interferograms) and
// packs the candidate mask into bits (one 64-bit word pair per 64 pixels).
// Candidate (or non-candidate) pixels are then compacted into a list in
// image order, so that every lane of a wavefront works on a pixel that
// needs the computation.  Per listed pixel, phase A walks the neighbour
// offsets in the reference's order and records the first nneigh candidate
// neighbours in LDS; phase B computes their similarities with the pixel's
// series held in registers: all lanes busy on the dot products, whatever
// the mask density.  Median: the smallest nneigh/2 + 1 similarities are
// kept sorted in registers by a compare-exchange chain (the median, element
// count/2, is among them); maximum: a running maximum.
//
// Arithmetic follows the reference's float sequence with contraction off
// (separately rounded products and sums), so results equal the float CPU
// restatement of the same loops.
#include <algorithm>
#include <climits>
#include <cstdlib>
#include <vector>

#include "s1ifg_internal.h"

using namespace s1ifg;

namespace {

constexpr int kTile = 64;        // pixels per wavefront tile (two mask words)
constexpr int kPrepChunk = 16;   // interferograms per LDS pass of the preparation
constexpr int kChunk = 32;       // interferograms held in registers per similarity pass
constexpr int kMaxRd = 2048;
constexpr int kXcds = 8;         // workgroups are spread round-robin over the XCDs
constexpr int kScan = 8;         // neighbour offsets tested per step of the walk (ov[] below)

// scan_array (phase_similarity.cu:295-366): rings r = 1 .. rdmax-1 of a
// midpoint circle walked from (r, 0) towards the diagonal; where the point
// one step inside at the same height is still unvisited the walk visits it
// first and steps back out, so the rings leave no holes.  Rings <= rdmin
// are walked (they mark the visited set) but not kept.  Each visited
// first-octant point is emitted with its mirror images in a fixed order.
// Packed as (dr << 16) | (dc & 0xffff).
void scan_offsets(int rdmin, int rdmax, std::vector<int32_t> &out) {
  out.clear();
  if (rdmax <= 0) return;
  std::vector<unsigned char> seen((size_t)rdmax * rdmax, 0);
  auto put = [&](int dr, int dc) { out.push_back((int32_t)(((uint32_t)dr << 16) | ((uint32_t)dc & 0xffffu))); };
  seen[0] = 1;
  for (int r = 1; r < rdmax; ++r) {
    const bool keep = r > rdmin;
    int x = r, y = 0, p = 1 - r, back = 0;
    if (keep) {
      put(r, 0);
      put(-r, 0);
      put(0, r);
      put(0, -r);
    }
    seen[(size_t)r * rdmax] = seen[r] = 1;
    while (x > y) {
      if (back == 0) {
        ++y;
        if (p <= 0) {
          p += 2 * y + 1;
        } else {
          --x;
          p += 2 * y - 2 * x + 1;
        }
      } else {
        --back;
      }
      if (x < y) break;
      while (x - 1 >= 0 && y < rdmax && !seen[(size_t)(x - 1) * rdmax + y]) {
        --x;
        ++back;
      }
      seen[(size_t)x * rdmax + y] = seen[(size_t)y * rdmax + x] = 1;
      if (keep) {
        put(x, y);
        put(-x, -y);
        put(x, -y);
        put(-x, y);
        if (x != y) {
          put(y, x);
          put(-y, -x);
          put(y, -x);
          put(-y, x);
        }
      }
      if (back > 0) ++x;
    }
  }
}

struct PsCache {
  int rdmin = 0, rdmax = -1, noff = 0, extent = 0;
  int32_t *dev = nullptr;
  size_t cap = 0;
  std::vector<int32_t> host;  // source of the last upload, kept until the next change
};

// The offset table for (rdmin, rdmax) on the device, built and uploaded
// when it changes (the stream is drained first: queued kernels may still
// read the previous table).
int ps_table(s1ifg_ctx *ctx, int rdmin, int rdmax, const int32_t **tab, int *noff,
             int *extent) {
  rdmin = std::max(rdmin, 0);
  PsCache *c = static_cast<PsCache *>(ctx->ps_cache);
  if (!c) {
    c = new PsCache();
    ctx->ps_cache = c;
  }
  if (c->rdmax != rdmax || c->rdmin != rdmin) {
    S1_HIP(hipStreamSynchronize(ctx->stream));
    scan_offsets(rdmin, rdmax, c->host);
    const size_t n0 = c->host.size();
    // padded with zero entries to a whole number of walk steps (kScan), so
    // that a step reads its entries in one go
    const size_t n = (n0 + kScan - 1) / kScan * kScan;
    c->host.resize(n, 0);
    if (n > c->cap) {
      if (c->dev) S1_HIP(hipFree(c->dev));
      c->dev = nullptr;
      c->cap = 0;
      S1_HIP(hipMalloc(&c->dev, n * sizeof(int32_t)));
      c->cap = n;
    }
    if (n)
      S1_HIP(hipMemcpyAsync(c->dev, c->host.data(), n * sizeof(int32_t), hipMemcpyHostToDevice,
                            ctx->stream));
    c->rdmin = rdmin;
    c->rdmax = rdmax;
    c->noff = (int)n0;
    c->extent = 0;
    for (int32_t o : c->host)
      c->extent = std::max(c->extent, std::max(std::abs(o >> 16), std::abs((int)(int16_t)(o & 0xffff))));
  }
  *tab = c->dev;
  *noff = c->noff;
  *extent = c->extent;
  return S1IFG_OK;
}

// normalize (phase_similarity.cu:116-134).
__device__ __forceinline__ float2 unit_phasor(float2 v) {
#pragma clang fp contract(off)
  const float mag = sqrtf(v.x * v.x + v.y * v.y);
  if (mag > 1e-8f) {
    const float inv = 1.0f / mag;
    return make_float2(v.x * inv, v.y * inv);
  }
  return make_float2(0.f, 0.f);
}

__global__ __launch_bounds__(256) void k_ps_normalize(const float2 *in, float2 *out, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = unit_phasor(in[i]);
}

// update_ps (phase_similarity.cu:145-157).
__global__ __launch_bounds__(256) void k_ps_update(int *ps, const float *__restrict__ sim, float th,
                                                   int overwrite, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (sim[i] >= th)
    ps[i] = 1;
  else if (overwrite)
    ps[i] = 0;
}

// One wavefront per 64-pixel tile: candidate bits and per-tile counts;
// fill[p] = value for the pixels whose candidacy equals fill_cand (optional);
// the (optionally normalised) stack rewritten pixel-major with row stride
// `stride` complex (even, >= nifg; the pad is zero).
__global__ __launch_bounds__(256) void k_ps_prep(const float2 *__restrict__ stack, int nifg, int npix,
                                                 int stride, int normalize,
                                                 const int *__restrict__ ps, float2 *__restrict__ zt,
                                                 unsigned *__restrict__ bits, int *__restrict__ counts,
                                                 float *__restrict__ fill, int fill_cand,
                                                 float fill_value) {
  __shared__ float2 sh[4][kTile][kPrepChunk + 1];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ntiles = (npix + kTile - 1) / kTile;
  const int t = blockIdx.x * 4 + w;
  const bool tile_ok = t < ntiles;
  const long p0 = (long)t * kTile;
  const long p = p0 + lane;
  const bool in = tile_ok && p < npix;
  const bool cand = in && ps[p] != 0;
  const unsigned long long m = __ballot(cand);
  if (tile_ok && lane == 0) {
    bits[2 * t] = (unsigned)m;
    bits[2 * t + 1] = (unsigned)(m >> 32);
    counts[t] = __popcll(m);
  }
  if (fill && in && (int)cand == fill_cand) fill[p] = fill_value;
  const int np = tile_ok ? (int)min((long)kTile, (long)npix - p0) : 0;
  for (int j0 = 0; j0 < nifg; j0 += kPrepChunk) {
    const int cw = min(kPrepChunk, nifg - j0);
    for (int jj = 0; jj < cw; ++jj) {
      float2 v = make_float2(0.f, 0.f);
      if (in) {
        v = stack[(size_t)(j0 + jj) * npix + p];
        if (normalize) v = unit_phasor(v);
      }
      sh[w][lane][jj] = v;
    }
    __syncthreads();
    for (int e = lane; e < np * cw; e += 64) {
      const int i = e / cw, jj = e - i * cw;
      zt[(size_t)(p0 + i) * stride + j0 + jj] = sh[w][i][jj];
    }
    __syncthreads();
  }
  if (in && stride > nifg) zt[(size_t)p * stride + nifg] = make_float2(0.f, 0.f);
}

// Exclusive scan of the per-tile counts (of set bits, or with invert of the
// clear ones inside the image) into list offsets; total = list length.  One
// block walks the counts 1024 at a time (coalesced), carrying the running
// sum: a wavefront prefix by shuffles, then across the 16 wavefronts.
__global__ __launch_bounds__(1024) void k_ps_scan(const int *__restrict__ counts, int ntiles, int npix,
                                                  int invert, int *__restrict__ base,
                                                  int *__restrict__ total) {
  __shared__ int wsum[16];
  __shared__ int carry_s;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  if (tid == 0) carry_s = 0;
  __syncthreads();
  for (int t0 = 0; t0 < ntiles; t0 += 1024) {
    const int t = t0 + tid;
    int c = 0;
    if (t < ntiles) {
      c = counts[t];
      if (invert) c = min(kTile, npix - t * kTile) - c;
    }
    int x = c;  // inclusive prefix within the wavefront
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(x, off);
      if (lane >= off) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int before = carry_s;
    for (int k = 0; k < w; ++k) before += wsum[k];
    if (t < ntiles) base[t] = before + x - c;
    __syncthreads();
    if (tid == 1023) carry_s = before + x;
    __syncthreads();
  }
  if (tid == 0) *total = carry_s;
}

// Pixels whose bit is set (clear with invert) into list, in image order.
__global__ __launch_bounds__(256) void k_ps_compact(const unsigned *__restrict__ bits,
                                                    const int *__restrict__ base, int npix, int invert,
                                                    int *__restrict__ list) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ntiles = (npix + kTile - 1) / kTile;
  const int t = blockIdx.x * 4 + w;
  if (t >= ntiles) return;
  unsigned long long m = ((unsigned long long)bits[2 * t + 1] << 32) | bits[2 * t];
  if (invert) {
    m = ~m;
    const int np = min(kTile, npix - t * kTile);
    if (np < kTile) m &= (1ull << np) - 1;
  }
  if ((m >> lane) & 1ull) {
    const int before = __popcll(m & ((1ull << lane) - 1));
    list[base[t] + before] = t * kTile + lane;
  }
}

// New candidates from the median similarity (update_ps with overwrite),
// their bits and counts, and max_sim = 1 for them (maximum_similarity's
// value for candidates, phase_similarity.cu:247-250).
__global__ __launch_bounds__(256) void k_ps_update_bits(int *__restrict__ ps, const float *__restrict__ med,
                                                        float th, int npix, unsigned *__restrict__ bits,
                                                        int *__restrict__ counts,
                                                        float *__restrict__ maxsim) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ntiles = (npix + kTile - 1) / kTile;
  const int t = blockIdx.x * 4 + w;
  const bool tile_ok = t < ntiles;
  const long p = (long)t * kTile + lane;
  const bool in = tile_ok && p < npix;
  bool hit = false;
  if (in) {
    hit = med[p] >= th;
    ps[p] = hit ? 1 : 0;
    if (hit) maxsim[p] = 1.f;
  }
  const unsigned long long m = __ballot(hit);
  if (tile_ok && lane == 0) {
    bits[2 * t] = (unsigned)m;
    bits[2 * t + 1] = (unsigned)(m >> 32);
    counts[t] = __popcll(m);
  }
}

template <int NN>
constexpr int waves_per_block() {
  return NN >= 128 ? 1 : (NN >= 64 ? 2 : 4);
}

// Similarities of each listed pixel to its first nmax candidate neighbours:
// MEDIAN -> element cnt/2 of the sorted values (median_similarity,
// phase_similarity.cu:174-233), else the maximum (maximum_similarity,
// :235-283); 0 without neighbours.  One wavefront per 64 list entries,
// grid-strided.  LDS per wavefront: the neighbour list [NN][64] (+ the
// partial sums [NN][64] when nifg > kChunk).
template <int NN, bool MEDIAN, bool MULTI>
__global__ __launch_bounds__(64 * waves_per_block<NN>()) void k_ps_sim(
    const float4 *__restrict__ zt, int nifg, int stride4, int nrow, int ncol,
    const unsigned *__restrict__ bits, const int *__restrict__ list, const int *__restrict__ total,
    const int *__restrict__ tab, int noff, int nmax, float *__restrict__ out) {
#pragma clang fp contract(off)
  constexpr int WPB = waves_per_block<NN>();
  constexpr int H = NN / 2 + 1;
  extern __shared__ int lds[];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int *Lq = lds + w * NN * 64 * (MULTI ? 2 : 1);
  float *Lp = reinterpret_cast<float *>(Lq + NN * 64);
  const int n = *total;
  const float fn = (float)nifg;
  const int nchunks = MULTI ? (nifg + kChunk - 1) / kChunk : 1;
  // Workgroups go round-robin to the 8 XCDs: XCD x takes the x-th eighth
  // of the list (a band of rows), so that its L2 serves the neighbourhoods.
  const long nlist = (n + 63) / 64;
  const int xcd = blockIdx.x % kXcds;
  const long band0 = nlist * xcd / kXcds, band1 = nlist * (xcd + 1) / kXcds;
  const long step = (long)(gridDim.x / kXcds) * WPB;
  for (long chunk = band0 + (long)(blockIdx.x / kXcds) * WPB + w; chunk < band1; chunk += step) {
    const long e = chunk * 64 + lane;
    const bool live = e < n;
    const int p = list[live ? e : chunk * 64];
    const int r0 = p / ncol, c0 = p - r0 * ncol;
    // phase A: the first nmax candidate neighbours in exploration order
    // (kScan offsets per step: their mask words are fetched together, then
    // taken in order)
    int cnt = live ? 0 : nmax;
    for (int i0 = 0; i0 < noff; i0 += kScan) {
      if (__ballot(cnt < nmax) == 0) break;
      int qv[kScan];
      unsigned wv[kScan];
#pragma unroll
      for (int u = 0; u < kScan; ++u) {
        qv[u] = -1;
        wv[u] = 0;
        if (i0 + u < noff) {
          const int o = tab[i0 + u];
          const int r = r0 + (o >> 16), c = c0 + (int)(short)(o & 0xffff);
          if ((unsigned)r < (unsigned)nrow && (unsigned)c < (unsigned)ncol) {
            qv[u] = r * ncol + c;
            wv[u] = bits[qv[u] >> 5];
          }
        }
      }
#pragma unroll
      for (int u = 0; u < kScan; ++u) {
        if (qv[u] >= 0 && cnt < nmax && ((wv[u] >> (qv[u] & 31)) & 1u)) {
          Lq[cnt * 64 + lane] = qv[u];
          ++cnt;
        }
      }
    }
    if (!live) cnt = 0;
    int kmax = cnt;
    for (int off = 32; off > 0; off >>= 1) kmax = max(kmax, __shfl_xor(kmax, off));
    // phase B: similarities, the pixel's series in registers
    float srt[MEDIAN ? H : 1];
#pragma unroll
    for (int h = 0; h < (MEDIAN ? H : 1); ++h) srt[h] = __builtin_inff();
    float best = -1.f;
    for (int ch = 0; ch < nchunks; ++ch) {
      const int j0 = ch * kChunk;
      const int cw = min(kChunk, nifg - j0);
      const bool last = ch == nchunks - 1;
      float2 cv[kChunk];
      const float4 *cp = zt + (size_t)p * stride4 + j0 / 2;
#pragma unroll
      for (int jj = 0; jj < kChunk / 2; ++jj) {
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (2 * jj < cw) v = cp[jj];
        cv[2 * jj] = make_float2(v.x, v.y);
        cv[2 * jj + 1] = make_float2(v.z, v.w);
      }
      for (int k = 0; k < kmax; ++k) {
        const bool act = k < cnt;
        const int q = act ? Lq[k * 64 + lane] : p;
        float acc = (MULTI && ch > 0) ? Lp[k * 64 + lane] : 0.f;
        const float4 *qp = zt + (size_t)q * stride4 + j0 / 2;
#pragma unroll
        for (int jj = 0; jj < kChunk / 2; ++jj) {
          if (2 * jj < cw) {
            const float4 v = qp[jj];
            acc += cv[2 * jj].x * v.x + cv[2 * jj].y * v.y;
            if (2 * jj + 1 < cw) acc += cv[2 * jj + 1].x * v.z + cv[2 * jj + 1].y * v.w;
          }
        }
        if (MULTI && !last) {
          Lp[k * 64 + lane] = acc;
          continue;
        }
        const float sim = acc / fn;
        if (act) {
          if (MEDIAN) {
            float v = sim;
#pragma unroll
            for (int h = 0; h < H; ++h) {
              const float lo = fminf(srt[h], v);
              v = fmaxf(srt[h], v);
              srt[h] = lo;
            }
          } else if (best < sim) {
            best = sim;
          }
        }
      }
    }
    if (live) {
      float res = 0.f;
      if (MEDIAN) {
        const int m = cnt / 2;
#pragma unroll
        for (int h = 0; h < H; ++h)
          if (h == m) res = srt[h];
        if (cnt == 0) res = 0.f;
      } else {
        res = cnt ? best : 0.f;
      }
      out[p] = res;
    }
  }
}

// LDS-tiled similarities (the form used when the series fit): a block owns
// a kPT x kPT tile of pixels.  It stages into LDS the series of the
// eligible neighbours (candidates / kept PS) inside the tile grown by a
// halo of up to kHalo pixels, in window order, up to `cap` of them (the
// rest are read from global memory), with a map window pixel -> slot.
// Phase A: each of the tile's centres (candidates for the median pass,
// non-PS pixels for the maximum pass) walks the offsets in the reference's
// order and records its first nmax eligible neighbours (an LDS slot, or a
// global pixel for those outside the staged window).  Phase B: all 256
// threads compute the (centre, neighbour) similarities, consecutive
// threads on one centre's neighbours, and store them over the references.
// Phase C: each centre reduces its similarities (median by the
// compare-exchange chain, or maximum).  Every similarity is the same float
// sequence as in k_ps_sim, whatever memory the series come from.
constexpr int kPT = 16;      // tile side (pixels)
constexpr int kHalo = 10;    // staged halo around the tile (pixels)
constexpr int kTileLds = 76 * 1024;  // LDS per block (two blocks per CU, with margin)
constexpr int kStageLoads = 8;       // staging loads in flight per thread

// Per-centre rows of NN + 1 words (odd stride: a wavefront's lanes on
// consecutive centres, or on consecutive neighbours of one centre, fall on
// different banks).
template <int NN>
__host__ __device__ constexpr int tile_fixed_lds(int ww) {
  return ((256 * (NN + 1) * 4 + 15) & ~15) + ((ww * ww * 2 + 15) & ~15);
}

template <bool CL, bool NL>
__device__ __forceinline__ float ps_dot(const float4 *__restrict__ a, const float4 *__restrict__ b,
                                        int nifg, int s4, float fn) {
#pragma clang fp contract(off)
  float acc = 0.f;
  for (int jj = 0; jj < s4; ++jj) {
    const float4 u = a[jj], v = b[jj];
    acc += u.x * v.x + u.y * v.y;
    if (2 * jj + 1 < nifg) acc += u.z * v.z + u.w * v.w;
  }
  return acc / fn;
}

template <int NN, bool MEDIAN>
__global__ __launch_bounds__(256) void k_ps_tile(const float4 *__restrict__ zt, int nifg, int s4,
                                                 int nrow, int ncol, const unsigned *__restrict__ bits,
                                                 const int *__restrict__ tab, int noff, int nmax,
                                                 int halo, int cap, int per_xcd,
                                                 float *__restrict__ out) {
#pragma clang fp contract(off)
  extern __shared__ __align__(16) unsigned char lds_raw[];
  __shared__ int cpix[256], ccnt[256], wsum[4], spix_dummy;
  const int ww = kPT + 2 * halo;
  constexpr int PS = NN + 1;
  int *pr = reinterpret_cast<int *>(lds_raw);  // [256][PS]
  short *smap = reinterpret_cast<short *>(lds_raw + ((256 * PS * 4 + 15) & ~15));  // [ww][ww]
  const int ss = s4 | 1;  // staged series stride (odd: slots spread over the banks)
  float4 *ser = reinterpret_cast<float4 *>(lds_raw + tile_fixed_lds<NN>(ww));  // [cap][ss]
  int *spix = reinterpret_cast<int *>(ser + (size_t)cap * ss);      // [cap]
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int tiles_c = (ncol + kPT - 1) / kPT, tiles_r = (nrow + kPT - 1) / kPT;
  // XCD-contiguous tiles: workgroup b runs on XCD b % 8
  const int t = (blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
  if (t >= tiles_r * tiles_c) return;  // whole block, before any barrier
  const int tr = t / tiles_c, tc = t - tr * tiles_c;
  const int r0 = tr * kPT - halo, c0 = tc * kPT - halo;
  (void)spix_dummy;
  // 1. window slot map, the staged pixels in window order
  int base = 0;
  for (int i0 = 0; i0 < ww * ww; i0 += 256) {
    const int i = i0 + tid;
    int q = -1;
    bool el = false;
    if (i < ww * ww) {
      const int r = r0 + i / ww, c = c0 + i % ww;
      if ((unsigned)r < (unsigned)nrow && (unsigned)c < (unsigned)ncol) {
        q = r * ncol + c;
        el = (bits[q >> 5] >> (q & 31)) & 1u;
      }
    }
    const unsigned long long m = __ballot(el);
    if (lane == 0) wsum[w] = __popcll(m);
    __syncthreads();
    int before = base;
    for (int k = 0; k < w; ++k) before += wsum[k];
    const int total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    const int slot = before + __popcll(m & ((1ull << lane) - 1));
    if (i < ww * ww) smap[i] = !el ? (short)-1 : (slot < cap ? (short)slot : (short)-2);
    if (el && slot < cap) spix[slot] = q;
    base += total;
    __syncthreads();
  }
  const int E = min(base, cap);
  for (int u0 = tid; u0 < E * s4; u0 += kStageLoads * 256) {
    float4 v[kStageLoads];  // loads in flight before their stores
#pragma unroll
    for (int k = 0; k < kStageLoads; ++k) {
      const int u = u0 + k * 256;
      if (u < E * s4) {
        const int sl = u / s4;
        v[k] = zt[(size_t)spix[sl] * s4 + (u - sl * s4)];
      }
    }
#pragma unroll
    for (int k = 0; k < kStageLoads; ++k) {
      const int u = u0 + k * 256;
      if (u < E * s4) {
        const int sl = u / s4;
        ser[sl * ss + (u - sl * s4)] = v[k];
      }
    }
  }
  // 2. the tile's centres, compacted
  {
    const int ty = tid / kPT, tx = tid - ty * kPT;
    const int r = tr * kPT + ty, c = tc * kPT + tx;
    bool isc = false;
    int q = -1;
    if (r < nrow && c < ncol) {
      q = r * ncol + c;
      const bool el = (bits[q >> 5] >> (q & 31)) & 1u;
      isc = MEDIAN ? el : !el;
    }
    const unsigned long long m = __ballot(isc);
    if (lane == 0) wsum[w] = __popcll(m);
    __syncthreads();
    int before = 0;
    for (int k = 0; k < w; ++k) before += wsum[k];
    if (isc) cpix[before + __popcll(m & ((1ull << lane) - 1))] = q;
  }
  __syncthreads();  // cpix written
  const int nc = wsum[0] + wsum[1] + wsum[2] + wsum[3];
  // 3. phase A: first nmax eligible neighbours in exploration order
  if (tid < nc) {
    const int q = cpix[tid];
    const int r = q / ncol, c = q - r * ncol;
    const int wr = r - r0, wc = c - c0;
    int cnt = 0;
    int *mine = pr + tid * PS;
    const int cw = wr * ww + wc;  // the centre's window index
    // kScan offsets per step, branch-free so that their table entries,
    // slots and mask words are all in flight together, then taken in
    // order.  An offset within the halo (a uniform test) stays inside the
    // window for every centre of the tile (the map holds -1 for window
    // pixels outside the image); beyond it the mask word decides.  Reads
    // that a step does not need use harmless addresses (the centre).
    for (int i0 = 0; i0 < noff && cnt < nmax; i0 += kScan) {
      int ref[kScan], qv[kScan], sv[kScan];
      unsigned wv[kScan];
      bool inh[kScan], inimg[kScan], ok[kScan];
      int ov[kScan];  // the step's table entries (the table is padded to whole steps)
      {
        const int4 *t4 = reinterpret_cast<const int4 *>(tab + i0);
        const int4 a = t4[0], b = t4[1];
        ov[0] = a.x; ov[1] = a.y; ov[2] = a.z; ov[3] = a.w;
        ov[4] = b.x; ov[5] = b.y; ov[6] = b.z; ov[7] = b.w;
      }
      bool allh = true;  // the whole step inside the halo (uniform)
#pragma unroll
      for (int u = 0; u < kScan; ++u) {
        const int dr = ov[u] >> 16, dc = (int)(short)(ov[u] & 0xffff);
        allh = allh && abs(dr) <= halo && abs(dc) <= halo;
      }
      if (allh) {
        // the window map alone (the usual case: the rings within the halo
        // come first)
#pragma unroll
        for (int u = 0; u < kScan; ++u) {
          const int dr = ov[u] >> 16, dc = (int)(short)(ov[u] & 0xffff);
          sv[u] = smap[cw + dr * ww + dc];
        }
#pragma unroll
        for (int u = 0; u < kScan; ++u) {
          const int dr = ov[u] >> 16, dc = (int)(short)(ov[u] & 0xffff);
          const int gref = -(q + dr * ncol + dc) - 1;
          ref[u] = (i0 + u >= noff) ? INT_MIN
                                    : (sv[u] >= 0 ? sv[u] : (sv[u] == -2 ? gref : INT_MIN));
        }
      } else {
#pragma unroll
        for (int u = 0; u < kScan; ++u) {
          ok[u] = i0 + u < noff;
          const int o = ov[u];
          const int dr = o >> 16, dc = (int)(short)(o & 0xffff);
          inh[u] = abs(dr) <= halo && abs(dc) <= halo;
          const int rr = r + dr, cc = c + dc;
          inimg[u] = (unsigned)rr < (unsigned)nrow && (unsigned)cc < (unsigned)ncol;
          qv[u] = inimg[u] ? rr * ncol + cc : q;
          sv[u] = smap[cw + (inh[u] ? dr * ww + dc : 0)];
          wv[u] = bits[qv[u] >> 5];
        }
#pragma unroll
        for (int u = 0; u < kScan; ++u) {
          const bool elg = inimg[u] && ((wv[u] >> (qv[u] & 31)) & 1u);
          const int gref = -qv[u] - 1;
          const int lref = sv[u] >= 0 ? sv[u] : (sv[u] == -2 ? gref : INT_MIN);
          ref[u] = !ok[u] ? INT_MIN : (inh[u] ? lref : (elg ? gref : INT_MIN));
        }
      }
      // record the step's eligible neighbours in order while fewer than
      // nmax: positions by a running count, the rest written to the row's
      // spare entry NN (never read)
      int pos = cnt;
#pragma unroll
      for (int u = 0; u < kScan; ++u) {
        const bool take = ref[u] != INT_MIN && pos < nmax;
        mine[take ? pos : NN] = ref[u];
        pos += take ? 1 : 0;
      }
      cnt = pos;
    }
    ccnt[tid] = cnt;
  }
  __syncthreads();
  const float fn = (float)nifg;
  if (!MEDIAN && s4 <= kChunk / 2) {
    // maximum pass, series of up to kChunk interferograms: one thread per
    // centre (non-PS pixels, not staged), its series in registers, read
    // once; the neighbours from LDS (or global memory past the window).
    // (The median pass has few centres per tile; it spreads the
    // (centre, neighbour) pairs over all threads below instead.)
    if (tid < nc) {
      const int q = cpix[tid];
      float4 cen[kChunk / 2];
#pragma unroll
      for (int jj = 0; jj < kChunk / 2; ++jj)
        cen[jj] = jj < s4 ? zt[(size_t)q * s4 + jj] : make_float4(0.f, 0.f, 0.f, 0.f);
      const int cnt = ccnt[tid];
      float best = -1.f;
      for (int k = 0; k < cnt; ++k) {
        const int ref = pr[tid * PS + k];
        float acc = 0.f;
        if (ref >= 0) {
          const float4 *nb = ser + (size_t)ref * ss;
#pragma unroll
          for (int jj = 0; jj < kChunk / 2; ++jj) {
            if (jj < s4) {
              const float4 v = nb[jj];
              acc += cen[jj].x * v.x + cen[jj].y * v.y;
              if (2 * jj + 1 < nifg) acc += cen[jj].z * v.z + cen[jj].w * v.w;
            }
          }
        } else {
          const float4 *nb = zt + (size_t)(-ref - 1) * s4;
#pragma unroll
          for (int jj = 0; jj < kChunk / 2; ++jj) {
            if (jj < s4) {
              const float4 v = nb[jj];
              acc += cen[jj].x * v.x + cen[jj].y * v.y;
              if (2 * jj + 1 < nifg) acc += cen[jj].z * v.z + cen[jj].w * v.w;
            }
          }
        }
        const float sim = acc / fn;
        if (best < sim) best = sim;
      }
      out[q] = cnt ? best : 0.f;
    }
    return;  // no barrier follows
  }
  // 4. phase B: the (centre, neighbour) similarities
  for (int pidx = tid; pidx < nc * NN; pidx += 256) {
    const int ci = pidx / NN, k = pidx - ci * NN;
    if (k >= ccnt[ci]) continue;
    const int q = cpix[ci];
    const int ref = pr[ci * PS + k];
    int cs = -1;  // the centre's slot (median pass: a staged candidate)
    if (MEDIAN) {
      const int r = q / ncol, c = q - r * ncol;
      cs = smap[(r - r0) * ww + (c - c0)];
    }
    float sim;
    if (ref >= 0) {
      if (cs >= 0)
        sim = ps_dot<true, true>(ser + (size_t)cs * ss, ser + (size_t)ref * ss, nifg, s4, fn);
      else
        sim = ps_dot<false, true>(zt + (size_t)q * s4, ser + (size_t)ref * ss, nifg, s4, fn);
    } else {
      const float4 *nb = zt + (size_t)(-ref - 1) * s4;
      if (cs >= 0)
        sim = ps_dot<true, false>(ser + (size_t)cs * ss, nb, nifg, s4, fn);
      else
        sim = ps_dot<false, false>(zt + (size_t)q * s4, nb, nifg, s4, fn);
    }
    pr[ci * PS + k] = __float_as_int(sim);
  }
  __syncthreads();
  // 5. phase C: median or maximum per centre
  if (tid < nc) {
    const int cnt = ccnt[tid];
    float res = 0.f;
    if (MEDIAN) {
      constexpr int H = NN / 2 + 1;
      float srt[H];
#pragma unroll
      for (int h = 0; h < H; ++h) srt[h] = __builtin_inff();
      for (int k = 0; k < cnt; ++k) {
        float v = __int_as_float(pr[tid * PS + k]);
#pragma unroll
        for (int h = 0; h < H; ++h) {
          const float lo = fminf(srt[h], v);
          v = fmaxf(srt[h], v);
          srt[h] = lo;
        }
      }
      const int m = cnt / 2;
#pragma unroll
      for (int h = 0; h < H; ++h)
        if (h == m) res = srt[h];
      if (cnt == 0) res = 0.f;
    } else {
      float best = -1.f;
      for (int k = 0; k < cnt; ++k) {
        const float v = __int_as_float(pr[tid * PS + k]);
        if (best < v) best = v;
      }
      res = cnt ? best : 0.f;
    }
    out[cpix[tid]] = res;
  }
}

struct PsWork {
  float2 *zt;
  unsigned *bits;
  int *counts, *base, *total, *list;
  int npix, ntiles, stride;
};

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

int ps_check(s1ifg_ctx *ctx, const void *stack, int nifg, int nrow, int ncol, const void *ps,
             int nneigh, int rdmax) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!stack || !ps) return set_error(S1IFG_EINVAL, "null stack or candidate mask");
  if (nifg < 1 || nrow < 1 || ncol < 1)
    return set_error(S1IFG_EINVAL, "stack of %d x %d x %d", nifg, nrow, ncol);
  if ((int64_t)nrow * ncol > ((int64_t)1 << 31) - 2 * kTile)
    return set_error(S1IFG_ERANGE, "image of %d x %d pixels exceeds 2^31", nrow, ncol);
  if (nneigh < 0 || nneigh > 128)
    return set_error(S1IFG_ERANGE, "nneigh %d outside [0, 128]", nneigh);
  if (rdmax > kMaxRd) return set_error(S1IFG_ERANGE, "rdmax %d exceeds %d", rdmax, kMaxRd);
  return S1IFG_OK;
}

int ps_workspace(s1ifg_ctx *ctx, int nifg, int npix, PsWork *wk) {
  wk->npix = npix;
  wk->ntiles = (npix + kTile - 1) / kTile;
  wk->stride = (nifg + 1) & ~1;
  const size_t bz = align256((size_t)npix * wk->stride * sizeof(float2));
  const size_t bb = align256((size_t)wk->ntiles * 2 * sizeof(unsigned));
  const size_t bt = align256((size_t)wk->ntiles * sizeof(int));
  const size_t bl = align256((size_t)npix * sizeof(int));
  void *ws = nullptr;
  int rc = ctx_workspace(ctx, bz + bb + 2 * bt + 256 + bl, &ws);
  if (rc) return rc;
  char *b = static_cast<char *>(ws);
  wk->zt = reinterpret_cast<float2 *>(b);
  b += bz;
  wk->bits = reinterpret_cast<unsigned *>(b);
  b += bb;
  wk->counts = reinterpret_cast<int *>(b);
  b += bt;
  wk->base = reinterpret_cast<int *>(b);
  b += bt;
  wk->total = reinterpret_cast<int *>(b);
  b += 256;
  wk->list = reinterpret_cast<int *>(b);
  return S1IFG_OK;
}

int launch_prep(s1ifg_ctx *ctx, const PsWork &wk, const s1ifg_cf32 *stack, int nifg, int normalize,
                const int32_t *ps, float *fill, int fill_cand, float fill_value) {
  const unsigned blocks = (unsigned)((wk.ntiles + 3) / 4);
  hipLaunchKernelGGL(k_ps_prep, dim3(blocks), dim3(256), 0, ctx->stream,
                     reinterpret_cast<const float2 *>(stack), nifg, wk.npix, wk.stride, normalize,
                     reinterpret_cast<const int *>(ps), wk.zt, wk.bits, wk.counts, fill, fill_cand,
                     fill_value);
  S1_LAUNCH_CHECK("k_ps_prep");
  return S1IFG_OK;
}

int launch_list(s1ifg_ctx *ctx, const PsWork &wk, int invert) {
  hipLaunchKernelGGL(k_ps_scan, dim3(1), dim3(1024), 0, ctx->stream, wk.counts, wk.ntiles, wk.npix,
                     invert, wk.base, wk.total);
  S1_LAUNCH_CHECK("k_ps_scan");
  hipLaunchKernelGGL(k_ps_compact, dim3((unsigned)((wk.ntiles + 3) / 4)), dim3(256), 0, ctx->stream,
                     wk.bits, wk.base, wk.npix, invert, wk.list);
  S1_LAUNCH_CHECK("k_ps_compact");
  return S1IFG_OK;
}

template <int NN, bool MEDIAN>
int launch_sim_nn(s1ifg_ctx *ctx, const PsWork &wk, int nifg, int nrow, int ncol, const int32_t *tab,
                  int noff, int nmax, float *out) {
  constexpr int WPB = waves_per_block<NN>();
  const bool multi = nifg > kChunk;
  const size_t lds = (size_t)WPB * NN * 64 * sizeof(int) * (multi ? 2 : 1);
  const long chunks = (wk.npix + 63) / 64;
  const long want = (chunks + WPB - 1) / WPB;
  const long cap = (long)device_cu_count(ctx->device) * 16;
  // a multiple of the XCD count (each XCD works on its own band of the list)
  const long g = std::max(1L, std::min(want, cap));
  const unsigned grid = (unsigned)((g + kXcds - 1) / kXcds * kXcds);
  auto kern = multi ? k_ps_sim<NN, MEDIAN, true> : k_ps_sim<NN, MEDIAN, false>;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * WPB), lds, ctx->stream,
                     reinterpret_cast<const float4 *>(wk.zt), nifg, wk.stride / 2, nrow, ncol,
                     wk.bits, wk.list, wk.total, reinterpret_cast<const int *>(tab), noff, nmax,
                     out);
  S1_LAUNCH_CHECK("k_ps_sim");
  return S1IFG_OK;
}

// The tiled form for nmax <= 64 when at least 32 series fit in LDS;
// *used = false leaves the pass to the list form.
template <int NN, bool MEDIAN>
int launch_tile_nn(s1ifg_ctx *ctx, const PsWork &wk, int nifg, int nrow, int ncol,
                   const int32_t *tab, int noff, int extent, int nmax, float *out, bool *used) {
  const int halo = std::min(extent, kHalo);
  const int ww = kPT + 2 * halo;
  const int s4 = wk.stride / 2;
  const int static_lds = 2 * 256 * 4 + 64;
  const int budget = (NN <= 32 ? kTileLds : 2 * kTileLds - 4096) - static_lds;
  const int fixed = tile_fixed_lds<NN>(ww);
  const int ss = s4 | 1;
  const int cap = std::min(1024, (budget - fixed) / (ss * 16 + 4));
  *used = false;
  if (cap < 32) return S1IFG_OK;
  const int dyn = fixed + cap * (ss * 16 + 4);
  auto kern = k_ps_tile<NN, MEDIAN>;
  S1_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                             hipFuncAttributeMaxDynamicSharedMemorySize, dyn));
  const int ntiles = ((nrow + kPT - 1) / kPT) * ((ncol + kPT - 1) / kPT);
  const int per_xcd = (ntiles + kXcds - 1) / kXcds;
  hipLaunchKernelGGL(kern, dim3((unsigned)(per_xcd * kXcds)), dim3(256), dyn, ctx->stream,
                     reinterpret_cast<const float4 *>(wk.zt), nifg, s4, nrow, ncol, wk.bits,
                     reinterpret_cast<const int *>(tab), noff, nmax, halo, cap, per_xcd, out);
  S1_LAUNCH_CHECK("k_ps_tile");
  *used = true;
  return S1IFG_OK;
}

template <bool MEDIAN>
int launch_sim(s1ifg_ctx *ctx, const PsWork &wk, int nifg, int nrow, int ncol, const int32_t *tab,
               int noff, int nneigh, float *out) {
  const int nmax = std::max(nneigh, 1);
  if (nmax <= 32) return launch_sim_nn<32, MEDIAN>(ctx, wk, nifg, nrow, ncol, tab, noff, nmax, out);
  if (nmax <= 64) return launch_sim_nn<64, MEDIAN>(ctx, wk, nifg, nrow, ncol, tab, noff, nmax, out);
  return launch_sim_nn<128, MEDIAN>(ctx, wk, nifg, nrow, ncol, tab, noff, nmax, out);
}

// One similarity pass over the pixels whose bit is set (median) or clear
// (maximum): the tiled form where it applies, else list + k_ps_sim.
template <bool MEDIAN>
int run_pass(s1ifg_ctx *ctx, const PsWork &wk, int nifg, int nrow, int ncol, const int32_t *tab,
             int noff, int extent, int nneigh, float *out) {
  const int nmax = std::max(nneigh, 1);
  bool used = false;
  int rc = S1IFG_OK;
  if (nmax <= 32)
    rc = launch_tile_nn<32, MEDIAN>(ctx, wk, nifg, nrow, ncol, tab, noff, extent, nmax, out, &used);
  else if (nmax <= 64)
    rc = launch_tile_nn<64, MEDIAN>(ctx, wk, nifg, nrow, ncol, tab, noff, extent, nmax, out, &used);
  if (rc || used) return rc;
  if ((rc = launch_list(ctx, wk, MEDIAN ? 0 : 1))) return rc;
  return launch_sim<MEDIAN>(ctx, wk, nifg, nrow, ncol, tab, noff, nneigh, out);
}

}  // namespace

namespace s1ifg {
void ps_cache_free(s1ifg_ctx *ctx) {
  PsCache *c = static_cast<PsCache *>(ctx->ps_cache);
  if (!c) return;
  if (c->dev) (void)hipFree(c->dev);
  delete c;
  ctx->ps_cache = nullptr;
}
}  // namespace s1ifg

int s1ifg_ps_scan_offsets(int rdmin, int rdmax, int32_t *offsets, int cap) {
  if (rdmax > kMaxRd) return set_error(S1IFG_ERANGE, "rdmax %d exceeds %d", rdmax, kMaxRd);
  if (cap < 0 || (cap > 0 && !offsets)) return set_error(S1IFG_EINVAL, "bad offset array");
  std::vector<int32_t> t;
  scan_offsets(std::max(rdmin, 0), rdmax, t);
  const int n = (int)t.size();
  for (int i = 0; i < std::min(n, cap); ++i) {
    offsets[2 * i] = t[i] >> 16;
    offsets[2 * i + 1] = (int32_t)(int16_t)(t[i] & 0xffff);
  }
  return n;
}

int s1ifg_ps_normalize(s1ifg_ctx *ctx, const s1ifg_cf32 *in, s1ifg_cf32 *out, int64_t n) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (n < 0 || (n > 0 && (!in || !out))) return set_error(S1IFG_EINVAL, "bad normalize arguments");
  if (n == 0) return S1IFG_OK;
  hipLaunchKernelGGL(k_ps_normalize, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                     reinterpret_cast<const float2 *>(in), reinterpret_cast<float2 *>(out), (long)n);
  S1_LAUNCH_CHECK("k_ps_normalize");
  return S1IFG_OK;
}

int s1ifg_ps_update(s1ifg_ctx *ctx, int32_t *ps, const float *sim, float threshold, int overwrite,
                    int64_t n) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (n < 0 || (n > 0 && (!ps || !sim))) return set_error(S1IFG_EINVAL, "bad update arguments");
  if (n == 0) return S1IFG_OK;
  hipLaunchKernelGGL(k_ps_update, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                     reinterpret_cast<int *>(ps), sim, threshold, overwrite, (long)n);
  S1_LAUNCH_CHECK("k_ps_update");
  return S1IFG_OK;
}

int s1ifg_ps_median_similarity(s1ifg_ctx *ctx, const s1ifg_cf32 *stack_norm, int nifg, int nrow,
                               int ncol, const int32_t *ps, int nneigh, int rdmin, int rdmax,
                               float *sim_median) {
  int rc = ps_check(ctx, stack_norm, nifg, nrow, ncol, ps, nneigh, rdmax);
  if (rc) return rc;
  if (!sim_median) return set_error(S1IFG_EINVAL, "null output");
  const int32_t *tab;
  int noff, extent;
  PsWork wk;
  if ((rc = ps_table(ctx, rdmin, rdmax, &tab, &noff, &extent))) return rc;
  if ((rc = ps_workspace(ctx, nifg, nrow * ncol, &wk))) return rc;
  if ((rc = launch_prep(ctx, wk, stack_norm, nifg, 0, ps, sim_median, 0, 0.f))) return rc;
  return run_pass<true>(ctx, wk, nifg, nrow, ncol, tab, noff, extent, nneigh, sim_median);
}

int s1ifg_ps_maximum_similarity(s1ifg_ctx *ctx, const s1ifg_cf32 *stack_norm, int nifg, int nrow,
                                int ncol, const int32_t *ps, int nneigh, int rdmin, int rdmax,
                                float *sim_max) {
  int rc = ps_check(ctx, stack_norm, nifg, nrow, ncol, ps, nneigh, rdmax);
  if (rc) return rc;
  if (!sim_max) return set_error(S1IFG_EINVAL, "null output");
  const int32_t *tab;
  int noff, extent;
  PsWork wk;
  if ((rc = ps_table(ctx, rdmin, rdmax, &tab, &noff, &extent))) return rc;
  if ((rc = ps_workspace(ctx, nifg, nrow * ncol, &wk))) return rc;
  if ((rc = launch_prep(ctx, wk, stack_norm, nifg, 0, ps, sim_max, 1, 1.f))) return rc;
  return run_pass<false>(ctx, wk, nifg, nrow, ncol, tab, noff, extent, nneigh, sim_max);
}

int s1ifg_ps_similarity(s1ifg_ctx *ctx, const s1ifg_cf32 *stack, int nifg, int nrow, int ncol,
                        int32_t *ps, int nneigh, int rdmin, int rdmax, float med_sim_th,
                        float *med_sim, float *max_sim) {
  int rc = ps_check(ctx, stack, nifg, nrow, ncol, ps, nneigh, rdmax);
  if (rc) return rc;
  if (!med_sim || !max_sim) return set_error(S1IFG_EINVAL, "null output");
  const int32_t *tab;
  int noff, extent;
  PsWork wk;
  if ((rc = ps_table(ctx, rdmin, rdmax, &tab, &noff, &extent))) return rc;
  if ((rc = ps_workspace(ctx, nifg, nrow * ncol, &wk))) return rc;
  if ((rc = launch_prep(ctx, wk, stack, nifg, 1, ps, med_sim, 0, 0.f))) return rc;
  if ((rc = run_pass<true>(ctx, wk, nifg, nrow, ncol, tab, noff, extent, nneigh, med_sim)))
    return rc;
  hipLaunchKernelGGL(k_ps_update_bits, dim3((unsigned)((wk.ntiles + 3) / 4)), dim3(256), 0,
                     ctx->stream, reinterpret_cast<int *>(ps), med_sim, med_sim_th, wk.npix, wk.bits,
                     wk.counts, max_sim);
  S1_LAUNCH_CHECK("k_ps_update_bits");
  return run_pass<false>(ctx, wk, nifg, nrow, ncol, tab, noff, extent, nneigh, max_sim);
}
