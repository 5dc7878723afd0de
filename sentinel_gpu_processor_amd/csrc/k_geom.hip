// Geometry-path kernels: grid tables, orbit evaluation, deramp, geo2rdr,
// geocode, valid-row scan, and the fused burst-pair kernel.
#include <mutex>

#include "geom_dev.h"
#include "pair_lds_dev.h"

using namespace s1ifg;

namespace {

// Row / column tables of the target grid: trigonometry of llh2xyz
// (orbit.cu:15-35) and the DEM sampling coordinates.
__global__ void k_grid_tables(GridDev gr, DemDev dem, RowTab *rows, ColTab *cols) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  grid_table_entry(gr, dem, i, rows, cols);
}

__global__ void k_orbit_eval(OrbitDev o, const double *t, long n, double *pos, double *vel,
                             double *acc) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x[3], v[3], a[3];
  orbit_eval(o, t[i], x, v, a);
  for (int k = 0; k < 3; ++k) {
    if (pos) pos[i * 3 + k] = x[k];
    if (vel) vel[i * 3 + k] = v[k];
    if (acc) acc[i * 3 + k] = a[k];
  }
}

// deramp_burst (geo2rdr.cu:138-158): phase in FP64 (eta as geo2rdr.cu:529-532),
// reduced to [-pi, pi] before a float sincos.  Block = 256 sample pairs x
// kDerampRows rows (deramp_tile).  in and out may alias (in-place deramp).
__global__ __launch_bounds__(256) void k_deramp(GeomDev g, const float2 *in, float2 *out, int vec) {
  deramp_tile(g, in, out, blockIdx.y * kDerampRows, blockIdx.x * blockDim.x, vec != 0);
}

// The deramp of a CInt16 burst into complex float (deramp_tile_ci16).
__global__ __launch_bounds__(256) void k_deramp_ci16(GeomDev g, const short *in, float2 *out,
                                                     int vec) {
  deramp_tile_ci16(g, in, out, blockIdx.y * kDerampRows, blockIdx.x * blockDim.x, vec != 0);
}

__global__ void k_geo2rdr(const GeomDev *__restrict__ gp, OffPoly op, GridDev gr, DemDev dem, const RowTab *rows,
                          const ColTab *cols, const float *dh, double *rgoff, double *azoff,
                          double *rng) {
  const GeomDev &g = const_ref(gp);
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  int r = blockIdx.y * blockDim.y + threadIdx.y;
  if (r >= gr.nrow || c >= gr.ncol) return;
  double X[3];
  pixel_xyz(gr, dem, rows, cols, dh, r, c, X);
  double t, dr[3];
  zero_doppler(g, X, t, dr);
  double R = sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
  double rg = (R - g.rngstart) * g.inv_dmrg;
  double az = (t - g.tstart) * g.inv_dtaz;
  if (op.on) {
    double xs = rg * g.inv_samples, ys = az * g.inv_lines;
    double drg = op.p[0] + op.p[1] * xs + op.p[2] * ys;
    double daz = op.p[3] + op.p[4] * xs + op.p[5] * ys;
    rg += drg;
    az += daz;
  }
  size_t o = (size_t)r * gr.ncol + c;
  if (rgoff) rgoff[o] = rg;
  if (azoff) azoff[o] = az;
  if (rng) rng[o] = R;
}

template <int INTERP>
__global__ void k_geocode(const GeomDev *__restrict__ gp, OffPoly op, GridDev gr, DemDev dem, const RowTab *rows,
                          const ColTab *cols, const float *dh, const float2 *__restrict__ z,
                          float2 *__restrict__ out) {
  const GeomDev &g = const_ref(gp);
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  int r = blockIdx.y * blockDim.y + threadIdx.y;
  if (r >= gr.nrow || c >= gr.ncol) return;
  double X[3];
  pixel_xyz(gr, dem, rows, cols, dh, r, c, X);
  out[(size_t)r * gr.ncol + c] = geocode_pixel<INTERP>(g, op, X, z);
}

// Fused burst pair: one thread per multilooked pixel; the look cell is
// accumulated as column means then a row mean (the order of
// cpx_col_look / cpx_row_look, sario.cu:241-286).
template <int INTERP, bool RAW>
__global__ void k_pair(const GeomDev *__restrict__ gpr, const GeomDev *__restrict__ gps, OffPoly ops, GridDev gr, DemDev dem,
                       const RowTab *rows, const ColTab *cols, const float *dh,
                       const float2 *__restrict__ zr, const float2 *__restrict__ zs, int rl,
                       int cl, int nr_ml, int nc_ml, float2 *__restrict__ ifg,
                       float *__restrict__ pr, float *__restrict__ ps) {
  int cm = blockIdx.x * blockDim.x + threadIdx.x;
  int rm = blockIdx.y * blockDim.y + threadIdx.y;
  if (rm >= nr_ml || cm >= nc_ml) return;
  pair_ml_pixel<INTERP, RAW>(const_ref(gpr), const_ref(gps), ops, gr, dem, rows, cols, dh, zr, zs, rl, cl, rm, cm, nc_ml,
                        ifg, pr, ps);
}

// Sinc pair kernel with LDS-staged burst windows: one 16 x 16 tile of look
// cells per block (rl * cl <= kPairMaxPix).
template <int NP, int CL, bool RAW>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_pair_tiles(
    const GeomDev *__restrict__ gpr, const GeomDev *__restrict__ gps, OffPoly ops, GridDev gr,
    DemDev dem, const RowTab *rows, const ColTab *cols, const float *dh,
    const float2 *__restrict__ zr, const float2 *__restrict__ zs, int rl, int cl, int nr_ml,
    int nc_ml, float2 *__restrict__ ifg, float *__restrict__ pr, float *__restrict__ ps,
    unsigned long long *stats) {
  __shared__ PairLds L;
  pair_tile_sinc<NP, CL, RAW>(const_ref(gpr), const_ref(gps), ops, gr, dem, rows, cols, dh, zr, zs, rl, cl, blockIdx.y,
                 blockIdx.x, nr_ml, nc_ml, ifg, pr, ps, L, stats);
}

__global__ void k_rows_init(int *fl, int nrow) {
  if (threadIdx.x == 0) {
    fl[0] = nrow;
    fl[1] = -1;
  }
}

// find_first_last_nonzero_rows_kernel (geo2rdr.cu:86-128): one block per row.
__global__ void k_rows_scan(const float2 *__restrict__ img, int nrow, int ncol, int *fl) {
  int row = blockIdx.x;
  const float2 *p = img + (size_t)row * ncol;
  int any = 0;
  for (int c = threadIdx.x; c < ncol; c += blockDim.x) {
    float2 v = p[c];
    if (v.x != 0.0f || v.y != 0.0f) {
      any = 1;
      break;
    }
  }
  any = __syncthreads_or(any);
  if (threadIdx.x == 0 && any) {
    atomicMin(&fl[0], row);
    atomicMax(&fl[1], row + 1);
  }
}

__global__ void k_rows_final(int *fl, int nrow) {
  if (threadIdx.x == 0 && fl[0] == nrow) {
    fl[0] = 0;
    fl[1] = 0;
  }
}

// Byte offsets of the column table and of what follows the tables.
size_t grid_tables_col_offset(const GridDev &gr) {
  return (sizeof(RowTab) * (size_t)gr.nrow + 255) & ~(size_t)255;
}
size_t grid_tables_bytes(const GridDev &gr) {
  return (grid_tables_col_offset(gr) + sizeof(ColTab) * (size_t)gr.ncol + 255) & ~(size_t)255;
}

// Fill the grid tables at base (grid_tables_bytes(gr) bytes of device memory).
int grid_tables_at(s1ifg_ctx *ctx, const GridDev &gr, const DemDev &dem, void *base,
                   RowTab **rows, ColTab **cols) {
  *rows = (RowTab *)base;
  *cols = (ColTab *)((char *)base + grid_tables_col_offset(gr));
  int n = gr.nrow > gr.ncol ? gr.nrow : gr.ncol;
  hipLaunchKernelGGL(k_grid_tables, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, gr, dem,
                     *rows, *cols);
  S1_LAUNCH_CHECK("k_grid_tables");
  return S1IFG_OK;
}

}  // namespace

namespace s1ifg {

// Grid tables in the context workspace, followed by `extra` bytes of
// scratch for the caller (256-byte aligned; *extra_ptr may be NULL).
int grid_tables(s1ifg_ctx *ctx, const GridDev &gr, const DemDev &dem, RowTab **rows,
                ColTab **cols, size_t extra, void **extra_ptr) {
  const size_t off_x = grid_tables_bytes(gr);
  void *ws = nullptr;
  int rc = ctx_workspace(ctx, off_x + extra + 256, &ws);
  if (rc) return rc;
  if (extra_ptr) *extra_ptr = (char *)ws + off_x;
  return grid_tables_at(ctx, gr, dem, ws, rows, cols);
}

}  // namespace s1ifg

namespace {

// Device memory of the allocation-free pair entry point: the grid tables of
// one call, in a file-scope device array of fixed size.  Each device holds
// its own copy for the life of the process; host code reaches it only
// through the address hipGetSymbolAddress returns on that device (the
// array's name is a host-side stand-in, not a device address).  Calls that
// overlap on one device share it and race.
constexpr size_t kStaticScratchBytes = (size_t)2 << 20;  // 2 MiB
__device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_dev[kMaxDevices];  // device address of `scratch`, per device

// The device address of `scratch` on the context's device (current after
// check_ctx), looked up on the first call on that device and kept.
int static_scratch(s1ifg_ctx *ctx, void **out) {
  if (ctx->device < 0 || ctx->device >= kMaxDevices)
    return set_error(S1IFG_EINVAL, "device %d beyond the %d this entry point supports",
                     ctx->device, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[ctx->device]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_error(e, "hipGetSymbolAddress(scratch)");
    if (!p) return set_error(S1IFG_EHIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_dev[ctx->device] = p;
  }
  *out = g_scratch_dev[ctx->device];
  return S1IFG_OK;
}

}  // namespace

namespace {

int grid_tables(s1ifg_ctx *ctx, const GridDev &gr, const DemDev &dem, RowTab **rows,
                ColTab **cols) {
  return s1ifg::grid_tables(ctx, gr, dem, rows, cols, 0, nullptr);
}

int common_setup(s1ifg_ctx *ctx, const s1ifg_geom *g, const s1ifg_grid *grid,
                 const s1ifg_dem *dem, const double *offpoly, GridDev *gr, DemDev *dd,
                 OffPoly *op) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (g && (rc = check_geom(ctx, g))) return rc;
  if ((rc = make_grid(grid, gr))) return rc;
  if ((rc = make_dem(dem, dd))) return rc;
  if ((rc = make_offpoly(offpoly, op))) return rc;
  return S1IFG_OK;
}

}  // namespace

namespace {

// The fused burst pair with its grid tables in the context workspace
// (static_tables = false), or in the file-scope `scratch` array
// (static_tables = true: nothing is allocated); raw: the bursts are the raw
// (not deramped) ones, deramped as their samples are read.
int pair_impl(s1ifg_ctx *ctx, const s1ifg_geom *g_ref, const s1ifg_geom *g_sec,
              const s1ifg_grid *grid, const s1ifg_dem *dem, const float *dh,
              const double *offpoly_sec, const s1ifg_cf32 *ref_deramped,
              const s1ifg_cf32 *sec_deramped, int interp, int rl, int cl, s1ifg_cf32 *ifg_ml,
              float *pow_ref, float *pow_sec, bool static_tables, bool raw = false) {
  GridDev gr;
  DemDev dd;
  OffPoly op;
  int rc = common_setup(ctx, g_ref, grid, dem, offpoly_sec, &gr, &dd, &op);
  if (rc) return rc;
  if ((rc = check_geom(ctx, g_sec))) return rc;
  if (!ref_deramped || !sec_deramped || !ifg_ml) return set_error(S1IFG_EINVAL, "null data pointer");
  if (rl < 1 || cl < 1) return set_error(S1IFG_EINVAL, "looks %d x %d", rl, cl);
  if (interp != S1IFG_INTERP_BILINEAR && interp != S1IFG_INTERP_SINC)
    return set_error(S1IFG_EINVAL, "interp %d", interp);
  int nr = gr.nrow / rl, nc = gr.ncol / cl;
  if (nr < 1 || nc < 1) return set_error(S1IFG_EINVAL, "grid smaller than one look cell");
  RowTab *rows;
  ColTab *cols;
  if (static_tables) {
    const size_t need = grid_tables_bytes(gr);
    if (need > kStaticScratchBytes)
      return set_error(S1IFG_ERANGE,
                       "grid %d x %d needs %zu bytes of tables, beyond the %zu bytes of the "
                       "allocation-free pair (32 * nrow + 24 * ncol, each rounded up to 256)",
                       gr.nrow, gr.ncol, need, kStaticScratchBytes);
    void *base = nullptr;
    if ((rc = static_scratch(ctx, &base))) return rc;
    if ((rc = grid_tables_at(ctx, gr, dd, base, &rows, &cols))) return rc;
  } else if ((rc = grid_tables(ctx, gr, dd, &rows, &cols))) {
    return rc;
  }
  dim3 blk(64, 4), grd((nc + 63) / 64, (nr + 3) / 4);
  if (interp == S1IFG_INTERP_SINC && rl * cl <= kPairMaxPix) {
    dim3 tg((nc + kPairTile - 1) / kPairTile, (nr + kPairTile - 1) / kPairTile);
    // the 1 x 4 looks of the workloads with the look shape compiled in
    const bool l14 = rl == 1 && cl == 4;
    auto kern = raw ? (l14 ? k_pair_tiles<4, 4, true> : k_pair_tiles<0, 0, true>)
                    : (l14 ? k_pair_tiles<4, 4, false> : k_pair_tiles<0, 0, false>);
    hipLaunchKernelGGL(kern, tg, dim3(256), 0, ctx->stream, g_ref->d_dev, g_sec->d_dev, op, gr,
                       dd, rows, cols, dh, (const float2 *)ref_deramped,
                       (const float2 *)sec_deramped, rl, cl, nr, nc, (float2 *)ifg_ml, pow_ref,
                       pow_sec, ctx->tile_stats);
  } else {
    auto kern = interp == S1IFG_INTERP_BILINEAR
                    ? (raw ? k_pair<S1IFG_INTERP_BILINEAR, true> : k_pair<S1IFG_INTERP_BILINEAR, false>)
                    : (raw ? k_pair<S1IFG_INTERP_SINC, true> : k_pair<S1IFG_INTERP_SINC, false>);
    hipLaunchKernelGGL(kern, grd, blk, 0, ctx->stream, g_ref->d_dev, g_sec->d_dev, op, gr, dd,
                       rows, cols, dh, (const float2 *)ref_deramped, (const float2 *)sec_deramped,
                       rl, cl, nr, nc, (float2 *)ifg_ml, pow_ref, pow_sec);
  }
  S1_LAUNCH_CHECK("k_pair");
  return S1IFG_OK;
}

}  // namespace

extern "C" {

int s1ifg_orbit_eval(s1ifg_ctx *ctx, const s1ifg_geom *g, const double *t, int64_t n,
                     double *pos, double *vel, double *acc) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if ((rc = check_geom(ctx, g))) return rc;
  if (n < 0 || (n > 0 && !t)) return set_error(S1IFG_EINVAL, "bad time array");
  if (n == 0) return S1IFG_OK;
  hipLaunchKernelGGL(k_orbit_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                     g->dev.orb, t, (long)n, pos, vel, acc);
  S1_LAUNCH_CHECK("k_orbit_eval");
  return S1IFG_OK;
}

int s1ifg_deramp(s1ifg_ctx *ctx, const s1ifg_geom *g, const s1ifg_cf32 *in, s1ifg_cf32 *out) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if ((rc = check_geom(ctx, g))) return rc;
  if (!in || !out) return set_error(S1IFG_EINVAL, "null burst pointer");
  const GeomDev &gd = g->dev;
  int pairs = (gd.samples + 1) / 2;
  int bx = (pairs + 255) / 256;
  // float4 path: even row length and 16-byte aligned rows; otherwise float2
  int vec = (gd.samples % 2 == 0) && (((uintptr_t)in | (uintptr_t)out) % 16 == 0);
  hipLaunchKernelGGL(k_deramp, dim3(bx, (gd.lines + kDerampRows - 1) / kDerampRows), dim3(256),
                     0, ctx->stream, gd, (const float2 *)in, (float2 *)out, vec);
  S1_LAUNCH_CHECK("k_deramp");
  return S1IFG_OK;
}

int s1ifg_deramp_ci16(s1ifg_ctx *ctx, const s1ifg_geom *g, const int16_t *in, s1ifg_cf32 *out) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if ((rc = check_geom(ctx, g))) return rc;
  if (!in || !out) return set_error(S1IFG_EINVAL, "null burst pointer");
  const GeomDev &gd = g->dev;
  const uintptr_t ib = (uintptr_t)in, ob = (uintptr_t)out;
  const uintptr_t ie = ib + (size_t)gd.lines * gd.samples * 4;
  const uintptr_t oe = ob + (size_t)gd.lines * gd.samples * 8;
  if (ib < oe && ob < ie) return set_error(S1IFG_EINVAL, "CInt16 input overlaps the output");
  int pairs = (gd.samples + 1) / 2;
  int bx = (pairs + 255) / 256;
  // 8-byte loads / 16-byte stores: even row length and aligned rows
  int vec = (gd.samples % 2 == 0) && (ib % 8 == 0) && (ob % 16 == 0);
  hipLaunchKernelGGL(k_deramp_ci16, dim3(bx, (gd.lines + kDerampRows - 1) / kDerampRows),
                     dim3(256), 0, ctx->stream, gd, (const short *)in, (float2 *)out, vec);
  S1_LAUNCH_CHECK("k_deramp_ci16");
  return S1IFG_OK;
}

int s1ifg_geo2rdr(s1ifg_ctx *ctx, const s1ifg_geom *g, const s1ifg_grid *grid,
                  const s1ifg_dem *dem, const float *dh, const double *offpoly, double *rgoff,
                  double *azoff, double *rng) {
  GridDev gr;
  DemDev dd;
  OffPoly op;
  int rc = common_setup(ctx, g, grid, dem, offpoly, &gr, &dd, &op);
  if (rc) return rc;
  RowTab *rows;
  ColTab *cols;
  if ((rc = grid_tables(ctx, gr, dd, &rows, &cols))) return rc;
  dim3 blk(64, 4), grd((gr.ncol + 63) / 64, (gr.nrow + 3) / 4);
  hipLaunchKernelGGL(k_geo2rdr, grd, blk, 0, ctx->stream, g->d_dev, op, gr, dd, rows, cols, dh,
                     rgoff, azoff, rng);
  S1_LAUNCH_CHECK("k_geo2rdr");
  return S1IFG_OK;
}

int s1ifg_geocode(s1ifg_ctx *ctx, const s1ifg_geom *g, const s1ifg_grid *grid,
                  const s1ifg_dem *dem, const float *dh, const double *offpoly,
                  const s1ifg_cf32 *deramped, int interp, s1ifg_cf32 *out) {
  GridDev gr;
  DemDev dd;
  OffPoly op;
  int rc = common_setup(ctx, g, grid, dem, offpoly, &gr, &dd, &op);
  if (rc) return rc;
  if (!deramped || !out) return set_error(S1IFG_EINVAL, "null data pointer");
  if (interp != S1IFG_INTERP_BILINEAR && interp != S1IFG_INTERP_SINC)
    return set_error(S1IFG_EINVAL, "interp %d", interp);
  RowTab *rows;
  ColTab *cols;
  if ((rc = grid_tables(ctx, gr, dd, &rows, &cols))) return rc;
  dim3 blk(64, 4), grd((gr.ncol + 63) / 64, (gr.nrow + 3) / 4);
  if (interp == S1IFG_INTERP_BILINEAR)
    hipLaunchKernelGGL(k_geocode<S1IFG_INTERP_BILINEAR>, grd, blk, 0, ctx->stream, g->d_dev, op, gr,
                       dd, rows, cols, dh, (const float2 *)deramped, (float2 *)out);
  else
    hipLaunchKernelGGL(k_geocode<S1IFG_INTERP_SINC>, grd, blk, 0, ctx->stream, g->d_dev, op, gr, dd,
                       rows, cols, dh, (const float2 *)deramped, (float2 *)out);
  S1_LAUNCH_CHECK("k_geocode");
  return S1IFG_OK;
}

int s1ifg_valid_rows(s1ifg_ctx *ctx, const s1ifg_cf32 *img, int nrow, int ncol,
                     int32_t *first_last) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!img || !first_last || nrow <= 0 || ncol <= 0)
    return set_error(S1IFG_EINVAL, "bad arguments");
  hipLaunchKernelGGL(k_rows_init, dim3(1), dim3(64), 0, ctx->stream, first_last, nrow);
  hipLaunchKernelGGL(k_rows_scan, dim3(nrow), dim3(256), 0, ctx->stream, (const float2 *)img, nrow,
                     ncol, first_last);
  hipLaunchKernelGGL(k_rows_final, dim3(1), dim3(64), 0, ctx->stream, first_last, nrow);
  S1_LAUNCH_CHECK("k_rows_scan");
  return S1IFG_OK;
}

int s1ifg_pair(s1ifg_ctx *ctx, const s1ifg_geom *g_ref, const s1ifg_geom *g_sec,
               const s1ifg_grid *grid, const s1ifg_dem *dem, const float *dh,
               const double *offpoly_sec, const s1ifg_cf32 *ref_deramped,
               const s1ifg_cf32 *sec_deramped, int interp, int rl, int cl, s1ifg_cf32 *ifg_ml,
               float *pow_ref, float *pow_sec) {
  return pair_impl(ctx, g_ref, g_sec, grid, dem, dh, offpoly_sec, ref_deramped, sec_deramped,
                   interp, rl, cl, ifg_ml, pow_ref, pow_sec, false);
}

int s1ifg_pair_raw(s1ifg_ctx *ctx, const s1ifg_geom *g_ref, const s1ifg_geom *g_sec,
                   const s1ifg_grid *grid, const s1ifg_dem *dem, const float *dh,
                   const double *offpoly_sec, const s1ifg_cf32 *ref_slc, const s1ifg_cf32 *sec_slc,
                   int interp, int rl, int cl, s1ifg_cf32 *ifg_ml, float *pow_ref,
                   float *pow_sec) {
  return pair_impl(ctx, g_ref, g_sec, grid, dem, dh, offpoly_sec, ref_slc, sec_slc, interp, rl,
                   cl, ifg_ml, pow_ref, pow_sec, false, true);
}

int s1ifg_pair_noalloc(s1ifg_ctx *ctx, const s1ifg_geom *g_ref, const s1ifg_geom *g_sec,
                       const s1ifg_grid *grid, const s1ifg_dem *dem, const float *dh,
                       const double *offpoly_sec, const s1ifg_cf32 *ref_deramped,
                       const s1ifg_cf32 *sec_deramped, int interp, int rl, int cl,
                       s1ifg_cf32 *ifg_ml, float *pow_ref, float *pow_sec) {
  return pair_impl(ctx, g_ref, g_sec, grid, dem, dh, offpoly_sec, ref_deramped, sec_deramped,
                   interp, rl, cl, ifg_ml, pow_ref, pow_sec, true);
}

}  // extern "C"
