// Internal declarations shared by the HIP translation units of libs1ifg.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/s1ifg.h"

namespace s1ifg {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 6.28318530717958647692;
constexpr double kSol = 299792458.0;   // geo2rdr.cu:26-28
constexpr double kRa = 6378137.0;      // orbit.hpp:6-8
constexpr double kRe2 = 0.00669437999015;  // orbit.hpp:9-11

int set_error(int code, const char *fmt, ...);
int hip_error(hipError_t e, const char *what);

#define S1_HIP(call)                                        \
  do {                                                      \
    hipError_t e_ = (call);                                 \
    if (e_ != hipSuccess) return ::s1ifg::hip_error(e_, #call); \
  } while (0)

#define S1_LAUNCH_CHECK(name)                               \
  do {                                                      \
    hipError_t e_ = hipGetLastError();                      \
    if (e_ != hipSuccess) return ::s1ifg::hip_error(e_, name); \
  } while (0)

// Orbit in polynomial form: for stencil index k in [1, n-3] (the clamped
// truncated index of orbit.cu:229-232) the 4-node Hermite interpolant is a
// degree-7 polynomial in s = (t - t_k)/h, h = t_1 - t_0.  coef layout:
// [(k-1)*24 + comp*8 + m].
struct OrbitDev {
  const double *coef;
  double t0;     // t_0
  double h;      // t_1 - t_0
  double inv_h;  // 1 / h
  int n;         // state vectors
  int node_m;    // node intervals per stencil interval (0: no node table)
  const double *node;  // node records, (node_m + 1) per stencil index
};

// Orbit node record: [tau, x[3], v[3], a[3], j[3], pad[3]] (position and
// its first three time derivatives at time tau).
constexpr int kNodeRec = 16;

// Everything a kernel needs about one burst, passed by value.
struct GeomDev {
  OrbitDev orb;
  double tstart, dtaz, rngstart, dmrg, wvl, ks, etac0;
  double fm[4];  // t0, c0, c1, c2
  double dc[4];
  double tmid, xmid[3], vmid[3], amid[3];
  double ati;    // azimuth_time_interval (deramp eta)
  double inv_dmrg, inv_dtaz, inv_samples, inv_lines;
  double k4pi_wvl;  // 4 pi / wvl (two-way phase per metre of range)
  double zd_warm_dt2;  // longest squared warm Newton step accepted (zero_doppler_near)
  int lines, samples;
  int flv, llv, fvs, lvs;
  const double *kt;      // [samples]
  const double *etaref;  // [samples]
};

struct GridDev {
  int nrow, ncol;
  double lat0, lon0, dlat, dlon;
};

struct DemDev {
  const void *data;
  int dtype, nrow, ncol;
  double lat0, lon0, dlat, dlon;
};

struct OffPoly {
  double p[6];
  int on;
};

// A read-only parameter block in device memory (a kernel's GeomDev, for
// instance), read through the constant address space: the wavefront-uniform
// fields then come in by scalar loads into SGPRs instead of one vector load
// per lane.
template <typename T>
__device__ __forceinline__ const T &const_ref(const T *p) {
  typedef const __attribute__((address_space(4))) T *CP;
  return *(const T *)(CP)p;
}

}  // namespace s1ifg

struct s1ifg_ctx {
  int device;
  hipStream_t stream;
  void *ws;            // workspace
  size_t ws_bytes;
  int64_t ws_limit;    // optional cap (S1IFG_WORKSPACE_LIMIT env), -1 = none
  unsigned long long *tile_stats;  // optional device counters of the pair kernel's tile classes
  void *ps_cache;      // PS neighbour-offset table on the device (k_ps.hip), or null
};

struct s1ifg_geom {
  int device;
  s1ifg::GeomDev dev;
  double *d_buf;       // device: orbit coef | kt | etaref | GeomDev copy
  const s1ifg::GeomDev *d_dev;  // device copy of `dev` (for kernels taking pointers)
  // host copies for bounds()
  int n_orb;
  double *h_t, *h_pos, *h_vel;
};

namespace s1ifg {
// Workspace of at least `bytes` on ctx (grown with hipMalloc, kept).
int ctx_workspace(s1ifg_ctx *ctx, size_t bytes, void **out);
int check_ctx(s1ifg_ctx *ctx);
int check_geom(s1ifg_ctx *ctx, const s1ifg_geom *g);
int make_grid(const s1ifg_grid *grid, GridDev *out);
int make_dem(const s1ifg_dem *dem, DemDev *out);
int make_offpoly(const double *p, OffPoly *out);
int device_cu_count(int device);
// Releases ctx->ps_cache (k_ps.hip).
void ps_cache_free(s1ifg_ctx *ctx);
}  // namespace s1ifg
