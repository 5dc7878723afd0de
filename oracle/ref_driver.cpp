// C entry points over the reference's own host geometry code
// (csrc/src/orbit.cu and csrc/src/bounds.cu, compiled as plain C++ from the
// reference tree by oracle/Makefile).  Used only to write golden vectors.
#include <cstddef>
#include <string>

#include "bounds.hpp"
#include "orbit.hpp"

extern "C" {

void ref_llh2xyz(const double *llh, double *xyz, long n) {
  for (long i = 0; i < n; ++i) {
    double l[3] = {llh[3 * i], llh[3 * i + 1], llh[3 * i + 2]};
    llh2xyz(l, xyz + 3 * i);
  }
}

void ref_xyz2llh(const double *xyz, double *llh, long n) {
  for (long i = 0; i < n; ++i) {
    double x[3] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    xyz2llh(x, llh + 3 * i);
  }
}

void ref_intp_orbit(long nsv, double *t, double *x, double *v, const double *tq, long n,
                    double *pos, double *vel) {
  for (long i = 0; i < n; ++i) intp_orbit((std::size_t)nsv, t, x, v, tq[i], pos + 3 * i, vel + 3 * i);
}

void ref_orbitrangetime(long nsv, double *t, double *x, double *v, const double *xyz, long n,
                        double tline0, double *satx0, double *satv0, double *tline, double *dr) {
  for (long i = 0; i < n; ++i) {
    double X[3] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    double tl;
    orbitrangetime((std::size_t)nsv, t, x, v, X, tline0, satx0, satv0, tl, dr + 3 * i);
    tline[i] = tl;
  }
}

void ref_bounds(double t0, double t1, double r0, double r1, double hmin, double hmax, long nsv,
                double *t, double *x, double *v, double *latlons) {
  bounds(t0, t1, r0, r1, hmin, hmax, t, x, v, (std::size_t)nsv, latlons, "RIGHT");
}
}
