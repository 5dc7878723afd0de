/*
 * s1cpu_ps.c -- C/OpenMP restatement of the reference's PS phase-similarity
 * executable (csrc/src/phase_similarity.cu) for whole images: the CPU
 * timing beside the GPU and the checker for stacks too large for the NumPy
 * loops of oracle/psps.py.  Test infrastructure (tests/ and bench.py's CPU
 * baseline load it through ctypes).
 *
 * Reference code each function follows (file:line):
 *   s1c_ps_scan       scan_array, phase_similarity.cu:295-366
 *   s1c_ps_normalize  normalize, phase_similarity.cu:116-134
 *   s1c_ps_median     median_similarity<N>, phase_similarity.cu:174-233
 *                     (pixels that are not candidates written 0)
 *   s1c_ps_update     update_ps, phase_similarity.cu:145-157
 *   s1c_ps_max        maximum_similarity, phase_similarity.cu:235-283
 * The stack is [nifg][nrow][ncol] interleaved complex float, the candidate
 * mask int32 [nrow][ncol]; offsets are (dr, dc) int32 pairs.  Similarities
 * are float sums over the interferograms in order, then / nifg, as the
 * reference's loops.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Writes at most cap offsets; returns how many the scan has (> cap: the
 * caller's array was too small and holds the first cap). */
int s1c_ps_scan(int rdmin, int rdmax, int *off, int cap) {
  if (rdmax <= 0) return 0;
  unsigned char *seen = calloc((size_t)rdmax * rdmax, 1);
  if (!seen) return -1;
  int n = 0;
#define EMIT(a, b)                  \
  do {                              \
    if (n < cap) {                  \
      off[2 * n] = (a);             \
      off[2 * n + 1] = (b);         \
    }                               \
    ++n;                            \
  } while (0)
  seen[0] = 1;
  for (int r = 1; r < rdmax; ++r) {
    const int keep = r > rdmin;
    int x = r, y = 0, p = 1 - r, back = 0;
    if (keep) {
      EMIT(r, 0);
      EMIT(-r, 0);
      EMIT(0, r);
      EMIT(0, -r);
    }
    seen[r * rdmax] = seen[r] = 1;
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
      while (x - 1 >= 0 && y < rdmax && !seen[(x - 1) * rdmax + y]) {
        --x;
        ++back;
      }
      seen[x * rdmax + y] = seen[y * rdmax + x] = 1;
      if (keep) {
        EMIT(x, y);
        EMIT(-x, -y);
        EMIT(x, -y);
        EMIT(-x, y);
        if (x != y) {
          EMIT(y, x);
          EMIT(-y, -x);
          EMIT(y, -x);
          EMIT(-y, x);
        }
      }
      if (back > 0) ++x;
    }
  }
#undef EMIT
  free(seen);
  return n;
}

void s1c_ps_normalize(const float *in, float *out, long n) {
#pragma omp parallel for schedule(static)
  for (long k = 0; k < n; ++k) {
    const float x = in[2 * k], y = in[2 * k + 1];
    const float mag = sqrtf(x * x + y * y);
    if (mag > 1e-8f) {
      const float inv = 1.0f / mag;
      out[2 * k] = x * inv;
      out[2 * k + 1] = y * inv;
    } else {
      out[2 * k] = 0.f;
      out[2 * k + 1] = 0.f;
    }
  }
}

static float pair_sim(const float *z, int nifg, size_t plane, size_t a, size_t b) {
  float s = 0.f;
  for (int j = 0; j < nifg; ++j) {
    const float *p = z + 2 * (j * plane + a), *q = z + 2 * (j * plane + b);
    s += p[0] * q[0] + p[1] * q[1];
  }
  return s / nifg;
}

static int cmp_float(const void *a, const void *b) {
  const float x = *(const float *)a, y = *(const float *)b;
  return (x > y) - (x < y);
}

void s1c_ps_median(const float *z, int nifg, int nrow, int ncol, const int *ps, int nneigh,
                   const int *off, int noff, float *out) {
  const size_t plane = (size_t)nrow * ncol;
  const int cap = nneigh > 0 ? nneigh : 1;
#pragma omp parallel
  {
    float *sims = malloc(sizeof(float) * cap);
#pragma omp for schedule(dynamic, 4)
    for (int r0 = 0; r0 < nrow; ++r0)
      for (int c0 = 0; c0 < ncol; ++c0) {
        const size_t a = (size_t)r0 * ncol + c0;
        if (!ps[a]) {
          out[a] = 0.f;
          continue;
        }
        int cnt = 0;
        for (int i = 0; i < noff && cnt < cap; ++i) {
          const int r = r0 + off[2 * i], c = c0 + off[2 * i + 1];
          if (r < 0 || r >= nrow || c < 0 || c >= ncol) continue;
          const size_t b = (size_t)r * ncol + c;
          if (!ps[b]) continue;
          sims[cnt++] = pair_sim(z, nifg, plane, a, b);
        }
        if (cnt) {
          qsort(sims, cnt, sizeof(float), cmp_float);
          out[a] = sims[cnt / 2];
        } else {
          out[a] = 0.f;
        }
      }
    free(sims);
  }
}

void s1c_ps_update(int *ps, const float *sim, float th, int overwrite, long n) {
#pragma omp parallel for schedule(static)
  for (long k = 0; k < n; ++k) {
    if (sim[k] >= th)
      ps[k] = 1;
    else if (overwrite)
      ps[k] = 0;
  }
}

void s1c_ps_max(const float *z, int nifg, int nrow, int ncol, const int *ps, int nneigh,
                const int *off, int noff, float *out) {
  const size_t plane = (size_t)nrow * ncol;
  const int cap = nneigh > 0 ? nneigh : 1;
#pragma omp parallel for schedule(dynamic, 4)
  for (int r0 = 0; r0 < nrow; ++r0)
    for (int c0 = 0; c0 < ncol; ++c0) {
      const size_t a = (size_t)r0 * ncol + c0;
      if (ps[a]) {
        out[a] = 1.f;
        continue;
      }
      float best = -1.f;
      int cnt = 0;
      for (int i = 0; i < noff && cnt < cap; ++i) {
        const int r = r0 + off[2 * i], c = c0 + off[2 * i + 1];
        if (r < 0 || r >= nrow || c < 0 || c >= ncol) continue;
        const size_t b = (size_t)r * ncol + c;
        if (!ps[b]) continue;
        const float s = pair_sim(z, nifg, plane, a, b);
        if (best < s) best = s;
        ++cnt;
      }
      out[a] = cnt ? best : 0.f;
    }
}
