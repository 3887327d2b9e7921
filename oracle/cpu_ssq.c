/* CPU restatement of the reference's parallel squeeze loops -- TEST AND
 * BASELINE INFRASTRUCTURE ONLY (used by tests/ and bench.py's cpu_baseline
 * leg; the product never links it).
 *
 * Follows _ssq_cwt_log_piecewise_par / _ssq_cwt_log_par / _ssq_cwt_lin_par /
 * _ssq_stft_par (ssqueezepy/algos.py:879-984): OpenMP over columns (numba
 * prange), serial ascending scale loop, float32 data with the phase
 * transform's denominator and the bin arithmetic in float64 as numba
 * promotes them, Python round (half to even) via nearbyint.
 * Complex arrays are interleaved float32 pairs, row-major [rows][N]. */
#include <math.h>
#include <stdint.h>

#define TWO_PI 6.283185307179586

static inline double phase_w(const float* W, const float* dW, int64_t o) {
  const float A = dW[2 * o], B = dW[2 * o + 1];
  const float C = W[2 * o], D = W[2 * o + 1];
  return fabs((double)(B * C - A * D) / ((double)(C * C + D * D) * TWO_PI));
}

/* kind: 0 linear (v0=vmin, d0=dv), 1 log (v0=vlmin, d0=dvl),
 *       2 log-piecewise (v0, v1, d0, d1, idx1).  Sfs non-NULL: stft form. */
void ssq_squeeze_par(const float* W, const float* dW, const float* Sfs,
                     float* out, const float* cst, int64_t na, int64_t N,
                     double gamma, int kind, double v0, double d0, double v1,
                     double d1, int64_t idx1, int64_t omax, int flipud) {
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < N; ++j) {
    for (int64_t i = 0; i < na; ++i) {
      const int64_t o = i * N + j;
      const float C = W[2 * o], D = W[2 * o + 1];
      if (!(sqrtf(C * C + D * D) > gamma)) continue;
      double w;
      if (Sfs) {
        const float A = dW[2 * o], B = dW[2 * o + 1];
        w = fabs((double)Sfs[i] - (double)(B * C - A * D) /
                 ((double)(C * C + D * D) * TWO_PI));
      } else {
        w = phase_w(W, dW, o);
      }
      double k;
      if (kind == 0) {
        k = fmin(nearbyint(fmax((w - v0) / d0, 0.0)), (double)omax);
      } else if (kind == 1) {
        k = fmin(nearbyint(fmax((log2(w) - v0) / d0, 0.0)), (double)omax);
      } else {
        const double wl = log2(w);
        if (wl > v1) k = fmin(nearbyint((wl - v1) / d1) + (double)idx1, (double)omax);
        else k = fmax(nearbyint((wl - v0) / d0), 0.0);
      }
      int64_t ki = (int64_t)k;
      if (flipud) ki = omax - ki;
      out[2 * (ki * N + j)] += C * cst[i];
      out[2 * (ki * N + j) + 1] += D * cst[i];
    }
  }
}
