// Inverses on the transform outputs: column sums of the real part of Tx or
// Wx (issq_cwt, issq_stft, the one-integral icwt), whole or over curve
// bands per column.
//
// Layout: A [batch][rows][N] complex64 (interleaved float32 pairs); one
// thread per column, rows in ascending order (the order of NumPy's axis-0
// reduction of a C-contiguous array, so the float32 sums equal the
// reference's CPU results bit for bit).  Only the real halves are used, but
// a complex row is read whole: the kernels are HBM-bound at 8 B per element.
#include "common.hpp"

using namespace ssqk;

namespace {

constexpr int kNT = 256;
constexpr int kU = 8;          // rows in flight per thread

// x[b][j] = f32(f32(scale * s) + add) (+ x[b][j] when accumulating), with
// s the float32 sum of Re A[b][i][j] / div[i - r0] over r0 <= i < r1
__global__ __launch_bounds__(kNT) void k_rowsum(const float2* __restrict__ A, int64_t rows,
                                                int64_t N, int r0, int r1,
                                                const float* __restrict__ div, double scale,
                                                double add, int accumulate,
                                                float* __restrict__ x) {
  const int64_t j = (int64_t)blockIdx.x * kNT + threadIdx.x;
  if (j >= N) return;
  const int b = blockIdx.y;
  const float2* a = A + (int64_t)b * rows * N + j;
  float s = 0.f;
  int i = r0;
  for (; i + kU <= r1; i += kU) {
    float v[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) v[u] = a[(int64_t)(i + u) * N].x;
#pragma unroll
    for (int u = 0; u < kU; ++u) s += div ? v[u] / div[i + u - r0] : v[u];
  }
  for (; i < r1; ++i) {
    const float v = a[(int64_t)i * N].x;
    s += div ? v / div[i - r0] : v;
  }
  float r = (float)((double)s * scale);
  if (add != 0.0) r = (float)((double)r + add);
  float* o = x + (int64_t)b * N + j;
  *o = accumulate ? *o + r : r;
}

// Curve bands: component n of column j covers rows lo..hi (inclusive),
// hi = clip(cc + cw, 0, rows), lo = clip(cc - cw, 0, rows), none where
// cc = -1.  x[n][j] = scale * (float64 sum over the band); x[K][j] = scale *
// (float32 sum over the rows outside every band), both ascending.
__global__ __launch_bounds__(kNT) void k_bandsum(const float2* __restrict__ A, int rows,
                                                 int64_t N, const int32_t* __restrict__ cc,
                                                 const int32_t* __restrict__ cw, int K,
                                                 double scale, double* __restrict__ x) {
  const int64_t j = (int64_t)blockIdx.x * kNT + threadIdx.x;
  if (j >= N) return;
  const float2* a = A + j;
  auto band = [&](int n, int& lo, int& hi) {
    const int c = cc[j * K + n], w = cw[j * K + n];
    hi = min(max(c + w, 0), rows);
    lo = min(max(c - w, 0), rows);
    if (c == -1) { lo = 1; hi = 0; }
    hi = min(hi, rows - 1);
  };
  for (int n = 0; n < K; ++n) {
    int lo, hi;
    band(n, lo, hi);
    double s = 0.0;
    for (int i = lo; i <= hi; ++i) s += (double)a[(int64_t)i * N].x;
    x[(int64_t)n * N + j] = s * scale;
  }
  float r = 0.f;
  for (int i = 0; i < rows; ++i) {
    bool covered = false;
    for (int n = 0; n < K && !covered; ++n) {
      int lo, hi;
      band(n, lo, hi);
      covered = i >= lo && i <= hi;
    }
    if (!covered) r += a[(int64_t)i * N].x;
  }
  x[(int64_t)K * N + j] = (double)r * scale;
}

}  // namespace

extern "C" int ssqk_real_rowsum(const void* A, int64_t batch, int64_t rows, int64_t N,
                                int64_t r0, int64_t r1, const float* div, double scale,
                                double add, int accumulate, float* x, void* stream) {
  if (!A || !x || batch < 1 || batch > 65535 || rows < 1 || N < 1 || r0 < 0 || r1 < r0 ||
      r1 > rows || rows >= ((int64_t)1 << 31)) {
    ssqk_host::set_error("bad arguments to ssqk_real_rowsum");
    return SSQK_ERR_ARG;
  }
  const int64_t nb = (N + kNT - 1) / kNT;
  if (nb >= ((int64_t)1 << 31)) {
    ssqk_host::set_error("ssqk_real_rowsum: N=%lld too long", (long long)N);
    return SSQK_ERR_ARG;
  }
  hipLaunchKernelGGL(k_rowsum, dim3((unsigned)nb, (unsigned)batch), dim3(kNT), 0,
                     (hipStream_t)stream, (const float2*)A, rows, N, (int)r0, (int)r1, div,
                     scale, add, accumulate, x);
  SSQK_CHECK(hipGetLastError());
  return SSQK_OK;
}

extern "C" int ssqk_real_bandsum(const void* A, int64_t rows, int64_t N, const int32_t* cc,
                                 const int32_t* cw, int64_t K, double scale, double* x,
                                 void* stream) {
  if (!A || !x || !cc || !cw || rows < 1 || N < 1 || K < 1 || K > 4096 ||
      rows >= ((int64_t)1 << 31)) {
    ssqk_host::set_error("bad arguments to ssqk_real_bandsum");
    return SSQK_ERR_ARG;
  }
  const int64_t nb = (N + kNT - 1) / kNT;
  if (nb >= ((int64_t)1 << 31)) {
    ssqk_host::set_error("ssqk_real_bandsum: N=%lld too long", (long long)N);
    return SSQK_ERR_ARG;
  }
  hipLaunchKernelGGL(k_bandsum, dim3((unsigned)nb), dim3(kNT), 0, (hipStream_t)stream,
                     (const float2*)A, (int)rows, N, cc, cw, (int)K, scale, x);
  SSQK_CHECK(hipGetLastError());
  return SSQK_OK;
}
