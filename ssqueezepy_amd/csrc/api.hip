// Library plumbing (error state, twiddle tables) and the elementwise and
// squeeze kernels of the synchrosqueezing path.
#include "common.hpp"
#include "fft_core.hpp"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

using namespace ssqk;

// ------------------------------------------------------------ errors
namespace {
thread_local char g_err[512] = "";
}

namespace ssqk_host {
void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
int check_hip(hipError_t e, const char* what) {
  set_error("%s: %s", what, hipGetErrorString(e));
  return SSQK_ERR_HIP;
}

// exp(+2 pi i t / M) for t in [0, M), built in float64 on the host and kept
// per (device, M) until the process ends.
const float2* twiddles(int M) {
  static std::mutex mu;
  static std::map<std::pair<int, int>, float2*> cache;
  int dev = 0;
  hipGetDevice(&dev);
  std::lock_guard<std::mutex> lock(mu);
  auto key = std::make_pair(dev, M);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  std::vector<float2> h(M);
  for (int t = 0; t < M; ++t) {
    // exact octant symmetry keeps the table's cos/sin pairs consistent
    const double a = 2.0 * M_PI * (double)t / (double)M;
    h[t] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  float2* d = nullptr;
  if (hipMalloc(&d, sizeof(float2) * M) != hipSuccess) return nullptr;
  if (hipMemcpy(d, h.data(), sizeof(float2) * M, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    return nullptr;
  }
  cache[key] = d;
  return d;
}

const float2* phase_table(int L) {
  static std::mutex mu;
  static std::map<std::pair<int, int>, float2*> cache;
  int dev = 0;
  hipGetDevice(&dev);
  std::lock_guard<std::mutex> lock(mu);
  auto key = std::make_pair(dev, L);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  const int S = ssqk::phase_tab_split(L);
  const int n_lo = 1 << S, n_hi = (L >> S) > 0 ? (L >> S) : 1;
  std::vector<float2> h(n_lo + n_hi);
  for (int j = 0; j < n_lo; ++j) {
    const double a = 2.0 * M_PI * (double)j / (double)L;
    h[j] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  for (int j = 0; j < n_hi; ++j) {
    const double a = 2.0 * M_PI * (double)((int64_t)j << S) / (double)L;
    h[n_lo + j] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  float2* d = nullptr;
  if (hipMalloc(&d, sizeof(float2) * h.size()) != hipSuccess) return nullptr;
  if (hipMemcpy(d, h.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    return nullptr;
  }
  cache[key] = d;
  return d;
}
}  // namespace ssqk_host

extern "C" const char* ssqk_last_error(void) { return g_err; }
extern "C" int ssqk_version(void) { return 1; }

namespace {

// --------------------------------------------------------------- psih
__global__ void k_psih(WavDev wav, const double* __restrict__ scales, int64_t L,
                       int nohalf, float* __restrict__ out) {
  const int i = blockIdx.y;
  const double s = scales[i];
  const double dxi = kTwoPi / (double)L;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < L;
       k += (int64_t)gridDim.x * blockDim.x) {
    // xi grid of wavelets.py:473-484: bins above L/2 are negative
    const int64_t kk = (k <= L / 2) ? k : k - L;
    double v = psih_eval(wav, s * ((double)kk * dxi));
    if (!nohalf && (L % 2 == 0) && k == L / 2) v *= 0.5;
    out[(int64_t)i * L + k] = (float)v;
  }
}

// ------------------------------------------------------- phase & bins
__global__ void k_phase(const float2* __restrict__ Wx, const float2* __restrict__ dWx,
                        const float* __restrict__ Sfs, int64_t N, int64_t total,
                        float gamma2, float* __restrict__ w) {
  for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < total;
       o += (int64_t)gridDim.x * blockDim.x) {
    const float2 z = Wx[o], zd = dWx[o];
    const float m2 = z.x * z.x + z.y * z.y;
    float v;
    if (m2 < gamma2) {
      v = INFINITY;
    } else {
      v = (zd.y * z.x - zd.x * z.y) / (m2 * kTwoPiF);
      v = Sfs ? fabsf(Sfs[o / N] - v) : fabsf(v);
    }
    w[o] = v;
  }
}

__global__ void k_bins_dwx(const float2* __restrict__ Wx, const float2* __restrict__ dWx,
                           const float* __restrict__ Sfs, int64_t N, int64_t total,
                           float gamma2, BinDev bp, uint16_t* __restrict__ bins) {
  for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < total;
       o += (int64_t)gridDim.x * blockDim.x) {
    const float2 z = Wx[o], zd = dWx[o];
    const float m2 = z.x * z.x + z.y * z.y;
    int k = kNoBin;
    if (m2 > gamma2) {
      if (Sfs) {
        // |Sfs[i] - Im(dWx/Wx)/2pi| (algos.py:971-984), linear bins
        const float sf = Sfs[o / N];
        const float wv = fabsf(sf - (zd.y * z.x - zd.x * z.y) / (m2 * kTwoPiF));
        k = bin_from_w(wv, bp);
      } else {
        k = bin_from_dwx(zd.x, zd.y, z.x, z.y, bp);
      }
    }
    bins[o] = (uint16_t)k;
  }
}

__global__ void k_bins_w(const float* __restrict__ w, int64_t total, BinDev bp,
                         uint16_t* __restrict__ bins) {
  for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < total;
       o += (int64_t)gridDim.x * blockDim.x) {
    const float v = w[o];
    bins[o] = (uint16_t)(isinf(v) || isnan(v) ? kNoBin : bin_from_w(v, bp));
  }
}

// -------------------------------------------------------------- squeeze
// One workgroup owns W consecutive columns of one signal for every scale.
// Wave v handles columns [v W/4, (v+1) W/4); the lanes of a column split the
// output rows into equal slices, so every (row, column) accumulator has
// exactly one owner lane and each column is summed in ascending scale order
// (deterministic, no atomics).  The lanes of a column load the same
// addresses, which the memory pipeline serves once.  Accumulators are
// interleaved complex [nf][W] in LDS; kSqU scales are loaded ahead so each
// wave keeps enough HBM requests in flight.
constexpr int kSqW = 64;
constexpr int kSqNT = 256;
constexpr int kSqU = 32;

__global__ __launch_bounds__(kSqNT) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_squeeze(const float2* __restrict__ Wx,
                                                   const uint16_t* __restrict__ bins,
                                                   const float* __restrict__ cst, int na,
                                                   int nf, int64_t N, int64_t in_stride,
                                                   int64_t out_stride, int W,
                                                   float2* __restrict__ Tx) {
  extern __shared__ float2 s_acc[];          // [nf][W]
  const int b = blockIdx.y;
  const int64_t j0 = (int64_t)blockIdx.x * W;
  const int tid = threadIdx.x;
  for (int e = tid; e < nf * W; e += kSqNT) s_acc[e] = make_float2(0.f, 0.f);
  __syncthreads();

  const int lanes_per_col = kSqNT / W;
  const int wave = tid >> 6, lane = tid & 63;
  const int cols_per_wave = W / 4;
  const int c = wave * cols_per_wave + (lane % cols_per_wave);
  const int q = lane / cols_per_wave;
  const int rlo = (int)(((int64_t)nf * q) / lanes_per_col);
  const int rhi = (int)(((int64_t)nf * (q + 1)) / lanes_per_col);
  const int64_t j = j0 + c;
  const bool live = j < N;
  const float2* wp = Wx + (int64_t)b * in_stride + (live ? j : 0);
  const uint16_t* bp_ = bins + (int64_t)b * in_stride + (live ? j : 0);

  // double-buffered: the loads of chunk i0 + kSqU are in flight while the
  // accumulators take chunk i0; loads are unconditional (clamped rows) so
  // the compiler issues them back to back, invalid rows are masked
  float2 wA[kSqU];
  uint32_t kA[kSqU];
#pragma unroll
  for (int u = 0; u < kSqU; ++u) {
    const int i = min(u, na - 1);
    kA[u] = bp_[(int64_t)i * N];
    wA[u] = wp[(int64_t)i * N];
  }
  for (int i0 = 0; i0 < na; i0 += kSqU) {
    float2 wB[kSqU];
    uint32_t kB[kSqU];
#pragma unroll
    for (int u = 0; u < kSqU; ++u) {
      const int i = min(i0 + kSqU + u, na - 1);
      kB[u] = bp_[(int64_t)i * N];
      wB[u] = wp[(int64_t)i * N];
    }
#pragma unroll
    for (int u = 0; u < kSqU; ++u) {
      const int k = (live && i0 + u < na) ? (int)kA[u] : kNoBin;
      if (k >= rlo && k < rhi) {
        const float cc = cst[i0 + u];
        float2 acc = s_acc[k * W + c];
        acc.x += wA[u].x * cc;
        acc.y += wA[u].y * cc;
        s_acc[k * W + c] = acc;
      }
    }
#pragma unroll
    for (int u = 0; u < kSqU; ++u) {
      kA[u] = kB[u];
      wA[u] = wB[u];
    }
  }
  __syncthreads();
  float2* tp = Tx + (int64_t)b * out_stride;
  for (int e = tid; e < nf * W; e += kSqNT) {
    const int k = e / W, cc = e - k * W;
    const int64_t jj = j0 + cc;
    if (jj < N) tp[(int64_t)k * N + jj] = s_acc[e];
  }
}

int grid1d(int64_t total, int nt) {
  int64_t g = (total + nt - 1) / nt;
  return (int)std::min<int64_t>(std::max<int64_t>(g, 1), 65536);
}

}  // namespace

// =================================================================== ABI
extern "C" int ssqk_psih(int wavelet_kind, const double* params, const double* scales,
                         int64_t na, int64_t L, int nohalf, float* out, void* stream) {
  if (!params || !scales || !out || na < 1 || na > 65535 || L < 1) {
    ssqk_host::set_error("bad arguments to ssqk_psih");
    return SSQK_ERR_ARG;
  }
  WavDev wav{};
  if (wavelet_kind == SSQK_WAVELET_GMW) {
    const double g = params[0], b = params[1];
    if (!(g > 0) || !(b > 0)) {
      ssqk_host::set_error("GMW needs gamma > 0 and beta > 0");
      return SSQK_ERR_ARG;
    }
    const double wc = std::pow(b / g, 1.0 / g);
    wav = WavDev{SSQK_WAVELET_GMW, g, b, std::log(wc), std::pow(wc, g)};
  } else if (wavelet_kind == SSQK_WAVELET_MORLET) {
    const double mu = params[0];
    const double cs = std::pow(1 + std::exp(-mu * mu) - 2 * std::exp(-0.75 * mu * mu), -0.5);
    wav = WavDev{SSQK_WAVELET_MORLET, mu, std::exp(-0.5 * mu * mu), -0.5,
                 std::sqrt(2.0) * cs * std::pow(M_PI, 0.25)};
  } else {
    ssqk_host::set_error("unknown wavelet kind %d", wavelet_kind);
    return SSQK_ERR_UNSUPPORTED;
  }
  hipStream_t st = (hipStream_t)stream;
  const int gx = (int)std::min<int64_t>((L + 255) / 256, 1024);
  hipLaunchKernelGGL(k_psih, dim3(gx, (unsigned)na), dim3(256), 0, st, wav, scales, L, nohalf,
                     out);
  SSQK_CHECK(hipGetLastError());
  return SSQK_OK;
}

extern "C" int ssqk_phase(const void* Wx, const void* dWx, const float* Sfs, int64_t rows,
                          int64_t N, double gamma, float* w, void* stream) {
  if (!Wx || !dWx || !w || rows < 1 || N < 1) {
    ssqk_host::set_error("bad arguments to ssqk_phase");
    return SSQK_ERR_ARG;
  }
  const int64_t total = rows * N;
  hipLaunchKernelGGL(k_phase, dim3(grid1d(total, 256)), dim3(256), 0, (hipStream_t)stream,
                     (const float2*)Wx, (const float2*)dWx, Sfs, N, total,
                     (float)(gamma * gamma), w);
  SSQK_CHECK(hipGetLastError());
  return SSQK_OK;
}

extern "C" int ssqk_bins_from_dwx(const void* Wx, const void* dWx, const float* Sfs,
                                  int64_t rows, int64_t N, const ssqk_bin_params* bp,
                                  double gamma, uint16_t* bins, void* stream) {
  if (!Wx || !dWx || !bp || !bins || rows < 1 || N < 1 || bp->omax < 0 ||
      bp->omax >= kNoBin) {
    ssqk_host::set_error("bad arguments to ssqk_bins_from_dwx");
    return SSQK_ERR_ARG;
  }
  const int64_t total = rows * N;
  hipLaunchKernelGGL(k_bins_dwx, dim3(grid1d(total, 256)), dim3(256), 0, (hipStream_t)stream,
                     (const float2*)Wx, (const float2*)dWx, Sfs, N, total,
                     (float)(gamma * gamma), make_bin_dev(bp), bins);
  SSQK_CHECK(hipGetLastError());
  return SSQK_OK;
}

extern "C" int ssqk_bins_from_w(const float* w, int64_t rows, int64_t N,
                                const ssqk_bin_params* bp, uint16_t* bins, void* stream) {
  if (!w || !bp || !bins || rows < 1 || N < 1 || bp->omax < 0 || bp->omax >= kNoBin) {
    ssqk_host::set_error("bad arguments to ssqk_bins_from_w");
    return SSQK_ERR_ARG;
  }
  const int64_t total = rows * N;
  hipLaunchKernelGGL(k_bins_w, dim3(grid1d(total, 256)), dim3(256), 0, (hipStream_t)stream, w,
                     total, make_bin_dev(bp), bins);
  SSQK_CHECK(hipGetLastError());
  return SSQK_OK;
}

extern "C" int ssqk_squeeze(const void* Wx, const uint16_t* bins, const float* cst,
                            int64_t batch, int64_t na, int64_t nf, int64_t N, void* Tx,
                            void* stream) {
  if (!Wx || !bins || !cst || !Tx || batch < 1 || batch > 65535 || na < 1 || nf < 1 ||
      nf >= kNoBin || N < 1) {
    ssqk_host::set_error("bad arguments to ssqk_squeeze");
    return SSQK_ERR_ARG;
  }
  // accumulators [2][nf][W] in LDS; W = 64 columns while it fits 160 KiB,
  // narrower for very many output rows
  int W = kSqW;
  while (W > 4 && (size_t)nf * W * sizeof(float2) > 160 * 1024) W /= 2;
  if ((size_t)nf * W * sizeof(float2) > 160 * 1024) {
    ssqk_host::set_error("nf=%lld output rows do not fit the LDS accumulator",
                         (long long)nf);
    return SSQK_ERR_UNSUPPORTED;
  }
  const size_t lds = (size_t)nf * W * sizeof(float2);
  static bool attr_set = false;
  if (!attr_set) {
    SSQK_CHECK(hipFuncSetAttribute((const void*)k_squeeze,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    attr_set = true;
  }
  dim3 grid((unsigned)((N + W - 1) / W), (unsigned)batch);
  hipLaunchKernelGGL(k_squeeze, grid, dim3(kSqNT), lds, (hipStream_t)stream,
                     (const float2*)Wx, bins, cst, (int)na, (int)nf, N, na * N, nf * N, W,
                     (float2*)Tx);
  SSQK_CHECK(hipGetLastError());
  return SSQK_OK;
}
