// Forward-backward ridge tracking on a time-frequency map
// (ridge_extraction.py:11-232).
//
// Layout: energy, E (the tracked -log normalized energy) and pen (the
// forward accumulated penalty) are float32 [nf][N] in the workspace; the
// time recursion runs in one workgroup per map (each step needs the whole
// previous column), every other pass one thread per column.  Arithmetic
// follows the reference's float32 NumPy evaluation operation by operation
// (no fused multiply-adds; logs and hypot in float64, rounded once), so
// the ridge indices equal its serial path's.
#include "common.hpp"

#include <algorithm>

// every float operation rounded on its own, as NumPy evaluates it: no
// multiply-add contraction for the operators of this file (functions from
// the HIP headers, __fmul_rn and the like, carry the default and would
// still fuse)
#pragma clang fp contract(off)

using namespace ssqk;

namespace {

constexpr int kNT = 256;

// penalty * (a - b)^2, rounded as penalty * np.subtract.outer(s, s)**2
__device__ __forceinline__ float jump(float penalty, float a, float b) {
  const float d = a - b;
  return penalty * (d * d);
}

// energy = |Tf|^2 (np.abs(Tf)**2: hypot rounded to float32, then squared)
__global__ __launch_bounds__(kNT) void k_abs2(const float2* __restrict__ Tf, int64_t n,
                                              float* __restrict__ energy) {
  const int64_t i = (int64_t)blockIdx.x * kNT + threadIdx.x;
  if (i >= n) return;
  const float2 z = Tf[i];
  const float h = (float)sqrt((double)z.x * z.x + (double)z.y * z.y);
  energy[i] = h * h;
}

// E[f][t] = -log(energy[f][t] / max_f energy[:, t] + eps), np.max's NaN
// propagation included
__global__ __launch_bounds__(kNT) void k_ridge_E(const float* __restrict__ energy, int nf,
                                                 int64_t N, float eps, float* __restrict__ E) {
  const int64_t t = (int64_t)blockIdx.x * kNT + threadIdx.x;
  if (t >= N) return;
  float mx = energy[t];
  for (int f = 1; f < nf; ++f) {
    const float v = energy[(int64_t)f * N + t];
    mx = (v > mx || isnan(v)) && !isnan(mx) ? v : mx;
  }
  for (int f = 0; f < nf; ++f) {
    const float q = energy[(int64_t)f * N + t] / mx + eps;
    E[(int64_t)f * N + t] = -(float)log((double)q);
  }
}

// pen[:, 0] = E[:, 0]; pen[f][t] = E[f][t] + min_j (pen[j][t-1] + jump(s_f, s_j))
// (np.amin: a NaN anywhere gives NaN)
__global__ void k_ridge_fw(const float* __restrict__ E, const float* __restrict__ s, int nf,
                           int64_t N, float penalty, float* __restrict__ pen) {
  extern __shared__ float sh[];              // s [nf], prev [2][nf]
  float* ss = sh;
  float* prev = sh + nf;
  const int nt = blockDim.x;
  for (int f = threadIdx.x; f < nf; f += nt) {
    ss[f] = s[f];
    const float e0 = E[(int64_t)f * N];
    prev[f] = e0;
    pen[(int64_t)f * N] = e0;
  }
  __syncthreads();
  for (int64_t t = 1; t < N; ++t) {
    const float* pv = prev + ((t - 1) & 1) * nf;
    float* nx = prev + (t & 1) * nf;
    for (int f = threadIdx.x; f < nf; f += nt) {
      const float sf = ss[f];
      float m = pv[0] + jump(penalty, sf, ss[0]);
      for (int j = 1; j < nf; ++j) {
        const float v = pv[j] + jump(penalty, sf, ss[j]);
        m = (v < m || isnan(v)) && !isnan(m) ? v : m;
      }
      const float p = E[(int64_t)f * N + t] + m;
      nx[f] = p;
      pen[(int64_t)f * N + t] = p;
    }
    __syncthreads();
  }
}

// ridge[t] = argmin_f pen[f][t] (first minimum; first NaN if any), reduced
// modulo N as np.unravel_index(..., pen.shape)[1] does
__global__ __launch_bounds__(kNT) void k_ridge_argmin(const float* __restrict__ pen, int nf,
                                                      int64_t N, int32_t* __restrict__ ridge) {
  const int64_t t = (int64_t)blockIdx.x * kNT + threadIdx.x;
  if (t >= N) return;
  float m = pen[t];
  int k = 0;
  for (int f = 1; f < nf && !isnan(m); ++f) {
    const float v = pen[(int64_t)f * N + t];
    if (v < m || isnan(v)) { m = v; k = f; }
  }
  ridge[t] = (int32_t)(k % N);
}

// backward: for t = N-2 .. 0, with r = ridge[t+1] and val = pen[r][t+1] -
// E[r][t+1], ridge[t] = the last f with |val - (pen[f][t] + jump(s_r, s_f))|
// < eps, else the forward ridge's value
__global__ void k_ridge_bw(const float* __restrict__ E, const float* __restrict__ pen,
                           const float* __restrict__ s, int nf, int64_t N, float penalty,
                           float eps, int32_t* __restrict__ ridge) {
  __shared__ int best[2];
  __shared__ int rnext;
  if (threadIdx.x == 0) {
    best[0] = best[1] = -1;
    rnext = ridge[N - 1];
  }
  __syncthreads();
  for (int64_t t = N - 2; t >= 0; --t) {
    const int k = (int)(t & 1);
    const int r = rnext;
    const float val = pen[(int64_t)r * N + t + 1] - E[(int64_t)r * N + t + 1];
    const float sr = s[r];
    for (int f = threadIdx.x; f < nf; f += blockDim.x) {
      const float c = pen[(int64_t)f * N + t] + jump(penalty, sr, s[f]);
      if (fabsf(val - c) < eps) atomicMax(&best[k], f);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const int b = best[k];
      const int rt = b >= 0 ? b : ridge[t];
      ridge[t] = rt;
      rnext = rt;
      best[k ^ 1] = -1;
    }
    __syncthreads();
  }
}

// energy along the ridge (before its band is cleared), then the band
// energy[r - bw : r + bw, t] = 0 with Python's slice semantics (negative
// bounds count from the end)
__global__ __launch_bounds__(kNT) void k_ridge_clear(float* __restrict__ energy, int nf,
                                                     int64_t N, const int32_t* __restrict__ ridge,
                                                     int bw, float* __restrict__ ridge_e) {
  const int64_t t = (int64_t)blockIdx.x * kNT + threadIdx.x;
  if (t >= N) return;
  const int r = ridge[t];
  if (ridge_e) ridge_e[t] = energy[(int64_t)r * N + t];
  auto bound = [&](int64_t v) {
    if (v < 0) v += nf;
    return (int)std::min<int64_t>(std::max<int64_t>(v, 0), nf);
  };
  const int a = bound((int64_t)r - bw), b = bound((int64_t)r + bw);
  for (int f = a; f < b; ++f) energy[(int64_t)f * N + t] = 0.f;
}

inline unsigned blocks(int64_t n) { return (unsigned)((n + kNT - 1) / kNT); }

}  // namespace

extern "C" int ssqk_ridge_workspace_size(int64_t nf, int64_t N, size_t* bytes) {
  if (!bytes || nf < 1 || N < 1) {
    ssqk_host::set_error("bad arguments to ssqk_ridge_workspace_size");
    return SSQK_ERR_ARG;
  }
  *bytes = (size_t)3 * nf * N * sizeof(float);
  return SSQK_OK;
}

extern "C" int ssqk_ridge_track(const void* Tf, int64_t nf, int64_t N, const float* s,
                                float penalty, float eps, int n_ridges, int bw,
                                int32_t* ridge, float* ridge_e, void* workspace,
                                void* stream) {
  if (!Tf || !s || !ridge || !workspace || nf < 1 || N < 1 || n_ridges < 1 ||
      nf > 13000 || nf * N >= ((int64_t)1 << 40) || N >= ((int64_t)1 << 31)) {
    // (nf <= 13000: s and two columns of the recursion, 12 nf bytes, in LDS)
    ssqk_host::set_error("bad arguments to ssqk_ridge_track");
    return SSQK_ERR_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  float* energy = (float*)workspace;
  float* E = energy + nf * N;
  float* pen = E + nf * N;
  const int nt = (int)std::min<int64_t>(1024, (nf + 63) / 64 * 64);
  const int lds = (int)(3 * nf * sizeof(float));
  int rc = ssqk_host::allow_dynamic_lds((const void*)k_ridge_fw, lds);
  if (rc) return rc;
  hipLaunchKernelGGL(k_abs2, dim3(blocks(nf * N)), dim3(kNT), 0, st, (const float2*)Tf,
                     nf * N, energy);
  SSQK_CHECK(hipGetLastError());
  for (int i = 0; i < n_ridges; ++i) {
    int32_t* ri = ridge + (int64_t)i * N;
    hipLaunchKernelGGL(k_ridge_E, dim3(blocks(N)), dim3(kNT), 0, st, energy, (int)nf, N, eps, E);
    SSQK_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_ridge_fw, dim3(1), dim3(nt), lds, st, E, s, (int)nf, N, penalty, pen);
    SSQK_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_ridge_argmin, dim3(blocks(N)), dim3(kNT), 0, st, pen, (int)nf, N, ri);
    SSQK_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_ridge_bw, dim3(1), dim3(nt), 0, st, E, pen, s, (int)nf, N, penalty,
                       eps, ri);
    SSQK_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_ridge_clear, dim3(blocks(N)), dim3(kNT), 0, st, energy, (int)nf, N,
                       ri, bw, ridge_e ? ridge_e + (int64_t)i * N : nullptr);
    SSQK_CHECK(hipGetLastError());
  }
  return SSQK_OK;
}
