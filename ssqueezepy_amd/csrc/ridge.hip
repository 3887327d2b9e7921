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
// (np.amin: a NaN anywhere gives NaN).  tpf threads share a frequency f,
// each taking jc consecutive j (a multiple of 4: float4 reads of the
// previous column, broadcast to the wave) with four running minima; the
// minimum of a set does not depend on the order it is taken in, so the
// split is exact.  LDS: s and two columns padded to nfp (+inf past nf,
// which never wins), and the partial minima [tpf][nf].
__global__ void k_ridge_fw(const float* __restrict__ E, const float* __restrict__ s, int nf,
                           int64_t N, float penalty, int tpf, int jc,
                           float* __restrict__ pen) {
  extern __shared__ float sh[];
  const int nfp = (nf + 3) & ~3;
  float* ss = sh;
  float* prev = sh + nfp;
  float* part = sh + 3 * nfp;
  const int nt = blockDim.x;
  for (int e = threadIdx.x; e < nfp; e += nt) {
    ss[e] = e < nf ? s[e] : 0.f;
    const float e0 = e < nf ? E[(int64_t)e * N] : INFINITY;
    prev[e] = e0;
    prev[nfp + e] = INFINITY;
    if (e < nf) pen[(int64_t)e * N] = e0;
  }
  __syncthreads();
  for (int64_t t = 1; t < N; ++t) {
    const float* pv = prev + ((t - 1) & 1) * nfp;
    float* nx = prev + (t & 1) * nfp;
    for (int w = threadIdx.x; w < tpf * nf; w += nt) {
      const int c = w / nf, f = w - c * nf;
      const float sf = ss[f];
      float m0 = INFINITY, m1 = INFINITY, m2 = INFINITY, m3 = INFINITY;
      bool nan = false;
      const int j0 = c * jc, j1 = min(j0 + jc, nfp);
      for (int j = j0; j < j1; j += 4) {
        const float4 p4 = *reinterpret_cast<const float4*>(pv + j);
        const float4 s4 = *reinterpret_cast<const float4*>(ss + j);
        const float v0 = p4.x + jump(penalty, sf, s4.x);
        const float v1 = p4.y + jump(penalty, sf, s4.y);
        const float v2 = p4.z + jump(penalty, sf, s4.z);
        const float v3 = p4.w + jump(penalty, sf, s4.w);
        nan = nan || isnan(v0) || isnan(v1) || isnan(v2) || isnan(v3);
        m0 = fminf(m0, v0);
        m1 = fminf(m1, v1);
        m2 = fminf(m2, v2);
        m3 = fminf(m3, v3);
      }
      part[c * nf + f] = nan ? NAN : fminf(fminf(m0, m1), fminf(m2, m3));
    }
    __syncthreads();
    for (int f = threadIdx.x; f < nf; f += nt) {
      float m = part[f];
      bool nan = isnan(m);
      for (int c = 1; c < tpf; ++c) {
        const float q = part[c * nf + f];
        nan = nan || isnan(q);
        m = fminf(m, q);
      }
      const float p = E[(int64_t)f * N + t] + (nan ? NAN : m);
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
      nf > 10000 || nf * N >= ((int64_t)1 << 40) || N >= ((int64_t)1 << 31)) {
    // (nf <= 10000: s, two columns and the partial minima of the
    // recursion, 16 (nf + 3) bytes at most, in 160 KB of LDS)
    ssqk_host::set_error("bad arguments to ssqk_ridge_track");
    return SSQK_ERR_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  float* energy = (float*)workspace;
  float* E = energy + nf * N;
  float* pen = E + nf * N;
  const int nt = (int)std::min<int64_t>(1024, (nf + 63) / 64 * 64);
  // forward recursion: up to 8 threads per frequency, j in chunks of a
  // multiple of 4
  const int nfp = (int)((nf + 3) & ~3);
  const int tpf = (int)std::max<int64_t>(1, std::min<int64_t>(8, 1024 / nf));
  const int jc = ((nfp + tpf - 1) / tpf + 3) & ~3;
  const int nt_fw = (int)std::min<int64_t>(1024, (tpf * nf + 63) / 64 * 64);
  const int lds = (int)((3 * nfp + tpf * nf) * sizeof(float));
  if (lds > 160 * 1024) {
    ssqk_host::set_error("ssqk_ridge_track: nf=%lld too large", (long long)nf);
    return SSQK_ERR_ARG;
  }
  int rc = ssqk_host::allow_dynamic_lds((const void*)k_ridge_fw, lds);
  if (rc) return rc;
  hipLaunchKernelGGL(k_abs2, dim3(blocks(nf * N)), dim3(kNT), 0, st, (const float2*)Tf,
                     nf * N, energy);
  SSQK_CHECK(hipGetLastError());
  for (int i = 0; i < n_ridges; ++i) {
    int32_t* ri = ridge + (int64_t)i * N;
    hipLaunchKernelGGL(k_ridge_E, dim3(blocks(N)), dim3(kNT), 0, st, energy, (int)nf, N, eps, E);
    SSQK_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_ridge_fw, dim3(1), dim3(nt_fw), lds, st, E, s, (int)nf, N, penalty,
                       tpf, jc, pen);
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
