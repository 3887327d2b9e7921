// Short-time Fourier transform with its time derivative, and the fused
// STFT phase transform + squeeze bins.
//
// Frame f covers padded samples xp[f*hop, f*hop + n).  The window and the
// frequency-domain derivative of the window are applied together as one
// complex sequence z[p] = xp[f*hop+p] * (win[p] + i fs dwin[p]), so a single
// complex FFT gives both real-input transforms:
//   Sx[k]  = (Z[k] + conj Z[n-k]) / 2,  dSx[k] = (Z[k] - conj Z[n-k]) / 2i
// times the modulation phase exp(2 pi i k s21 / n) of the frame ifftshift
// (utils/stft_utils.py:69-98, _stft.py:127-147).
#include "common.hpp"
#include "fft_core.hpp"

using namespace ssqk;

namespace {

constexpr int kP = 8192;
constexpr int kNT = 512;

struct StftArgs {
  const float* x;
  int64_t N, n1;
  int pad, modulated, s21;
  int64_t hop, n_frames;
  const float* win;
  const float* dwin;
  float fs;
  float2* Sx;
  float2* dSx;
  const float* Sfs;
  uint16_t* bins;
  float gamma2;
  BinDev bp;
};

__device__ __forceinline__ int64_t pad_src_s(int64_t n, int64_t N, int64_t n1, int type) {
  int64_t t = n - n1;
  if (t >= 0 && t < N) return t;
  if (type == SSQK_PAD_REFLECT) {
    if (N == 1) return 0;
    const int64_t P = 2 * (N - 1);
    int64_t m = t % P;
    if (m < 0) m += P;
    return m < N ? m : P - m;
  }
  return -1;
}

template <int NF>
__global__ __launch_bounds__(kNT) void k_stft(StftArgs a, const float2* __restrict__ tw) {
  constexpr int F = kP / NF;          // frames per workgroup
  constexpr int NR = NF / 2 + 1;      // output rows
  __shared__ float2 s[Lds<NF>::elems(F)];
  const int b = blockIdx.y;
  const int f0 = blockIdx.x * F;
  const int nfr = (int)a.n_frames, hop = (int)a.hop, N = (int)a.N, n1 = (int)a.n1;
  const float* x = a.x + (int64_t)b * a.N;
  // frames wholly inside the signal skip the padding map (positions and
  // lengths are below 2^31: checked on the host)
  const bool interior = f0 + F <= nfr && f0 * hop - n1 >= 0 &&
                        (f0 + F - 1) * hop + NF - 1 - n1 < N;
  for (int e = threadIdx.x; e < F * NF; e += kNT) {
    const int g = e / NF, p = e - g * NF;
    const int f = f0 + g;
    float2 z = make_float2(0.f, 0.f);
    if (interior) {
      const float v = x[f * hop + p - n1];
      z = make_float2(v * a.win[p], a.dwin ? v * a.fs * a.dwin[p] : 0.f);
    } else if (f < nfr) {
      const int64_t src = pad_src_s((int64_t)f * hop + p, a.N, a.n1, a.pad);
      const float v = src >= 0 ? x[src] : 0.f;
      z = make_float2(v * a.win[p], a.dwin ? v * a.fs * a.dwin[p] : 0.f);
    }
    s[Lds<NF>::idx(g, p)] = z;
  }
  float2* const bufs[1] = {s};
  lds_fft<NF, F, -1, kNT, 1>(bufs, tw);

  // Sx / dSx rows of this signal; row k of frame f at k * n_frames + f
  float2* const Sx = a.Sx + (int64_t)b * NR * a.n_frames;
  float2* const dSx = a.dSx ? a.dSx + (int64_t)b * NR * a.n_frames : nullptr;
  uint16_t* const bins = a.bins ? a.bins + (int64_t)b * NR * a.n_frames : nullptr;
  auto spectra = [&](int g, int k, float2& S, float2& dS) {
    const float2 Zk = s[Lds<NF>::idx(g, k)];
    const float2 Zm = s[Lds<NF>::idx(g, (NF - k) & (NF - 1))];
    S = make_float2(0.5f * (Zk.x + Zm.x), 0.5f * (Zk.y - Zm.y));
    dS = make_float2(0.5f * (Zk.y + Zm.y), -0.5f * (Zk.x - Zm.x));
    if (a.modulated && (k & 1)) {
      // frame ifftshift by s21 = NF / 2 (NF a power of two): the phase
      // exp(2 pi i k s21 / NF) = (-1)^k
      S = make_float2(-S.x, -S.y);
      dS = make_float2(-dS.x, -dS.y);
    }
  };
  static_assert((F * NR + kNT - 1) / kNT <= 32, "edge mask");
  uint32_t edge = 0;
  for (int it = 0, e = threadIdx.x; e < F * NR; ++it, e += kNT) {
    const int g = e % F, k = e / F;
    const int f = f0 + g;
    if (f >= nfr) continue;
    float2 S, dS;
    spectra(g, k, S, dS);
    const int o = k * nfr + f;
    Sx[o] = S;
    if (dSx) dSx[o] = dS;
    if (bins) {
      const float m2 = S.x * S.x + S.y * S.y;
      int kb = kNoBin;
      if (m2 > a.gamma2) kb = bin_fast_w(phase_fast_sfs(a.Sfs[k], dS.x, dS.y, S.x, S.y), a.bp);
      if (kb < 0) edge |= 1u << it;
      else bins[o] = (uint16_t)kb;
    }
  }
  while (edge) {                             // points at a bin rounding edge:
    const int it = __builtin_ctz(edge);      // float64 from the same inputs
    edge &= edge - 1;
    const int e = threadIdx.x + it * kNT;
    const int g = e % F, k = e / F;
    float2 S, dS;
    spectra(g, k, S, dS);
    bins[k * nfr + f0 + g] = (uint16_t)bin_edge_sfs(a.Sfs[k], dS.x, dS.y, S.x, S.y, a.bp);
  }
}

}  // namespace

extern "C" int ssqk_stft_exec(const float* x, int64_t batch, int64_t N, int64_t n_fft,
                              int64_t hop, int pad_type, int modulated, const float* window,
                              const float* diff_window, double fs, void* Sx, void* dSx,
                              const float* Sfs, uint16_t* bins, const ssqk_bin_params* bp,
                              double gamma, void* stream) {
  if (!x || !window || !Sx || batch < 1 || batch > 65535 || N < 1 || hop < 1 ||
      n_fft < 2 || n_fft > kP || (n_fft & (n_fft - 1)) != 0) {
    ssqk_host::set_error("bad arguments to ssqk_stft_exec (n_fft=%lld must be a power of "
                         "two in [2, %d])", (long long)n_fft, kP);
    return n_fft > 1 && (n_fft & (n_fft - 1)) != 0 ? SSQK_ERR_UNSUPPORTED : SSQK_ERR_ARG;
  }
  if (pad_type != SSQK_PAD_REFLECT && pad_type != SSQK_PAD_ZERO) {
    ssqk_host::set_error("stft pad type must be reflect or zero");
    return SSQK_ERR_ARG;
  }
  if (bins && (!bp || !Sfs || !diff_window || bp->omax < 0 || bp->omax >= kNoBin)) {
    ssqk_host::set_error("stft bins need diff_window, Sfs and bin parameters");
    return SSQK_ERR_ARG;
  }
  // 32-bit positions and output offsets in the kernel
  const int64_t n_frames = (N + n_fft - 1 - n_fft) / hop + 1;
  if (N + n_fft >= ((int64_t)1 << 31) || (n_fft / 2 + 1) * n_frames >= ((int64_t)1 << 31)) {
    ssqk_host::set_error("stft: N=%lld too long for one signal", (long long)N);
    return SSQK_ERR_ARG;
  }
  StftArgs a;
  a.x = x;
  a.N = N;
  // padsignal(padlength=N+n_fft-1): odd total pad puts the extra sample left
  const int64_t tot = n_fft - 1;
  a.n1 = (tot % 2 == 0) ? tot / 2 : tot / 2 + 1;
  a.pad = pad_type;
  a.modulated = modulated;
  a.s21 = (int)(n_fft / 2);            // n_fft is even (a power of two)
  a.hop = hop;
  a.n_frames = (N + n_fft - 1 - n_fft) / hop + 1;
  a.win = window;
  a.dwin = diff_window;
  a.fs = (float)fs;
  a.Sx = (float2*)Sx;
  a.dSx = (float2*)dSx;
  a.Sfs = Sfs;
  a.bins = bins;
  a.gamma2 = (float)(gamma * gamma);
  ssqk_bin_params zero{};
  a.bp = make_bin_dev(bp ? bp : &zero);
  hipStream_t st = (hipStream_t)stream;
  SSQK_GET_TW(tw, (int)n_fft);
  switch (n_fft) {
#define C(m)                                                                          \
  case m:                                                                             \
    hipLaunchKernelGGL(k_stft<m>,                                                     \
                       dim3((unsigned)((a.n_frames + kP / m - 1) / (kP / m)), (unsigned)batch), \
                       dim3(kNT), 0, st, a, tw);                  \
    break;
    C(2) C(4) C(8) C(16) C(32) C(64) C(128) C(256) C(512) C(1024) C(2048) C(4096) C(8192)
#undef C
    default:
      return SSQK_ERR_UNSUPPORTED;
  }
  SSQK_CHECK(hipGetLastError());
  return SSQK_OK;
}
