"""Gradients of ssq_cwt and cwt through the hand-written adjoint
(ssqk_cwt_adjoint) against a float64 torch.fft restatement of the forward
(with the squeeze rows of the GPU forward held fixed) and against float64
central finite differences of that restatement."""
import numpy as np
import pytest

from oracle import ssq_oracle as O
from test_gpu_cwt import _psih, _wav, chirp_tone_noise

pytestmark = pytest.mark.gpu


def _ref_forward(x64, kind, scales64, padtype='reflect'):
    """float64 torch (CPU) Wx of x64 [B, N] (differentiable): the oracle's
    cwt (pad, fft, psih(scale*xi) with the Nyquist bin halved, ifft, unpad)
    written with torch ops."""
    import torch
    N = x64.shape[-1]
    L, n1, n2 = O.pad_lengths(N)
    mode = {'reflect': 'reflect', 'zero': 'constant'}[padtype]
    if mode == 'constant':
        xp = torch.nn.functional.pad(x64, (n1, n2))
    else:
        idx = torch.from_numpy(np.pad(np.arange(N), (n1, n2), mode='reflect'))
        xp = x64[..., idx]
    xh = torch.fft.fft(xp)
    xi = O.xi_grid(L)
    P = _psih(kind)(np.asarray(scales64, dtype=np.float64).reshape(-1, 1) * xi)
    P[:, L // 2] /= 2
    P = torch.from_numpy(P)
    return torch.fft.ifft(P * xh[:, None, :])[..., n1:n1 + N]


def _ref_squeeze(Wx, bins, const, nf):
    """Tx[b, k, j] += const[i] Wx[b, i, j] for bins k >= 0 (differentiable)."""
    import torch
    B, na, N = Wx.shape
    out = []
    c = torch.as_tensor(np.asarray(const, dtype=np.float64)).reshape(-1, 1)
    cols = torch.arange(N).expand(na, N)
    for b in range(B):
        k = torch.as_tensor(bins[b])
        m = k >= 0
        T = torch.zeros(nf, N, dtype=torch.complex128)
        T = T.index_put((k[m], cols[m]), (Wx[b] * c)[m], accumulate=True)
        out.append(T)
    return torch.stack(out)


def _loss(Tx, target):
    return ((Tx.abs() - target) ** 2).mean()


@pytest.mark.parametrize('N,kind,B', [(700, 'morlet', 1), (5000, 'gmw', 1),
                                      (3000, 'morlet', 2), (24000, 'gmw', 1)])
def test_ssq_cwt_gradient(N, kind, B):
    import torch
    import ssqueezepy_amd as sq
    from ssqueezepy_amd._ssq_cwt import ssq_cwt_grid
    x = np.stack([chirp_tone_noise(N, 3 + b) for b in range(B)])
    xt = np.stack([chirp_tone_noise(N, 30 + b) for b in range(B)])
    wav = _wav(kind)
    print(f"ssq_cwt gradient N={N} {kind} B={B}", flush=True)
    target = sq.ssq_cwt(torch.as_tensor(xt, device='cuda'), wav)[0].abs().detach()
    xg = torch.tensor(x, device='cuda', requires_grad=True)
    Tx, Wx, *_ = sq.ssq_cwt(xg, wav)
    loss = _loss(Tx, target)
    loss.backward()
    g32 = xg.grad.cpu().numpy().astype(np.float64)
    # the rows of the GPU forward, held fixed in the float64 restatement
    from ssqueezepy_amd import _lib
    g = ssq_cwt_grid(N, wav)
    bins_t = torch.empty(Wx.shape, dtype=torch.int16, device='cuda')
    from ssqueezepy_amd._cwt import run_cwt
    Wtmp = torch.empty_like(Wx)
    run_cwt(torch.as_tensor(x, device='cuda'), g['wavelet'], g['scales64'], g['dt'],
            'reflect', Wtmp, bins=bins_t, bp=g['bp'], gamma=10 * np.finfo(np.float32).eps)
    bins = bins_t.cpu().numpy().astype(np.int64) & 0xFFFF
    bins[bins == 0xFFFF] = -1
    tgt64 = target.cpu().to(torch.float64)
    x64 = torch.tensor(x.astype(np.float64), requires_grad=True)

    def ref_loss(xv):
        W = _ref_forward(xv, kind, g['scales64'])
        return _loss(_ref_squeeze(W, bins, g['const'], g['na']), tgt64)

    l64 = ref_loss(x64)
    l64.backward()
    g64 = x64.grad.numpy()
    err = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    print(f"  loss fp32 {loss.item():.6e} fp64 {l64.item():.6e}; grad rel-L2 {err:.2e}",
          flush=True)
    assert err < 1e-4
    # float64 central differences of the restatement at 16 seeded coordinates
    rng = np.random.default_rng(5)
    h = 1e-6
    worst = 0.
    for _ in range(16):
        b, j = int(rng.integers(B)), int(rng.integers(N))
        e = torch.zeros_like(x64)
        e[b, j] = h
        with torch.no_grad():
            fd = (ref_loss(x64 + e) - ref_loss(x64 - e)).item() / (2 * h)
        worst = max(worst, abs(fd - g64[b, j]) / np.abs(g64).max())
    print(f"  central differences vs float64 adjoint: max |d| / max|g| {worst:.2e}",
          flush=True)
    assert worst < 1e-6


@pytest.mark.parametrize('N,kind,padtype', [(1000, 'gmw', 'reflect'), (9000, 'morlet', 'zero'),
                                            (40000, 'gmw', 'reflect')])
def test_cwt_gradient(N, kind, padtype):
    import torch
    import ssqueezepy_amd as sq
    from ssqueezepy_amd._scales import process_scales
    x = chirp_tone_noise(N, 11)[None]
    wav = _wav(kind)
    print(f"cwt gradient N={N} {kind} {padtype}", flush=True)
    xg = torch.tensor(x, device='cuda', requires_grad=True)
    Wx, scales = sq.cwt(xg, wav, padtype=padtype)
    rng = np.random.default_rng(2)
    wts = rng.random(Wx.shape[-2:])
    loss = (Wx.abs() ** 2 * torch.as_tensor(wts, device='cuda', dtype=torch.float32)).sum()
    loss.backward()
    g32 = xg.grad.cpu().numpy().astype(np.float64)
    s64 = np.asarray(process_scales('log-piecewise', N, wav, nv=32), dtype=np.float64).ravel()
    x64 = torch.tensor(x.astype(np.float64), requires_grad=True)
    W = _ref_forward(x64, kind, s64, padtype)
    l64 = (W.abs() ** 2 * torch.as_tensor(wts)).sum()
    l64.backward()
    g64 = x64.grad.numpy()
    err = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    print(f"  grad rel-L2 {err:.2e}", flush=True)
    assert err < 1e-4


def test_gradient_deterministic():
    import torch
    import ssqueezepy_amd as sq
    x = chirp_tone_noise(6000, 4)
    grads = []
    for _ in range(2):
        xg = torch.tensor(x, device='cuda', requires_grad=True)
        Tx = sq.ssq_cwt(xg, _wav('gmw'))[0]
        (Tx.abs() ** 2).mean().backward()
        grads.append(xg.grad.clone())
    assert torch.equal(grads[0], grads[1])
