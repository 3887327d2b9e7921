"""cwt(rpadded=True) on the GPU (_cwt.py:293-305): Wx and dWx over the
whole padded length, the signal padded on the device (ssqk_pad_signal) and
the unpadded plan run over it; against the oracle's padded transform."""
import numpy as np
import pytest

from oracle import ssq_oracle as O
from test_gpu_cwt import chirp_tone_noise, rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('N', [256, 3000, 40000])
@pytest.mark.parametrize('padtype', ['reflect', 'zero', 'symmetric', 'wrap'])
def test_cwt_rpadded_against_oracle(N, padtype):
    """All L columns on every path of the plan (band, overlap-save with its
    circular wrap, four-step); the unpadded columns equal cwt's own."""
    import ssqueezepy_amd as sq
    from ssqueezepy_amd._common import p2up
    x = chirp_tone_noise(N, 11)
    print(f"cwt rpadded N={N} padtype={padtype}", flush=True)
    Wx, scales, dWx = sq.cwt(x, 'gmw', derivative=True, rpadded=True, padtype=padtype)
    L, n1, _ = p2up(N)
    assert Wx.shape == (len(scales), L) and dWx.shape == Wx.shape
    s64 = scales.cpu().numpy().astype(np.float64)
    Wo, dWo = O.cwt(x.astype(np.float64), O.psih_fn('gmw'), s64, padtype=padtype,
                    rpadded=True)
    assert rel_l2(Wx.cpu().numpy(), Wo) < 1e-5
    assert rel_l2(dWx.cpu().numpy(), dWo) < 1e-5
    W0, _ = sq.cwt(x, 'gmw', padtype=padtype)
    assert rel_l2(Wx[:, n1:n1 + N].cpu().numpy(), W0.cpu().numpy()) < 1e-5


def test_cwt_rpadded_batched_and_pad_signal():
    """A [B, N] batch gives [B, na, L]; ssqk_pad_signal equals numpy's pad
    bit for bit for every mode."""
    import torch
    import ssqueezepy_amd as sq
    from ssqueezepy_amd import _lib
    from ssqueezepy_amd._common import p2up, padsignal_np
    rng = np.random.default_rng(3)
    x = rng.standard_normal((3, 1000)).astype(np.float32)
    L, n1, _ = p2up(1000)
    for pad in ('reflect', 'zero', 'symmetric', 'replicate', 'wrap'):
        print(f"pad_signal {pad}", flush=True)
        xd = torch.as_tensor(x, device='cuda')
        out = torch.empty((3, L), dtype=torch.float32, device='cuda')
        _lib.pad_signal(xd, L, n1, pad, out)
        assert np.array_equal(out.cpu().numpy(), padsignal_np(x, pad)[0])
    print("cwt rpadded batch", flush=True)
    Wb, scales = sq.cwt(x, 'gmw', rpadded=True)
    assert Wb.shape == (3, len(scales), L)
    W1, _ = sq.cwt(x[1], 'gmw', rpadded=True)
    assert rel_l2(Wb[1].cpu().numpy(), W1.cpu().numpy()) < 1e-6
