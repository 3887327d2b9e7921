"""Higher-order GMW transforms on the GPU (cwt / ssq_cwt with order > 0,
_cwt.py:517-610, _ssq_cwt.py:227-243) against the oracle's float64 CWT with
the order-k wavelet (_gmw.py:267-394)."""
import numpy as np
import pytest

from oracle import ssq_oracle as O
from test_gpu_cwt import chirp_tone_noise, rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('N', [256, 3000, 40000])
@pytest.mark.parametrize('k', [1, 2, 3])
def test_cwt_order_k(N, k):
    """Every path of the plan (band, overlap-save, four-step) with the
    Laguerre-modulated wavelet, whose lobes cross zero."""
    import ssqueezepy_amd as sq
    x = chirp_tone_noise(N, k)
    print(f"cwt order={k} N={N}", flush=True)
    from ssqueezepy_amd._ssq_cwt import ssq_cwt_grid
    Wx, scales, dWx = sq.cwt(x, 'gmw', order=k, derivative=True)
    s64 = ssq_cwt_grid(N, sq.Wavelet('gmw'))['scales64']      # the zeroth order's
    Wo, dWo = O.cwt(x.astype(np.float64), O.psih_fn('gmw_k', k=k), s64)
    assert rel_l2(Wx.cpu().numpy(), Wo) < 1e-5
    assert rel_l2(dWx.cpu().numpy(), dWo) < 1e-5


def test_cwt_orders_averaged_and_listed():
    """order=range(3), average=False gives one transform per order
    (z_all_test.py:322-341); order=(0, 1, 2) their mean."""
    import ssqueezepy_amd as sq
    from ssqueezepy_amd._ssq_cwt import ssq_cwt_grid
    N = 2000
    x = chirp_tone_noise(N, 7)
    x = x + x[::-1]
    print("cwt orders averaged", flush=True)
    Wk, _ = sq.cwt(x, 'gmw', order=range(3), average=False)
    Wa, _ = sq.cwt(x, 'gmw', order=(0, 1, 2))
    s64 = ssq_cwt_grid(N, sq.Wavelet('gmw'))['scales64']
    Wo = [O.cwt(x.astype(np.float64), O.psih_fn('gmw_k', k=k), s64)[0] for k in range(3)]
    assert isinstance(Wk, list) and len(Wk) == 3
    for a, b in zip(Wk, Wo):
        assert rel_l2(a.cpu().numpy(), b) < 1e-5
    assert rel_l2(Wa.cpu().numpy(), np.mean(Wo, axis=0)) < 1e-5
    W1, _, dW1 = sq.cwt(x, 'gmw', order=1, average=False, derivative=True)
    assert rel_l2(W1.cpu().numpy(), Wo[1]) < 1e-5


@pytest.mark.parametrize('order', [2, (0, 1)])
def test_ssq_cwt_orders(order):
    """ssq_cwt(order): Wx and dWx are the orders' average; Tx squeezes that
    Wx into the rows of its phase transform (index-matched against the
    oracle's float64 Wx with the same rows)."""
    import ssqueezepy_amd as sq
    from ssqueezepy_amd._ssq_cwt import ssq_cwt_grid
    N = 5000
    x = chirp_tone_noise(N, 3)
    print(f"ssq_cwt order={order}", flush=True)
    Tx, Wx, f, s, w, dWx = sq.ssq_cwt(x, 'gmw', order=order, get_w=True, get_dWx=True)
    g = ssq_cwt_grid(N, sq.Wavelet('gmw'))
    ks = [order] if isinstance(order, int) else list(order)
    outs = [O.cwt(x.astype(np.float64), O.psih_fn('gmw_k', k=k), g['scales64']) for k in ks]
    Wo = np.mean([o[0] for o in outs], axis=0)
    dWo = np.mean([o[1] for o in outs], axis=0)
    assert rel_l2(Wx.cpu().numpy(), Wo) < 1e-5
    assert rel_l2(dWx.cpu().numpy(), dWo) < 1e-5
    k = O.squeeze_bins(w.cpu().numpy().astype(np.float64), O.bin_params(g['ssq_freqs'], True),
                       len(g['scales64']) - 1, True)
    Tx_o = O.ssqueeze(Wo, k, g['const'])
    assert rel_l2(Tx.cpu().numpy(), Tx_o) < 1e-5
    Tx2 = sq.ssq_cwt(x, 'gmw', order=order)[0]
    assert rel_l2(Tx2.cpu().numpy(), Tx.cpu().numpy()) < 1e-2
