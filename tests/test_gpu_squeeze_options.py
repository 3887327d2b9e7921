"""ssqueeze on a user transform with the reference's other options
(ssqueezing.py:13-224, experimental.py:145-189): squeezing 'abs' |
'lebesgue' | a function, maprange tuple / 'maximal', flipud; against the
oracle's scatter of the same values into the bins of the same phase
transform."""
import numpy as np
import pytest

from oracle import ssq_oracle as O
from test_gpu_cwt import chirp_tone_noise, rel_l2

pytestmark = pytest.mark.gpu


def _inputs(N=3000, seed=4):
    import ssqueezepy_amd as sq
    x = chirp_tone_noise(N, seed)
    Wx, scales, dWx = sq.cwt(x, 'gmw', derivative=True)
    w = sq.phase_cwt(Wx, dWx, gamma=1e-5)
    return x, Wx, scales, dWx, w


def _oracle_tx(vals, w, ssq_freqs_asc, log, flipud, const):
    k = O.squeeze_bins(w.astype(np.float64), O.bin_params(ssq_freqs_asc, log),
                       vals.shape[0] - 1, flipud)
    return O.ssqueeze(vals, k, const)


@pytest.mark.parametrize('squeezing', ['abs', 'lebesgue', 'square', 'real'])
@pytest.mark.parametrize('flipud', [False, True])
def test_squeezing_options(squeezing, flipud):
    """Tx of each squeezing against the oracle on the same w; 'abs' and a
    real-valued function give a real Tx (the reference's zeros of |Wx|'s
    dtype)."""
    import torch
    import ssqueezepy_amd as sq
    from ssqueezepy_amd.ssqueezing import cwt_const
    from ssqueezepy_amd._scales import process_scales
    x, Wx, scales, dWx, w = _inputs()
    s = scales.cpu().numpy()
    na = len(s)
    fn = {'square': lambda W: W * W, 'real': lambda W: W.real}.get(squeezing, squeezing)
    f_asc = np.geomspace(0.004, 0.45, na)
    print(f"ssqueeze squeezing={squeezing} flipud={flipud}", flush=True)
    Tx, f_out = sq.ssqueeze(Wx, w, ssq_freqs=f_asc, scales=s, squeezing=fn, flipud=flipud)
    W64 = Wx.cpu().numpy().astype(np.complex128)
    vals = {'abs': np.abs(W64) + 0j, 'lebesgue': np.ones_like(W64) / len(W64),
            'square': W64 * W64, 'real': W64.real + 0j}[squeezing]
    _, stype, _, nv = process_scales(s, Wx.shape[-1], get_params=True)
    const = cwt_const(s, stype, nv, na)
    Tx_o = _oracle_tx(vals, w.cpu().numpy(), f_asc, True, flipud, const)
    T = Tx.cpu().numpy()
    if squeezing in ('abs', 'real'):
        assert Tx.dtype == torch.float32
        Tx_o = Tx_o.real
    else:
        assert Tx.dtype == torch.complex64
    assert rel_l2(T, Tx_o) < 1e-6
    assert np.array_equal(f_out, f_asc[::-1])      # cwt: returned descending


def test_maprange_tuple_and_maximal():
    """maprange (fm, fM) spans exactly fm..fM geometrically; 'maximal' is
    1/(N dt)..1/(2 dt) (ssqueezing.py:228-244); Tx against the oracle."""
    import ssqueezepy_amd as sq
    from ssqueezepy_amd.ssqueezing import cwt_const
    from ssqueezepy_amd._scales import process_scales
    x, Wx, scales, dWx, w = _inputs()
    s = scales.cpu().numpy()
    na, N = len(s), Wx.shape[-1]
    for maprange, (fm, fM) in [((0.01, 0.3), (0.01, 0.3)), ('maximal', (1 / N, 0.5))]:
        print(f"ssqueeze maprange={maprange}", flush=True)
        Tx, f_out = sq.ssqueeze(Wx, w, ssq_freqs='log', scales=s, maprange=maprange,
                                wavelet='gmw')
        f_asc = f_out[::-1]
        assert np.isclose(f_asc[0], fm, rtol=1e-12) and np.isclose(f_asc[-1], fM, rtol=1e-12)
        r = f_asc[1:] / f_asc[:-1]
        assert np.allclose(r, r[0], rtol=1e-10)
        _, stype, _, nv = process_scales(s, N, get_params=True)
        Tx_o = _oracle_tx(Wx.cpu().numpy().astype(np.complex128), w.cpu().numpy(), f_asc,
                          True, False, cwt_const(s, stype, nv, na))
        assert rel_l2(Tx.cpu().numpy(), Tx_o) < 1e-6


def test_phase_ssqueeze_equals_ssqueeze():
    """phase_ssqueeze = phase_transform then ssqueeze (experimental.py:
    145-189): the same Tx bit for bit as ssqueeze on phase_cwt's w, and
    the (Wx, dWx) route equals ssqueeze's."""
    import torch
    import ssqueezepy_amd as sq
    from ssqueezepy_amd.experimental import phase_ssqueeze
    x, Wx, scales, dWx, w = _inputs(seed=5)
    s = scales.cpu().numpy()
    print("phase_ssqueeze", flush=True)
    Tx1, _, f1, _, _, w1, _ = phase_ssqueeze(Wx, dWx, scales=s, wavelet='gmw', get_w=True,
                                             gamma=1e-5)
    Tx2, f2 = sq.ssqueeze(Wx, w, scales=s, wavelet='gmw', maprange='peak')
    assert torch.equal(w1, w) and torch.equal(Tx1, Tx2) and np.array_equal(f1, f2)
    Tx3, _, f3, *_ = phase_ssqueeze(Wx, dWx, scales=s, wavelet='gmw', gamma=1e-5)
    Tx4, f4 = sq.ssqueeze(Wx, None, scales=s, wavelet='gmw', maprange='peak', dWx=dWx,
                          gamma=1e-5)
    assert torch.equal(Tx3, Tx4) and np.array_equal(f3, f4)
