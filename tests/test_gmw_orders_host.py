"""Higher-order GMWs on the host (no GPU): the oracle's order-k wavelet is
the zeroth-order GMW at k = 0, its orders are mutually orthogonal, and the
product's device constants are the oracle's."""
import numpy as np
from scipy import integrate

from oracle import ssq_oracle as O


def test_order_zero_is_gmw():
    w = np.linspace(0, 8, 2001)
    for beta in (60., 8.):
        assert np.allclose(O.gmw_k_psih(w, 3., beta, 0), O.gmw_psih(w, 3., beta),
                           rtol=1e-13, atol=0)


def test_orders_are_orthogonal():
    """psih_j psih_k integrates to 0 for j != k (Laguerre polynomials of
    2 w^gamma under the weight x^c e^-x, c = (2 beta + 1) / gamma - 1)."""
    for gamma, beta in ((3., 60.), (3., 8.), (2., 5.)):
        f = [lambda v, k=k: O.gmw_k_psih(np.array([v]), gamma, beta, k)[0] for k in range(4)]
        hi = 4 * (beta / gamma)**(1 / gamma) + 10
        G = np.array([[integrate.quad(lambda v: f[j](v) * f[k](v), 0, hi, limit=400)[0]
                       for k in range(4)] for j in range(4)])
        d = np.sqrt(np.diag(G))
        off = G / np.outer(d, d) - np.eye(4)
        assert np.abs(off).max() < 1e-8, (gamma, beta, off)


def test_device_constants_match_the_oracle():
    import ssqueezepy_amd as sq
    from ssqueezepy_amd._cwt import gmw_order_params, KIND_GMW_K
    for beta in (60., 8.):
        wav = sq.Wavelet(('gmw', {'beta': beta}))
        for k in range(5):
            kind, prm = gmw_order_params(wav, k)
            assert kind == KIND_GMW_K and prm[:3] == (3., beta, k + 1.)
            assert np.allclose(prm[3:], O.gmw_k_consts(3., beta, k), rtol=1e-13, atol=0)


def test_average_constants_give_the_mean_wavelet():
    """One wavelet with the orders' mean constants is the mean of the
    orders' wavelets (the transform is linear in psih), for orders with
    and without zero, against the oracle's per-order psih."""
    import ssqueezepy_amd as sq
    from ssqueezepy_amd._cwt import gmw_average_params
    w = np.linspace(0.01, 12, 2001)
    wav = sq.Wavelet('gmw')
    wc = (60 / 3)**(1 / 3)
    base = np.exp(-60 * np.log(wc) + wc**3 + 60 * np.log(w) - w**3)
    for orders in [(0, 1), (0, 1, 2), (1, 3), (2,)]:
        kind, prm = gmw_average_params(wav, orders)
        assert kind == 2 and int(prm[2]) == max(orders) + 1 == len(prm) - 3
        psih = np.polyval(np.asarray(prm[3:])[::-1], 2 * w**3) * base
        ref = np.mean([O.gmw_k_psih(w, 3., 60., k) if k else O.gmw_psih(w, 3., 60.)
                       for k in orders], axis=0)
        assert np.allclose(psih, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
