"""CPU oracle for the ssq_cwt / ssq_stft / cwt / stft path -- TEST
INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, and only as the checker (or as the timed CPU baseline); the product
package never imports it.

A float64 NumPy restatement of the reference's CPU algorithm.  Scale and
frequency grids are inputs here: they come from the product's host code,
which the golden fixtures written from the reference pin separately.  Each
function cites the reference lines it follows (paths relative to the
ssqueezepy package root).  How each function is pinned:

  pad_signal        np.pad (library) + golden e2e fixtures
  gmw_psih,
  morlet_psih       closed forms (_gmw.py:203-210, wavelets.py:521-523),
                    tests: peak value 2 at wc, L1 time norm, golden e2e
  cwt               brute-force DFT convolution at tiny N; golden e2e Wx/dWx;
                    l1_norm=False: equal time-domain energy of every scale's
                    impulse response, and the reconstruction threshold;
                    rpadded: brute-force circular convolution of the
                    padded signal, unpadded columns = the regular cwt
  phase_transform   closed form Im(dWx/Wx)/2pi (tests/fft_test.py:159-161)
  squeeze_bins,
  ssqueeze          brute-force per-point loop; golden squeeze.npz
                    (ssqueeze_fast / indexed_sum_onfly outputs)
  stft              direct per-frame DFT definition; golden e2e Sx/dSx
  ssq_cwt, ssq_stft golden e2e fixtures (Tx, Wx, w, dWx) and C1 samples
  gmw_adm_ssq,
  gmw_adm_cwt       closed forms (Gamma-function integrals of the GMW)
  gmw_k_consts,
  gmw_k_psih        order 0 is gmw_psih; orders are mutually orthogonal
                    (Laguerre polynomials under the weight x^c e^-x)
  issq_full,
  invert_components,
  icwt_1int,
  icwt_2int         the reference's reconstruction thresholds
                    (tests/reconstruction_test.py:65-123) on the float64
                    chain; components + remainder = full inversion
  ridge_track       known answer (tests/ridge_extraction_test.py:17-26);
                    the tracked path has the least total cost among all
                    paths (brute force on tiny maps)
"""
import numpy as np

TWO_PI = 6.283185307179586


# ---------------------------------------------------------------- padding
def pad_lengths(N, padlength=None):
    """utils/common.py:32-51 (p2up) and :100-111."""
    if padlength is None:
        up = int(2**(1 + np.round(np.log2(N))))
        n2 = (up - N) // 2
        return up, up - N - n2, n2
    if abs(padlength - N) % 2 == 0:
        n1 = n2 = (padlength - N) // 2
    else:
        n2 = (padlength - N) // 2
        n1 = n2 + 1
    return padlength, n1, n2


def pad_signal(x, padtype='reflect', padlength=None):
    """utils/common.py:54-158 (NumPy branch)."""
    N = x.shape[-1]
    up, n1, n2 = pad_lengths(N, padlength)
    width = (n1, n2) if x.ndim == 1 else [(0, 0), (n1, n2)]
    mode = {'zero': 'constant', 'reflect': 'reflect', 'replicate': 'edge',
            'wrap': 'wrap', 'symmetric': 'symmetric'}[padtype]
    return np.pad(x, width, mode=mode), up, n1, n2


# --------------------------------------------------------------- wavelets
def xi_grid(N):
    """wavelets.py:473-484: [0..N/2, -(N/2-1)..-1] * 2pi/N (Nyquist +pi)."""
    i = np.arange(N, dtype=np.float64)
    i[N // 2 + 1:] -= N
    return i * (2 * np.pi / N)


def gmw_psih(w, gamma=3., beta=60.):
    """_gmw.py:203-210: 2 exp(-b ln wc + wc^g + b ln w - w^g) [w >= 0]."""
    w = np.asarray(w, dtype=np.float64)
    wc = (beta / gamma)**(1 / gamma)
    out = np.zeros_like(w)
    m = w > 0
    out[m] = 2 * np.exp(-beta * np.log(wc) + wc**gamma
                        + beta * np.log(w[m]) - w[m]**gamma)
    return out


def morlet_psih(w, mu=13.4):
    """wavelets.py:498-523."""
    w = np.asarray(w, dtype=np.float64)
    cs = (1 + np.exp(-mu**2) - 2 * np.exp(-3 / 4 * mu**2)) ** (-.5)
    ks = np.exp(-.5 * mu**2)
    C1 = np.sqrt(2) * cs * np.pi**.25
    return C1 * (np.exp(-.5 * (w - mu)**2) - ks * np.exp(-.5 * w**2))


def gmw_k_consts(gamma=3., beta=60., k=1):
    """_gmw.py:366-394 (bandpass): 2 * coeff * (-1)^m Gamma(k+c+1) /
    (Gamma(c+m+1) Gamma(k-m+1) m!), r = (2 beta + 1) / gamma, c = r - 1,
    coeff = sqrt(Gamma(r) k! / Gamma(k + r))."""
    from scipy.special import gammaln
    r = (2 * beta + 1) / gamma
    c = r - 1
    coeff = np.sqrt(np.exp(gammaln(r) + gammaln(k + 1) - gammaln(k + r)))
    m = np.arange(k + 1)
    L = (-1.)**m * np.exp(gammaln(k + c + 1) - gammaln(c + m + 1) - gammaln(k - m + 1)
                          - gammaln(m + 1))
    return 2 * coeff * L


def gmw_k_psih(w, gamma=3., beta=60., k=1):
    """_gmw.py:267-293: the order-k bandpass GMW, sum_m kc[m] (2 w^gamma)^m
    times exp(-beta ln wc + wc^gamma + beta ln w - w^gamma), 0 for w < 0."""
    w = np.asarray(w, dtype=np.float64)
    kc = gmw_k_consts(gamma, beta, k)
    wc = (beta / gamma)**(1 / gamma)
    out = np.zeros_like(w)
    m = w > 0
    u = 2 * w[m]**gamma
    C = np.polyval(kc[::-1], u)
    out[m] = C * np.exp(-beta * np.log(wc) + wc**gamma + beta * np.log(w[m]) - w[m]**gamma)
    return out


def psih_fn(kind, **params):
    if kind == 'gmw_k':
        return lambda w: gmw_k_psih(w, params.get('gamma', 3.), params.get('beta', 60.),
                                    params.get('k', 1))
    if kind == 'gmw':
        return lambda w: gmw_psih(w, params.get('gamma', 3.), params.get('beta', 60.))
    return lambda w: morlet_psih(w, params.get('mu', 13.4))


# -------------------------------------------------------------------- cwt
def cwt(x, psih, scales, dt=1., padtype='reflect', derivative=True, l1_norm=True,
        rpadded=False):
    """_cwt.py:167-177, 246-320 (vectorized, float64): pad, fft,
    psih(scale*xi) with the Nyquist bin halved, ifft, unpad (all L columns
    with rpadded, _cwt.py:293-305); with l1_norm=False Wx and dWx times
    sqrt(scale) (_cwt.py:307-311).  x 1D or 2D."""
    x = np.asarray(x, dtype=np.float64)
    N = x.shape[-1]
    xp, L, n1, _ = pad_signal(x, padtype)
    xh = np.fft.fft(xp, axis=-1)
    xi = xi_grid(L)
    scales = np.asarray(scales, dtype=np.float64).reshape(-1, 1)
    P = psih(scales * xi)
    P[:, L // 2] /= 2                                  # wavelets.py:86-95
    if x.ndim == 2:
        xh = xh[:, None]
    if not l1_norm:
        P = P * np.sqrt(scales)
    cols = slice(None) if rpadded else slice(n1, n1 + N)
    Wx = np.fft.ifft(P * xh, axis=-1)[..., cols]
    if not derivative:
        return Wx, None
    dWx = np.fft.ifft(P * xh * (1j * xi / dt), axis=-1)[..., cols]
    return Wx, dWx


# ------------------------------------------------------------ phase / bins
def phase_transform(Wx, dWx, gamma, Sfs=None):
    """algos.py:720-741 (cwt), 794-816 (stft): |Im(dWx conj Wx)| / (|Wx|^2 2pi),
    inf where |Wx| < gamma; stft form |Sfs[i] - ...|."""
    A, B = dWx.real, dWx.imag
    C, D = Wx.real, Wx.imag
    with np.errstate(divide='ignore', invalid='ignore'):
        v = (B * C - A * D) / ((C**2 + D**2) * TWO_PI)
    w = np.abs(v if Sfs is None else Sfs.reshape(-1, 1) - v)
    w[np.abs(Wx) < gamma] = np.inf
    return w


def bin_params(ssq_freqs, logscale):
    """algos.py:83-90, 356-374 (including the EPS64 step clamp)."""
    v = np.asarray(ssq_freqs, dtype=np.float64)
    eps = np.finfo(np.float64).eps
    if not logscale:
        return dict(kind='lin', vmin=float(v[0]), dv=max(float(v[1] - v[0]), eps))
    d2 = np.abs(np.diff(np.log(v), 2))
    idx = int(np.argmax(d2) + 2)
    mx = d2.max()
    d2[idx - 2] = 0
    th = 1e-14 if v.dtype == np.float64 else 1e-6
    piecewise = bool(np.any(mx > 100 * np.abs(d2).mean()) and np.all(np.abs(d2) < th))
    vlmin = float(np.log2(v[0]))
    dvl0 = max(float(np.log2(v[1]) - np.log2(v[0])), eps)
    if not piecewise:
        return dict(kind='log', vlmin=vlmin, dvl=dvl0)
    return dict(kind='piecewise', vlmin0=vlmin, vlmin1=float(np.log2(v[idx - 1])),
                dvl0=dvl0, dvl1=max(float(np.log2(v[idx]) - np.log2(v[idx - 1])), eps),
                idx1=idx - 1)


def squeeze_bins(w, params, omax, flipud):
    """Row index of every finite w (algos.py:172-250, 859-953): round half to
    even, clamps in the reference's order; -1 where w is inf."""
    w = np.asarray(w, dtype=np.float64)
    fin = np.isfinite(w)
    wf = np.where(fin, w, 1.)
    with np.errstate(divide='ignore', invalid='ignore'):
        if params['kind'] == 'lin':
            k = np.minimum(np.rint(np.maximum((wf - params['vmin']) / params['dv'], 0)), omax)
        elif params['kind'] == 'log':
            k = np.minimum(np.rint(np.maximum(
                (np.log2(wf) - params['vlmin']) / params['dvl'], 0)), omax)
        else:
            wl = np.log2(wf)
            hi = np.minimum(np.rint((wl - params['vlmin1']) / params['dvl1'])
                            + params['idx1'], omax)
            lo = np.maximum(np.rint((wl - params['vlmin0']) / params['dvl0']), 0)
            k = np.where(wl > params['vlmin1'], hi, lo)
    k = k.astype(np.int64)
    if flipud:
        k = omax - k
    return np.where(fin, k, -1)


def ssqueeze(Wx, k, const, nf=None):
    """Tx[k[i, j], j] += Wx[i, j] * const[i], rows i in ascending order
    (algos.py:172-250, 859-984); k == -1 is skipped."""
    Wx = np.asarray(Wx)
    na, N = Wx.shape
    nf = nf or na
    Tx = np.zeros((nf, N), dtype=np.complex128)
    const = np.broadcast_to(np.asarray(const, dtype=np.float64).ravel(), (na,))
    cols = np.arange(N)
    for i in range(na):
        m = k[i] >= 0
        Tx[k[i, m], cols[m]] += Wx[i, m] * const[i]
    return Tx


def ssq_cwt(x, psih, scales, ssq_freqs, const, logscale, gamma, dt=1.,
            padtype='reflect', flipud=True, fused=True):
    """_ssq_cwt.py:190-310 in float64 with given grids.  ``ssq_freqs``
    ascending (before the flip); returns (Tx, Wx, dWx, w, bins).  ``fused``
    uses the |Wx| > gamma test of ssqueeze_fast, else the phase_cwt +
    indexed_sum_onfly route (|Wx| >= gamma)."""
    Wx, dWx = cwt(x, psih, scales, dt, padtype)
    params = bin_params(ssq_freqs, logscale)
    na = Wx.shape[-2]
    w = phase_transform(Wx, dWx, gamma)
    if fused:
        wf = np.where(np.abs(Wx) > gamma, np.abs(
            (dWx.imag * Wx.real - dWx.real * Wx.imag)
            / ((Wx.real**2 + Wx.imag**2) * TWO_PI)), np.inf)
    else:
        wf = w
    k = squeeze_bins(wf, params, na - 1, flipud)
    if Wx.ndim == 2:
        Tx = ssqueeze(Wx, k, const)
    else:
        Tx = np.stack([ssqueeze(a, b, const) for a, b in zip(Wx, k)])
    return Tx, Wx, dWx, w, k


# ------------------------------------------------------------------- stft
def stft(x, window, diff_window, n_fft, hop=1, fs=1., padtype='reflect',
         modulated=True, chunk=4096):
    """_stft.py:127-147 and utils/stft_utils.py:69-98 by the DFT definition:
    frames xp[f hop + p], modulated = frame ifftshift, times the ifftshifted
    window; rfft over the frame.  Returns Sx, dSx [n_fft//2+1, n_frames]."""
    x = np.asarray(x, dtype=np.float64)
    N = len(x)
    xp, _, _, _ = pad_signal(x, padtype, padlength=N + n_fft - 1)
    n_frames = (len(xp) - n_fft) // hop + 1
    s20 = int(np.ceil(n_fft / 2))
    s21 = s20 - 1 if (n_fft % 2 == 1) else s20
    perm = np.arange(n_fft)
    if modulated:
        perm = np.concatenate([np.arange(s21, s21 + s20), np.arange(0, s21)])
        win = np.fft.ifftshift(window)
        dwin = np.fft.ifftshift(diff_window) * fs
    else:
        win, dwin = window, diff_window
    rows = n_fft // 2 + 1
    Sx = np.zeros((rows, n_frames), dtype=np.complex128)
    dSx = np.zeros_like(Sx)
    for f0 in range(0, n_frames, chunk):
        f = np.arange(f0, min(f0 + chunk, n_frames))
        fr = xp[f[None, :] * hop + perm[:, None]]
        Sx[:, f] = np.fft.rfft(fr * win[:, None], axis=0)
        dSx[:, f] = np.fft.rfft(fr * dwin[:, None], axis=0)
    return Sx, dSx


def ssq_stft(x, window, diff_window, n_fft, gamma, hop=1, fs=1.,
             padtype='reflect', modulated=True, flipud=False, fused=True):
    """_ssq_stft.py:88-136 in float64: linear bins on Sfs, const = dSfs."""
    Sx, dSx = stft(x, window, diff_window, n_fft, hop, fs, padtype, modulated)
    rows = Sx.shape[0]
    Sfs = np.linspace(0, .5 * fs, rows)
    w = phase_transform(Sx, dSx, gamma, Sfs=Sfs)
    if fused:
        A, B, C, D = dSx.real, dSx.imag, Sx.real, Sx.imag
        with np.errstate(divide='ignore', invalid='ignore'):
            wf = np.abs(Sfs.reshape(-1, 1) - (B * C - A * D) / ((C**2 + D**2) * TWO_PI))
        wf[~(np.abs(Sx) > gamma)] = np.inf
    else:
        wf = w
    params = bin_params(Sfs, False)
    k = squeeze_bins(wf, params, rows - 1, flipud)
    Tx = ssqueeze(Sx, k, Sfs[1] - Sfs[0])
    return Tx, Sx, dSx, w, k, Sfs


# --------------------------------------------------------------- inverses
def gmw_adm_ssq(gamma=3., beta=60.):
    """integral(psih(w) / w, 0..inf) of the L1-normed GMW (utils/cwt_utils.py:
    28-47 integrates it numerically): 2 e^(wc^g) wc^(-b) Gamma(b/g) / g."""
    from scipy.special import gammaln
    wc = (beta / gamma)**(1 / gamma)
    return 2 * np.exp(wc**gamma - beta * np.log(wc) + gammaln(beta / gamma)) / gamma


def gmw_adm_cwt(gamma=3., beta=60.):
    """integral(psih(w)^2 / w, 0..inf) (utils/cwt_utils.py:50-63):
    4 e^(2 wc^g) wc^(-2b) Gamma(2b/g) / (g 2^(2b/g))."""
    from scipy.special import gammaln
    wc = (beta / gamma)**(1 / gamma)
    return 4 * np.exp(2 * wc**gamma - 2 * beta * np.log(wc) + gammaln(2 * beta / gamma)
                      - (2 * beta / gamma) * np.log(2)) / gamma


def issq_full(Tx, scale):
    """issq_cwt / issq_stft full inversion (_ssq_cwt.py:364-378,
    _ssq_stft.py:192-198) with NumPy's semantics: the axis-0 sum of Re Tx
    in Tx's real dtype, then times the float64 scale in place."""
    x = Tx.real.sum(axis=0)
    x *= scale
    return x


def invert_components(Tx, cc, cw, scale):
    """_invert_components (_ssq_cwt.py:381-403), then times scale: rows
    clip(cc - cw) .. clip(cc + cw) of each column per component (float64
    sums), the rest (Tx's dtype) last; cc = -1 marks no curve."""
    rows, N = Tx.shape
    K = cc.shape[1]
    x = np.zeros((K + 1, N))
    rem = Tx.copy()
    for n in range(K):
        up = np.clip(cc[:, n] + cw[:, n], 0, rows)
        lo = np.clip(cc[:, n] - cw[:, n], 0, rows)
        up[cc[:, n] == -1] = 0
        lo[cc[:, n] == -1] = 1
        mask = np.zeros(Tx.shape, dtype='complex128')
        for m in range(N):
            sl = slice(lo[m], up[m] + 1)
            mask[sl, m] = Tx[sl, m]
            rem[sl, m] = 0
        x[n] = mask.real.sum(axis=0)
    x[K] = rem.real.sum(axis=0)
    x *= scale
    return x


def icwt_1int(Wx, segments):
    """One-integral icwt (_cwt.py:437-494) with l1_norm: per (r0, r1, k,
    div) segment, k * sum of Re Wx[r0:r1] / div (div None = norm 1), the
    segments added."""
    x = None
    for r0, r1, k, div in segments:
        W = Wx[r0:r1].real
        if div is not None:
            W = W / div.reshape(-1, 1)
        xs = W.sum(axis=0) * k
        x = xs if x is None else x + xs
    return x


# ------------------------------------------------------------------ ridges

def icwt_2int(Wx, psih, scales, row_factor, padtype='reflect'):
    """Double-integral icwt (_cwt.py:456-476) in float64, each row times its
    factor (k / norm(scale)): pad the rows of Wx, fft, times psih(scale xi)
    (Nyquist bin whole: wavelet(scale=, N=) has nohalf=True) times
    (-1)^n, ifft, ifftshift, real part, sum over rows, unpad."""
    Wx = np.asarray(Wx, dtype=np.complex128)
    N = Wx.shape[-1]
    Wp, L, n1, _ = pad_signal(Wx, padtype)
    P = psih(np.asarray(scales, dtype=np.float64).reshape(-1, 1) * xi_grid(L))
    pn = (-1.) ** np.arange(L)
    xa = np.fft.ifftshift(np.fft.ifft(np.fft.fft(Wp, axis=-1) * P * pn, axis=-1), axes=-1)
    x = (xa.real * np.asarray(row_factor, dtype=np.float64)[:, None]).sum(axis=0)
    return x[n1:n1 + N]


def ridge_track(Tf, scales, penalty=2., n_ridges=1, bw=15, transform='cwt', cr=True):
    """extract_ridges (ridge_extraction.py:11-232), serial path, float32 as
    the reference evaluates a complex64 map: -log(energy / max + eps),
    forward accumulated penalty (min over jumps penalty * (s_f - s_j)^2),
    backward pass keeping the last frequency within eps, then the band
    energy[r - bw : r + bw] cleared per column.  ``cr``: |Tf| and the log
    correctly rounded (the GPU's evaluation), else NumPy's own float32
    ones.  Returns (ridge [N, K], ridge energies [N, K])."""
    eps = np.float32(np.finfo(np.float32).eps)
    scales = np.asarray(scales, dtype=np.float32)
    s = (np.log(scales) if transform == 'cwt' else scales).squeeze()
    P = np.float32(penalty) * np.subtract.outer(s, s)**2
    T = np.asarray(Tf, dtype=np.complex64)
    if cr:
        # |Tf| and log correctly rounded to float32 (NumPy's SIMD float32
        # log and complex abs are within 1 ulp of it, platform-dependent)
        h = np.sqrt(T.real.astype(np.float64)**2 + T.imag.astype(np.float64)**2)
        energy = h.astype(np.float32)**2
    else:
        energy = np.abs(T)**2
    nf, N = energy.shape
    ridges = np.zeros((N, n_ridges), dtype=int)
    ridge_e = np.zeros((N, n_ridges), dtype=np.float32)
    for i in range(n_ridges):
        q = energy / energy.max(axis=0) + eps
        E = -(np.log(q.astype(np.float64)).astype(np.float32) if cr else np.log(q))
        pen = E.copy()
        for t in range(1, N):
            pen[:, t] += np.min(pen[:, t - 1][None, :] + P, axis=1)
        r = np.argmin(pen, axis=0) % N
        for t in range(N - 2, -1, -1):
            rn = r[t + 1]
            val = pen[rn, t + 1] - E[rn, t + 1]
            hit = np.nonzero(np.abs(val - (pen[:, t] + P[rn, :])) < eps)[0]
            if len(hit):
                r[t] = hit[-1]
        ridges[:, i] = r
        ridge_e[:, i] = energy[r, np.arange(N)]
        for t in range(N):
            energy[int(r[t] - bw):int(r[t] + bw), t] = 0
    return ridges, ridge_e
