"""Default parameters of the transforms.

Values mirror the reference's configuration table (ssqueezepy/configs.ini:1-45):
wavelet defaults, STFT dtype and the log-piecewise downsampling factor.  A
keyword left as ``None`` by the caller takes the value from here, as the
reference's ``gdefaults`` does (ssqueezepy/configs.py:27-82).
"""

DEFAULTS = {
    'wavelets.morlet': dict(mu=13.4, dtype='float32'),
    '_gmw.gmw': dict(gamma=3., beta=60., norm='bandpass', order=0,
                     centered_scale=False, dtype='float32'),
    '_stft.stft': dict(dtype='float32'),
    'utils.cwt_utils.make_scales': dict(downsample=4),
}


def gdefaults(module_and_obj, get_all=False, **kw):
    """Fill ``None``-valued keywords from ``DEFAULTS[module_and_obj]``.

    With ``get_all`` also adds every default not passed.  Returns a dict.
    Follows ssqueezepy/configs.py:27-82 (dict form only).
    """
    table = DEFAULTS.get(module_and_obj, {})
    out = dict(kw)
    for k, v in out.items():
        if v is None and k in table:
            out[k] = table[k]
    if get_all:
        ordered = {k: out.get(k, v) for k, v in table.items()}
        for k, v in out.items():
            if k not in ordered:
                ordered[k] = v
        out = ordered
    return out
