"""The C-ABI library loads and exports every entry point include/ssqk.h
declares (no GPU needed: nothing is called)."""
import ctypes
import os
import re

import conftest
from ssqueezepy_amd import _lib

HEADER = os.path.join(conftest.ROOT, 'include', 'ssqk.h')


def _declared():
    src = open(HEADER).read()
    src = re.sub(r'/\*.*?\*/', '', src, flags=re.S)
    return sorted(set(re.findall(r'\b(ssqk_[a-z0-9_]+)\s*\(', src)))


def test_header_symbols_exported():
    names = _declared()
    assert 'ssqk_cwt_exec' in names and 'ssqk_squeeze' in names
    lib = _lib.load()
    for n in names:
        assert hasattr(lib, n), n
    # the Python binding declares exactly the header's functions
    assert sorted(_lib.EXPORTED) == names


def test_bin_params_layout():
    # int32 type, int32 flipud, int64 omax, int64 idx1, 4 doubles
    assert ctypes.sizeof(_lib.BinParams) == 56
    assert _lib.BinParams.v0.offset == 24
