"""The C-ABI library: loads, exports every declared symbol, rejects bad input (CPU only)."""

import ctypes as ct
import os
import re

import numpy as np
import pytest

import heyoka_amd as hy
from heyoka_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "hyx.h")).read()
    return sorted(set(re.findall(r"\b(hyx_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = _declared()
    assert len(names) >= 20
    for nm in names:
        assert hasattr(L, nm), nm
    assert set(names) == set(_lib.EXPORTED)


def test_invalid_oplist_rejected():
    L = _lib.lib()
    words = np.array([0xDEADBEEF, 1, 0, 1, 0, 6, 0], dtype=np.uint32)
    h = ct.c_void_p()
    rc = L.hyx_sys_create(_lib.dptr(words, ct.c_uint32), len(words), None, 0, 20, _lib.FLAG_COMPILE_ONLY,
                          ct.byref(h))
    assert rc == -1 and b"magic" in L.hyx_last_error()


def test_generated_source_and_compile_only(tmp_path, monkeypatch):
    monkeypatch.setenv("HYX_CACHE_DIR", str(tmp_path))
    dec = hy.taylor_decompose(hy.model.pendulum())
    s = _lib.System(dec, 20, compile_only=True)
    src = s.source()
    assert "hyx_taylor_step" in src and "hyx_persist_k" in src
    assert not s.compact
    assert any(f.suffix == ".co" for f in tmp_path.iterdir())
    sc = _lib.System(dec, 20, compile_only=True, compact=True)
    assert sc.compact and "HYX_COMPACT" in sc.source()


def test_no_device_fails_loudly():
    if _lib.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(_lib.HyxError, match="no HIP device"):
        hy.taylor_adaptive_batch(hy.model.pendulum(), [[0.1], [0.0]])
