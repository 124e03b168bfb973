"""The C-ABI library loads and exports every entry point include/td1c_hprod.h declares
(no compute: runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    text = open(os.path.join(ROOT, "include", "td1c_hprod.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:int|void|double|const char\*)\s+\**\s*([A-Za-z_][A-Za-z0-9_]*)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_lists_expected_functions():
    from td1c_amd.abi import HEADER_FUNCTIONS

    assert _header_functions() == sorted(HEADER_FUNCTIONS)


def test_library_exports_every_header_symbol():
    import td1c_amd

    path = td1c_amd.library_path()
    if not os.path.exists(path):
        pytest.skip("library not built")
    lib = ctypes.CDLL(path)
    for name in _header_functions():
        assert hasattr(lib, name), name


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    import td1c_amd

    monkeypatch.setattr(td1c_amd, "_LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(td1c_amd, "_lib", None)
    with pytest.raises(RuntimeError):
        td1c_amd.load_library()


def test_field_values_without_gpu():
    """td1c_field_value is host code: check it against the closed forms of
    clfield1 (src/field/clfield1.cpp:263-300) without touching a GPU."""
    import math

    import numpy as np
    import td1c_amd

    if not os.path.exists(td1c_amd.library_path()):
        pytest.skip("library not built")
    f = td1c_amd.Field(env_type="sin2", gauge="velocity1", fint=1e14, wlen=800.0, numcyc=4.0)
    period = td1c_amd.field_period(f)
    freq = 1239.84190 / (800.0 * 27.2113845)
    assert abs(period - 2 * math.pi / freq) < 1e-12
    famp = math.sqrt(1e14 / 3.50944506e16)
    for t in [0.3 * period, 1.7 * period, 3.2 * period]:
        lf = td1c_amd.field_value(f, t)
        env = math.sin(math.pi * t / (4 * period)) ** 2
        a = env * famp / freq * math.sin(freq * t)
        assert abs(lf[2] - a) < 1e-15 and abs(lf[8] - a) < 1e-15
        # E = -dA/dt (analytic derivative vs the 6-point stencil)
        denv = 2 * math.sin(math.pi * t / (4 * period)) * math.cos(math.pi * t / (4 * period)) * math.pi / (4 * period)
        e = -(denv * famp / freq * math.sin(freq * t) + env * famp * math.cos(freq * t))
        assert abs(lf[5] - e) < 1e-9
    assert np.all(td1c_amd.field_value(f, -1.0) == 0.0)
    assert np.all(td1c_amd.field_value(f, 4 * period + 1.0) == 0.0)
