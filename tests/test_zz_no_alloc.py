"""hprod_htot_static, the allocation-free variant of hprod_htot_: the same
results as hprod_htot_ (and the oracle) on two calls in a row through the
C ABI, the second on a smaller problem, and a clear error for a wave
function beyond its static staging array."""
import ctypes

import numpy as np
import pytest

from tests import cases as C

pytestmark = pytest.mark.gpu


def _call(fn, dt, lf, orb, cic):
    hw = np.zeros_like(orb)
    hc = np.zeros_like(cic)
    d = ctypes.c_double(dt)
    args = (ctypes.byref(d), lf.ctypes.data_as(ctypes.c_void_p), orb.ctypes.data_as(ctypes.c_void_p),
            cic.ctypes.data_as(ctypes.c_void_p), hw.ctypes.data_as(ctypes.c_void_p),
            hc.ctypes.data_as(ctypes.c_void_p))
    return fn(*args), hw, hc


def test_static_entry_point_matches_hprod_htot(gpu_lib):
    import td1c_amd
    from oracle.oracle import Oracle

    gpu_lib.hprod_htot_static.restype = ctypes.c_int
    lf = np.ascontiguousarray(C.LF_LEN)
    keep = []
    for name in ["cas6_small", "he_small"]:  # the second call is on a smaller wave function
        prob = getattr(C, name)()
        o = Oracle(prob)
        w = C.wfn_for(prob, o.lcic)
        o.set_wfn0(w)
        ref = o.htot(0.5, C.LF_LEN, w)
        hp = td1c_amd.HProd(prob)
        keep.append(hp)
        hp.set_wfn0(w)
        hp.bind()
        orb = np.ascontiguousarray(w[:o.norb])
        cic = np.ascontiguousarray(w[o.norb:])
        _, hw0, hc0 = _call(gpu_lib.hprod_htot_, 0.5, lf, orb, cic)
        rc, hw1, hc1 = _call(gpu_lib.hprod_htot_static, 0.5, lf, orb, cic)
        assert rc == 0, td1c_amd.load_library().td1c_last_error()
        # the same kernels on the same inputs: bit-identical to hprod_htot_
        assert np.array_equal(hw1, hw0) and np.array_equal(hc1, hc0), name
        rel = np.linalg.norm(hw1 - ref[:o.norb]) / np.linalg.norm(ref[:o.norb])
        assert rel < 1e-11, (name, rel)
        if o.lcic > 1:
            assert np.linalg.norm(hc1 - ref[o.norb:]) / np.linalg.norm(ref[o.norb:]) < 1e-11


def test_static_entry_point_rejects_a_too_large_wave_function(gpu_lib):
    import td1c_amd
    from td1c_amd.abi import Problem

    gpu_lib.hprod_htot_static.restype = ctypes.c_int
    gpu_lib.td1c_last_error.restype = ctypes.c_char_p
    # 160 elements x 20 = 3199 radial points, lmax1 63, 11 orbitals:
    # 2252096 complex values, above the 2097152 the static array stages
    prob = Problem(nfe=160, ndvr=20, rmax=320.0, lmax1=63, lmax2=126, znuc=22.0, ndcore=11,
                   nel_alpha=11, nel_beta=11)
    hp = td1c_amd.HProd(prob)
    assert hp.wfn_size > 2097152
    hp.bind()
    orb = np.zeros(hp.norb, np.complex128)
    cic = np.zeros(max(hp.lcic, 1), np.complex128)
    rc, _, _ = _call(gpu_lib.hprod_htot_static, 0.5, np.ascontiguousarray(C.LF_LEN), orb, cic)
    assert rc == -2
    assert b"static staging limit" in gpu_lib.td1c_last_error()
