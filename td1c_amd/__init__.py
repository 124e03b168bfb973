"""td1c_amd -- MI355X (gfx950) orbital Hamiltonian product and RK step of td1c.

Python mirror of the reference's C++ interface for the hot path:

  HProd        ~ clhprod (src/hprod/clhprod.cpp): set_wfn0, htot, energies
  CRK          ~ clcrk   (src/prop/clcrk.cpp):   prop (rk_order 1..4)
  Field        ~ clfield (src/field/clfield.cpp): get_value(t) -> lfield[9]

Wave functions are complex128 torch tensors on the GPU holding the orbital
coefficients (td1c layout [orbital][l][r]) followed by the CI vector.  All
compute runs in the HIP library libtd1c_hprod.so through its C-ABI
(include/td1c_hprod.h); torch only provides device memory and streams.  The
library must be built (td1c_amd/build.py); there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

from .abi import HEADER_FUNCTIONS, Field, Problem, TD1CField, TD1CProblem, TD1CSizes

__all__ = ["Problem", "Field", "HProd", "CRK", "load_library", "library_path", "field_value", "field_period",
           "wfn_write", "wfn_read", "ci_layout"]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libtd1c_hprod.so")
_lib = None


def library_path() -> str:
    return _LIB_PATH


def load_library():
    """Loads libtd1c_hprod.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise RuntimeError(
            f"{_LIB_PATH} is missing: build it with `python td1c_amd/build.py` "
            "(the hprod has no CPU fallback)")
    L = ctypes.CDLL(_LIB_PATH)
    vp = ctypes.c_void_p
    L.td1c_ctx_create.argtypes = [ctypes.POINTER(TD1CProblem), ctypes.POINTER(vp)]
    L.td1c_ctx_create.restype = ctypes.c_int
    L.td1c_ctx_destroy.argtypes = [vp]
    L.td1c_ctx_destroy.restype = ctypes.c_int
    L.td1c_get_sizes.argtypes = [vp, ctypes.POINTER(TD1CSizes)]
    L.td1c_get_sizes.restype = ctypes.c_int
    L.td1c_set_wfn0.argtypes = [vp, vp, vp, vp]
    L.td1c_set_wfn0.restype = ctypes.c_int
    L.td1c_htot.argtypes = [vp, ctypes.c_double, vp, vp, vp, vp, vp, vp]
    L.td1c_htot.restype = ctypes.c_int
    L.td1c_get_energies.argtypes = [vp, vp]
    L.td1c_get_energies.restype = ctypes.c_int
    L.td1c_rk_prop.argtypes = [vp, ctypes.c_int, ctypes.c_double, vp, vp, ctypes.c_int, vp]
    L.td1c_rk_prop.restype = ctypes.c_int
    L.td1c_propagate.argtypes = [vp, ctypes.POINTER(TD1CField), ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_int, vp, ctypes.c_int, vp]
    L.td1c_propagate.restype = ctypes.c_int
    L.td1c_etdrb_gen.argtypes = [vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double]
    L.td1c_etdrb_gen.restype = ctypes.c_int
    L.td1c_etdrb_prop.argtypes = [vp, vp, vp, ctypes.c_int, vp]
    L.td1c_etdrb_prop.restype = ctypes.c_int
    L.td1c_etdrb_propagate.argtypes = [vp, ctypes.POINTER(TD1CField), ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                       vp, ctypes.c_int, vp]
    L.td1c_etdrb_propagate.restype = ctypes.c_int
    L.td1c_etdrk4_gen.argtypes = [vp, ctypes.c_double, ctypes.c_int, ctypes.c_double]
    L.td1c_etdrk4_gen.restype = ctypes.c_int
    L.td1c_etdrk4_prop.argtypes = [vp, vp, vp, ctypes.c_int, vp]
    L.td1c_etdrk4_prop.restype = ctypes.c_int
    L.td1c_etdrk4_propagate.argtypes = [vp, ctypes.POINTER(TD1CField), ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int, vp, ctypes.c_int, vp]
    L.td1c_etdrk4_propagate.restype = ctypes.c_int
    L.td1c_lawrk_gen.argtypes = [vp, ctypes.c_double, ctypes.c_int, ctypes.c_double]
    L.td1c_lawrk_gen.restype = ctypes.c_int
    L.td1c_lawrk_prop.argtypes = [vp, vp, vp, ctypes.c_int, vp]
    L.td1c_lawrk_prop.restype = ctypes.c_int
    L.td1c_lawrk_propagate.argtypes = [vp, ctypes.POINTER(TD1CField), ctypes.c_double, ctypes.c_double,
                                       ctypes.c_int, vp, ctypes.c_int, vp]
    L.td1c_lawrk_propagate.restype = ctypes.c_int
    L.td1c_expab_gen.argtypes = [vp, ctypes.c_double, ctypes.c_int, ctypes.c_double]
    L.td1c_expab_gen.restype = ctypes.c_int
    L.td1c_expab_prop.argtypes = [vp, ctypes.c_double, vp, vp, ctypes.c_int, vp]
    L.td1c_expab_prop.restype = ctypes.c_int
    L.td1c_expab_propagate.argtypes = [vp, ctypes.POINTER(TD1CField), ctypes.c_double, ctypes.c_double,
                                       ctypes.c_int, vp, ctypes.c_int, vp]
    L.td1c_expab_propagate.restype = ctypes.c_int
    L.td1c_chkconv.argtypes = [vp, vp, vp, vp, vp]
    L.td1c_chkconv.restype = ctypes.c_int
    L.td1c_init_prop12.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, vp, vp,
                                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int), vp]
    L.td1c_init_prop12.restype = ctypes.c_int
    L.td1c_set_graph.argtypes = [vp, ctypes.c_int]
    L.td1c_set_graph.restype = ctypes.c_int
    L.td1c_set_transform_path.argtypes = [vp, ctypes.c_int]
    L.td1c_set_transform_path.restype = ctypes.c_int
    L.td1c_dipole.argtypes = [vp, vp, vp, vp, vp, vp]
    L.td1c_dipole.restype = ctypes.c_int
    L.td1c_dipole3.argtypes = [vp, vp, vp, vp, vp, vp]
    L.td1c_dipole3.restype = ctypes.c_int
    L.td1c_dipole_all.argtypes = [vp, vp, vp, vp, vp, vp]
    L.td1c_dipole_all.restype = ctypes.c_int
    L.td1c_ipx.argtypes = [vp, ctypes.c_int, ctypes.c_double, vp, vp, vp, vp, vp]
    L.td1c_ipx.restype = ctypes.c_int
    L.td1c_ipd.argtypes = [vp, ctypes.c_int, ctypes.c_double, vp, vp, vp, vp, vp]
    L.td1c_ipd.restype = ctypes.c_int
    L.td1c_orth.argtypes = [vp, vp, vp]
    L.td1c_orth.restype = ctypes.c_int
    L.td1c_init_split2.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, vp, vp,
                                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int), vp]
    L.td1c_init_split2.restype = ctypes.c_int
    cp = ctypes.c_char_p
    ci = ctypes.c_int
    L.td1c_wfn_write.argtypes = [cp, cp, ci, ci, ci, ci, ci, vp]
    L.td1c_wfn_write.restype = ci
    L.td1c_wfn_read.argtypes = [cp, cp, ci, ci, ci, ci, ci, vp]
    L.td1c_wfn_read.restype = ci
    L.td1c_wfn_proj.argtypes = [ctypes.POINTER(TD1CProblem), vp, ctypes.POINTER(TD1CProblem), vp]
    L.td1c_wfn_proj.restype = ci
    L.td1c_ci_layout.argtypes = [ctypes.POINTER(TD1CProblem), vp, vp, vp, vp, vp, vp, vp, vp]
    L.td1c_ci_layout.restype = ci
    L.td1c_htotx.argtypes = [vp, ctypes.c_double, vp, vp, vp, vp]
    L.td1c_htotx.restype = ci
    for name in ("td1c_get_den1", "td1c_get_den2", "td1c_get_int1e", "td1c_get_int2e", "td1c_get_xmat"):
        getattr(L, name).argtypes = [vp, vp]
        getattr(L, name).restype = ci
    L.td1c_mask.argtypes = [vp, vp, vp]
    L.td1c_mask.restype = ctypes.c_int
    L.td1c_field_value.argtypes = [ctypes.POINTER(TD1CField), ctypes.c_double, vp]
    L.td1c_field_value.restype = ctypes.c_int
    L.td1c_field_der.argtypes = [ctypes.POINTER(TD1CField), ctypes.c_double, vp]
    L.td1c_field_der.restype = ctypes.c_int
    L.td1c_field_period.argtypes = [ctypes.POINTER(TD1CField)]
    L.td1c_field_period.restype = ctypes.c_double
    L.td1c_bind.argtypes = [vp]
    L.td1c_bind.restype = ctypes.c_int
    L.hprod_htot_.argtypes = [vp, vp, vp, vp, vp, vp]
    L.hprod_htot_.restype = None
    L.td1c_set_timing.argtypes = [vp, ctypes.c_int]
    L.td1c_set_timing.restype = ctypes.c_int
    L.td1c_get_timing.argtypes = [vp, ctypes.c_char_p, ctypes.c_int]
    L.td1c_get_timing.restype = ctypes.c_int
    L.td1c_set_timing_kernel.argtypes = [vp, ctypes.c_char_p]
    L.td1c_set_timing_kernel.restype = ctypes.c_int
    L.td1c_h1itr_gen.argtypes = [vp, ctypes.c_double, ctypes.c_int, ctypes.c_double]
    L.td1c_h1itr_gen.restype = ctypes.c_int
    L.td1c_h1itr_prop.argtypes = [vp, vp, vp, vp]
    L.td1c_h1itr_prop.restype = ctypes.c_int
    L.hprod_htot_static.argtypes = [vp, vp, vp, vp, vp, vp]
    L.hprod_htot_static.restype = ctypes.c_int
    L.td1c_set_onebody_shape.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.td1c_set_onebody_shape.restype = ctypes.c_int
    L.td1c_get_shapes.argtypes = [vp, ctypes.c_char_p, ctypes.c_int]
    L.td1c_get_shapes.restype = ctypes.c_int
    L.td1c_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(vp)]
    L.td1c_batch_create.restype = ctypes.c_int
    L.td1c_batch_destroy.argtypes = [vp]
    L.td1c_batch_destroy.restype = ctypes.c_int
    L.td1c_batch_propagate.argtypes = [vp, ctypes.POINTER(TD1CField), ctypes.c_int, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_int, ctypes.POINTER(vp), ctypes.c_int, vp]
    L.td1c_batch_propagate.restype = ctypes.c_int
    L.td1c_batch_set_graph.argtypes = [vp, ctypes.c_int]
    L.td1c_batch_set_graph.restype = ctypes.c_int
    L.td1c_batch_set_shapes.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.td1c_batch_set_shapes.restype = ctypes.c_int
    L.td1c_batch_autotune.argtypes = [vp, ctypes.POINTER(TD1CField), ctypes.c_int, ctypes.c_double,
                                      ctypes.POINTER(vp), vp]
    L.td1c_batch_autotune.restype = ctypes.c_int
    L.td1c_batch_get_info.argtypes = [vp, ctypes.c_char_p, ctypes.c_int]
    L.td1c_batch_get_info.restype = ctypes.c_int
    L.td1c_last_error.argtypes = []
    L.td1c_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def _check(rc: int, what: str):
    if rc != 0:
        msg = load_library().td1c_last_error().decode()
        raise RuntimeError(f"{what} failed ({rc}): {msg}")


def _ptr(a) -> ctypes.c_void_p:
    if hasattr(a, "data_ptr"):
        return ctypes.c_void_p(a.data_ptr())
    return a.ctypes.data_as(ctypes.c_void_p)


def field_value(field: Field, t: float) -> np.ndarray:
    """clfield::get_value: lfield[0..2] gauge value, [3..5] E, [6..8] A."""
    out = np.zeros(9)
    c = field.to_c()
    _check(load_library().td1c_field_value(ctypes.byref(c), float(t), _ptr(out)), "td1c_field_value")
    return out


def field_der(field: Field, t: float) -> np.ndarray:
    """clfield::get_der: d lfield[0..2] / dt (the rest zero)."""
    out = np.zeros(9)
    c = field.to_c()
    _check(load_library().td1c_field_der(ctypes.byref(c), float(t), _ptr(out)), "td1c_field_der")
    return out


def etd_fields(field: Field, t: float, dt: float) -> np.ndarray:
    """The six field arrays of one cletdrb step from t (etd_time_of_h 0):
    lfield(t), lfield(t + dt/2), lfield(t + dt), d lfield/dt (t),
    lfield(t + dt/2) - lfield(t), lfield(t + dt) - lfield(t)."""
    f0, f1, f3 = field_value(field, t), field_value(field, t + 0.5 * dt), field_value(field, t + dt)
    return np.concatenate([f0, f1, f3, field_der(field, t), f1 - f0, f3 - f0])


def field_period(field: Field) -> float:
    c = field.to_c()
    return float(load_library().td1c_field_period(ctypes.byref(c)))


def wfn_write(orb_path: str, cic_path: str, prob: Problem, lcic: int, wfn: np.ndarray):
    """clwfn::write: the state (host complex128, orbitals then CI) to the
    reference's .orb / .cic text files."""
    w = np.ascontiguousarray(wfn, dtype=np.complex128)
    nrad = prob.nr + 1
    if w.size != prob.nfun * prob.nbas + lcic:
        raise ValueError("wave function size does not match the problem")
    mmax1 = prob.mmax1 if prob.mmax1 >= 0 else max([abs(m) for m in (prob.mval or [0])])
    _check(load_library().td1c_wfn_write(orb_path.encode(), cic_path.encode(), nrad, prob.lmax1, mmax1,
                                         prob.nfun, lcic, _ptr(w)), "td1c_wfn_write")


def ci_layout(prob: Problem):
    """td1c_ci_layout: the CI vector's layout in the reference's ORMAS order.
    Returns a dict with the ordered alpha / beta string occupation bit
    patterns (bit k = active orbital k), the string pair of each determinant
    (det_alpha, det_beta), lcic and cic0 (the determinant ormas_cic0 sets,
    -1 without active orbitals).  Host computation, no device."""
    L = load_library()
    c, keep = prob.to_c()
    na, nb, lc, c0 = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _check(L.td1c_ci_layout(ctypes.byref(c), ctypes.byref(na), ctypes.byref(nb), ctypes.byref(lc),
                            ctypes.byref(c0), None, None, None, None), "td1c_ci_layout")
    sa = np.zeros(max(na.value, 1), np.uint32)
    sb = np.zeros(max(nb.value, 1), np.uint32)
    da = np.zeros(lc.value, np.int32)
    db = np.zeros(lc.value, np.int32)
    _check(L.td1c_ci_layout(ctypes.byref(c), None, None, None, None, _ptr(sa), _ptr(sb), _ptr(da), _ptr(db)),
           "td1c_ci_layout")
    del keep
    return {"str_alpha": sa[:na.value], "str_beta": sb[:nb.value], "det_alpha": da, "det_beta": db,
            "lcic": lc.value, "cic0": c0.value}


def wfn_read(orb_path: str, cic_path: str, prob: Problem, lcic: int):
    """clwfn::read: returns (state, ci_read) for the problem's sizes."""
    nrad = prob.nr + 1
    w = np.zeros(prob.nfun * prob.nbas + lcic, np.complex128)
    cic0 = ci_layout(prob)["cic0"] if prob.nact > 0 else -1
    rc = load_library().td1c_wfn_read(orb_path.encode(), cic_path.encode(), nrad, prob.lmax1, prob.nfun, lcic,
                                      cic0, _ptr(w))
    if rc not in (0, 1):
        _check(rc, "td1c_wfn_read")
    return w, rc == 0


def wfn_proj(prob_from: Problem, wfn_from: np.ndarray, prob_to: Problem) -> np.ndarray:
    """clbas::proj with the CI copy of guess_type "read_proj": the state of
    prob_from (host complex128, orbitals then CI) on prob_to's grid and sizes.
    Host computation, no device; orthonormalise the device copy with
    HProd.orth as clwfn::orth follows in the reference."""
    lc_from = ci_layout(prob_from)["lcic"]
    lc_to = ci_layout(prob_to)["lcic"]
    w1 = np.ascontiguousarray(wfn_from, dtype=np.complex128)
    if w1.size != prob_from.nfun * prob_from.nbas + lc_from:
        raise ValueError("wave function size does not match the source problem")
    w2 = np.zeros(prob_to.nfun * prob_to.nbas + lc_to, np.complex128)
    a, ka = prob_from.to_c()
    b, kb = prob_to.to_c()
    _check(load_library().td1c_wfn_proj(ctypes.byref(a), _ptr(w1), ctypes.byref(b), _ptr(w2)), "td1c_wfn_proj")
    del ka, kb
    return w2


class HProd:
    """GPU hprod for one problem (clhprod)."""

    def __init__(self, prob: Problem):
        import torch  # device memory only

        self.torch = torch
        self.prob = prob
        L = load_library()
        c, self._keep = prob.to_c()
        h = ctypes.c_void_p()
        _check(L.td1c_ctx_create(ctypes.byref(c), ctypes.byref(h)), "td1c_ctx_create")
        self.h = h
        s = TD1CSizes()
        _check(L.td1c_get_sizes(h, ctypes.byref(s)), "td1c_get_sizes")
        self.sizes = s
        self.nbas, self.nfun, self.lcic = s.nbas, s.nfun, s.lcic
        self.norb = s.nbas * s.nfun
        self.wfn_size = int(s.wfn_size)
        self.npair = s.npair

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            load_library().td1c_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _stream(self, stream):
        if stream is None:
            return ctypes.c_void_p(self.torch.cuda.current_stream().cuda_stream)
        return ctypes.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else stream)

    def set_wfn0(self, wfn: np.ndarray, v2xfc: Optional[np.ndarray] = None):
        """clhprod::set_wfn0 (host arrays): frozen-core set-up."""
        w = np.ascontiguousarray(wfn, dtype=np.complex128)
        if w.size != self.wfn_size:
            raise ValueError(f"wave function must have {self.wfn_size} complex entries")
        cic = np.ascontiguousarray(w[self.norb:])
        vx = None if v2xfc is None else np.ascontiguousarray(v2xfc, dtype=np.complex128)
        _check(load_library().td1c_set_wfn0(self.h, _ptr(w), _ptr(cic), None if vx is None else _ptr(vx)),
               "td1c_set_wfn0")

    def to_device(self, wfn: np.ndarray):
        t = self.torch.from_numpy(np.ascontiguousarray(wfn, dtype=np.complex128))
        return t.to("cuda")

    def htot(self, dtime: float, lfield: Sequence[float], wfn, hwfn=None, stream=None):
        """clhprod::htot: hwfn = dt * d(wfn)/dt for a device wave function."""
        torch = self.torch
        if wfn.dtype != torch.complex128 or not wfn.is_cuda or wfn.numel() != self.wfn_size:
            raise ValueError("wfn must be a complex128 CUDA tensor of wfn_size elements")
        wfn = wfn.contiguous()
        if hwfn is None:
            hwfn = torch.empty_like(wfn)
        lf = np.ascontiguousarray(lfield, dtype=np.float64)
        if lf.size != 9:
            raise ValueError("lfield must have 9 entries")
        base = wfn.data_ptr()
        hbase = hwfn.data_ptr()
        off = self.norb * 16
        _check(load_library().td1c_htot(self.h, float(dtime), _ptr(lf), ctypes.c_void_p(base),
                                        ctypes.c_void_p(base + off), ctypes.c_void_p(hbase),
                                        ctypes.c_void_p(hbase + off), self._stream(stream)), "td1c_htot")
        return hwfn

    def h1itr_gen(self, dtime: float, maxcyc: int = 20, thresh: float = 1.0e-10):
        """clh1itr::gen for split_type = 1: the half-step factors of a step
        of dtime."""
        _check(load_library().td1c_h1itr_gen(self.h, float(dtime), int(maxcyc), float(thresh)), "td1c_h1itr_gen")

    def h1itr_prop(self, lfield: Sequence[float], wfn, stream=None):
        """One clh1itr half step (h1itr_prop) on a device wave function, in place."""
        torch = self.torch
        if wfn.dtype != torch.complex128 or not wfn.is_cuda or wfn.numel() != self.wfn_size:
            raise ValueError("wfn must be a complex128 CUDA tensor of wfn_size elements")
        lf = np.ascontiguousarray(lfield, dtype=np.float64)
        if lf.size != 9:
            raise ValueError("lfield must have 9 entries")
        _check(load_library().td1c_h1itr_prop(self.h, _ptr(lf), ctypes.c_void_p(wfn.data_ptr()),
                                              self._stream(stream)), "td1c_h1itr_prop")
        return wfn

    def energies(self) -> np.ndarray:
        """ene_fcore, ene_dcore, ene_core, ene_act, ene_tot of the last htot."""
        e = np.zeros(5)
        _check(load_library().td1c_get_energies(self.h, _ptr(e)), "td1c_get_energies")
        return e

    def htotx(self, dtime: float, lfield: Sequence[float], wfn, stream=None):
        """clhprod::htotx: the hprod without the orbital derivative; the
        RDMs, integrals and X matrix are then read with get()."""
        torch = self.torch
        if wfn.dtype != torch.complex128 or not wfn.is_cuda or wfn.numel() != self.wfn_size:
            raise ValueError("wfn must be a complex128 CUDA tensor of wfn_size elements")
        lf = np.ascontiguousarray(lfield, dtype=np.float64)
        if lf.size != 9:
            raise ValueError("lfield must have 9 entries")
        base = wfn.data_ptr()
        _check(load_library().td1c_htotx(self.h, float(dtime), _ptr(lf), ctypes.c_void_p(base),
                                         ctypes.c_void_p(base + 16 * self.norb), self._stream(stream)), "td1c_htotx")

    def get(self, name: str) -> np.ndarray:
        """den1, den2, int1e, int2e (active indices) or xmat (all orbitals)
        of the last htot / htotx, as column-major numpy arrays."""
        na, nf = self.prob.nact, self.nfun
        shape = {"den1": (na, na), "int1e": (na, na), "den2": (na,) * 4, "int2e": (na,) * 4,
                 "xmat": (nf, nf)}[name]
        out = np.zeros(int(np.prod(shape)), np.complex128)
        _check(getattr(load_library(), "td1c_get_" + name)(self.h, _ptr(out)), "td1c_get_" + name)
        return out.reshape(shape, order="F")

    def dipole(self, lfield: Sequence[float], wfn, out=None, stream=None):
        """Length-form dipole <z> (hprod_dipole dip(1)) of a device wave
        function; returns a float (synchronising) unless `out`, a float64
        CUDA tensor of one element, is given."""
        torch = self.torch
        lf = np.ascontiguousarray(lfield, dtype=np.float64)
        if lf.size != 9:
            raise ValueError("lfield must have 9 entries")
        ret = out is None
        if out is None:
            out = torch.zeros(1, dtype=torch.float64, device=wfn.device)
        base = wfn.data_ptr()
        _check(load_library().td1c_dipole(self.h, _ptr(lf), ctypes.c_void_p(base),
                                          ctypes.c_void_p(base + 16 * self.norb), ctypes.c_void_p(out.data_ptr()),
                                          self._stream(stream)), "td1c_dipole")
        return float(out.item()) if ret else out

    def dipole3(self, lfield: Sequence[float], wfn, out=None, stream=None):
        """hprod_dipole's Ehrenfest terms [dip(1), vel(1), acc(1)] (length,
        velocity and acceleration forms; frozen-core corrections excluded) of
        a device wave function; returns a numpy array (synchronising) unless
        `out`, a float64 CUDA tensor of three elements, is given."""
        torch = self.torch
        lf = np.ascontiguousarray(lfield, dtype=np.float64)
        if lf.size != 9:
            raise ValueError("lfield must have 9 entries")
        if wfn.dtype != torch.complex128 or not wfn.is_cuda or wfn.numel() != self.wfn_size:
            raise ValueError("wfn must be a complex128 CUDA tensor of wfn_size elements")
        ret = out is None
        if out is None:
            out = torch.zeros(3, dtype=torch.float64, device=wfn.device)
        elif out.dtype != torch.float64 or not out.is_cuda or out.numel() < 3:
            raise ValueError("out must be a float64 CUDA tensor of three elements")
        base = wfn.data_ptr()
        _check(load_library().td1c_dipole3(self.h, _ptr(lf), ctypes.c_void_p(base),
                                           ctypes.c_void_p(base + 16 * self.norb), ctypes.c_void_p(out.data_ptr()),
                                           self._stream(stream)), "td1c_dipole3")
        return out.cpu().numpy() if ret else out

    def dipole_all(self, lfield: Sequence[float], wfn, out=None, stream=None):
        """hprod_dipole in full: [dip(1..3), vel(1..3), acc(1..3)], (2) the
        frozen-core corrections, (3) = (1) + (2); returns a numpy array
        (synchronising) unless `out`, a float64 CUDA tensor of nine elements,
        is given."""
        torch = self.torch
        lf = np.ascontiguousarray(lfield, dtype=np.float64)
        if lf.size != 9:
            raise ValueError("lfield must have 9 entries")
        if wfn.dtype != torch.complex128 or not wfn.is_cuda or wfn.numel() != self.wfn_size:
            raise ValueError("wfn must be a complex128 CUDA tensor of wfn_size elements")
        ret = out is None
        if out is None:
            out = torch.zeros(9, dtype=torch.float64, device=wfn.device)
        elif out.dtype != torch.float64 or not out.is_cuda or out.numel() < 9:
            raise ValueError("out must be a float64 CUDA tensor of nine elements")
        base = wfn.data_ptr()
        _check(load_library().td1c_dipole_all(self.h, _ptr(lf), ctypes.c_void_p(base),
                                              ctypes.c_void_p(base + 16 * self.norb),
                                              ctypes.c_void_p(out.data_ptr()), self._stream(stream)),
               "td1c_dipole_all")
        if ret:
            torch.cuda.synchronize()
            return out.cpu().numpy()
        return out

    def ipx(self, max_ipx: int, rad_ipx: float, lfield: Sequence[float], wfn, stream=None) -> np.ndarray:
        """hprod_ipx: probabilities that exactly k = 0..max_ipx electrons are
        outside rad_ipx, for a device wave function (synchronising)."""
        torch = self.torch
        lf = np.ascontiguousarray(lfield, dtype=np.float64)
        if lf.size != 9:
            raise ValueError("lfield must have 9 entries")
        if wfn.dtype != torch.complex128 or not wfn.is_cuda or wfn.numel() != self.wfn_size:
            raise ValueError("wfn must be a complex128 CUDA tensor of wfn_size elements")
        out = np.zeros(max(int(max_ipx), 0) + 1)
        base = wfn.data_ptr()
        _check(load_library().td1c_ipx(self.h, int(max_ipx), float(rad_ipx), _ptr(lf), ctypes.c_void_p(base),
                                       ctypes.c_void_p(base + 16 * self.norb), _ptr(out), self._stream(stream)),
               "td1c_ipx")
        return out

    def ipd(self, max_ipd: int, rad_ipx: float, lfield: Sequence[float], wfn, stream=None) -> np.ndarray:
        """hprod_ipd: the factorial-moment estimate of the probabilities that
        k = 0..max_ipd (at most 2) electrons are outside rad_ipx (synchronising)."""
        torch = self.torch
        lf = np.ascontiguousarray(lfield, dtype=np.float64)
        if lf.size != 9:
            raise ValueError("lfield must have 9 entries")
        if wfn.dtype != torch.complex128 or not wfn.is_cuda or wfn.numel() != self.wfn_size:
            raise ValueError("wfn must be a complex128 CUDA tensor of wfn_size elements")
        out = np.zeros(max(int(max_ipd), 0) + 1)
        base = wfn.data_ptr()
        _check(load_library().td1c_ipd(self.h, int(max_ipd), float(rad_ipx), _ptr(lf), ctypes.c_void_p(base),
                                       ctypes.c_void_p(base + 16 * self.norb), _ptr(out), self._stream(stream)),
               "td1c_ipd")
        return out

    def orth(self, wfn, stream=None):
        """clwfn::orth (schmidt) in place on a device wave function."""
        torch = self.torch
        if wfn.dtype != torch.complex128 or not wfn.is_cuda or wfn.numel() != self.wfn_size:
            raise ValueError("wfn must be a complex128 CUDA tensor of wfn_size elements")
        _check(load_library().td1c_orth(self.h, ctypes.c_void_p(wfn.data_ptr()), self._stream(stream)), "td1c_orth")
        return wfn

    def chkconv(self, lfield, wfn, stream=None) -> np.ndarray:
        """hprod_chkconv: (energy, |hcic| / lcic, grad(2), grad(3))."""
        lf = np.ascontiguousarray(lfield, dtype=np.float64)
        out = np.zeros(4)
        _check(load_library().td1c_chkconv(self.h, _ptr(lf), ctypes.c_void_p(wfn.data_ptr()), _ptr(out),
                                           self._stream(stream)), "td1c_chkconv")
        return out

    def init_prop12(self, wfn, dstep: float = 0.01, tolene: float = 1e-12, tolwfn: float = 1e-6,
                    maxcyc: int = 1000000, lfield: Optional[Sequence[float]] = None, stream=None):
        """init_prop12 with cletdrb (imaginary-time relaxation, the reference's
        default) in place on a device wave function; returns (energy, last
        cycle).  The problem must have icomp = 0 and split_type = 1."""
        lf = np.ascontiguousarray(lfield if lfield is not None else np.zeros(9), dtype=np.float64)
        e, n = ctypes.c_double(), ctypes.c_int()
        _check(load_library().td1c_init_prop12(self.h, float(dstep), float(tolene), float(tolwfn), int(maxcyc),
                                               _ptr(lf), ctypes.c_void_p(wfn.data_ptr()), ctypes.byref(e),
                                               ctypes.byref(n), self._stream(stream)), "td1c_init_prop12")
        return e.value, n.value

    def init_split2(self, wfn, dstep: float = 0.01, tolene: float = 1e-13, maxcyc: int = 1000000,
                    lfield: Optional[Sequence[float]] = None, rk_order: int = 4, stream=None):
        """init_split2 (imaginary-time relaxation with the RK propagator) in
        place on a device wave function; returns (energy, last cycle).  The
        problem must have icomp = 0."""
        torch = self.torch
        if wfn.dtype != torch.complex128 or not wfn.is_cuda or wfn.numel() != self.wfn_size:
            raise ValueError("wfn must be a complex128 CUDA tensor of wfn_size elements")
        lf = np.zeros(9) if lfield is None else np.ascontiguousarray(lfield, dtype=np.float64)
        if lf.size != 9:
            raise ValueError("lfield must have 9 entries")
        e = ctypes.c_double()
        n = ctypes.c_int()
        _check(load_library().td1c_init_split2(self.h, int(rk_order), float(dstep), float(tolene), int(maxcyc),
                                               _ptr(lf), ctypes.c_void_p(wfn.data_ptr()), ctypes.byref(e),
                                               ctypes.byref(n), self._stream(stream)), "td1c_init_split2")
        return e.value, n.value

    def mask(self, wfn, stream=None):
        _check(load_library().td1c_mask(self.h, ctypes.c_void_p(wfn.data_ptr()), self._stream(stream)), "td1c_mask")
        return wfn

    def set_timing(self, on: bool, only: Optional[str] = None):
        """Per-kernel HIP-event timing on the launch stream; `only` restricts
        the events to the kernel of that name."""
        L = load_library()
        _check(L.td1c_set_timing_kernel(self.h, only.encode() if only else None), "td1c_set_timing_kernel")
        _check(L.td1c_set_timing(self.h, int(on)), "td1c_set_timing")

    def get_timing(self) -> dict:
        """{kernel: (total_ms, launches)} collected since set_timing(True)."""
        import json

        buf = ctypes.create_string_buffer(1 << 16)
        _check(load_library().td1c_get_timing(self.h, buf, len(buf)), "td1c_get_timing")
        return {k: tuple(v) for k, v in json.loads(buf.value.decode()).items()}

    def shapes(self) -> dict:
        """The launch shapes the autotuner chose for this context."""
        import json

        buf = ctypes.create_string_buffer(8192)
        _check(load_library().td1c_get_shapes(self.h, buf, 8192), "td1c_get_shapes")
        return json.loads(buf.value.decode())

    def set_onebody_shape(self, elements_per_block: int, l_per_block: int):
        _check(load_library().td1c_set_onebody_shape(self.h, int(elements_per_block), int(l_per_block)),
               "td1c_set_onebody_shape")

    def set_graph(self, on: bool):
        """td1c_propagate replays one captured step (True) or enqueues every
        kernel (False, the default)."""
        _check(load_library().td1c_set_graph(self.h, int(on)), "td1c_set_graph")

    def set_transform_path(self, path: int):
        """Form of the theta <-> l transforms: 0 by grid size (default), 1 the
        vector form, 2 the matrix-core (MFMA) form."""
        _check(load_library().td1c_set_transform_path(self.h, int(path)), "td1c_set_transform_path")

    def bind(self):
        """Makes this context the one the legacy hprod_htot_ entry point uses."""
        _check(load_library().td1c_bind(self.h), "td1c_bind")


class CRK:
    """Explicit Runge-Kutta propagator (clcrk, rk_formula = "standard")."""

    NODES = {1: [0.0], 2: [0.0, 0.5], 3: [0.0, 0.5, 1.0], 4: [0.0, 0.5, 0.5, 1.0]}

    def __init__(self, rk_order: int = 4):
        if rk_order not in self.NODES:
            raise ValueError("rk_order must be 1..4")
        self.rk_order = rk_order

    def stage_fields(self, field: Field, time: float, dtime: float) -> np.ndarray:
        return np.stack([field_value(field, time + dtime * c) for c in self.NODES[self.rk_order]])

    def prop(self, hprod: HProd, wfn, time: float, dtime: float, field: Optional[Field] = None,
             lfield_stages: Optional[np.ndarray] = None, apply_mask: bool = True, stream=None):
        """One step: wfn <- wfn + dt sum_s b_s k_s (in place), then the mask
        (tdse_split2 applies clwfn::mask after every step)."""
        if lfield_stages is None:
            if field is None:
                raise ValueError("need a field or explicit lfield_stages")
            lfield_stages = self.stage_fields(field, time, dtime)
        lf = np.ascontiguousarray(lfield_stages, dtype=np.float64)
        if lf.shape != (len(self.NODES[self.rk_order]), 9):
            raise ValueError("lfield_stages must be (nstage, 9)")
        _check(load_library().td1c_rk_prop(hprod.h, self.rk_order, float(dtime), _ptr(lf),
                                           ctypes.c_void_p(wfn.data_ptr()), int(apply_mask),
                                           hprod._stream(stream)), "td1c_rk_prop")
        return wfn

    def propagate(self, hprod: HProd, wfn, time: float, dtime: float, nstep: int, field: Field,
                  apply_mask: bool = True, stream=None):
        """nstep steps from `time` with the stage fields evaluated on the
        device (tdse_split2's loop): every kernel enqueued on the stream, or
        with hprod.set_graph(True) one captured step replayed per step.  For
        split_type = 1 a step is the clh1itr half step, the RK step of the
        split hprod, the second half step, then the mask."""
        cf = field.to_c()
        _check(load_library().td1c_propagate(hprod.h, ctypes.byref(cf), self.rk_order, float(time), float(dtime),
                                             int(nstep), ctypes.c_void_p(wfn.data_ptr()), int(apply_mask),
                                             hprod._stream(stream)), "td1c_propagate")
        return wfn


class ETDRB:
    """Exponential Rosenbrock RK4 with Pade phi functions (cletdrb, the
    reference's default prop_type; td1c_etdrb_*).  Needs split_type = 1.
    A context holds the tables of one exponential scheme at a time."""

    _fn = "etdrb"

    def __init__(self, hprod: "HProd", dtime: float, rosenbrock: bool = True, maxcyc: int = 20,
                 thresh: float = 1e-10):
        self.hprod = hprod
        self.dtime = float(dtime)
        self._gen(rosenbrock, maxcyc, thresh)

    def _gen(self, rosenbrock, maxcyc, thresh):
        _check(load_library().td1c_etdrb_gen(self.hprod.h, self.dtime, int(rosenbrock), int(maxcyc),
                                             float(thresh)), "td1c_etdrb_gen")

    def prop(self, wfn, time: float = 0.0, field: Optional[Field] = None, lfields: Optional[np.ndarray] = None,
             apply_mask: bool = True, stream=None):
        """One step in place with host field arrays (etd_fields of the field
        at `time`, or lfields given), then the mask."""
        if lfields is None:
            if field is None:
                raise ValueError("need a field or explicit lfields")
            lfields = etd_fields(field, time, self.dtime)
        lf = np.ascontiguousarray(lfields, dtype=np.float64)
        if lf.size != 54:
            raise ValueError("lfields must hold 6 x 9 values")
        fn = getattr(load_library(), f"td1c_{self._fn}_prop")
        _check(fn(self.hprod.h, _ptr(lf), ctypes.c_void_p(wfn.data_ptr()), int(apply_mask),
                  self.hprod._stream(stream)), f"td1c_{self._fn}_prop")
        return wfn

    def propagate(self, wfn, time: float, nstep: int, field: Field, apply_mask: bool = True, stream=None):
        """nstep steps of tdse_prop12 from `time`, the fields evaluated on the
        device."""
        cf = field.to_c()
        fn = getattr(load_library(), f"td1c_{self._fn}_propagate")
        _check(fn(self.hprod.h, ctypes.byref(cf), float(time), self.dtime, int(nstep),
                  ctypes.c_void_p(wfn.data_ptr()), int(apply_mask), self.hprod._stream(stream)),
               f"td1c_{self._fn}_propagate")
        return wfn


class ETDRK4(ETDRB):
    """The Cox-Matthews exponential RK4 with [2/2] Pade forms (cletdrk4,
    prop_type "etdrk4"; td1c_etdrk4_*).  Needs split_type = 1."""

    _fn = "etdrk4"

    def __init__(self, hprod: "HProd", dtime: float, maxcyc: int = 20, thresh: float = 1e-10):
        super().__init__(hprod, dtime, False, maxcyc, thresh)

    def _gen(self, rosenbrock, maxcyc, thresh):
        _check(load_library().td1c_etdrk4_gen(self.hprod.h, self.dtime, int(maxcyc), float(thresh)),
               "td1c_etdrk4_gen")


class LAWRK(ETDRB):
    """Lawson RK4 with the (1,1) Pade of exp(-i h1 dt/2) (cllawrk,
    prop_type "lawrk"; td1c_lawrk_*).  Needs split_type = 1."""

    _fn = "lawrk"

    def __init__(self, hprod: "HProd", dtime: float, maxcyc: int = 20, thresh: float = 1e-10):
        super().__init__(hprod, dtime, False, maxcyc, thresh)

    def _gen(self, rosenbrock, maxcyc, thresh):
        _check(load_library().td1c_lawrk_gen(self.hprod.h, self.dtime, int(maxcyc), float(thresh)),
               "td1c_lawrk_gen")


class EXPAB(ETDRB):
    """Exponential Adams-Bashforth of order 4 (clexpab, prop_type "expab";
    td1c_expab_*).  The history lives in the context: consecutive calls
    continue one propagation; constructing restarts it."""

    _fn = "expab"

    def __init__(self, hprod: "HProd", dtime: float, maxcyc: int = 20, thresh: float = 1e-10):
        super().__init__(hprod, dtime, False, maxcyc, thresh)
        self.nstep = 0

    def _gen(self, rosenbrock, maxcyc, thresh):
        _check(load_library().td1c_expab_gen(self.hprod.h, self.dtime, int(maxcyc), float(thresh)),
               "td1c_expab_gen")

    @staticmethod
    def history_fields(field: Field, t0: float, dtime: float, n: int) -> np.ndarray:
        """lfield(t_n) and lfield at the history times t_max(n-k, 0), k = 1..3."""
        return np.concatenate([field_value(field, t0 + max(n - k, 0) * dtime) for k in range(4)])

    def prop(self, wfn, time: float = 0.0, field: Optional[Field] = None, lfields: Optional[np.ndarray] = None,
             apply_mask: bool = True, stream=None, t0: Optional[float] = None):
        """One step at `time`; lfields (4 x 9) as history_fields gives them
        (from the field and the propagation start t0 when not given)."""
        if lfields is None:
            if field is None or t0 is None:
                raise ValueError("need lfields, or a field and the start time t0")
            lfields = self.history_fields(field, t0, self.dtime, self.nstep)
        lf = np.ascontiguousarray(lfields, dtype=np.float64)
        if lf.size != 36:
            raise ValueError("lfields must hold 4 x 9 values")
        _check(load_library().td1c_expab_prop(self.hprod.h, float(time), _ptr(lf), ctypes.c_void_p(wfn.data_ptr()),
                                              int(apply_mask), self.hprod._stream(stream)), "td1c_expab_prop")
        self.nstep += 1
        return wfn

    def propagate(self, wfn, time: float, nstep: int, field: Field, apply_mask: bool = True, stream=None):
        super().propagate(wfn, time, nstep, field, apply_mask, stream)
        self.nstep += nstep
        return wfn


class Batch:
    """Independent propagations run together (td1c_batch_*): the members are
    HProd contexts (dynamical cores only, ndvr = 20, one nfun, precision,
    split_type and gauge); their grids, lmax and pulses may differ.  Each
    step issues one launch per kernel site for all members.  Keep the
    members alive while the batch exists."""

    def __init__(self, members: Sequence["HProd"]):
        import torch

        self.torch = torch
        self.members = list(members)
        arr = (ctypes.c_void_p * len(self.members))(*[m.h.value for m in self.members])
        self.h = ctypes.c_void_p()
        _check(load_library().td1c_batch_create(arr, len(self.members), ctypes.byref(self.h)), "td1c_batch_create")

    def close(self):
        if self.h:
            load_library().td1c_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_graph(self, on: bool):
        _check(load_library().td1c_batch_set_graph(self.h, int(on)), "td1c_batch_set_graph")

    def set_shapes(self, elements_per_block: int, l_per_block: int, transform_variant: int):
        _check(load_library().td1c_batch_set_shapes(self.h, int(elements_per_block), int(l_per_block),
                                                    int(transform_variant)), "td1c_batch_set_shapes")

    def autotune(self, scratch, dtime: float, fields: Sequence[Field], rk_order: int = 4, stream=None):
        """Times the batch's launch shapes on the scratch device wave
        functions (overwritten) and keeps the fastest."""
        n = len(self.members)
        cf = (TD1CField * n)(*[f.to_c() for f in fields])
        ptrs = (ctypes.c_void_p * n)(*[w.data_ptr() for w in scratch])
        _check(load_library().td1c_batch_autotune(self.h, cf, int(rk_order), float(dtime), ptrs, self._stream(stream)),
               "td1c_batch_autotune")

    def _stream(self, stream):
        if stream is None:
            return ctypes.c_void_p(self.torch.cuda.current_stream().cuda_stream)
        return ctypes.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else stream)

    def info(self) -> dict:
        import json

        buf = ctypes.create_string_buffer(8192)
        _check(load_library().td1c_batch_get_info(self.h, buf, 8192), "td1c_batch_get_info")
        return json.loads(buf.value.decode())

    def propagate(self, wfns, time: float, dtime: float, nstep: int, fields: Sequence[Field], rk_order: int = 4,
                  apply_mask: bool = True, stream=None):
        """nstep steps of every member from `time`: member m's device wave
        function wfns[m] under fields[m], as CRK(rk_order).propagate would
        step each one alone."""
        n = len(self.members)
        if len(wfns) != n or len(fields) != n:
            raise ValueError("need one wave function and one field per member")
        cf = (TD1CField * n)(*[f.to_c() for f in fields])
        ptrs = (ctypes.c_void_p * n)(*[w.data_ptr() for w in wfns])
        _check(load_library().td1c_batch_propagate(self.h, cf, int(rk_order), float(time), float(dtime), int(nstep),
                                                   ptrs, int(apply_mask), self._stream(stream)), "td1c_batch_propagate")
        return wfns
