"""TEST INFRASTRUCTURE ONLY -- Python loader of the FP64 CPU oracle.

Loads oracle/_build/liborc.so (built from oracle/td1c_oracle.c by
`build()`, single-threaded: the checker of every test) and exposes the
reference-shaped operations on numpy arrays.  `build(openmp=True)` makes
oracle/_build/liborc_omp.so from the same source with OpenMP (threads over
independent output rows only, so its results are bit-identical), which the
benchmark's CPU-baseline leg times.  Only tests/, the smoke entry point and
that leg use this module; the product package never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional

import numpy as np

from td1c_amd.abi import Problem

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD_DIR = os.path.join(HERE, "_build")
LIB_PATH = os.path.join(BUILD_DIR, "liborc.so")
LIB_OMP_PATH = os.path.join(BUILD_DIR, "liborc_omp.so")
SRC = os.path.join(HERE, "td1c_oracle.c")

_lib = None
_lib_omp = None


def build(openmp: bool = False) -> str:
    """Compiles the oracle with gcc (-O2, no fast-math: IEEE FP64); with
    openmp the OpenMP variant liborc_omp.so."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    out = LIB_OMP_PATH if openmp else LIB_PATH
    flags = ["-fopenmp"] if openmp else ["-Wno-unknown-pragmas"]
    cmd = ["gcc", "-O2", "-std=gnu11", "-fPIC", "-shared", "-Wall"] + flags + ["-o", out, SRC, "-lm"]
    subprocess.run(cmd, check=True)
    return out


def lib(openmp: bool = False):
    global _lib, _lib_omp
    if (_lib_omp if openmp else _lib) is None:
        path = LIB_OMP_PATH if openmp else LIB_PATH
        if not os.path.exists(path) or os.path.getmtime(path) < os.path.getmtime(SRC):
            build(openmp)
        L = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [vp]
        L.orc_destroy.argtypes = [vp]
        L.orc_htot.argtypes = [vp, ctypes.c_double, vp, vp, vp, vp, vp]
        L.orc_htotx.argtypes = [vp, ctypes.c_double, vp, vp, vp]
        L.orc_set_wfn0.argtypes = [vp, vp, vp, vp]
        L.orc_mask.argtypes = [vp, vp]
        L.orc_rk_prop.argtypes = [vp, ctypes.c_int, ctypes.c_double, vp, vp, ctypes.c_int]
        L.orc_rk_prop.restype = ctypes.c_int
        L.orc_sizes.argtypes = [vp, vp]
        L.orc_table.argtypes = [vp, ctypes.c_char_p]
        L.orc_table.restype = vp
        L.orc_energies.argtypes = [vp, vp]
        L.orc_dipole.argtypes = [vp, vp, vp, vp]
        L.orc_dipole.restype = ctypes.c_double
        L.orc_dipole3.argtypes = [vp, vp, vp, vp, vp]
        L.orc_dipole3.restype = None
        L.orc_dipole_all.argtypes = [vp, vp, vp, vp, vp]
        L.orc_dipole_all.restype = None
        L.orc_ipx.argtypes = [vp, ctypes.c_int, ctypes.c_double, vp, vp, vp, vp]
        L.orc_ipx.restype = ctypes.c_int
        L.orc_ipd.argtypes = [vp, ctypes.c_int, ctypes.c_double, vp, vp, vp, vp]
        L.orc_ipd.restype = ctypes.c_int
        L.orc_orth.argtypes = [vp, vp, vp]
        L.orc_orth.restype = None
        L.orc_wfn_proj.argtypes = [vp, vp, vp, vp, vp, vp]
        L.orc_wfn_proj.restype = None
        L.orc_etdrb_gen.argtypes = [vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.orc_etdrb_gen.restype = ctypes.c_int
        L.orc_etdrb_step.argtypes = [vp, vp, vp, ctypes.c_int]
        L.orc_etdrb_step.restype = ctypes.c_int
        L.orc_etd_rat_scalar.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, vp]
        L.orc_etd_rat_scalar.restype = None
        L.orc_etdrk4_gen.argtypes = [vp, ctypes.c_double, ctypes.c_int, ctypes.c_double]
        L.orc_etdrk4_gen.restype = ctypes.c_int
        L.orc_etdrk4_step.argtypes = [vp, vp, vp, ctypes.c_int]
        L.orc_etdrk4_step.restype = ctypes.c_int
        L.orc_lawrk_gen.argtypes = [vp, ctypes.c_double, ctypes.c_int, ctypes.c_double]
        L.orc_lawrk_gen.restype = ctypes.c_int
        L.orc_lawrk_step.argtypes = [vp, vp, vp, ctypes.c_int]
        L.orc_lawrk_step.restype = ctypes.c_int
        L.orc_expab_gen.argtypes = [vp, ctypes.c_double, ctypes.c_int, ctypes.c_double]
        L.orc_expab_gen.restype = ctypes.c_int
        L.orc_expab_step.argtypes = [vp, vp, vp, ctypes.c_int]
        L.orc_expab_step.restype = ctypes.c_int
        L.orc_chkconv.argtypes = [vp, vp, vp, vp, vp]
        L.orc_chkconv.restype = None
        L.orc_band_chol.argtypes = [ctypes.c_int, ctypes.c_int, vp]
        L.orc_band_chol.restype = ctypes.c_int
        L.orc_band_chol_solve.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp]
        L.orc_herm_eig.argtypes = [ctypes.c_int, vp, vp, vp]
        L.orc_ci_sigma.argtypes = [vp, vp, vp, vp, vp]
        L.orc_ci_mkden.argtypes = [vp, vp]
        L.orc_ci_layout.argtypes = [vp, vp, vp, vp, vp, vp]
        L.orc_ci_layout.restype = ctypes.c_int
        L.orc_nstr.argtypes = [vp, ctypes.c_int]
        L.orc_nstr.restype = ctypes.c_int
        L.orc_band_lu.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp]
        L.orc_band_lu.restype = ctypes.c_int
        L.orc_band_lu_solve.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp, vp]
        L.orc_h1itr_gen.argtypes = [vp, ctypes.c_double]
        L.orc_h1itr_gen.restype = ctypes.c_int
        L.orc_h1itr_prop.argtypes = [vp, vp, ctypes.c_double, ctypes.c_int, ctypes.c_double, vp]
        L.orc_h1itr_prop.restype = ctypes.c_int
        L.orc_split2_step.argtypes = [vp, ctypes.c_int, ctypes.c_double, vp, vp, vp, ctypes.c_int,
                                      ctypes.c_double, vp, ctypes.c_int, vp]
        L.orc_split2_step.restype = ctypes.c_int
        if openmp:
            _lib_omp = L
        else:
            _lib = L
    return _lib_omp if openmp else _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    """FP64 restatement of td1c's hprod for one problem."""

    def __init__(self, prob: Problem, openmp: bool = False):
        self.prob = prob
        self._L = lib(openmp)
        c, self._keep = prob.to_c()
        self._c = c
        self.h = self._L.orc_create(ctypes.byref(c))
        if not self.h:
            raise RuntimeError("oracle set-up failed")
        s = np.zeros(10, dtype=np.int32)
        self._L.orc_sizes(self.h, _p(s))
        (self.nrad, self.nbas, self.nfun, self.lcic, self.nlat, self.L1, self.L2,
         self.nr, self.mmax1, self.mmax2) = [int(v) for v in s]

    def __del__(self):
        if getattr(self, "h", None):
            self._L.orc_destroy(self.h)
            self.h = None

    @property
    def norb(self) -> int:
        return self.nbas * self.nfun

    @property
    def wfn_size(self) -> int:
        return self.norb + self.lcic

    def set_wfn0(self, wfn: np.ndarray, v2xfc: Optional[np.ndarray] = None):
        wfn = np.ascontiguousarray(wfn, dtype=np.complex128)
        orb, cic = wfn[: self.norb], np.ascontiguousarray(wfn[self.norb:])
        if cic.size == 0:
            cic = np.zeros(self.lcic, np.complex128)
        vx = None if v2xfc is None else np.ascontiguousarray(v2xfc, dtype=np.complex128)
        self._L.orc_set_wfn0(self.h, _p(orb), _p(cic), None if vx is None else _p(vx))

    def htot(self, dtime: float, lfield, wfn: np.ndarray) -> np.ndarray:
        """Returns hwfn (orbitals followed by CI) for a full wave function."""
        wfn = np.ascontiguousarray(wfn, dtype=np.complex128)
        lf = np.ascontiguousarray(lfield, dtype=np.float64)
        out = np.zeros(self.wfn_size, np.complex128)
        orb = wfn[: self.norb]
        cic = np.ascontiguousarray(wfn[self.norb:])
        hw = out[: self.norb]
        hc = np.zeros(self.lcic, np.complex128)
        self._L.orc_htot(self.h, dtime, _p(lf), _p(orb), _p(cic), _p(hw), _p(hc))
        out[self.norb:] = hc
        return out

    def rk_prop(self, order: int, dtime: float, lfield_stages: np.ndarray, wfn: np.ndarray,
                apply_mask: bool = True) -> np.ndarray:
        w = np.array(wfn, dtype=np.complex128, copy=True)
        lf = np.ascontiguousarray(lfield_stages, dtype=np.float64)
        rc = self._L.orc_rk_prop(self.h, order, dtime, _p(lf), _p(w), int(apply_mask))
        if rc:
            raise ValueError("bad rk_order")
        return w

    def h1itr_gen(self, dt_half: float):
        """clh1itr::gen: LU factors of 1 + i T dt_half/2 per l (dt_half is the
        reference's dt2 = dt/2)."""
        rc = self._L.orc_h1itr_gen(self.h, dt_half)
        if rc:
            raise RuntimeError(f"h1itr_gen: singular factor at l = {-rc - 1}")

    def h1itr_prop(self, lfield, dt_half: float, wfn: np.ndarray, maxcyc: int = 20, thresh: float = 1e-10):
        """h1itr_prop on a copy of wfn; returns (wfn, cycles)."""
        w = np.array(wfn, dtype=np.complex128, copy=True)
        lf = np.ascontiguousarray(lfield, dtype=np.float64)
        n = self._L.orc_h1itr_prop(self.h, _p(lf), dt_half, maxcyc, thresh, _p(w))
        return w, n

    def split2_step(self, order: int, dtime: float, lf_quarter, lfield_stages, lf_three_quarter,
                    wfn: np.ndarray, maxcyc: int = 20, thresh: float = 1e-10, apply_mask: bool = True):
        """One tdse_split2 step (h1itr half step, RK step of the split hprod,
        h1itr half step, mask) on a copy of wfn; returns (wfn, (cycles1, cycles2))."""
        w = np.array(wfn, dtype=np.complex128, copy=True)
        q1 = np.ascontiguousarray(lf_quarter, dtype=np.float64)
        q3 = np.ascontiguousarray(lf_three_quarter, dtype=np.float64)
        st = np.ascontiguousarray(lfield_stages, dtype=np.float64)
        nc = np.zeros(2, np.int32)
        rc = self._L.orc_split2_step(self.h, order, dtime, _p(q1), _p(st), _p(q3), maxcyc, thresh, _p(w),
                                   int(apply_mask), _p(nc))
        if rc:
            raise ValueError("bad rk_order")
        return w, (int(nc[0]), int(nc[1]))

    def mask(self, wfn: np.ndarray) -> np.ndarray:
        w = np.array(wfn, dtype=np.complex128, copy=True)
        self._L.orc_mask(self.h, _p(w))
        return w

    def dipole(self, lfield, wfn: np.ndarray) -> float:
        """Length-form dipole <z> of a wave function (hprod_dipole dip(1))."""
        w = np.ascontiguousarray(wfn, dtype=np.complex128)
        lf = np.ascontiguousarray(lfield, dtype=np.float64)
        return float(self._L.orc_dipole(self.h, _p(lf), _p(w[: self.norb]), _p(w[self.norb:])))

    def dipole3(self, lfield, wfn: np.ndarray) -> np.ndarray:
        """hprod_dipole's Ehrenfest terms [dip(1), vel(1), acc(1)]: length,
        velocity and acceleration forms (frozen-core corrections excluded)."""
        w = np.ascontiguousarray(wfn, dtype=np.complex128)
        lf = np.ascontiguousarray(lfield, dtype=np.float64)
        out = np.zeros(3)
        self._L.orc_dipole3(self.h, _p(lf), _p(w[: self.norb]), _p(w[self.norb:]), _p(out))
        return out

    def dipole_all(self, lfield, wfn: np.ndarray) -> np.ndarray:
        """hprod_dipole with the frozen-core corrections: [dip(1..3),
        vel(1..3), acc(1..3)], (2) the corrections, (3) = (1) + (2)."""
        w = np.ascontiguousarray(wfn, dtype=np.complex128)
        lf = np.ascontiguousarray(lfield, dtype=np.float64)
        out = np.zeros(9)
        self._L.orc_dipole_all(self.h, _p(lf), _p(w[: self.norb]), _p(w[self.norb:]), _p(out))
        return out

    def ipx(self, max_ipx: int, rad_ipx: float, lfield, wfn: np.ndarray) -> np.ndarray:
        """hprod_ipx: probabilities that exactly k = 0..max_ipx electrons are
        outside rad_ipx."""
        w = np.ascontiguousarray(wfn, dtype=np.complex128)
        lf = np.ascontiguousarray(lfield, dtype=np.float64)
        out = np.zeros(max_ipx + 1)
        cic = np.ascontiguousarray(w[self.norb:])
        if cic.size == 0:
            cic = np.zeros(max(1, self.lcic), np.complex128)
        rc = self._L.orc_ipx(self.h, max_ipx, rad_ipx, _p(lf), _p(w[: self.norb]), _p(cic), _p(out))
        if rc:
            raise ValueError("bad max_ipx")
        return out

    def ipd(self, max_ipd: int, rad_ipx: float, lfield, wfn: np.ndarray) -> np.ndarray:
        """hprod_ipd: the 1-/2-RDM estimate of the probabilities that k = 0..max_ipd
        (at most 2) electrons are outside rad_ipx."""
        w = np.ascontiguousarray(wfn, dtype=np.complex128)
        lf = np.ascontiguousarray(lfield, dtype=np.float64)
        out = np.zeros(max_ipd + 1)
        cic = np.ascontiguousarray(w[self.norb:])
        if cic.size == 0:
            cic = np.zeros(max(1, self.lcic), np.complex128)
        rc = self._L.orc_ipd(self.h, max_ipd, rad_ipx, _p(lf), _p(w[: self.norb]), _p(cic), _p(out))
        if rc:
            raise ValueError("bad max_ipd")
        return out

    def orth(self, wfn: np.ndarray) -> np.ndarray:
        """clwfn::orth (schmidt): CI normalised, dynamical orbitals by modified
        Gram-Schmidt; returns a new wave function."""
        w = np.array(wfn, dtype=np.complex128, copy=True)
        cic = w[self.norb:]
        self._L.orc_orth(self.h, _p(w[: self.norb]), _p(cic) if cic.size else None)
        return w

    # the rational functions of cletdrb, in orc_etdrb_gen's order
    ETD_RATS = ("phi1dt1", "phi1dt2", "phid1dt1", "phid1dt2", "phi2dt1", "phi2dt2", "phi3dt1")

    def etdrb_gen(self, dtime: float, rosenbrock: bool = True, maxcyc: int = 20, thresh: float = 1e-10):
        """cletdrb::gen for step dtime (the band factors use tmat as it is, so
        call after set_wfn0 with frozen cores)."""
        rc = self._L.orc_etdrb_gen(self.h, dtime, int(rosenbrock), maxcyc, thresh)
        if rc:
            raise RuntimeError(f"orc_etdrb_gen failed ({rc})")

    def etdrb_step(self, lfields: np.ndarray, wfn: np.ndarray, apply_mask: bool = True) -> np.ndarray:
        """cletdrb::prop4_orb then the mask; lfields: the six arrays of
        td1c_amd.etd_fields.  Returns the new wave function."""
        w = np.array(wfn, dtype=np.complex128, copy=True)
        lf = np.ascontiguousarray(lfields, dtype=np.float64)
        assert lf.size == 54
        rc = self._L.orc_etdrb_step(self.h, _p(lf), _p(w), int(apply_mask))
        if rc:
            raise RuntimeError(f"orc_etdrb_step failed ({rc})")
        return w

    # the rational functions of cletdrk4, in orc_etdrk4_gen's order
    ETK_RATS = ("exp1", "exp2", "gfun", "f0fun", "f1fun", "f2fun")

    def etd_rat_scalar(self, name: str, lam: complex) -> complex:
        """A rational function of the last etdrb / etdrk4 gen at a scalar
        eigenvalue lam of h1."""
        out = np.zeros(2)
        if name == "expab_phi":  # clexpab's form, slot 7
            idx = 7
        else:
            idx = self.ETD_RATS.index(name) if name in self.ETD_RATS else self.ETK_RATS.index(name)
        self._L.orc_etd_rat_scalar(self.h, idx, lam.real, lam.imag, _p(out))
        return complex(out[0], out[1])

    def etdrk4_gen(self, dtime: float, maxcyc: int = 20, thresh: float = 1e-10):
        """cletdrk4::gen for step dtime (defaults: the [2/2] forms)."""
        rc = self._L.orc_etdrk4_gen(self.h, dtime, maxcyc, thresh)
        if rc:
            raise RuntimeError(f"orc_etdrk4_gen failed ({rc})")

    def lawrk_gen(self, dtime: float, maxcyc: int = 20, thresh: float = 1e-10):
        """cllawrk::gen for step dtime (lawrk_order 4, lawrk_npade 11)."""
        rc = self._L.orc_lawrk_gen(self.h, dtime, maxcyc, thresh)
        if rc:
            raise RuntimeError(f"orc_lawrk_gen failed ({rc})")

    def lawrk_step(self, lfields: np.ndarray, wfn: np.ndarray, apply_mask: bool = True) -> np.ndarray:
        """cllawrk::prop4 then the mask; lfields as for etdrb_step."""
        w = np.array(wfn, dtype=np.complex128, copy=True)
        lf = np.ascontiguousarray(lfields, dtype=np.float64)
        assert lf.size == 54
        rc = self._L.orc_lawrk_step(self.h, _p(lf), _p(w), int(apply_mask))
        if rc:
            raise RuntimeError(f"orc_lawrk_step failed ({rc})")
        return w

    def chkconv(self, lfield, wfn: np.ndarray) -> np.ndarray:
        """hprod_chkconv: (energy, |hcic| / lcic, grad(2), grad(3))."""
        w = np.ascontiguousarray(wfn, dtype=np.complex128)
        lf = np.ascontiguousarray(lfield, dtype=np.float64)
        cic = np.ascontiguousarray(w[self.norb:])
        out = np.zeros(4)
        self._L.orc_chkconv(self.h, _p(lf), _p(w[: self.norb]), _p(cic), _p(out))
        return out

    def expab_gen(self, dtime: float, maxcyc: int = 20, thresh: float = 1e-10):
        """clexpab::gen for step dtime (order 4, ndpade 12); restarts the history."""
        rc = self._L.orc_expab_gen(self.h, dtime, maxcyc, thresh)
        if rc:
            raise RuntimeError(f"orc_expab_gen failed ({rc})")

    def expab_step(self, lfields: np.ndarray, wfn: np.ndarray, apply_mask: bool = True) -> np.ndarray:
        """clexpab::prop_general then the mask; lfields: lfield(t), then
        lfield at the history times timeP[1..3] (4 x 9)."""
        w = np.array(wfn, dtype=np.complex128, copy=True)
        lf = np.ascontiguousarray(lfields, dtype=np.float64)
        assert lf.size == 36
        rc = self._L.orc_expab_step(self.h, _p(lf), _p(w), int(apply_mask))
        if rc:
            raise RuntimeError(f"orc_expab_step failed ({rc})")
        return w

    def etdrk4_step(self, lfields: np.ndarray, wfn: np.ndarray, apply_mask: bool = True) -> np.ndarray:
        """cletdrk4::prop then the mask; lfields as for etdrb_step."""
        w = np.array(wfn, dtype=np.complex128, copy=True)
        lf = np.ascontiguousarray(lfields, dtype=np.float64)
        assert lf.size == 54
        rc = self._L.orc_etdrk4_step(self.h, _p(lf), _p(w), int(apply_mask))
        if rc:
            raise RuntimeError(f"orc_etdrk4_step failed ({rc})")
        return w

    def proj_to(self, other: "Oracle", wfn: np.ndarray) -> np.ndarray:
        """clbas::proj from this problem's grid to other's, with the CI copy of
        guess read_proj; returns other's wave function (not orthonormalised)."""
        w = np.ascontiguousarray(wfn, dtype=np.complex128)
        if w.size != self.wfn_size:
            raise ValueError("wave function size does not match the source problem")
        out = np.zeros(other.wfn_size, np.complex128)
        cic1 = np.ascontiguousarray(w[self.norb:])
        self._L.orc_wfn_proj(self.h, _p(w[: self.norb]), _p(cic1), other.h, _p(out[: other.norb]),
                             _p(out[other.norb:]))
        return out

    def energies(self) -> np.ndarray:
        e = np.zeros(5)
        self._L.orc_energies(self.h, _p(e))
        return e

    _SHAPES = None

    def table(self, name: str) -> np.ndarray:
        """Copy of an internal table / scratch array, flattened, Fortran order."""
        n = self.prob.ndvr
        nr, L1, L2, nl, nf = self.nr, self.L1, self.L2, self.nlat, self.nfun
        M1, M2 = 2 * self.mmax1 + 1, 2 * self.mmax2 + 1
        bw = 2 * n + 1
        na = max(self.prob.nact, 1)
        nfe = self.prob.nfe
        shapes = {
            "xrad": (self.nrad + 1, np.float64), "wrad": (self.nrad + 1, np.float64),
            "mask": (self.nrad + 1, np.float64),
            "radk": (bw * (self.nrad + 1), np.float64), "radk0": (bw * (self.nrad + 1), np.float64),
            "radp": (bw * (self.nrad + 1), np.float64),
            "cost": (nl, np.float64), "wlat": (nl, np.float64),
            "legf1": (L1 * nl * M1, np.float64), "legb1": (L1 * nl * M1, np.float64),
            "legf2": (L2 * nl * M2, np.float64), "legb2": (L2 * nl * M2, np.float64),
            "alph": (L1 * M1, np.float64),
            "tmat": (bw * nr * L1, np.complex128), "pmat": (bw * nr, np.complex128),
            "d2ll": ((n + 1) * nr * L2, np.float64),
            "zfac": (nr * L1 * M1, np.complex128), "pzfac1": (L1 * M1, np.complex128),
            "pzfac2": (nr * L1 * M1, np.complex128),
            "d2fac1": (L2, np.float64), "d2fac2": (nr * L2, np.float64),
            "d2invr": (nr * L2, np.float64), "d2rpl0": (nr * L2, np.float64),
            "d2rpl1": (nr * L2, np.float64),
            "orb": (self.nbas * nf, np.complex128), "h0orb": (self.nbas * nf, np.complex128),
            "h1orb": (self.nbas * nf, np.complex128), "gorb": (self.nbas * nf, np.complex128),
            "v2orb": (self.nbas * nf, np.complex128), "orb0": (self.nbas * nf, np.complex128),
            "orbg": (nr * nl * nf, np.complex128), "gorbg": (nr * nl * nf, np.complex128),
            "v2orbg": (nr * nl * nf, np.complex128),
            "rho2": (nr * L2 * nf * nf, np.complex128), "v2sph": (nr * L2 * nf * nf, np.complex128),
            "v2ang": (nr * nl * nf * nf, np.complex128),
            "rhofc": (nr, np.complex128), "v2jfc": (nr, np.complex128),
            "den1": (na * na, np.complex128), "rden": (na * na, np.complex128),
            "rrden": (na * na, np.complex128), "int1e": (na * na, np.complex128),
            "den2": (na ** 4, np.complex128), "int2e": (na ** 4, np.complex128),
            "xmat": (nf * nf, np.complex128), "dcic": (self.lcic, np.complex128),
            "dvr_x": (nfe * (n + 1), np.float64), "dvr_w": (nfe * (n + 1), np.float64),
            "dshape": (nfe * (n + 1) ** 2, np.float64),
        }
        size, dt = shapes[name]
        ptr = self._L.orc_table(self.h, name.encode())
        if not ptr:
            raise KeyError(name)
        buf = (ctypes.c_char * (size * np.dtype(dt).itemsize)).from_address(ptr)
        return np.frombuffer(buf, dtype=dt).copy()

    # helpers for pin tests -------------------------------------------------
    def ci_sigma(self, cic, int1e, int2e):
        cic = np.ascontiguousarray(cic, np.complex128)
        sig = np.zeros(self.lcic, np.complex128)
        i1 = np.ascontiguousarray(int1e, np.complex128).ravel(order="F")
        i2 = np.ascontiguousarray(int2e, np.complex128).ravel(order="F")
        i1 = np.ascontiguousarray(i1)
        i2 = np.ascontiguousarray(i2)
        self._L.orc_ci_sigma(self.h, _p(cic), _p(sig), _p(i1), _p(i2))
        return sig

    def ci_layout(self):
        """(alpha strings, beta strings, det_alpha, det_beta, cic0): the
        occupation bit patterns of the ordered strings, the string pair of each
        determinant and the determinant ormas_cic0 sets."""
        na, nb = self._L.orc_nstr(self.h, 0), self._L.orc_nstr(self.h, 1)
        sa = np.zeros(max(na, 1), np.uint32)
        sb = np.zeros(max(nb, 1), np.uint32)
        da = np.zeros(self.lcic, np.int32)
        db = np.zeros(self.lcic, np.int32)
        c0 = ctypes.c_int(0)
        self._L.orc_ci_layout(self.h, _p(sa), _p(sb), _p(da), _p(db), ctypes.byref(c0))
        return sa[:na], sb[:nb], da, db, c0.value

    def ci_mkden(self, cic):
        cic = np.ascontiguousarray(cic, np.complex128)
        self._L.orc_ci_mkden(self.h, _p(cic))
        na = self.prob.nact
        d1 = self.table("den1").reshape((na, na), order="F")
        d2 = self.table("den2").reshape((na,) * 4, order="F")
        return d1, d2


def band_chol(ab: np.ndarray, kd: int) -> int:
    """In-place LAPACK-'L' band Cholesky of ab[(kd+1)*j + (i-j)]."""
    n = ab.size // (kd + 1)
    return lib().orc_band_chol(n, kd, _p(ab))


def band_chol_solve(ab: np.ndarray, kd: int, x: np.ndarray) -> None:
    n = ab.size // (kd + 1)
    lib().orc_band_chol_solve(n, kd, _p(ab), _p(x))


def band_lu(ab: np.ndarray, kd: int):
    """In-place zgbtrf (kl = ku = kd) of ab[(3kd+1)*j + row]; returns (info, ipiv)."""
    n = ab.size // (3 * kd + 1)
    ipiv = np.zeros(n, np.int32)
    info = lib().orc_band_lu(n, kd, _p(ab), _p(ipiv))
    return info, ipiv


def band_lu_solve(ab: np.ndarray, kd: int, ipiv: np.ndarray, b: np.ndarray) -> None:
    n = ab.size // (3 * kd + 1)
    lib().orc_band_lu_solve(n, kd, _p(ab), _p(ipiv), _p(b))


def herm_eig(a: np.ndarray):
    n = a.shape[0]
    af = np.ascontiguousarray(a.ravel(order="F"), np.complex128)
    w = np.zeros(n)
    u = np.zeros(n * n, np.complex128)
    lib().orc_herm_eig(n, _p(af), _p(w), _p(u))
    return w, u.reshape((n, n), order="F")
