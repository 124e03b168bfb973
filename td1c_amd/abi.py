"""ctypes view of the C-ABI declared in include/td1c_hprod.h.

The structs mirror the header field for field; `Problem` is the Python-side
description of one td1c problem (the values the reference reads from its
input deck: grid card, lmax1/lmax2, orbital card, control flags), and
`Problem.to_c()` produces the `td1c_problem` the library consumes.
"""
from __future__ import annotations

import ctypes
import dataclasses
from typing import List, Optional, Sequence

import numpy as np

c_int = ctypes.c_int
c_double = ctypes.c_double
c_int_p = ctypes.POINTER(ctypes.c_int)
c_double_p = ctypes.POINTER(ctypes.c_double)


class TD1CProblem(ctypes.Structure):
    _fields_ = [
        ("nfe", c_int), ("ndvr", c_int), ("edges", c_double_p),
        ("lmax1", c_int), ("lmax2", c_int), ("mmax1", c_int), ("mmax2", c_int),
        ("nlat", c_int), ("znuc", c_double),
        ("nfcore2", c_int), ("nfcore1", c_int), ("ndcore", c_int), ("nact", c_int),
        ("mval", c_int_p), ("lval", c_int_p),
        ("nel_alpha", c_int), ("nel_beta", c_int), ("mtot", c_int), ("nradfc", c_int),
        ("icomp", c_int), ("igauge", c_int), ("isplit", c_int), ("ioorot", c_int),
        ("iprojfc", c_int), ("jfc_implicit", c_int), ("xfc_implicit", c_int),
        ("throcc1", c_double), ("throcc2", c_double),
        ("rmask", c_double), ("mask_cos_order", c_double),
        ("precision", c_int), ("lconst", c_int),
    ]


class TD1CSizes(ctypes.Structure):
    _fields_ = [
        ("nrad", c_int), ("nbas", c_int), ("nfun", c_int), ("lcic", c_int),
        ("nlat", c_int), ("lmax1", c_int), ("lmax2", c_int), ("npair", c_int),
        ("wfn_size", ctypes.c_longlong),
    ]


class TD1CField(ctypes.Structure):
    _fields_ = [
        ("env_type", c_int), ("gauge", c_int), ("fint", c_double), ("wlen", c_double),
        ("cep", c_double), ("numcyc", c_double), ("delay", c_double),
    ]


@dataclasses.dataclass
class Problem:
    """One td1c problem: FEDVR grid, angular cut-offs, orbital classes, flags.

    Orbitals are ordered as in ORMAS (src/ormas/ormas_info.F90:18-44):
    frozen core type -2, frozen core type -1, dynamical core, active.
    """

    nfe: int
    ndvr: int
    rmax: float
    lmax1: int
    lmax2: int
    znuc: float
    nfcore2: int = 0
    nfcore1: int = 0
    ndcore: int = 0
    nact: int = 0
    nel_alpha: int = 1
    nel_beta: int = 1
    mval: Optional[Sequence[int]] = None
    lval: Optional[Sequence[int]] = None
    rmin: float = 0.0
    edges: Optional[Sequence[float]] = None
    nlat: int = 0
    mmax1: int = -1
    mmax2: int = -1
    mtot: int = 0
    nradfc: int = 0
    icomp: int = 1
    igauge: int = 0
    isplit: int = 0
    ioorot: int = 0
    iprojfc: int = 2
    jfc_implicit: int = 1
    xfc_implicit: int = 1
    throcc1: float = 1.0e-15
    throcc2: float = 1.0e-15
    rmask: float = 0.0
    mask_cos_order: float = 4.0
    precision: int = 0
    lconst: int = 0          # 0 off, 1 lconst (all orbitals), 2 lconst_core (cores)

    @property
    def nfun(self) -> int:
        return self.nfcore2 + self.nfcore1 + self.ndcore + self.nact

    @property
    def nfcore(self) -> int:
        return self.nfcore2 + self.nfcore1

    @property
    def ncore(self) -> int:
        return self.nfcore + self.ndcore

    @property
    def nr(self) -> int:
        return self.nfe * self.ndvr - 1

    @property
    def nbas(self) -> int:
        return self.nr * (self.lmax1 + 1)

    def element_edges(self) -> np.ndarray:
        if self.edges is not None:
            e = np.asarray(self.edges, dtype=np.float64)
            if e.shape != (self.nfe + 1,):
                raise ValueError("edges must hold nfe+1 values")
            return e
        # clfedvr::read_x: x0 + dfe*ife with dfe = (x1-x0)/nfe
        dfe = (self.rmax - self.rmin) / self.nfe
        e = np.array([self.rmin + dfe * i for i in range(self.nfe)] + [0.0])
        e[self.nfe] = e[self.nfe - 1] + dfe
        return e

    def to_c(self):
        """Returns (TD1CProblem, keepalive) -- keep the second alive while the
        struct is in use, it owns the arrays the struct points to."""
        nf = self.nfun
        mval = list(self.mval) if self.mval is not None else [0] * nf
        lval = list(self.lval) if self.lval is not None else [0] * nf
        if len(mval) != nf or len(lval) != nf:
            raise ValueError("mval/lval must have nfun entries")
        edges = np.ascontiguousarray(self.element_edges())
        m_arr = (c_int * nf)(*mval)
        l_arr = (c_int * nf)(*lval)
        p = TD1CProblem()
        p.nfe, p.ndvr = self.nfe, self.ndvr
        p.edges = edges.ctypes.data_as(c_double_p)
        p.lmax1, p.lmax2 = self.lmax1, self.lmax2
        p.mmax1, p.mmax2, p.nlat = self.mmax1, self.mmax2, self.nlat
        p.znuc = self.znuc
        p.nfcore2, p.nfcore1, p.ndcore, p.nact = self.nfcore2, self.nfcore1, self.ndcore, self.nact
        p.mval = ctypes.cast(m_arr, c_int_p)
        p.lval = ctypes.cast(l_arr, c_int_p)
        p.nel_alpha, p.nel_beta, p.mtot, p.nradfc = self.nel_alpha, self.nel_beta, self.mtot, self.nradfc
        p.icomp, p.igauge, p.isplit, p.ioorot = self.icomp, self.igauge, self.isplit, self.ioorot
        p.iprojfc, p.jfc_implicit, p.xfc_implicit = self.iprojfc, self.jfc_implicit, self.xfc_implicit
        p.throcc1, p.throcc2 = self.throcc1, self.throcc2
        p.rmask, p.mask_cos_order = self.rmask, self.mask_cos_order
        p.precision = self.precision
        p.lconst = self.lconst
        return p, (edges, m_arr, l_arr)


@dataclasses.dataclass
class Field:
    """One laser pulse (src/field/clfield1.cpp:14-116)."""

    env_type: str = "sin2"   # "zero", "sin2", "flat"
    gauge: str = "length1"   # "length1", "velocity1"
    fint: float = 0.0        # W/cm^2
    wlen: float = 800.0      # nm
    cep: float = 0.0         # degrees
    numcyc: float = 1.0
    delay: float = 0.0       # attoseconds

    def to_c(self) -> TD1CField:
        env = {"zero": 0, "sin2": 1, "flat": 2}[self.env_type]
        gauge = {"length1": 0, "velocity1": 1}[self.gauge]
        return TD1CField(env, gauge, self.fint, self.wlen, self.cep, self.numcyc, self.delay)


HEADER_FUNCTIONS: List[str] = [
    "td1c_ctx_create", "td1c_ctx_destroy", "td1c_get_sizes", "td1c_set_wfn0",
    "td1c_htot", "td1c_get_energies", "td1c_rk_prop", "td1c_mask",
    "td1c_field_value", "td1c_field_der", "td1c_field_period", "td1c_bind", "hprod_htot_",
    "td1c_last_error", "td1c_set_timing", "td1c_get_timing", "td1c_set_timing_kernel", "td1c_propagate", "td1c_dipole", "td1c_dipole3", "td1c_ipx", "td1c_ipd", "td1c_orth", "td1c_init_split2", "td1c_wfn_write", "td1c_wfn_read", "td1c_wfn_proj", "td1c_etdrb_gen", "td1c_etdrb_prop", "td1c_etdrb_propagate", "td1c_etdrk4_gen", "td1c_etdrk4_prop", "td1c_etdrk4_propagate", "td1c_lawrk_gen", "td1c_lawrk_prop", "td1c_lawrk_propagate", "td1c_expab_gen", "td1c_expab_prop", "td1c_expab_propagate", "td1c_chkconv", "td1c_init_prop12", "td1c_htotx", "td1c_get_den1", "td1c_get_den2", "td1c_get_int1e",
    "td1c_get_int2e", "td1c_get_xmat", "td1c_set_graph", "td1c_set_transform_path", "td1c_ci_layout",
    "hprod_htot_static", "td1c_h1itr_gen", "td1c_h1itr_prop", "td1c_get_shapes", "td1c_set_onebody_shape",
    "td1c_batch_create", "td1c_batch_destroy", "td1c_batch_propagate", "td1c_batch_set_graph",
    "td1c_batch_set_shapes", "td1c_batch_get_info", "td1c_batch_autotune", "td1c_dipole_all",
]
