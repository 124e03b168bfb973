"""Small problems shared by the parity tests (sizes the oracle finishes in
well under a second)."""
from __future__ import annotations

import numpy as np

from td1c_amd.abi import Problem
from td1c_amd import workloads as W


def he_small(**kw):
    return Problem(nfe=10, ndvr=20, rmax=30.0, lmax1=5, lmax2=10, znuc=2.0, ndcore=1,
                   nel_alpha=1, nel_beta=1, **kw)


def dc3_small(**kw):
    # three doubly occupied orbitals, one of them m = 1 paired with m = -1
    kw.setdefault("mval", [0, 1, -1])
    return Problem(nfe=8, ndvr=20, rmax=24.0, lmax1=6, lmax2=12, znuc=6.0, ndcore=3,
                   nel_alpha=3, nel_beta=3, **kw)


def ne_small(**kw):
    kw.setdefault("igauge", 1)
    return Problem(nfe=10, ndvr=20, rmax=30.0, lmax1=7, lmax2=14, znuc=10.0, nfcore1=1, nact=4,
                   nel_alpha=5, nel_beta=5, lval=[0, 0, 0, 0, 0], **kw)


def cas_small(**kw):
    # 1 dynamical core + CAS(2e, 2o): 4 determinants, D not idempotent
    return Problem(nfe=8, ndvr=20, rmax=24.0, lmax1=5, lmax2=10, znuc=4.0, ndcore=1, nact=2,
                   nel_alpha=2, nel_beta=2, **kw)


def cas_fc_small(**kw):
    # frozen core + 1 dynamical core + CAS(2e, 3o)
    kw.setdefault("igauge", 1)
    return Problem(nfe=8, ndvr=20, rmax=24.0, lmax1=5, lmax2=10, znuc=6.0, nfcore1=1, ndcore=1, nact=3,
                   nel_alpha=3, nel_beta=3, lval=[0, 0, 0, 0, 0], **kw)


def he_n16(**kw):
    # ndvr = 16 runs the band one-body kernel instead of the element kernel
    return Problem(nfe=10, ndvr=16, rmax=30.0, lmax1=5, lmax2=10, znuc=2.0, ndcore=1,
                   nel_alpha=1, nel_beta=1, **kw)


def he_long(**kw):
    # 140 elements (2799 radial rows): a radial row beyond the fused Poisson
    # kernel's LDS, so the bridge + element kernel pair runs
    return Problem(nfe=140, ndvr=20, rmax=280.0, lmax1=2, lmax2=4, znuc=2.0, ndcore=1,
                   nel_alpha=1, nel_beta=1, **kw)


def dc4_small(**kw):
    # four doubly occupied orbitals: the generic theta-grid mean-field kernel
    return Problem(nfe=6, ndvr=20, rmax=18.0, lmax1=5, lmax2=10, znuc=8.0, ndcore=4,
                   nel_alpha=4, nel_beta=4, **kw)


def fc2_small(**kw):
    # one type -2 frozen core (exchange only through v2xfc) + two dynamical cores
    kw.setdefault("igauge", 1)
    return Problem(nfe=6, ndvr=20, rmax=18.0, lmax1=5, lmax2=10, znuc=6.0, nfcore2=1, ndcore=2,
                   nel_alpha=3, nel_beta=3, lval=[0, 0, 0], **kw)


def wfn_for(prob: Problem, lcic: int, seed: int = 11) -> np.ndarray:
    orbs = W.synth_orbitals(prob, seed)
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    cic = rng.standard_normal(lcic) + 1j * rng.standard_normal(lcic)
    if lcic == 1:
        cic[:] = 1.0
    cic /= np.linalg.norm(cic)
    return np.concatenate([orbs.ravel(), cic])


LF_LEN = np.array([0.0, 0.0, 0.05, 0.0, 0.0, 0.05, 0.0, 0.0, 0.0])
LF_VEL = np.array([0.0, 0.0, 0.3, 0.0, 0.0, 0.04, 0.0, 0.0, 0.3])


def he_itr(**kw):
    # He for imaginary-time relaxation (icomp = 0): a grid coarse enough for an
    # explicit RK4 step of 4e-3
    kw.setdefault("icomp", 0)
    return Problem(nfe=12, ndvr=10, rmax=30.0, lmax1=1, lmax2=2, znuc=2.0, ndcore=1, nel_alpha=1, nel_beta=1, **kw)


def cas6_small(**kw):
    # 1 dynamical core + CAS(4e, 6o) with m = 0, +-1 active orbitals (M = 0 determinants only)
    kw.setdefault("mval", [0, 0, 0, 1, -1, 0, 0])
    return Problem(nfe=6, ndvr=20, rmax=18.0, lmax1=4, lmax2=8, znuc=6.0, ndcore=1, nact=6,
                   nel_alpha=3, nel_beta=3, **kw)


def cas7_small(**kw):
    # CAS(6e, 7o), no core: 1225 determinants, the CI work in global memory
    # (beyond the island's LDS budget)
    return Problem(nfe=6, ndvr=20, rmax=18.0, lmax1=4, lmax2=8, znuc=6.0, ndcore=0, nact=7,
                   nel_alpha=3, nel_beta=3, **kw)


def cas8_small(**kw):
    # CAS(6e, 8o), no core: 3136 determinants and 36 active pairs (1296 int2e
    # Gram entries), the largest active space of the theta-grid mean-field kernel
    return Problem(nfe=6, ndvr=20, rmax=18.0, lmax1=4, lmax2=8, znuc=6.0, ndcore=0, nact=8,
                   nel_alpha=3, nel_beta=3, **kw)


def he_scan_max(**kw):
    # the largest He problem of the scan workload: 120 x 21 on [0, 240]
    # (2399 points), lmax1 63 / lmax2 126
    kw.setdefault("igauge", 1)
    return Problem(nfe=120, ndvr=20, rmax=240.0, lmax1=63, lmax2=126, znuc=2.0, ndcore=1,
                   nel_alpha=1, nel_beta=1, **kw)


def smooth_wfn_for(prob: Problem, lcic: int, seed: int = 11) -> np.ndarray:
    """workloads.bound_wfn: every l component of the dynamical orbitals goes
    as r^(l+1) at the origin, as a bound orbital does.  The one-body
    operator is then bounded on the span of the orbitals, so the split hprod
    (whose projector terms carry <phi_j|h0|phi_i>) is not stiff and an
    explicit RK4 step of dt = 0.01 is stable; wfn_for's r^p exp(-zeta r) for
    every l is not."""
    return W.bound_wfn(prob, seed, lcic)
