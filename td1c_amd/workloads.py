"""Workload definitions and seeded synthetic inputs.

Workloads (sizes as quoted for the hprod / RK4 benchmark):
  he_l15   : He, 40 elements x 21 GLL points on [0,40] (799 radial points),
             lmax1 15 / lmax2 30, Z=2, one doubly occupied orbital (TDHF)
  ne_l31   : Ne, 80 x 21 on [0,160] (1599 points), lmax1 31 / lmax2 62, Z=10,
             1 frozen core (type -1) + 4 active orbitals, 800 nm sin^2 pulse,
             velocity gauge, dt 0.01
  ar_l63   : Ar, 120 x 21 on [0,240] (2399 points), lmax1 63 / lmax2 126,
             Z=18, 9 doubly occupied orbitals
  mf16_l63 : 16 doubly occupied random orbitals, 100 x 21 on [0,200],
             lmax1 63 / lmax2 126 (136 pair densities)
Inputs are seeded (numpy PCG64, seed = index*1000 + 7): smooth radial
shapes r^p exp(-zeta r) times complex l-coefficients with |c_l| ~ 2^-l,
orthonormalised; frozen cores are pure hydrogenic 1s functions.  For
propagations at the production step (dt = 0.01 with split_type = 1),
bound_wfn gives each l component a nodeless hydrogenic shape, r^(l+1) at
the origin as a bound orbital (make_wfn's r^p for every l puts large
one-body matrix elements <phi_j|h0|phi_i> into the split hprod and makes
it stiff).
"""
from __future__ import annotations

import math
from typing import Dict, Tuple

import numpy as np

from .abi import Field, Problem


def gll_grid(prob: Problem) -> Tuple[np.ndarray, np.ndarray]:
    """Interior FEDVR points r_i and bridged weights w_i (numpy restatement
    for input synthesis only)."""
    n = prob.ndvr
    p = np.polynomial.legendre.Legendre.basis(n)
    inner = np.sort(p.deriv().roots().real)
    x = np.concatenate([[-1.0], inner, [1.0]])
    pn = p(x)
    w = 2.0 / (n * (n + 1) * pn * pn)
    e = prob.element_edges()
    xs, ws = [], []
    for i in range(prob.nfe):
        half, mid = 0.5 * (e[i + 1] - e[i]), 0.5 * (e[i + 1] + e[i])
        xs.append(x * half + mid)
        ws.append(w * half)
    nrad = n * prob.nfe
    xr = np.zeros(nrad + 1)
    wr = np.zeros(nrad + 1)
    for i in range(prob.nfe):
        xr[i * n:i * n + n + 1] = xs[i]
        wr[i * n:i * n + n] += ws[i][:n]
        wr[i * n + n] += ws[i][n]
    return xr[1:nrad], wr[1:nrad]


def hydrogenic_1s(prob: Problem, z: float) -> np.ndarray:
    """FEDVR coefficients (l = 0 block) of u(r) = 2 z^1.5 r exp(-z r)."""
    r, w = gll_grid(prob)
    return 2.0 * z ** 1.5 * r * np.exp(-z * r) * np.sqrt(w)


def synth_orbitals(prob: Problem, seed: int, fc_z: float = 0.0) -> np.ndarray:
    """Seeded orthonormal orbitals, shape (nfun, L1, nr)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    r, w = gll_grid(prob)
    nr, L1, nf = prob.nr, prob.lmax1 + 1, prob.nfun
    orbs = np.zeros((nf, L1, nr), np.complex128)
    mval = list(prob.mval) if prob.mval is not None else [0] * nf
    for f in range(nf):
        if f < prob.nfcore:
            lv = prob.lval[f] if prob.lval is not None else 0
            orbs[f, lv] = hydrogenic_1s(prob, fc_z if fc_z > 0 else prob.znuc)
            continue
        k = f - prob.nfcore
        zeta = max(0.5, prob.znuc / (1.0 + 0.75 * k)) if prob.znuc > 0 else 1.0
        pw = 1 + (k % 3)
        rad = r ** pw * np.exp(-zeta * r) * np.sqrt(w)
        am = abs(mval[f])
        c = (rng.standard_normal(L1) + 1j * rng.standard_normal(L1)) * (2.0 ** -np.arange(L1))
        c[:am] = 0.0
        orbs[f] = c[:, None] * rad[None, :]
    # Gram-Schmidt within equal-m blocks (frozen cores kept as they are)
    for f in range(nf):
        v = orbs[f].ravel()
        if f >= prob.nfcore:
            for g in range(f):
                if mval[g] != mval[f]:
                    continue
                u = orbs[g].ravel()
                v = v - u * np.vdot(u, v)
        orbs[f] = (v / np.linalg.norm(v)).reshape(L1, nr)
    return orbs


def he_l15(**kw) -> Problem:
    return Problem(nfe=40, ndvr=20, rmax=40.0, lmax1=15, lmax2=30, znuc=2.0, ndcore=1,
                   nel_alpha=1, nel_beta=1, **kw)


def ne_l31(**kw) -> Problem:
    kw.setdefault("igauge", 1)
    return Problem(nfe=80, ndvr=20, rmax=160.0, lmax1=31, lmax2=62, znuc=10.0, nfcore1=1, nact=4,
                   nel_alpha=5, nel_beta=5, lval=[0, 0, 0, 0, 0], **kw)


def ne_l31_cas(**kw) -> Problem:
    """The Ne grid with a correlated active space: frozen 1s + CAS(8e, 6o)
    (225 determinants), so that the RDMs and the CI derivative are not those
    of one determinant (the 4-orbital Ne case fills its active space)."""
    kw.setdefault("igauge", 1)
    return Problem(nfe=80, ndvr=20, rmax=160.0, lmax1=31, lmax2=62, znuc=10.0, nfcore1=1, nact=6,
                   nel_alpha=5, nel_beta=5, lval=[0] * 7, **kw)


def ar_l63(**kw) -> Problem:
    return Problem(nfe=120, ndvr=20, rmax=240.0, lmax1=63, lmax2=126, znuc=18.0, ndcore=9,
                   nel_alpha=9, nel_beta=9, **kw)


def mf16_l63(**kw) -> Problem:
    return Problem(nfe=100, ndvr=20, rmax=200.0, lmax1=63, lmax2=126, znuc=16.0, ndcore=16,
                   nel_alpha=16, nel_beta=16, **kw)


WORKLOADS = {"he_l15": (he_l15, 1), "ne_l31": (ne_l31, 2), "ar_l63": (ar_l63, 3), "mf16_l63": (mf16_l63, 5)}


def he_scan(n: int = 64, seed: int = 4007, **kw):
    """The scan workload: n He TDHF problems (one doubly occupied orbital,
    Z = 2) on grids of nfe elements x 21 points, nfe uniform in [40, 120],
    rmax = 2 nfe, lmax1 uniform in {15, 31, 47, 63}, lmax2 = 2 lmax1;
    velocity-gauge sin^2 pulses of 4 cycles, wavelengths uniform in
    [400, 1600] nm, intensities log-spaced over 1e13 - 1e15 W/cm^2 (in
    member order), carrier-envelope phases uniform in [0, 360) degrees.  Returns
    (problems, fields); kw goes to every Problem (precision, isplit, ...)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    kw.setdefault("igauge", 1)
    lmaxes = (15, 31, 47, 63)
    ints = np.logspace(13.0, 15.0, n) if n > 1 else np.array([1.0e14])
    probs, fields = [], []
    for i in range(n):
        nfe = int(rng.integers(40, 121))
        l1 = lmaxes[int(rng.integers(0, 4))]
        probs.append(Problem(nfe=nfe, ndvr=20, rmax=2.0 * nfe, lmax1=l1, lmax2=2 * l1, znuc=2.0, ndcore=1,
                             nel_alpha=1, nel_beta=1, **kw))
        fields.append(Field(env_type="sin2", gauge="velocity1", fint=float(ints[i]),
                            wlen=float(rng.uniform(400.0, 1600.0)), cep=float(rng.uniform(0.0, 360.0)),
                            numcyc=4.0))
    return probs, fields


def ne_field() -> Field:
    """800 nm, 4-cycle sin^2 pulse at 1e14 W/cm^2, velocity gauge."""
    return Field(env_type="sin2", gauge="velocity1", fint=1.0e14, wlen=800.0, numcyc=4.0)


def bound_orbitals(prob: Problem, seed: int) -> np.ndarray:
    """Seeded orthonormal orbitals (nfun, L1, nr) whose l components are
    c_l u_l(r), u_l = r^(l+1) exp(-zeta r) normalised on the grid (|c_l| ~
    2^-l, zeta = Z / (1 + 0.75 k) for the k-th dynamical orbital): each
    component lives where the orbital does, near r = (l+1) / zeta, and goes
    as r^(l+1) at the origin, so the one-body operator stays of order
    zeta^2 on them; frozen cores as in synth_orbitals."""
    rng = np.random.Generator(np.random.PCG64(seed))
    r, w = gll_grid(prob)
    nr, L1, nf = prob.nr, prob.lmax1 + 1, prob.nfun
    orbs = synth_orbitals(prob, seed)
    mval = list(prob.mval) if prob.mval is not None else [0] * nf
    ls = np.arange(L1)[:, None]
    for f in range(prob.nfcore, nf):
        k = f - prob.nfcore
        zeta = max(0.5, prob.znuc / (1.0 + 0.75 * k)) if prob.znuc > 0 else 1.0
        c = (rng.standard_normal(L1) + 1j * rng.standard_normal(L1)) * (2.0 ** -np.arange(L1))
        c[:abs(mval[f])] = 0.0
        # r^(l+1) exp(-zeta r) as exp((l+1) log (r / rp) - zeta (r - rp)), rp = (l+1) / zeta
        # (1 at the peak: no overflow at l = 63), then normalised per l
        rp = (ls + 1) / zeta
        rad = np.exp((ls + 1) * np.log(r[None, :] / rp) - zeta * (r[None, :] - rp)) * np.sqrt(w)[None, :]
        rad /= np.linalg.norm(rad, axis=1, keepdims=True)
        orbs[f] = c[:, None] * rad
    for f in range(prob.nfcore, nf):
        v = orbs[f].ravel()
        for g in range(f):
            if mval[g] == mval[f]:
                u = orbs[g].ravel()
                v = v - u * np.vdot(u, v)
        orbs[f] = (v / np.linalg.norm(v)).reshape(L1, nr)
    return orbs


def bound_wfn(prob: Problem, seed: int, lcic: int = 1) -> np.ndarray:
    """bound_orbitals followed by a seeded normalised complex CI vector (1
    for a single determinant)."""
    orbs = bound_orbitals(prob, seed)
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    cic = rng.standard_normal(lcic) + 1j * rng.standard_normal(lcic)
    if lcic == 1:
        cic[:] = 1.0
    cic /= np.linalg.norm(cic)
    return np.concatenate([orbs.ravel(), cic])


def make_wfn(prob: Problem, seed: int, lcic: int = 1) -> np.ndarray:
    """Full wave function: orbitals then CI (first determinant = 1)."""
    orbs = synth_orbitals(prob, seed)
    cic = np.zeros(lcic, np.complex128)
    cic[0] = 1.0
    return np.concatenate([orbs.ravel(), cic])


def he_l15_wfn(prob: Problem, seed: int = 1007) -> np.ndarray:
    """He orbital r exp(-r) times seeded per-l complex Gaussian coefficients
    with |c_l| ~ 2^-l, normalised."""
    rng = np.random.Generator(np.random.PCG64(seed))
    r, w = gll_grid(prob)
    L1 = prob.lmax1 + 1
    c = (rng.standard_normal(L1) + 1j * rng.standard_normal(L1)) * (2.0 ** -np.arange(L1))
    orb = c[:, None] * (r * np.exp(-r) * np.sqrt(w))[None, :]
    orb /= np.linalg.norm(orb)
    return np.concatenate([orb.ravel(), np.ones(1, np.complex128)])
