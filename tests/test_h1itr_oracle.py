"""Pins of the oracle's split-operator one-body propagator (clh1itr:
h1itr_gen_cninv, h1itr_cninv, h1itr_prop; tdse_split2) and of the split
hprod (hprod_htot_dtorb with isplit = 1, 2).  CPU only."""
import numpy as np
import pytest
import scipy.linalg
import scipy.linalg.lapack as lapack

from td1c_amd.abi import Problem
from td1c_amd import workloads as W
from oracle import oracle as O


def _to_lapack(a, kd):
    """Dense n x n -> LAPACK gbtrf band storage with kl = ku = kd (ldab = 3kd+1)."""
    n = a.shape[0]
    ab = np.zeros((3 * kd + 1, n), np.complex128)
    for j in range(n):
        for i in range(max(0, j - kd), min(n, j + kd + 1)):
            ab[2 * kd + i - j, j] = a[i, j]
    return ab


@pytest.mark.parametrize("seed,weak", [(1, False), (2, True), (3, True)])
def test_band_lu_equals_lapack_zgbtrf(seed, weak):
    """band_lu / band_lu_solve (zgbtrf / zgbtrs, kl = ku = kd): the factor and
    the pivots equal LAPACK's zgbtrf (scipy's build) and the solution equals
    a dense solve; `weak` diagonals force row interchanges."""
    rng = np.random.default_rng(seed)
    n, kd = 40, 4
    a = np.zeros((n, n), np.complex128)
    for i in range(n):
        for j in range(max(0, i - kd), min(n, i + kd + 1)):
            a[i, j] = rng.standard_normal() + 1j * rng.standard_normal()
    if not weak:
        a += 20 * np.eye(n)
    ab = _to_lapack(a, kd)
    lu_ref, piv_ref, info_ref = lapack.zgbtrf(ab.copy(), kd, kd)
    assert info_ref == 0
    ours = np.ascontiguousarray(ab.T).ravel()  # column j at [(3kd+1) j]
    info, ipiv = O.band_lu(ours, kd)
    assert info == 0
    np.testing.assert_array_equal(ipiv, piv_ref)
    if weak:
        assert np.any(ipiv != np.arange(n))
    np.testing.assert_allclose(ours.reshape(n, 3 * kd + 1).T, lu_ref, rtol=0, atol=1e-12 * np.abs(lu_ref).max())
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x = b.copy()
    O.band_lu_solve(ours, kd, ipiv, x)
    np.testing.assert_allclose(x, np.linalg.solve(a, b), rtol=0, atol=1e-11 * np.abs(np.linalg.solve(a, b)).max())


def _small(isplit, igauge=0, **kw):
    base = dict(nfe=6, ndvr=8, rmax=12.0, lmax1=3, lmax2=6, znuc=2.0, ndcore=1, nel_alpha=1, nel_beta=1,
                isplit=isplit, igauge=igauge)
    base.update(kw)
    return Problem(**base)


def _dense_T(o, l):
    n, nr = o.prob.ndvr, o.nr
    t = o.table("tmat").reshape(o.L1, nr, 2 * n + 1)[l]
    T = np.zeros((nr, nr), np.complex128)
    for i in range(nr):
        for j in range(max(0, i - n), min(nr, i + n + 1)):
            T[i, j] = t[i, n + j - i]
    return T


def _alpha(l, m):
    return np.sqrt(((l + 1) ** 2 - m * m) / ((2 * l + 1) * (2 * l + 3)))


def _dense_H(o, zfield, m=0):
    """T + E z over all (l, r) of one m = 0 orbital: T from the bound tmat
    (pinned by the hydrogenic spectrum), z from the closed form
    <l+1,m|cos|l,m> r."""
    nr, L1 = o.nr, o.L1
    r = o.table("xrad")[1:nr + 1]
    H = np.zeros((nr * L1, nr * L1), np.complex128)
    for l in range(L1):
        H[l * nr:(l + 1) * nr, l * nr:(l + 1) * nr] = _dense_T(o, l)
    for l in range(L1 - 1):
        c = zfield * r * _alpha(l, m)
        idx = np.arange(nr)
        H[l * nr + idx, (l + 1) * nr + idx] += c
        H[(l + 1) * nr + idx, l * nr + idx] += c
    return H


def test_h1itr_isplit2_is_the_cayley_step_of_T():
    """isplit = 2: h1itr_prop = (1 + i T dt/4)^-1 (1 - i T dt/4) psi per l for
    the dynamical orbitals (dense solve), unitary."""
    prob = _small(2)
    o = O.Oracle(prob)
    w = W.make_wfn(prob, 5, o.lcic)
    dt = 0.05
    o.h1itr_gen(dt / 2)
    out, ncyc = o.h1itr_prop(np.zeros(9), dt / 2, w)
    assert ncyc == 0
    nr = o.nr
    for l in range(o.L1):
        T = _dense_T(o, l)
        I = np.eye(nr)
        x = w[l * nr:(l + 1) * nr]
        ref = np.linalg.solve(I + 0.25j * dt * T, (I - 0.25j * dt * T) @ x)
        np.testing.assert_allclose(out[l * nr:(l + 1) * nr], ref, rtol=0, atol=1e-13)
    assert abs(np.vdot(out[:o.norb], out[:o.norb]).real - np.vdot(w[:o.norb], w[:o.norb]).real) < 1e-13


@pytest.mark.parametrize("thresh", [1e-10, 1e-14])
def test_h1itr_isplit1_converges_to_the_cayley_step_of_T_plus_V(thresh):
    """isplit = 1, length gauge: the Neumann series of h1itr_prop converges to
    (1 + i H dt/4)^-1 (1 - i H dt/4) psi, H = T + E z (dense solve), to the
    series' stopping threshold, in a few cycles."""
    prob = _small(1)
    o = O.Oracle(prob)
    w = W.make_wfn(prob, 6, o.lcic)
    dt, E = 0.05, 0.3
    o.h1itr_gen(dt / 2)
    lf = np.zeros(9)
    lf[2] = lf[5] = E
    out, ncyc = o.h1itr_prop(lf, dt / 2, w, maxcyc=40, thresh=thresh)
    assert 2 <= ncyc < 40
    H = _dense_H(o, E)
    I = np.eye(H.shape[0])
    ref = np.linalg.solve(I + 0.25j * dt * H, (I - 0.25j * dt * H) @ w[:o.norb])
    err = np.linalg.norm(out[:o.norb] - ref)
    assert err < 50 * thresh + 1e-14, err


def test_h1itr_velocity_gauge_is_unitary_and_reversible():
    """isplit = 1, velocity gauge (H = T + A p_z, Hermitian; pzprod is pinned
    by the 1s-2p velocity matrix element): the step conserves the norm, and
    the step with -dt (factors from h1itr_gen(-dt/2)) undoes it, as the
    Cayley transform's inverse is the Cayley transform of -dt."""
    prob = _small(1, igauge=1)
    o = O.Oracle(prob)
    w = W.make_wfn(prob, 7, o.lcic)
    dt, A = 0.05, 0.6
    lf = np.zeros(9)
    lf[2] = lf[8] = A
    o.h1itr_gen(dt / 2)
    out, ncyc = o.h1itr_prop(lf, dt / 2, w, maxcyc=40, thresh=1e-13)
    assert 2 <= ncyc < 40
    n0 = np.vdot(w[:o.norb], w[:o.norb]).real
    assert abs(np.vdot(out[:o.norb], out[:o.norb]).real - n0) < 1e-11
    assert np.linalg.norm(out[:o.norb] - w[:o.norb]) > 1e-3
    o.h1itr_gen(-dt / 2)
    back, _ = o.h1itr_prop(lf, -dt / 2, out, maxcyc=40, thresh=1e-13)
    assert np.linalg.norm(back[:o.norb] - w[:o.norb]) < 1e-11


def test_h1itr_without_field_equals_isplit2():
    """isplit = 1 with a zero field stops after the first check of x1 = 0 and
    gives the Cayley step of T alone, as isplit = 2 does."""
    o1, o2 = O.Oracle(_small(1)), O.Oracle(_small(2))
    w = W.make_wfn(_small(1), 8, o1.lcic)
    o1.h1itr_gen(0.02)
    o2.h1itr_gen(0.02)
    a, n1 = o1.h1itr_prop(np.zeros(9), 0.02, w)
    b, _ = o2.h1itr_prop(np.zeros(9), 0.02, w)
    assert n1 == 2
    np.testing.assert_array_equal(a, b)


def test_split2_step_is_second_order():
    """orc_split2_step (tdse_split2 with clh1itr + clcrk, isplit = 1): the
    global error against a fine RK4 propagation of the unsplit hprod falls
    by ~4 when dt halves (Strang splitting), in a static length-gauge field."""
    p1 = Problem(nfe=10, ndvr=6, rmax=20.0, lmax1=2, lmax2=4, znuc=2.0, ndcore=1, nel_alpha=1, nel_beta=1,
                 isplit=1, rmask=20.0)
    p0 = Problem(nfe=10, ndvr=6, rmax=20.0, lmax1=2, lmax2=4, znuc=2.0, ndcore=1, nel_alpha=1, nel_beta=1,
                 isplit=0, rmask=20.0)
    o1, o0 = O.Oracle(p1), O.Oracle(p0)
    w = W.make_wfn(p1, 9, o1.lcic)
    T, E = 0.16, 0.05
    lf = np.zeros(9)
    lf[2] = lf[5] = E
    st = np.tile(lf, (4, 1))

    ref = w.copy()
    for _ in range(128):
        ref = o0.rk_prop(4, T / 128, st, ref, apply_mask=False)

    def run(nstep):
        x = w.copy()
        o1.h1itr_gen(T / nstep / 2)
        for _ in range(nstep):
            x, _ = o1.split2_step(4, T / nstep, lf, st, lf, x, apply_mask=False, thresh=1e-14, maxcyc=60)
        return x

    e1 = np.linalg.norm(run(8) - ref)
    e2 = np.linalg.norm(run(16) - ref)
    assert 3.0 < e1 / e2 < 5.0, (e1, e2, e1 / e2)
