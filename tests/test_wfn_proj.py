"""Grid projection of a state (clbas::proj, src/bas/clbas.cpp:280-332, with
the CI copy of guess_type "read_proj", src/main/guess.cpp:37-47).

The oracle's orc_wfn_proj is pinned by what interpolation fixes: on the same
grid the Lagrange polynomials are Kronecker deltas at the nodes, so the
projection is sqrt(w) times the coefficients bit for bit; a piecewise
polynomial of degree <= ndvr that vanishes at r = 0 and r = R (the points the
state does not store) is reproduced exactly on any other grid inside the
source box; points beyond the source box get zero.  The library's
td1c_wfn_proj (host code, no device) is compared with the oracle on random
grids, sizes and CI lengths; the GPU test follows it with clwfn::orth on the
device, as read_proj does."""
import os

import numpy as np
import pytest

import td1c_amd
from td1c_amd.abi import Problem

pytestmark = pytest.mark.skipif(not os.path.exists(td1c_amd.library_path()), reason="library not built")


def _prob(nfe, ndvr, rmax, lmax1=3, edges=None, **kw):
    base = dict(nfe=nfe, ndvr=ndvr, rmax=rmax, lmax1=lmax1, lmax2=2 * lmax1, znuc=2.0, ndcore=1,
                nel_alpha=1, nel_beta=1)
    base.update(kw)
    return Problem(edges=edges, **base)


def _edges(rng, nfe, rmax):
    w = rng.uniform(0.5, 1.5, nfe)
    e = np.concatenate([[0.0], np.cumsum(w)])
    return list(e * (rmax / e[-1]))


def _oracle(p):
    from oracle.oracle import Oracle

    return Oracle(p)


def _state(o, rng):
    n = o.wfn_size
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def test_same_grid_is_sqrt_weight_times_the_coefficients():
    """Target = source grid: every interior point, bridges included, gets
    sqrt(w) * c exactly (L_m(x_k) = delta_mk), the CI vector is copied."""
    rng = np.random.default_rng(1)
    p = _prob(5, 6, 15.0, edges=_edges(rng, 5, 15.0), ndcore=1, nact=2, nel_alpha=2, nel_beta=2)
    o = _oracle(p)
    w = _state(o, rng)
    got = o.proj_to(o, w)
    sw = np.sqrt(o.table("wrad")[1:o.nrad])
    exp_orb = (w[: o.norb].reshape(o.nfun, o.L1, o.nr) * sw).ravel()
    assert np.array_equal(got[: o.norb], exp_orb)
    assert np.array_equal(got[o.norb:], w[o.norb:])


def _poly_values(r, rmax, coef):
    """p(r) = (r/R)(1 - r/R) q(r/R), q with the given coefficients (complex)."""
    x = r / rmax
    return x * (1.0 - x) * np.polyval(coef, x)


@pytest.mark.parametrize("n1,n2,nfe1,nfe2,r2", [(8, 5, 4, 7, 20.0), (6, 9, 6, 3, 12.0), (10, 10, 3, 5, 19.5)])
def test_piecewise_polynomials_are_reproduced(n1, n2, nfe1, nfe2, r2):
    """A state whose coefficients are p(r_i) at the source points, p of degree
    <= ndvr with p(0) = p(R) = 0, projects to sqrt(w2) p(r2) on a target grid
    of other elements and orders inside the source box (1e-12 of max |p|)."""
    rng = np.random.default_rng(n1 * 100 + n2)
    r1 = 20.0
    p1 = _prob(nfe1, n1, r1, lmax1=2, edges=_edges(rng, nfe1, r1))
    p2 = _prob(nfe2, n2, r2, lmax1=2, edges=_edges(rng, nfe2, r2))
    o1, o2 = _oracle(p1), _oracle(p2)
    x1 = o1.table("xrad")[1:o1.nrad]
    x2 = o2.table("xrad")[1:o2.nrad]
    sw2 = np.sqrt(o2.table("wrad")[1:o2.nrad])
    coefs = [rng.standard_normal(n1 - 1) + 1j * rng.standard_normal(n1 - 1) for _ in range(o1.L1)]
    orb1 = np.concatenate([_poly_values(x1, r1, c) for c in coefs])
    w = np.concatenate([orb1, [1.0 + 0.5j]])
    got = o1.proj_to(o2, w)
    exp = np.concatenate([_poly_values(x2, r1, c) * sw2 for c in coefs])
    scale = np.max(np.abs(exp))
    assert np.max(np.abs(got[: o2.norb] - exp)) < 1e-12 * scale
    assert got[o2.norb] == 1.0 + 0.5j
    # the library: the same numbers (same node products and accumulation order)
    lib = td1c_amd.wfn_proj(p1, w, p2)
    assert np.max(np.abs(lib - got)) <= 1e-15 * scale


def test_points_beyond_the_source_box_are_zero():
    """A target box twice the source's: the points past R (and within 1e-10 of
    it) get zero, those inside interpolate the polynomial."""
    rng = np.random.default_rng(7)
    p1 = _prob(4, 7, 10.0, lmax1=1)
    p2 = _prob(8, 5, 20.0, lmax1=1, edges=_edges(rng, 8, 20.0))
    o1, o2 = _oracle(p1), _oracle(p2)
    x1 = o1.table("xrad")[1:o1.nrad]
    x2 = o2.table("xrad")[1:o2.nrad]
    sw2 = np.sqrt(o2.table("wrad")[1:o2.nrad])
    c = rng.standard_normal(6)
    w = np.concatenate([_poly_values(x1, 10.0, c), _poly_values(x1, 10.0, -c), [1.0]])
    got = o1.proj_to(o2, w)
    out = x2 + 1e-10 >= 10.0
    assert out.any() and (~out).any()
    g = got[: o2.norb].reshape(2, o2.nr)
    assert np.all(g[:, out] == 0)
    exp = _poly_values(x2, 10.0, c) * sw2
    assert np.max(np.abs(g[0, ~out] - exp[~out])) < 1e-12 * np.max(np.abs(exp))
    assert np.max(np.abs(g[1, ~out] + exp[~out])) < 1e-12 * np.max(np.abs(exp))
    assert np.array_equal(td1c_amd.wfn_proj(p1, w, p2), got)


def test_orbital_and_l_blocks_map_to_their_own():
    """Projection acts on each (orbital, l) radial block alone: a state with
    one nonzero block projects into that block only (linearity), for source
    and target of other orbital counts and lmax1; orbitals and l that only one
    side holds stay zero; the CI copy keeps min(lcic) coefficients."""
    rng = np.random.default_rng(3)
    big = _prob(5, 6, 15.0, lmax1=4, ndcore=1, nact=3, nel_alpha=2, nel_beta=2)
    small = _prob(4, 8, 12.0, lmax1=2, ndcore=1, nact=2, nel_alpha=2, nel_beta=2)
    ob, osm = _oracle(big), _oracle(small)
    assert ob.lcic > osm.lcic > 1
    for a, b, pa, pb in ((ob, osm, big, small), (osm, ob, small, big)):
        full = _state(a, rng)
        ref_full = a.proj_to(b, full)
        assert np.array_equal(td1c_amd.wfn_proj(pa, full, pb), ref_full)
        nci = min(a.lcic, b.lcic)
        assert np.array_equal(ref_full[b.norb:b.norb + nci], full[a.norb:a.norb + nci])
        assert np.all(ref_full[b.norb + nci:] == 0)
        blocks = ref_full[: b.norb].reshape(b.nfun, b.L1, b.nr)
        nf, nl = min(a.nfun, b.nfun), min(a.L1, b.L1)
        assert np.all(blocks[nf:] == 0) and np.all(blocks[:, nl:] == 0)
        total = np.zeros(b.nfun * b.nbas, np.complex128)
        for f in range(a.nfun):
            for l in range(a.L1):
                one = np.zeros(a.wfn_size, np.complex128)
                sl = slice((f * a.L1 + l) * a.nr, (f * a.L1 + l + 1) * a.nr)
                one[sl] = full[sl]
                g = a.proj_to(b, one)[: b.norb].reshape(b.nfun, b.L1, b.nr)
                mask = np.ones_like(g, dtype=bool)
                if f < b.nfun and l < b.L1:
                    mask[f, l] = False
                assert np.all(g[mask] == 0), (f, l)
                total += g.ravel()
        assert np.max(np.abs(total - ref_full[: b.norb])) < 1e-13 * np.max(np.abs(total))


def test_library_matches_oracle_on_random_grids():
    """Random element edges, orders and boxes both ways (finer, coarser,
    smaller and larger boxes; frozen-core and CAS problems): the library's
    projection equals the oracle's to 1e-15 of the largest coefficient."""
    rng = np.random.default_rng(11)
    for trial in range(6):
        nfe1, nfe2 = rng.integers(2, 9, 2)
        n1, n2 = rng.integers(3, 13, 2)
        r1, r2 = rng.uniform(5, 40, 2)
        kw = dict(nfcore2=1, ndcore=0, nact=2, nel_alpha=2, nel_beta=2, lval=[0, 0, 0]) if trial % 2 else {}
        p1 = _prob(int(nfe1), int(n1), float(r1), lmax1=int(rng.integers(0, 5)),
                   edges=_edges(rng, int(nfe1), float(r1)), **kw)
        p2 = _prob(int(nfe2), int(n2), float(r2), lmax1=int(rng.integers(0, 5)),
                   edges=_edges(rng, int(nfe2), float(r2)), **kw)
        o1, o2 = _oracle(p1), _oracle(p2)
        w = _state(o1, rng)
        ref = o1.proj_to(o2, w)
        got = td1c_amd.wfn_proj(p1, w, p2)
        assert got.shape == ref.shape
        assert np.max(np.abs(got - ref)) <= 1e-15 * max(np.max(np.abs(ref)), 1.0), trial


def test_rejects_bad_input():
    p = _prob(4, 6, 10.0)
    w = np.zeros(p.nfun * p.nbas + 1, np.complex128)
    with pytest.raises(ValueError):
        td1c_amd.wfn_proj(p, w[:-1], p)
    bad = _prob(3, 6, 10.0, edges=[0.0, 4.0, 3.0, 10.0])
    with pytest.raises(RuntimeError):
        td1c_amd.wfn_proj(bad, np.zeros(bad.nfun * bad.nbas + 1), p)


@pytest.mark.gpu
def test_read_proj_then_orth_on_the_device(gpu_lib):
    """guess read_proj: projection on the host, then clwfn::orth on the device
    copy, against the oracle's projection and orthonormalisation (1e-13)."""
    import torch

    from td1c_amd import workloads as W

    rng = np.random.default_rng(5)
    p1 = _prob(6, 8, 18.0, lmax1=3, ndcore=1, nact=2, nel_alpha=2, nel_beta=2, edges=_edges(rng, 6, 18.0))
    p2 = _prob(9, 7, 24.0, lmax1=4, ndcore=1, nact=2, nel_alpha=2, nel_beta=2, edges=_edges(rng, 9, 24.0))
    o1, o2 = _oracle(p1), _oracle(p2)
    w = W.bound_wfn(p1, 21, o1.lcic)
    proj = td1c_amd.wfn_proj(p1, w, p2)
    ref = o2.orth(o1.proj_to(o2, w))
    hp = td1c_amd.HProd(p2)
    d = hp.to_device(proj)
    hp.orth(d)
    torch.cuda.synchronize()
    got = d.cpu().numpy()
    assert np.max(np.abs(got - ref)) < 1e-13 * np.max(np.abs(ref))
    hp.close()
