"""Batched propagation (td1c_batch_*): every member of a batch, stepped in
one launch per kernel site for the whole batch, against the same member
stepped alone by td1c_propagate (the same kernels on the same arguments:
the same bits), and one member against the oracle's tdse_split2 step."""
import numpy as np
import pytest

from td1c_amd.abi import Problem
from td1c_amd import workloads as W
from tests import cases as C

pytestmark = pytest.mark.gpu

DT = 0.01


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _he(nfe, lmax1, **kw):
    return Problem(nfe=nfe, ndvr=20, rmax=2.0 * nfe, lmax1=lmax1, lmax2=2 * lmax1, znuc=2.0, ndcore=1,
                   nel_alpha=1, nel_beta=1, igauge=1, isplit=1, **kw)


def _dc3(nfe, lmax1, znuc, **kw):
    return Problem(nfe=nfe, ndvr=20, rmax=3.0 * nfe, lmax1=lmax1, lmax2=2 * lmax1, znuc=znuc, ndcore=3,
                   nel_alpha=3, nel_beta=3, mval=[0, 1, -1], igauge=1, isplit=1, **kw)


def _fields(n, seed=5):
    rng = np.random.default_rng(seed)
    import td1c_amd

    return [td1c_amd.Field(env_type="sin2", gauge="velocity1", fint=float(10 ** rng.uniform(13, 15)),
                           wlen=float(rng.uniform(400, 1600)), cep=float(rng.uniform(0, 360)), numcyc=2.0)
            for _ in range(n)]


def _members(probs):
    import td1c_amd

    hps, ws = [], []
    for i, p in enumerate(probs):
        hp = td1c_amd.HProd(p)
        w = W.bound_wfn(p, 100 + i, hp.lcic)
        hp.set_wfn0(w)
        hps.append(hp)
        ws.append(w)
    return hps, ws


def _run_pair(probs, graph, steps=(2, 3), t0=20.0, tune=False, info=None):
    """The batch and the members alone from the same states; returns the two
    lists of final states (info: gets the batch's get_info)."""
    import torch
    import td1c_amd

    hps, ws = _members(probs)
    fields = _fields(len(probs))
    b = td1c_amd.Batch(hps)
    b.set_graph(graph)
    if tune:
        b.autotune([hp.to_device(w) for hp, w in zip(hps, ws)], DT, fields)
    db = [hp.to_device(w) for hp, w in zip(hps, ws)]
    t = t0
    for n in steps:
        b.propagate(db, t, DT, n, fields)
        t += n * DT
    torch.cuda.synchronize()
    got = [d.cpu().numpy() for d in db]
    if info is not None:
        info.update(b.info())
    b.close()
    rk = td1c_amd.CRK(4)
    ref = []
    for hp, w, f in zip(hps, ws, fields):
        hp.set_transform_path(2)  # the batch's transform form
        d = hp.to_device(w)
        t = t0
        for n in steps:
            rk.propagate(hp, d, t, DT, n, f)
            t += n * DT
        torch.cuda.synchronize()
        ref.append(d.cpu().numpy())
    return hps, ws, fields, got, ref


@pytest.mark.parametrize("graph", [True, False])
@pytest.mark.parametrize("precision", [0, 1])
def test_batch_equals_members_alone(gpu_lib, precision, graph):
    """Four He problems of different grids and lmax under different pulses:
    two batch calls (2 + 3 steps) equal each member's own td1c_propagate."""
    probs = [_he(6, 3, precision=precision), _he(9, 7, precision=precision), _he(4, 5, precision=precision),
             _he(12, 2, precision=precision)]
    info = {}
    hps, ws, fields, got, ref = _run_pair(probs, graph, info=info)
    for m, (g, r, w) in enumerate(zip(got, ref, ws)):
        assert _rel(g, r) < 1e-13, (m, _rel(g, r))
        assert _rel(g, w) > 1e-6  # the state moved
    # the same element boundaries and charge: every member reads the shared
    # interior inverses (the one-body element matrices differ in their last
    # bits with the element's position, so they are shared only where equal)
    assert info["shared_tables"]["cn"] == 4 and info["shared_tables"]["poisson"] == 4, info
    # and the Poisson row tables: the interior pass runs over groups of
    # (member, pair) items reading one copy of the tables
    assert info["shared_tables"]["poisson_rows"] == 4 and info["poisson_groups"] == 1, info
    # one dynamical orbital each: the half-step interior pass is grouped too
    assert info["cn_groups"] == 1, info
    for hp in hps:
        hp.close()


def test_batch_of_members_of_other_element_widths(gpu_lib):
    """Members whose elements differ in width (rmax = 2 nfe and 3 nfe):
    their element tables differ, so only some members share them and the
    interior passes run per member (no item groups); the batch still equals
    the members alone."""
    probs = [_he(6, 3), _he(8, 4), Problem(nfe=7, ndvr=20, rmax=21.0, lmax1=5, lmax2=10, znuc=2.0, ndcore=1,
                                         nel_alpha=1, nel_beta=1, igauge=1, isplit=1)]
    info = {}
    hps, ws, fields, got, ref = _run_pair(probs, True, info=info)
    for m, (g, r) in enumerate(zip(got, ref)):
        assert _rel(g, r) < 1e-13, (m, _rel(g, r))
    assert info["shared_tables"]["poisson"] < 3 and info["poisson_groups"] == 0, info
    assert info["shared_tables"]["cn"] < 3 and info["cn_groups"] == 0, info
    for hp in hps:
        hp.close()


def test_batch_autotuned_shapes_give_the_same_bits(gpu_lib):
    """The batch autotuner (every element one-body shape, both transform
    variants) on scratch copies, then two batch calls equal the members
    alone."""
    probs = [_he(7, 4), _he(10, 6), _he(5, 3)]
    info = {}
    hps, ws, fields, got, ref = _run_pair(probs, True, tune=True, info=info)
    assert len(info["tune"]["ms_per_step"]) == 13, info
    for m, (g, r) in enumerate(zip(got, ref)):
        assert _rel(g, r) < 1e-13, (m, _rel(g, r))
    for hp in hps:
        hp.close()


def test_batch_of_three_orbital_members(gpu_lib):
    """Three doubly occupied orbitals (m = 0, 1, -1) per member, two grids
    and nuclear charges: the pair densities, exchange and X matrix of each
    member stay within it."""
    probs = [_dc3(8, 6, 6.0), _dc3(5, 4, 4.0), _dc3(10, 3, 8.0)]
    info = {}
    hps, ws, fields, got, ref = _run_pair(probs, True, info=info)
    for m, (g, r) in enumerate(zip(got, ref)):
        assert _rel(g, r) < 1e-13, (m, _rel(g, r))
    # the Crank-Nicolson inverses depend on the charge: only the first member
    # reads the shared copy; the Poisson inverses do not
    assert info["shared_tables"]["cn"] == 1 and info["shared_tables"]["poisson"] == 3, info
    # several pairs per member (m_j - m_i up to 2: the L < |m| rows are zero)
    assert info["shared_tables"]["poisson_rows"] == 3 and info["poisson_groups"] >= 3, info
    assert info["cn_groups"] == 0, info  # three dynamical orbitals: each member's own half-step blocks
    for hp in hps:
        hp.close()


def test_batch_member_matches_oracle(gpu_lib):
    """One batch step of a two-member batch (He on two grids): each member
    against the oracle's tdse_split2 step with host field values."""
    import torch
    import td1c_amd
    from oracle.oracle import Oracle

    probs = [_he(6, 3), _he(9, 5)]
    hps, ws = _members(probs)
    fields = _fields(2, seed=9)
    b = td1c_amd.Batch(hps)
    db = [hp.to_device(w) for hp, w in zip(hps, ws)]
    t0 = 25.0
    b.propagate(db, t0, DT, 1, fields)
    torch.cuda.synchronize()
    rk = td1c_amd.CRK(4)
    for m, (p, w, f) in enumerate(zip(probs, ws, fields)):
        o = Oracle(p)
        o.set_wfn0(w)
        o.h1itr_gen(DT / 2)
        ref, _ = o.split2_step(4, DT, td1c_amd.field_value(f, t0 + 0.25 * DT), rk.stage_fields(f, t0, DT),
                               td1c_amd.field_value(f, t0 + 0.75 * DT), w, apply_mask=True)
        assert _rel(db[m].cpu().numpy(), ref) < 1e-10, (m, _rel(db[m].cpu().numpy(), ref))
    b.close()
    for hp in hps:
        hp.close()


def test_batch_scan_extremes_full_size(gpu_lib):
    """The largest and the smallest member shapes of the scan workload (120
    elements with lmax1 63, 40 elements with lmax1 15; complex64 mean field)
    in one batch: two steps equal each member alone."""
    probs = [_he(120, 63, precision=1), _he(40, 15, precision=1)]
    hps, ws, fields, got, ref = _run_pair(probs, True, steps=(2,))
    for m, (g, r) in enumerate(zip(got, ref)):
        assert _rel(g, r) < 1e-13, (m, _rel(g, r))
    for hp in hps:
        hp.close()


def test_batch_rejects_what_is_not_batched(gpu_lib):
    import td1c_amd

    a = td1c_amd.HProd(_he(6, 3))
    mixed = td1c_amd.HProd(_he(6, 3, precision=1))
    fc = td1c_amd.HProd(C.ne_small(isplit=1))
    for members in ([a, mixed], [a, fc], [a, a]):
        with pytest.raises(RuntimeError):
            td1c_amd.Batch(members)
    for hp in (a, mixed, fc):
        hp.close()
