"""split_type = 1 on the GPU against the oracle: the split hprod
(hprod_htot_dtorb with isplit = 1), the clh1itr half step (h1itr_prop) and
the tdse_split2 step (half step, RK4 of the split hprod, half step, mask),
at the production step dt = 0.01.  FP64 tolerance as the hprod parity
tests (relative L2 1e-11 per call, 1e-10 over several steps)."""
import numpy as np
import pytest

from tests import cases as C

pytestmark = pytest.mark.gpu

DT = 0.01


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


CASES = [("he_small", C.LF_LEN), ("dc3_small", C.LF_LEN), ("ne_small", C.LF_VEL), ("cas_fc_small", C.LF_VEL),
         ("cas_small", C.LF_LEN), ("he_n16", C.LF_VEL)]  # ndvr = 16: the band right-hand-side kernel


@pytest.mark.parametrize("name,lf", CASES)
def test_split_hprod_matches_oracle(gpu_lib, name, lf):
    """The hprod of a split_type = 1 context: the orbital derivative without
    h0 + h1 in its Q part (X still holds the full one-body blocks)."""
    import torch
    import td1c_amd
    from oracle.oracle import Oracle

    prob = getattr(C, name)(isplit=1)
    o = Oracle(prob)
    w = C.wfn_for(prob, o.lcic)
    o.set_wfn0(w)
    ref = o.htot(DT, lf, w)
    hp = td1c_amd.HProd(prob)
    hp.set_wfn0(w)
    got = hp.htot(DT, lf, hp.to_device(w))
    torch.cuda.synchronize()
    assert _rel(got.cpu().numpy(), ref) < 1e-11


@pytest.mark.parametrize("name,lf", CASES)
def test_h1itr_half_step_matches_oracle(gpu_lib, name, lf):
    """td1c_h1itr_prop (element Schur solve + series with a device
    convergence test) against h1itr_prop (band LU), twice in a row on the
    same context (the flag and the ticket are reused)."""
    import torch
    import td1c_amd
    from oracle.oracle import Oracle

    prob = getattr(C, name)(isplit=1)
    o = Oracle(prob)
    w = C.wfn_for(prob, o.lcic)
    o.set_wfn0(w)
    o.h1itr_gen(DT / 2)
    hp = td1c_amd.HProd(prob)
    hp.set_wfn0(w)
    hp.h1itr_gen(DT)
    d = hp.to_device(w)
    ref = w
    for k in range(2):
        ref, ncyc = o.h1itr_prop(lf, DT / 2, ref)
        assert 1 <= ncyc < 20
        hp.h1itr_prop(lf, d)
        torch.cuda.synchronize()
        got = d.cpu().numpy()
        assert _rel(got, ref) < 1e-11, (k, _rel(got, ref))
        # frozen cores and the CI vector are untouched
        nb = hp.nbas
        assert np.array_equal(got[:nb * prob.nfcore], w[:nb * prob.nfcore])
        assert np.array_equal(got[hp.norb:], w[hp.norb:])


@pytest.mark.parametrize("name,lf,scale,ncyc", [("ne_small", C.LF_VEL, 10.0, 12), ("ne_small", C.LF_VEL, 30.0, 20),
                                                ("dc3_small", C.LF_LEN, 100.0, 7), ("he_small", C.LF_VEL, 100.0, 20)])
def test_long_series_match_oracle(gpu_lib, name, lf, scale, ncyc):
    """Half steps whose series runs 7 to 20 terms (fields scaled up; 20 is
    h1itr_maxcyc, a series that stops unconverged): the device convergence
    test stops where h1itr_prop's does, and the state agrees (1e-11)."""
    import torch
    import td1c_amd
    from oracle.oracle import Oracle

    prob = getattr(C, name)(isplit=1)
    lfs = [x * scale for x in lf]
    o = Oracle(prob)
    w = C.wfn_for(prob, o.lcic)
    o.set_wfn0(w)
    o.h1itr_gen(DT / 2)
    ref, n = o.h1itr_prop(lfs, DT / 2, w)
    assert n == ncyc
    hp = td1c_amd.HProd(prob)
    hp.set_wfn0(w)
    hp.h1itr_gen(DT)
    d = hp.to_device(w)
    hp.h1itr_prop(lfs, d)
    torch.cuda.synchronize()
    got = d.cpu().numpy()
    hp.close()
    assert _rel(got, ref) < 1e-11, _rel(got, ref)


@pytest.mark.parametrize("name,lf,scale", [("ne_small", C.LF_VEL, 10.0), ("dc3_small", C.LF_LEN, 1.0)])
def test_half_step_same_bits_for_every_onebody_shape(gpu_lib, name, lf, scale):
    """The convergence test inside the right-hand side launch sums the
    partial norms in the same order for 64-, 128- and 256-thread blocks (3,
    6, 12 elements per block): every element one-body shape gives the same
    half step bit for bit, and the oracle's (1e-11)."""
    import torch
    import td1c_amd
    from oracle.oracle import Oracle

    prob = getattr(C, name)(isplit=1)
    lfs = [x * scale for x in lf]
    o = Oracle(prob)
    w = C.wfn_for(prob, o.lcic)
    o.set_wfn0(w)
    o.h1itr_gen(DT / 2)
    ref, _ = o.h1itr_prop(lfs, DT / 2, w)
    hp = td1c_amd.HProd(prob)
    hp.set_wfn0(w)
    hp.h1itr_gen(DT)
    base = None
    for ge in (3, 6, 12):
        for lc in (1, 4, 16):
            hp.set_onebody_shape(ge, lc)
            d = hp.to_device(w)
            hp.h1itr_prop(lfs, d)
            torch.cuda.synchronize()
            got = d.cpu().numpy()
            if base is None:
                base = got
                assert _rel(got, ref) < 1e-11
            else:
                assert np.array_equal(got, base), (ge, lc)
    hp.close()


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("name", ["ne_small", "cas_fc_small", "dc3_small"])
def test_split2_propagation_matches_oracle(gpu_lib, name, graph):
    """td1c_propagate with split_type = 1 (device stage and half-step
    fields) against the oracle's tdse_split2 step with host field values:
    three calls of two steps at dt = 0.01 in a velocity-gauge pulse, eager
    hprods between the calls."""
    import torch
    import td1c_amd
    from oracle.oracle import Oracle

    prob = getattr(C, name)(isplit=1, igauge=1)
    o = Oracle(prob)
    w = C.smooth_wfn_for(prob, o.lcic)  # bound-state behaviour at r -> 0: dt = 0.01 is stable
    o.set_wfn0(w)
    o.h1itr_gen(DT / 2)
    hp = td1c_amd.HProd(prob)
    hp.set_wfn0(w)
    hp.set_graph(graph)
    field = td1c_amd.Field(env_type="sin2", gauge="velocity1", fint=1e14, wlen=800.0, numcyc=2.0)
    rk = td1c_amd.CRK(4)
    t = 30.0
    d = hp.to_device(w)
    scratch = hp.to_device(w)
    ref = w.copy()
    step = 0
    for call in range(3):
        hp.htot(1.0, C.LF_VEL, scratch)
        torch.cuda.synchronize()
        rk.propagate(hp, d, t + step * DT, DT, 2, field, apply_mask=True)
        for _ in range(2):
            t0 = t + step * DT
            ref, _ = o.split2_step(4, DT, td1c_amd.field_value(field, t0 + 0.25 * DT), rk.stage_fields(field, t0, DT),
                                   td1c_amd.field_value(field, t0 + 0.75 * DT), ref, apply_mask=True)
            step += 1
        torch.cuda.synchronize()
        got = d.cpu().numpy()
        assert _rel(got, ref) < 1e-10, (call, _rel(got, ref))
    assert abs(np.vdot(got[:hp.norb], got[:hp.norb]).real - np.vdot(w[:hp.norb], w[:hp.norb]).real) < 1e-3


def test_split_rejects_what_is_not_provided(gpu_lib):
    import td1c_amd

    with pytest.raises(RuntimeError):
        td1c_amd.HProd(C.ne_small(isplit=1, iprojfc=1))
    with pytest.raises(RuntimeError):
        td1c_amd.HProd(C.he_small(isplit=2))
    hp = td1c_amd.HProd(C.he_small())
    with pytest.raises(RuntimeError):
        hp.h1itr_gen(DT)
