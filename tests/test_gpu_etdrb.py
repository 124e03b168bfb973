"""cletdrb (exponential Rosenbrock RK4 with Pade phi functions, the
reference's default prop_type) on the GPU against the oracle's
orc_etdrb_step: He, three-orbital, frozen-core + CAS and CAS problems in both
gauges, the element (ndvr = 20) and band (ndvr = 16, 10) one-body paths, real
and imaginary time, FP64 and the complex64 mean field.  Two steps at dt =
0.01 from a bound-state-like state in a sin^2 pulse: FP64 relative L2 1e-10,
mixed 1e-6."""
import numpy as np
import pytest

from tests import cases as C

pytestmark = pytest.mark.gpu

DT = 0.01


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


CASES = [("he_small", 0), ("dc3_small", 1), ("ne_small", 1), ("cas_fc_small", 1), ("cas_small", 0),
         ("he_n16", 1)]


def _field(igauge):
    import td1c_amd

    return td1c_amd.Field(env_type="sin2", gauge="velocity1" if igauge else "length1", fint=1e14, wlen=800.0,
                          cep=25.0, numcyc=2.0)


def _setup(prob):
    import td1c_amd
    from oracle.oracle import Oracle

    o = Oracle(prob)
    w = C.smooth_wfn_for(prob, o.lcic)
    o.set_wfn0(w)
    o.etdrb_gen(DT)
    hp = td1c_amd.HProd(prob)
    hp.set_wfn0(w)
    return o, w, hp, td1c_amd.ETDRB(hp, DT)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name,igauge", CASES)
def test_etdrb_steps_match_oracle(gpu_lib, name, igauge, precision):
    import torch
    import td1c_amd

    prob = getattr(C, name)(isplit=1, igauge=igauge, precision=precision)
    o, w, hp, etd = _setup(prob)
    f = _field(igauge)
    d = hp.to_device(w)
    ref = w.copy()
    t = 30.0
    for k in range(2):
        lf = td1c_amd.etd_fields(f, t + k * DT, DT)
        etd.prop(d, lfields=lf)
        ref = o.etdrb_step(lf, ref, apply_mask=True)
    torch.cuda.synchronize()
    got = d.cpu().numpy()
    tol = 1e-10 if precision == 0 else 1e-6
    assert _rel(got, ref) < tol, _rel(got, ref)
    assert _rel(got, w) > 1e-5  # the state moved
    hp.close()


def test_etdrb_imaginary_time_matches_oracle(gpu_lib):
    """icomp = 0 (no Neumann series in the denominators; tunit = -1), band
    one-body path (ndvr = 10)."""
    import torch
    import td1c_amd

    prob = C.he_itr(isplit=1)
    o, w, hp, etd = _setup(prob)
    lf = np.zeros(54)
    d = hp.to_device(w)
    ref = w.copy()
    for _ in range(2):
        etd.prop(d, lfields=lf, apply_mask=False)
        ref = o.etdrb_step(lf, ref, apply_mask=False)
    torch.cuda.synchronize()
    assert _rel(d.cpu().numpy(), ref) < 1e-10
    hp.close()


@pytest.mark.parametrize("name,igauge", [("cas_fc_small", 1), ("he_small", 0)])
def test_etdrb_propagate_with_device_fields(gpu_lib, name, igauge):
    """td1c_etdrb_propagate (fields on the device, tdse_prop12's loop) equals
    the same steps with host field arrays."""
    import torch
    import td1c_amd

    prob = getattr(C, name)(isplit=1, igauge=igauge)
    o, w, hp, etd = _setup(prob)
    f = _field(igauge)
    t = 40.0
    a = hp.to_device(w)
    b = hp.to_device(w)
    etd.propagate(a, t, 3, f)
    for k in range(3):
        etd.prop(b, time=t + k * DT, field=f)
    torch.cuda.synchronize()
    assert _rel(a.cpu().numpy(), b.cpu().numpy()) < 1e-12
    hp.close()


def test_etdrb_rejects_what_is_not_provided(gpu_lib):
    import td1c_amd

    for kw in (dict(isplit=0), dict(isplit=1, lconst=1)):
        hp = td1c_amd.HProd(C.he_small(**kw))
        with pytest.raises(RuntimeError):
            td1c_amd.ETDRB(hp, DT)
        hp.close()
    import ctypes

    hp = td1c_amd.HProd(C.he_small(isplit=1))
    etd = td1c_amd.ETDRB(hp, DT)
    d = hp.to_device(C.smooth_wfn_for(C.he_small(isplit=1), 1))
    cf = _field(0).to_c()
    rc = td1c_amd.load_library().td1c_etdrb_propagate(hp.h, ctypes.byref(cf), 0.0, 2 * DT, 1,
                                                      ctypes.c_void_p(d.data_ptr()), 1, None)
    assert rc != 0  # dtime other than the generated one
    del etd
    hp.close()


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name,igauge", [("he_small", 0), ("cas_fc_small", 1), ("dc3_small", 1), ("he_n16", 1)])
def test_etdrk4_steps_match_oracle(gpu_lib, name, igauge, precision):
    """cletdrk4: two steps against orc_etdrk4_step (FP64 1e-10, mixed 1e-6)."""
    import torch
    import td1c_amd
    from oracle.oracle import Oracle

    prob = getattr(C, name)(isplit=1, igauge=igauge, precision=precision)
    o = Oracle(prob)
    w = C.smooth_wfn_for(prob, o.lcic)
    o.set_wfn0(w)
    o.etdrk4_gen(DT)
    hp = td1c_amd.HProd(prob)
    hp.set_wfn0(w)
    etk = td1c_amd.ETDRK4(hp, DT)
    f = _field(igauge)
    d = hp.to_device(w)
    ref = w.copy()
    for k in range(2):
        lf = td1c_amd.etd_fields(f, 30.0 + k * DT, DT)
        etk.prop(d, lfields=lf)
        ref = o.etdrk4_step(lf, ref, apply_mask=True)
    torch.cuda.synchronize()
    got = d.cpu().numpy()
    tol = 1e-10 if precision == 0 else 1e-6
    assert _rel(got, ref) < tol, _rel(got, ref)
    assert _rel(got, w) > 1e-5
    # the same steps with the fields on the device
    e = hp.to_device(w)
    etk.propagate(e, 30.0, 2, f)
    torch.cuda.synchronize()
    assert _rel(e.cpu().numpy(), got) < 1e-12
    # the etdrb entry points refuse the etdrk4 tables
    with pytest.raises(RuntimeError):
        td1c_amd.ETDRB.prop(type("E", (), {"hprod": hp, "dtime": DT, "_fn": "etdrb"})(), e,
                            lfields=td1c_amd.etd_fields(f, 30.0, DT))
    hp.close()


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name,igauge", [("he_small", 0), ("cas_fc_small", 1), ("ne_small", 1), ("he_n16", 1)])
def test_lawrk_steps_match_oracle(gpu_lib, name, igauge, precision):
    """cllawrk: two steps against orc_lawrk_step (FP64 1e-10, mixed 1e-6), the
    device-field propagation equal to them, and the factors rebuilt after a
    split2 call of another dt replaced them."""
    import torch
    import td1c_amd
    from oracle.oracle import Oracle

    prob = getattr(C, name)(isplit=1, igauge=igauge, precision=precision)
    o = Oracle(prob)
    w = C.smooth_wfn_for(prob, o.lcic)
    o.set_wfn0(w)
    o.lawrk_gen(DT)
    hp = td1c_amd.HProd(prob)
    hp.set_wfn0(w)
    law = td1c_amd.LAWRK(hp, DT)
    f = _field(igauge)
    d = hp.to_device(w)
    ref = w.copy()
    for k in range(2):
        lf = td1c_amd.etd_fields(f, 30.0 + k * DT, DT)
        law.prop(d, lfields=lf)
        ref = o.lawrk_step(lf, ref, apply_mask=True)
    torch.cuda.synchronize()
    got = d.cpu().numpy()
    tol = 1e-10 if precision == 0 else 1e-6
    assert _rel(got, ref) < tol, _rel(got, ref)
    assert _rel(got, w) > 1e-5
    td1c_amd.CRK(4).propagate(hp, hp.to_device(w), 0.0, 2 * DT, 1, f)  # split2 factors of 2 dt
    e = hp.to_device(w)
    law.propagate(e, 30.0, 2, f)
    torch.cuda.synchronize()
    assert _rel(e.cpu().numpy(), got) < 1e-12
    hp.close()


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name,igauge", [("he_small", 0), ("cas_fc_small", 1), ("ne_small", 1), ("he_n16", 1)])
def test_expab_steps_match_oracle(gpu_lib, name, igauge, precision):
    """clexpab: five steps (the history filled at the start, then shifted)
    against orc_expab_step (FP64 1e-10, mixed 1e-6), and td1c_expab_propagate
    over two calls (2 + 3 steps, the history carried between them) equal to
    them."""
    import torch
    import td1c_amd
    from oracle.oracle import Oracle

    prob = getattr(C, name)(isplit=1, igauge=igauge, precision=precision)
    o = Oracle(prob)
    w = C.smooth_wfn_for(prob, o.lcic)
    o.set_wfn0(w)
    o.expab_gen(DT)
    hp = td1c_amd.HProd(prob)
    hp.set_wfn0(w)
    xab = td1c_amd.EXPAB(hp, DT)
    f = _field(igauge)
    t0 = 30.0
    d = hp.to_device(w)
    ref = w.copy()
    for n in range(5):
        lf = td1c_amd.EXPAB.history_fields(f, t0, DT, n)
        xab.prop(d, time=t0 + n * DT, lfields=lf)
        ref = o.expab_step(lf, ref, apply_mask=True)
    torch.cuda.synchronize()
    got = d.cpu().numpy()
    tol = 1e-10 if precision == 0 else 1e-6
    assert _rel(got, ref) < tol, _rel(got, ref)
    assert _rel(got, w) > 1e-5
    xab2 = td1c_amd.EXPAB(hp, DT)  # restarts the history
    e = hp.to_device(w)
    xab2.propagate(e, t0, 2, f)
    xab2.propagate(e, t0 + 2 * DT, 3, f)
    torch.cuda.synchronize()
    assert _rel(e.cpu().numpy(), got) < 1e-12
    hp.close()


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name,igauge", [("he_small", 0), ("cas_small", 0), ("cas_fc_small", 1), ("dc3_small", 1),
                                         ("ne_small", 1)])
def test_chkconv_matches_oracle(gpu_lib, name, igauge, precision):
    """td1c_chkconv against orc_chkconv at a seeded state and a field:
    energy, CI residual and both orbital-gradient norms (FP64 1e-10,
    mixed 1e-6, relative to max(|value|, 1): the CI residual of a single
    CAS root is near zero)."""
    import torch
    import td1c_amd
    from oracle.oracle import Oracle

    prob = getattr(C, name)(isplit=1, igauge=igauge, precision=precision)
    o = Oracle(prob)
    w = C.smooth_wfn_for(prob, o.lcic)
    o.set_wfn0(w)
    lf = C.LF_VEL if igauge else C.LF_LEN
    ref = o.chkconv(lf, w)
    hp = td1c_amd.HProd(prob)
    hp.set_wfn0(w)
    got = hp.chkconv(lf, hp.to_device(w))
    torch.cuda.synchronize()
    tol = 1e-10 if precision == 0 else 1e-6
    for k in range(4):
        assert abs(got[k] - ref[k]) <= tol * max(abs(ref[k]), 1.0), (k, got, ref)
    assert ref[3] > 1e-3
    hp.close()


@pytest.mark.parametrize("cas", [False, True])
def test_init_prop12_relaxes_as_the_oracle(gpu_lib, cas):
    """td1c_init_prop12 (chkconv, cletdrb with icomp = 0, orth) against the
    same loop on the oracle: the same final energy (1e-10) and cycle count
    (within 2), the gradient below 1e-6."""
    import torch
    import td1c_amd
    from oracle.oracle import Oracle
    from td1c_amd.abi import Problem

    if cas:
        prob = Problem(nfe=12, ndvr=10, rmax=30.0, lmax1=1, lmax2=2, znuc=2.0, ndcore=0, nact=2, nel_alpha=1,
                       nel_beta=1, icomp=0, isplit=1)
    else:
        prob = C.he_itr(isplit=1)
    o = Oracle(prob)
    w = C.smooth_wfn_for(prob, o.lcic)
    o.set_wfn0(w)
    o.etdrb_gen(0.01)
    e1, ref, ncyc_ref = -1e10, w.copy(), None
    for icyc in range(4000):
        c = o.chkconv(np.zeros(9), ref)
        if abs(c[0] - e1) < 1e-12 and c[3] < 1e-6:
            ncyc_ref = icyc
            break
        e1 = c[0]
        ref = o.orth(o.etdrb_step(np.zeros(54), ref, apply_mask=False))
    assert ncyc_ref is not None
    hp = td1c_amd.HProd(prob)
    hp.set_wfn0(w)
    d = hp.to_device(w)
    e, n = hp.init_prop12(d, dstep=0.01, maxcyc=4000)
    torch.cuda.synchronize()
    assert abs(e - c[0]) < 1e-10 and abs(n - ncyc_ref) <= 2, (e, c[0], n, ncyc_ref)
    assert hp.chkconv(np.zeros(9), d)[3] < 1e-6
    hp.close()
