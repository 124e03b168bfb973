"""Wave-function files (td1c_wfn_write / td1c_wfn_read, clwfn::write / read,
src/wfn/clwfn.cpp:156-302).  Host-only functions of the library: CPU tests."""
import os

import numpy as np
import pytest

import td1c_amd
from td1c_amd.abi import Problem

pytestmark = pytest.mark.skipif(not os.path.exists(td1c_amd.library_path()), reason="library not built")


def _prob(**kw):
    base = dict(nfe=3, ndvr=4, rmax=6.0, lmax1=2, lmax2=4, znuc=2.0, ndcore=1, nact=2, nel_alpha=2, nel_beta=2)
    base.update(kw)
    return Problem(**base)


def _state(prob, lcic, seed=0):
    rng = np.random.default_rng(seed)
    n = prob.nfun * prob.nbas + lcic
    return rng.standard_normal(n) * 10.0 ** rng.integers(-12, 3, n) + 1j * rng.standard_normal(n)


def test_write_format_and_round_trip(tmp_path):
    """The text is the reference's: four "%10d" header lines (nrad, lmax1,
    mmax1, nfun), then "%25.15le%25.15le" lines in [orbital][l][r] order; the
    .cic file starts with lcic.  Reading it back restores the state to the
    15 significant digits written."""
    prob = _prob(mval=[0, 1, 0])
    lcic = 4
    w = _state(prob, lcic)
    po, pc = str(tmp_path / "x.orb"), str(tmp_path / "x.cic")
    td1c_amd.wfn_write(po, pc, prob, lcic, w)
    lines = open(po).read().splitlines()
    nrad = prob.nr + 1
    assert lines[:4] == ["%10d" % nrad, "%10d" % prob.lmax1, "%10d" % 1, "%10d" % prob.nfun]
    assert len(lines) == 4 + prob.nfun * prob.nbas
    q = 7
    assert lines[4 + q] == "%25.15e%25.15e" % (w[q].real, w[q].imag)
    cl = open(pc).read().splitlines()
    assert cl[0] == "%10d" % lcic and len(cl) == 1 + lcic
    assert cl[1 + 2] == "%25.15e%25.15e" % (w[prob.nfun * prob.nbas + 2].real, w[prob.nfun * prob.nbas + 2].imag)
    r, ci_read = td1c_amd.wfn_read(po, pc, prob, lcic)
    assert ci_read
    assert np.all(np.abs(r - w) <= 1e-14 * np.abs(w) + 1e-300)


def test_read_truncates_and_pads(tmp_path):
    """A file from a larger basis (more l, more radial points, more orbitals)
    keeps the leading block; one from a smaller basis leaves zeros; a CI vector
    of another length is replaced by the first determinant and reported."""
    big = _prob(nfe=4, lmax1=3, nact=3)
    small = _prob()
    lb, ls = 9, 4
    w = _state(big, lb, seed=1)
    po, pc = str(tmp_path / "b.orb"), str(tmp_path / "b.cic")
    td1c_amd.wfn_write(po, pc, big, lb, w)
    r, ci_read = td1c_amd.wfn_read(po, pc, small, ls)
    assert not ci_read
    wb = w[: big.nfun * big.nbas].reshape(big.nfun, big.lmax1 + 1, big.nr)
    rs = r[: small.nfun * small.nbas].reshape(small.nfun, small.lmax1 + 1, small.nr)
    exp = wb[: small.nfun, : small.lmax1 + 1, : small.nr]
    assert np.all(np.abs(rs - exp) <= 1e-14 * np.abs(exp) + 1e-300)
    cic = r[small.nfun * small.nbas:]
    assert cic[0] == 1.0 and np.all(cic[1:] == 0)

    ws = _state(small, ls, seed=2)
    td1c_amd.wfn_write(po, pc, small, ls, ws)
    r, ci_read = td1c_amd.wfn_read(po, pc, big, lb)
    rb = r[: big.nfun * big.nbas].reshape(big.nfun, big.lmax1 + 1, big.nr)
    wsr = ws[: small.nfun * small.nbas].reshape(small.nfun, small.lmax1 + 1, small.nr)
    assert np.all(np.abs(rb[: small.nfun, : small.lmax1 + 1, : small.nr] - wsr) <= 1e-14 * np.abs(wsr) + 1e-300)
    mask = np.ones(rb.shape, bool)
    mask[: small.nfun, : small.lmax1 + 1, : small.nr] = False
    assert np.all(rb[mask] == 0)
    assert not ci_read


def test_missing_file_is_an_error(tmp_path):
    prob = _prob()
    with pytest.raises(RuntimeError):
        td1c_amd.wfn_read(str(tmp_path / "none.orb"), str(tmp_path / "none.cic"), prob, 4)
