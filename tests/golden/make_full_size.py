"""Writes tests/golden/full_size.json: fixtures of FP64 oracle results at
full workload sizes, each too slow for the oracle inside a GPU test.

1. One hprod (split_type 0, workloads.make_wfn) of the Ar workload (9
   orbitals, 2399 radial points, lmax1 63 / lmax2 126) and of the 16-orbital
   mean-field workload (136 pairs, 1999 points).
2. One hprod and one tdse_split2 step (split_type 1, dt = 0.01, velocity
   gauge, workloads.bound_wfn with a seeded CI vector) of the correlated Ne
   workload (frozen 1s + CAS(8e, 6o), 225 determinants) and of the Ar
   workload.

Stored per case: the seed of the input, the field, dt, the five energies
(of the hprod), ||h|| and the projections <u_k|h> of the output onto eight
seeded random unit vectors u_k (numpy PCG64, seed 1000 + k), and for the
split step the same for the stepped state.  Any h' with ||h' - h|| = d moves
each projection by at most d, so the fixtures pin the output to the 2-norm
tolerance of the parity tests.  The oracle's OpenMP build (bit-identical to
the serial one) computes them.  Run from the repository root:
python tests/golden/make_full_size.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from td1c_amd import workloads as W  # noqa: E402
from td1c_amd.abi import Field  # noqa: E402

LF = [0.0, 0.0, 0.05, 0.0, 0.0, 0.05, 0.0, 0.0, 0.0]
CASES = {"ar_l63": 3007, "mf16_l63": 5007}
# split2 step cases: (workload factory, seed, peak intensity, step start time)
SPLIT_CASES = {"ne_l31_cas": ("ne_l31_cas", 2107, 1.0e14, 100.0), "ar_l63_split": ("ar_l63", 3007, 1.6e14, 100.0)}
SPLIT_DT = 0.01
NPROJ = 8


def projection_vectors(n: int):
    for k in range(NPROJ):
        rng = np.random.Generator(np.random.PCG64(1000 + k))
        u = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        yield u / np.linalg.norm(u)


def split_field(fint: float) -> Field:
    return Field(env_type="sin2", gauge="velocity1", fint=fint, wlen=800.0, numcyc=4.0)


def split_problem(name: str):
    factory, seed, fint, t0 = SPLIT_CASES[name]
    return getattr(W, factory)(isplit=1, igauge=1), seed, fint, t0


def field_values(fint: float, t0: float, dt: float):
    """(lfield at t0 + dt/4, the four RK4 stage fields, lfield at t0 + 3dt/4)
    from the library's clfield restatement (host code, no device)."""
    import td1c_amd

    f = split_field(fint)
    stages = np.stack([td1c_amd.field_value(f, t0 + c * dt) for c in (0.0, 0.5, 0.5, 1.0)])
    return td1c_amd.field_value(f, t0 + 0.25 * dt), stages, td1c_amd.field_value(f, t0 + 0.75 * dt)


def summary(h: np.ndarray) -> dict:
    proj = [complex(np.vdot(u, h)) for u in projection_vectors(h.size)]
    return {"norm": float(np.linalg.norm(h)), "proj_re": [p.real for p in proj], "proj_im": [p.imag for p in proj]}


def split_reference(name: str, openmp: bool = True) -> dict:
    prob, seed, fint, t0 = split_problem(name)
    o = O.Oracle(prob, openmp=openmp)
    w = W.bound_wfn(prob, seed, o.lcic)
    o.set_wfn0(w)
    q1, st, q3 = field_values(fint, t0, SPLIT_DT)
    h = o.htot(1.0, st[0], w)
    ene = o.energies().tolist()
    o.h1itr_gen(0.5 * SPLIT_DT)
    w1, ncyc = o.split2_step(4, SPLIT_DT, q1, st, q3, w, apply_mask=True)
    return {"seed": seed, "fint": fint, "t0": t0, "dtime": SPLIT_DT, "energies": ene, "hprod": summary(h),
            "step": summary(w1), "h1itr_cycles": list(ncyc)}


def main():
    O.build()
    O.build(openmp=True)
    out = {}
    for name, seed in CASES.items():
        prob = getattr(W, name)()
        o = O.Oracle(prob, openmp=True)
        w = W.make_wfn(prob, seed, o.lcic)
        o.set_wfn0(w)
        h = o.htot(1.0, np.array(LF), w)
        out[name] = dict({"seed": seed, "lfield": LF, "dtime": 1.0, "energies": o.energies().tolist()}, **summary(h))
        print(name, out[name]["norm"], out[name]["energies"], flush=True)
    for name in SPLIT_CASES:
        out[name] = split_reference(name)
        print(name, out[name]["hprod"]["norm"], out[name]["step"]["norm"], out[name]["h1itr_cycles"], flush=True)
    with open(os.path.join(ROOT, "tests", "golden", "full_size.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
