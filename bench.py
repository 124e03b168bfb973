"""Benchmark: RK4 propagation of the Ne-like workload with the gfx950 hprod.

Workload ne_l31: 80 FEDVR elements x 21 Gauss-Lobatto points on [0,160] a.u.
(1599 radial points), lmax1 31 / lmax2 62, Z = 10, one frozen 1s core
(type -1) + 4 active orbitals (one determinant), 800 nm sin^2 pulse of 4
cycles at 1e14 W/cm^2, velocity gauge, complex128 throughout.

A step is one classic RK4 step of the full wave function (4 hprod
evaluations + stage combinations) followed by the radial mask, i.e.
clcrk::prop + clwfn::mask as tdse_split2 runs them.  An explicit RK4 step
of the full hprod is stable only while dt * rho < 2.8, and the spectral
radius rho of the FEDVR one-body operator is 1.63e6 Ha here (the
l(l+1)/2r^2 term at l = 31 on the first Gauss-Lobatto node), so the step
uses dt = 1e-6 a.u.; the work per step does not depend on dt.
One propagation = 2000 steps.

The timed region runs the steps the way a propagation runs them
(td1c_propagate): the stage fields are evaluated on the device and no host
round trip happens between steps.  The per-kernel breakdown comes from an
untimed pass with HIP events around every kernel; the dominant kernel's
average launch time comes from a second timed region with events around
that kernel only.  Both the FP64 path and the mixed path (complex64 mean
field) are measured; `value` is the mixed path's rate when its dipole stays
within 1e-6 of the FP64 propagation (checked over 200 steps), else FP64's.

Prints one JSON line (rank 0).  --gpus above 1 is not provided.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
NSTEP_PROP = 2000
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "ne_l31_traffic.json")
# device kernels behind each timing name (HBM counters are per device kernel)
TIMING_KERNELS = {
    "poisson": ["k_poisson_bridge_lds", "k_poisson_interior"],
    "onebody": ["k_onebody_el"],
    "mfprod": ["k_mfprod_quad"],
    "island_a": ["k_island_a"],
    "island_b": ["k_island_b"],
    "dtorb_rk": ["k_dtorb"],
}


def source_hash() -> str:
    """Hash of the HIP sources, to tell whether recorded counters still
    describe the kernels being run."""
    h = hashlib.sha256()
    d = os.path.join(ROOT, "td1c_amd", "csrc")
    for name in sorted(os.listdir(d)):
        if name.endswith((".hip", ".hpp", ".cpp")):
            with open(os.path.join(d, name), "rb") as f:
                h.update(name.encode())
                h.update(f.read())
    return h.hexdigest()[:16]


def recorded_traffic(timing_name: str, path: str):
    """HBM bytes per launch of a timed kernel of the fp64 or mixed path from
    the recorded rocprofv3 FETCH_SIZE / WRITE_SIZE counters (scaled by the
    copy-kernel calibration), or None when no record matches the current
    sources."""
    try:
        with open(TRAFFIC_FILE) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    if rec.get("source_hash") != source_hash() or timing_name not in TIMING_KERNELS:
        return None
    total = 0.0
    for k in TIMING_KERNELS[timing_name]:
        if k not in rec.get(path, {}):
            return None
        e = rec[path][k]
        total += e["fetch_kib"] * 1024.0 * rec["copy_factor"]["FETCH_SIZE"]
        total += e["write_kib"] * 1024.0 * rec["copy_factor"]["WRITE_SIZE"]
    return total


def copy_probe(nbytes: int, reps: int = 5):
    """A plain device copy of nbytes, the calibration of the HBM counters."""
    import torch

    n = nbytes // 16
    src = torch.ones(n, dtype=torch.complex128, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 16, "reps": reps}), flush=True)


def kernel_bytes(name: str, s: dict, cm: int = 16) -> float:
    """Algorithmic HBM bytes of one launch of a named kernel: every input
    read once, every output written once, operator tables excluded.  State
    and l-space arrays are complex128 (16 B); the theta-grid and pair arrays
    of the mean field take cm bytes per complex value (16, or 8 when they
    are complex64)."""
    nr, L1, L2, nl = s["nr"], s["L1"], s["L2"], s["nlat"]
    nf, nd, npair, nfc = s["nfun"], s["ndyn"], s["npair"], s["nfcore"]
    nb, ng, c = nr * L1, nr * nl, 16
    table = {
        # with the velocity-gauge frozen-core phase carried by the same launch
        "l2theta": nd * (nb * c + ng * cm) + nfc * ng * (c + cm),
        "theta2l_fc": nfc * (ng * cm + nb * c),
        "onebody": 2 * nf * nb * c,
        "pair_theta2L": (nf * ng + npair * L2 * nr) * cm,
        "poisson": 2 * npair * L2 * nr * cm,
        "pair_L2theta": npair * (L2 * nr + ng) * cm,
        "mfprod": (nf * ng + npair * ng + 2 * nd * ng) * cm,
        "theta2l_mf": 2 * nd * (ng * cm + nb * c),
        "gram_l": (nf + nfc + 4 * nd) * nb * c,
        "gram_pair": 2 * s["nactpair"] * L2 * nr * cm,
        # k: read h01, g, v2, orb of nf and all orbitals for X; acc r+w; w0 r; next w; k w; earlier k r
        "dtorb_rk": (nd * 3 + nf + nf * 2 + nf + nf + nf + 1.5 * nf) * nb * c,
        "dtorb": (nd * 3 + nf + nf) * nb * c,
        "island_a": 0.0,
        "island_b": 0.0,
    }
    return float(table.get(name, 0.0))


def hprod_bytes(s: dict, c: int = 16) -> float:
    """Compulsory traffic of one hprod in the stage-separated dataflow:
    2 No 16 + 2 No c + 4 Ng c + 4 Np c + 4 lcic 16."""
    No = s["nr"] * s["L1"] * s["nfun"]
    Ng = s["nr"] * s["nlat"] * s["ndyn"]
    Np = s["nr"] * s["L2"] * s["npair"]
    return 2 * No * 16 + 2 * No * c + 4 * Ng * c + 4 * Np * c + 4 * s["lcic"] * 16


def cpu_baseline(prob, wfn, lfield, seconds: float = 12.0, name: str = "ne_l31") -> dict:
    """The FP64 oracle (single-threaded C restatement of hprod_htot) timed on
    this host on a bounded sample of the same workload."""
    from oracle.oracle import Oracle

    o = Oracle(prob)
    o.set_wfn0(wfn)
    o.htot(1.0, lfield, wfn)  # warm
    n, t0 = 0, time.perf_counter()
    while True:
        o.htot(1.0, lfield, wfn)
        n += 1
        el = time.perf_counter() - t0
        if el > seconds or n >= 50:
            break
    return {"value": n / el, "unit": "hprod evals/s", "cores": 1, "kind": "port",
            "sample": f"{n} hprod evaluations of {name} (FP64 oracle, same inputs) in {el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # ~0.5 s of timed steps per path after ~0.1 s of warm-up, so that the
    # clocks have settled and one slow launch does not move the rate
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--dt", type=float, default=1.0e-6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--only", choices=["fp64", "mixed"], default=None,
                    help="measure one path only (for profiler runs; no dipole check)")
    ap.add_argument("--workload", default="ne_l31",
                    help="ne_l31 (default, the headline), he_l15, ar_l63 or mf16_l63; the l63 grids need "
                         "--dt 2e-7 for a stable explicit step")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a device copy of this many bytes (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not provided; this build runs one propagation set on one GPU",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
        return

    import torch

    import td1c_amd
    from td1c_amd import workloads as W

    if args.workload not in W.WORKLOADS:
        print(f"bench.py: unknown workload {args.workload}", file=sys.stderr)
        sys.exit(2)
    make_prob, windex = W.WORKLOADS[args.workload]
    seed = windex * 1000 + 7
    field = W.ne_field()
    if args.workload != "ne_l31" and make_prob().igauge == 0:
        field = td1c_amd.Field(env_type="sin2", gauge="length1", fint=field.fint, wlen=field.wlen,
                               numcyc=field.numcyc)
    rk = td1c_amd.CRK(4)

    def measure(precision: int) -> dict:
        prob = make_prob(precision=precision)
        hp = td1c_amd.HProd(prob)
        wfn = W.make_wfn(prob, seed, hp.lcic)
        hp.set_wfn0(wfn)
        d = hp.to_device(wfn)
        nprof = max(10, min(args.steps, 50))
        t = 0.0
        # warm-up: the first call captures one RK step as a HIP graph
        rk.propagate(hp, d, t, args.dt, args.warmup, field)
        t += args.warmup * args.dt
        torch.cuda.synchronize()
        # untimed profiling pass (stream launches): events around every kernel
        hp.set_timing(True)
        for _ in range(nprof):
            rk.prop(hp, d, t, args.dt, field=field)
            t += args.dt
        breakdown = hp.get_timing()
        hp.set_timing(False)
        dname = max(breakdown.items(), key=lambda kv: kv[1][0])[0]
        # timed region: device-driven steps, graph replays
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rk.propagate(hp, d, t, args.dt, args.steps, field)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        t += args.steps * args.dt
        # timed region with HIP events around the dominant kernel only (stream
        # launches on the same stream), for its average launch time
        lfs = [rk.stage_fields(field, t + i * args.dt, args.dt) for i in range(args.steps)]
        hp.set_timing(True, only=dname)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        for i in range(args.steps):
            rk.prop(hp, d, t + i * args.dt, args.dt, lfield_stages=lfs[i])
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        times = hp.get_timing()
        hp.set_timing(False)
        norm = float(torch.linalg.vector_norm(d[: hp.norb]).item() ** 2)
        if not np.isfinite(norm):
            raise RuntimeError("propagation produced non-finite values")
        return dict(prob=prob, hp=hp, wfn=wfn, lf0=lfs[0][0], ms_step=(t1 - t0) * 1e3 / args.steps,
                    ms_stream=(t3 - t2) * 1e3 / args.steps, breakdown=breakdown, nprof=nprof, dname=dname,
                    dtime=times[dname], norm=norm)

    def dipole_series(precision: int, nchunk: int = 10, chunk: int = 20) -> np.ndarray:
        """<z>(t) every `chunk` steps of a propagation from the same state."""
        prob = make_prob(precision=precision)
        hp = td1c_amd.HProd(prob)
        wfn = W.make_wfn(prob, seed, hp.lcic)
        hp.set_wfn0(wfn)
        d = hp.to_device(wfn)
        out, t = [], 0.0
        for _ in range(nchunk):
            rk.propagate(hp, d, t, args.dt, chunk, field)
            t += chunk * args.dt
            out.append(hp.dipole(td1c_amd.field_value(field, t), d))
        return np.array(out)

    if args.only is not None:
        r = measure(0 if args.only == "fp64" else 1)
        print(json.dumps({"only": args.only, "ms_per_step": r["ms_step"],
                          "hprod_evals_per_s": 4.0 / (r["ms_step"] * 1e-3)}), flush=True)
        return
    fp64 = measure(0)
    mixed = measure(1)
    dseries = {0: dipole_series(0), 1: dipole_series(1)}
    dip_err = float(np.abs(dseries[1] - dseries[0]).max())
    use = mixed if dip_err <= 1e-6 else fp64
    cm = 8 if use is mixed else 16
    prob, hp = use["prob"], use["hp"]

    s = dict(nr=prob.nr, L1=prob.lmax1 + 1, L2=prob.lmax2 + 1, nlat=prob.lmax2 + 1, nfun=prob.nfun,
             ndyn=prob.nfun - prob.nfcore, npair=hp.npair, nfcore=prob.nfcore, lcic=hp.lcic,
             nactpair=prob.nact * (prob.nact + 1) // 2)
    dname = use["dname"]
    dms, dcount = use["dtime"]
    avg_s = dms / dcount * 1e-3
    abytes = kernel_bytes(dname, s, cm)
    achieved = abytes / avg_s / 1e9 if abytes > 0 else 0.0
    ms_step = use["ms_step"]
    hprods_per_s = 4.0 / (ms_step * 1e-3)
    step_bytes = 4 * hprod_bytes(s, cm) + 15 * (hp.wfn_size * 16)
    nprof = use["nprof"]
    result = {
        "metric": "hprod evals/s + propagations/s, mixed vs FP64 at dipole err <=1e-6; %HBM peak",
        "value": hprods_per_s,
        "unit": "hprod evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 mean field / c128 state" if use is mixed else "c128",
        "data": "synthetic (seeded orbitals, one CI determinant)",
        "config": {"workload": ("ne_l31: Ne-like RK4, 80x21 FEDVR (1599 pts) on [0,160], lmax1 31/lmax2 62, "
                                "Z=10, 1 frozen core + 4 active, 800 nm sin2 1e14 W/cm2 velocity gauge")
                   if args.workload == "ne_l31" else
                   (f"{args.workload}: RK4, {prob.nfe}x{prob.ndvr + 1} FEDVR ({prob.nr} pts) on [0,{prob.rmax:g}], "
                    f"lmax1 {prob.lmax1}/lmax2 {prob.lmax2}, Z={prob.znuc:g}, {prob.nfun} orbitals "
                    f"({prob.nfcore} frozen, {prob.nact} active), 800 nm sin2 1e14 W/cm2 "
                    f"{'velocity' if prob.igauge == 1 else 'length'} gauge"),
                   "precision": "mixed" if use is mixed else "fp64", "rk_order": 4, "dt": args.dt,
                   "steps_per_propagation": NSTEP_PROP, "parallelism": "single GPU"},
        "propagations_per_s": 1.0 / (NSTEP_PROP * ms_step * 1e-3),
        "hbm_frac_step": step_bytes / (ms_step * 1e-3) / (HBM_PEAK_GBS * 1e9),
        "fp64": {"hprod_evals_per_s": 4.0 / (fp64["ms_step"] * 1e-3), "ms_per_step": fp64["ms_step"],
                 "propagations_per_s": 1.0 / (NSTEP_PROP * fp64["ms_step"] * 1e-3)},
        "mixed": {"hprod_evals_per_s": 4.0 / (mixed["ms_step"] * 1e-3), "ms_per_step": mixed["ms_step"],
                  "propagations_per_s": 1.0 / (NSTEP_PROP * mixed["ms_step"] * 1e-3),
                  "dipole_max_abs_err_vs_fp64": dip_err,
                  "dipole_check": "<z> every 20 steps over 200 steps, mixed vs fp64, same initial state"},
        "roofline": {"bound": "hbm", "kernel": dname, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (recorded_traffic(dname, "mixed" if use is mixed else "fp64")
                                 if args.workload == "ne_l31" else None),
                     "algorithmic_bytes_per_launch": abytes, "avg_launch_us": avg_s * 1e6},
        "ms_per_step_stream_launch": use["ms_stream"],
        "kernel_ms_per_step": {k: v[0] / nprof for k, v in sorted(use["breakdown"].items(), key=lambda kv: -kv[1][0])},
        "kernel_us_per_launch": {k: 1e3 * v[0] / v[1]
                                 for k, v in sorted(use["breakdown"].items(), key=lambda kv: -kv[1][0])},
        "norm_after": use["norm"],
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(fp64["prob"], fp64["wfn"], fp64["lf0"], args.cpu_seconds,
                                              args.workload)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
