"""Benchmark: split-operator RK4 propagation of the Ar-like workload with the
gfx950 hprod, FP64 and mixed precision.

Workload ar_l63 (the default): 120 FEDVR elements x 21 Gauss-Lobatto points
on [0,240] a.u. (2399 radial points), lmax1 63 / lmax2 126, Z = 18, nine
doubly occupied orbitals (TDHF-like), 800 nm sin^2 pulse of 4 cycles at
1.6e14 W/cm^2 in the velocity gauge, dt = 0.01, 5000 steps per propagation.

A step is one step of td1c's split2 propagation with split_type = 1: the
clh1itr one-body half step (field at t + dt/4), a classic RK4 step of the
split hprod (4 hprod evaluations), the second half step (t + 3dt/4), then
the radial mask (tdse_split2, src/main/tdse.cpp:225-266).  The timed region
runs the steps as a propagation runs them (td1c_propagate: fields on the
device, no host round trip), with the state resident in HBM.

`value` is the FP64 path's hprod evaluations per second inside that
propagation (4 per step; the step time includes the two half steps).  Beside
it: the mixed path (complex64 mean field) with its hprod residual
||k_mixed - k_fp64|| / ||k_fp64|| on the same state, the dipole <z>(t) of
both paths over a 2000-step run from the same state (max difference,
absolute and relative to max |<z>|) and the ionisation probabilities P(k
electrons beyond 20 a.u.) after it; the 16-orbital mean-field workload's
hprod rate (FP64 and mixed, with its residual); the per-kernel breakdown;
the roofline of the dominant kernel; the OpenMP oracle on the host cores.

Prints one JSON line.  --gpus above 1 is not provided.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import math
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# MI355X dense matrix-core peaks (TFLOP/s)
MFMA_PEAK = {"f64": 78.6, "f32": 157.3}
# kernels on the main stream (the critical path of an hprod); the others run
# beside them on the side streams, where concurrency inflates their durations
MAIN_STREAM = {"l2theta", "pair_theta2L", "poisson_bridge", "poisson_interior", "pair_L2theta", "mfprod",
               "theta2l_mf", "gram_l", "dtorb_rk", "dtorb"}
# steps of one propagation per workload (the configurations' run lengths)
STEPS_PER_PROP = {"he_l15": 3000, "ne_l31": 2000, "ar_l63": 5000, "mf16_l63": 5000, "he_scan": 3000}
FINT = {"ne_l31": 1.0e14, "ar_l63": 1.6e14}
# timing names that are a sequence of launches, not one kernel
COMPOUND = {"h1itr"}


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


def recorded_traffic(workload: str, timing_name: str, path: str):
    """HBM bytes per launch of a timed kernel from the recorded rocprofv3
    FETCH_SIZE / WRITE_SIZE counters (scaled by the copy-kernel
    calibration), or None when no record matches the current sources."""
    fname = os.path.join(ROOT, "profiles", f"{workload}_traffic.json")
    try:
        with open(fname) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    if rec.get("source_hash") != source_hash():
        return None
    kern = rec.get("timing_kernels", {}).get(timing_name)
    if not kern:
        return None
    total = 0.0
    for k in kern:
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
    """Algorithmic HBM bytes of one launch of a named kernel: each of its
    inputs read once, each output written once, operator tables excluded.
    State and l-space arrays are complex128 (16 B); the theta-grid and pair
    arrays of the mean field take cm bytes per complex value (16, or 8 when
    they are complex64)."""
    nr, L1, L2, nl = s["nr"], s["L1"], s["L2"], s["nlat"]
    nf, nd, npair, nfc = s["nfun"], s["ndyn"], s["npair"], s["nfcore"]
    nb, ng, c = nr * L1, nr * nl, 16
    table = {
        # with the velocity-gauge frozen-core phase carried by the same launch
        "l2theta": nd * (nb * c + ng * cm) + nfc * ng * (c + cm),
        "theta2l_fc": nfc * (ng * cm + nb * c),
        "onebody": 2 * nf * nb * c,
        "pair_theta2L": (nf * ng + npair * L2 * nr) * cm,
        "poisson_bridge": npair * L2 * nr * cm,
        "poisson_interior": 2 * npair * L2 * nr * cm,
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


def kernel_flops(name: str, s: dict) -> float:
    """Arithmetic of one launch (real flops; a complex-by-real multiply-add
    is 4): the Legendre transforms on the matrix cores and the Poisson
    element products."""
    nr, L1, L2, nl = s["nr"], s["L1"], s["L2"], s["nlat"]
    nd, npair, nfc = s["ndyn"], s["npair"], s["nfcore"]
    table = {
        "l2theta": 4.0 * nr * L1 * nl * (nd + nfc),
        "pair_theta2L": 4.0 * nr * nl * L2 * npair + 6.0 * nr * nl * npair,
        "pair_L2theta": 4.0 * nr * L2 * nl * npair,
        "theta2l_mf": 2 * 4.0 * nr * nl * L1 * nd,
        "poisson_interior": 4.0 * (s["ndvr"] - 1) * npair * L2 * nr,
    }
    return float(table.get(name, 0.0))


def hprod_bytes(s: dict, c: int = 16) -> float:
    """Compulsory traffic of one hprod in the stage-separated dataflow:
    2 No 16 + 2 No c + 4 Ng c + 4 Np c + 4 lcic 16."""
    No = s["nr"] * s["L1"] * s["nfun"]
    Ng = s["nr"] * s["nlat"] * s["ndyn"]
    Np = s["nr"] * s["L2"] * s["npair"]
    return 2 * No * 16 + 2 * No * c + 4 * Ng * c + 4 * Np * c + 4 * s["lcic"] * 16


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_threads() -> int:
    v = os.environ.get("OMP_NUM_THREADS")
    if v and v.isdigit() and int(v) > 0:
        return int(v)
    return os.cpu_count() or 1


def cpu_baseline(prob, wfn, lfield, seconds: float, name: str) -> dict:
    """The FP64 oracle (C restatement of hprod_htot, built with OpenMP over
    independent output rows) timed on this host's cores on a bounded sample
    of the same workload: hprod evaluations of the same state."""
    from oracle import oracle as O

    threads = cpu_threads()
    os.environ["OMP_NUM_THREADS"] = str(threads)
    o = O.Oracle(prob, openmp=True)
    o.set_wfn0(wfn)
    o.htot(1.0, lfield, wfn)  # warm
    n, t0 = 0, time.perf_counter()
    while True:
        o.htot(1.0, lfield, wfn)
        n += 1
        el = time.perf_counter() - t0
        if el > seconds or n >= 50:
            break
    return {"value": n / el, "unit": "hprod evals/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(),
            "sample": f"{n} hprod evaluations of {name} (FP64 oracle, OpenMP, {threads} threads, "
                      f"same state and split_type) in {el:.1f} s"}


def scan_batch(n: int, dt: float, warmup: int, steps: int, gate_steps: int) -> dict:
    """The scan workload (n He TDHF problems of uneven grids, lmax and pulses)
    propagated as one batch (td1c_batch_propagate, one launch per kernel site
    for all members, one captured graph per step) in both precisions, against
    the plain FP64 path: each member alone through td1c_propagate."""
    import torch

    import td1c_amd
    from td1c_amd import workloads as W

    tune = os.environ.get("TD1C_AUTOTUNE")
    os.environ["TD1C_AUTOTUNE"] = "0"  # the batch chooses its own launch shapes
    rk = td1c_amd.CRK(4)
    out = {}
    members = {}
    try:
        for p in (1, 0):
            probs, fields = W.he_scan(n, precision=p, isplit=1)
            t0 = time.perf_counter()
            hps = [td1c_amd.HProd(pr) for pr in probs]
            wfns = [W.bound_wfn(pr, 4007 + i, hp.lcic) for i, (pr, hp) in enumerate(zip(probs, hps))]
            for hp, w in zip(hps, wfns):
                hp.set_wfn0(w)
            create_s = time.perf_counter() - t0
            b = td1c_amd.Batch(hps)
            t0 = time.perf_counter()
            b.autotune([hp.to_device(w) for hp, w in zip(hps, wfns)], dt, fields)
            tune_s = time.perf_counter() - t0
            d = [hp.to_device(w) for hp, w in zip(hps, wfns)]
            t = 0.0
            b.propagate(d, t, dt, warmup, fields)
            t += warmup * dt
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            b.propagate(d, t, dt, steps, fields)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3 / steps
            t += steps * dt
            if gate_steps > steps + warmup:
                b.propagate(d, t, dt, gate_steps - steps - warmup, fields)
                t = gate_steps * dt
            dip = np.array([hp.dipole(td1c_amd.field_value(f, t), x) for hp, f, x in zip(hps, fields, d)])
            info = b.info()
            step_bytes = 0.0
            for pr, hp in zip(probs, hps):
                s = dict(nr=pr.nr, L1=pr.lmax1 + 1, L2=pr.lmax2 + 1, nlat=pr.lmax2 + 1, nfun=pr.nfun,
                         ndyn=pr.nfun, npair=hp.npair, nfcore=0, lcic=hp.lcic)
                step_bytes += 4 * hprod_bytes(s, 8 if p == 1 else 16) + 15 * hp.wfn_size * 16
            out[p] = {"ms_per_step": ms, "propagations_per_s": n / (STEPS_PER_PROP["he_scan"] * ms * 1e-3),
                      "hbm_frac_step": step_bytes / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9),
                      "bytes_per_step": step_bytes, "create_s": create_s, "autotune_s": tune_s, "info": info,
                      "dipole": dip}
            progress(f"scan batch precision {p}: {ms:.3f} ms per batch step ({info['sites_per_step']} launches)")
            b.close()
            members[p] = (probs, fields, hps, wfns)
        # the plain FP64 path: every member alone, stream launches, its own
        # autotuned shapes
        os.environ["TD1C_AUTOTUNE"] = tune if tune is not None else "1"
        probs, fields, _, wfns = members[0]
        alone = []
        for i, (pr, f, w) in enumerate(zip(probs, fields, wfns)):
            hp = td1c_amd.HProd(pr)
            hp.set_wfn0(w)
            x = hp.to_device(w)
            rk.propagate(hp, x, 0.0, dt, 2, f)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rk.propagate(hp, x, 2 * dt, dt, 4, f)
            torch.cuda.synchronize()
            alone.append((time.perf_counter() - t0) * 1e3 / 4)
            hp.close()
        for p in (0, 1):
            for hp in members[p][2]:
                hp.close()
    finally:
        if tune is None:
            os.environ.pop("TD1C_AUTOTUNE", None)
        else:
            os.environ["TD1C_AUTOTUNE"] = tune
    alone_ms = float(np.sum(alone))
    d0, d1 = out[0]["dipole"], out[1]["dipole"]
    err = np.abs(d1 - d0)
    return {
        "workload": (f"he_scan: {n} He TDHF problems, nfe 40-120 x 21 points (rmax = 2 nfe), lmax1 in "
                     "{15, 31, 47, 63} (lmax2 = 2 lmax1), velocity-gauge sin2 pulses of 4 cycles, 400-1600 nm, "
                     "1e13-1e15 W/cm2, random CEP; split2 steps at dt " + f"{dt:g}"),
        "members": n,
        "steps_per_propagation": STEPS_PER_PROP["he_scan"],
        "mixed_batched": {k: v for k, v in out[1].items() if k != "dipole"},
        "fp64_batched": {k: v for k, v in out[0].items() if k != "dipole"},
        "fp64_alone": {"ms_per_step_all_members": alone_ms,
                       "propagations_per_s": n / (STEPS_PER_PROP["he_scan"] * alone_ms * 1e-3),
                       "ms_per_step_min_max": [float(np.min(alone)), float(np.max(alone))]},
        "speedup_mixed_batched_vs_fp64_alone": alone_ms / out[1]["ms_per_step"],
        "speedup_fp64_batched_vs_fp64_alone": alone_ms / out[0]["ms_per_step"],
        "dipole_gate": {"steps": max(gate_steps, steps + warmup), "dt": dt,
                        "max_abs_err": float(err.max()),
                        "max_err_over_max_dipole": float(err.max() / max(np.abs(d0).max(), 1e-300)),
                        "max_rel_err_per_member": float(np.max(err / np.maximum(np.abs(d0), 1e-12)))},
    }


def strict(x):
    """x with every non-finite float replaced by None, so that the line is
    strict JSON."""
    if isinstance(x, dict):
        return {k: strict(v) for k, v in x.items()}
    if isinstance(x, (list, tuple)):
        return [strict(v) for v in x]
    if isinstance(x, float) and not math.isfinite(x):
        return None
    return x


def progress(msg: str):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--dt", type=float, default=0.01)
    ap.add_argument("--workload", default="ar_l63", help="ar_l63 (default), ne_l31 or he_l15")
    ap.add_argument("--gate-steps", type=int, default=2000,
                    help="steps of the mixed-vs-FP64 dipole comparison (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the 16-orbital mean-field hprod rates")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--scan-members", type=int, default=64, help="problems of the batched scan (0: skip it)")
    ap.add_argument("--scan-steps", type=int, default=100, help="timed batch steps of the scan")
    ap.add_argument("--scan-gate-steps", type=int, default=1000,
                    help="steps after which the scan's mixed and FP64 dipoles are compared")
    ap.add_argument("--etd-steps", type=int, default=20, help="timed cletdrb steps on the workload (0: skip)")
    ap.add_argument("--only", choices=["fp64", "mixed", "batch", "etdrb"], default=None,
                    help="measure one path only (for profiler runs; no comparisons)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a device copy of this many bytes (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not provided; this build runs on one GPU "
              "(independent propagations shard over GPUs with no communication)", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
        return

    import torch

    import td1c_amd
    from td1c_amd import workloads as W

    if args.workload not in ("ar_l63", "ne_l31", "he_l15"):
        print(f"bench.py: unknown workload {args.workload}", file=sys.stderr)
        sys.exit(2)
    make_prob, windex = W.WORKLOADS[args.workload]
    seed = windex * 1000 + 7
    field = td1c_amd.Field(env_type="sin2", gauge="velocity1", fint=FINT.get(args.workload, 1.0e14), wlen=800.0,
                           numcyc=4.0)
    rk = td1c_amd.CRK(4)
    dt = args.dt
    nprop = STEPS_PER_PROP[args.workload]

    def context(precision: int):
        prob = make_prob(precision=precision, isplit=1, igauge=1)
        hp = td1c_amd.HProd(prob)
        wfn = W.bound_wfn(prob, seed, hp.lcic)
        hp.set_wfn0(wfn)
        return prob, hp, wfn

    def measure(precision: int) -> dict:
        prob, hp, wfn = context(precision)
        d = hp.to_device(wfn)
        nprof = max(5, min(args.steps, 20))
        t = 0.0
        rk.propagate(hp, d, t, dt, args.warmup, field)
        t += args.warmup * dt
        torch.cuda.synchronize()
        # untimed pass with HIP events around every kernel (stream launches)
        hp.set_timing(True)
        rk.propagate(hp, d, t, dt, nprof, field)
        t += nprof * dt
        torch.cuda.synchronize()
        breakdown = hp.get_timing()
        hp.set_timing(False)
        single = {k: v for k, v in breakdown.items() if k in MAIN_STREAM}
        dname = max(single.items(), key=lambda kv: kv[1][0])[0]
        # the timed region
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rk.propagate(hp, d, t, dt, args.steps, field)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        t += args.steps * dt
        # a second pass with events around the dominant kernel only, on the
        # stream it is launched on, for its average launch time
        hp.set_timing(True, only=dname)
        rk.propagate(hp, d, t, dt, nprof, field)
        torch.cuda.synchronize()
        times = hp.get_timing()
        hp.set_timing(False)
        norm = float(torch.linalg.vector_norm(d[: hp.norb]).item() ** 2)
        if not np.isfinite(norm):
            raise RuntimeError("propagation produced non-finite values")
        return dict(prob=prob, hp=hp, wfn=wfn, ms_step=(t1 - t0) * 1e3 / args.steps, breakdown=breakdown,
                    nprof=nprof, dname=dname, dtime=times[dname], norm=norm)

    def measure_etdrb(hp, wfn, steps: int, compare_steps: int = 0, scheme=None) -> dict:
        """An exponential propagator (cletdrb, the reference's default
        prop_type, unless scheme gives another) on the same context: ms per
        step after two warm-up steps, the event breakdown of one step, and
        the dipole after compare_steps steps against split2's."""
        etd = (scheme or td1c_amd.ETDRB)(hp, dt)
        d = hp.to_device(wfn)
        etd.propagate(d, 0.0, 2, field)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        etd.propagate(d, 2 * dt, steps, field)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        hp.set_timing(True)
        etd.propagate(d, (2 + steps) * dt, 1, field)
        torch.cuda.synchronize()
        bd = hp.get_timing()
        hp.set_timing(False)
        nh = 1.0 if scheme is td1c_amd.EXPAB else 4.0  # split hprods per step
        out = {"ms_per_step": ms, "hprod_evals_per_s": nh / (ms * 1e-3),
               "propagations_per_s": 1.0 / (nprop * ms * 1e-3),
               "event_ms_one_step": {k: round(v[0], 4) for k, v in bd.items()}}
        if compare_steps > 0:
            a = hp.to_device(wfn)
            b = hp.to_device(wfn)
            etd = (scheme or td1c_amd.ETDRB)(hp, dt)  # a fresh history for clexpab
            etd.propagate(a, 0.0, compare_steps, field)
            rk.propagate(hp, b, 0.0, dt, compare_steps, field)
            torch.cuda.synchronize()
            lf = td1c_amd.field_value(field, compare_steps * dt)
            da, db = float(hp.dipole(lf, a)), float(hp.dipole(lf, b))
            fin = bool(np.isfinite(da) and torch.isfinite(a).all().item())
            # non-finite values are reported as null (the line stays strict JSON)
            out["dipole_vs_split2"] = {"steps": compare_steps, "finite": fin, "this": da if fin else None,
                                       "split2": db, "abs_diff": abs(da - db) if fin else None}
        return out

    if args.only == "batch":
        r = scan_batch(args.scan_members, dt, args.warmup, args.scan_steps, args.scan_gate_steps)
        print(json.dumps({"only": "batch", **r}), flush=True)
        return

    if args.only == "etdrb":
        _, hp, wfn = context(0)
        print(json.dumps({"only": "etdrb", "workload": args.workload, **measure_etdrb(hp, wfn, args.etd_steps)}),
              flush=True)
        return

    if args.only is not None:
        r = measure(0 if args.only == "fp64" else 1)
        n = r["nprof"]
        print(json.dumps({"only": args.only, "workload": args.workload, "ms_per_step": r["ms_step"],
                          "hprod_evals_per_s": 4.0 / (r["ms_step"] * 1e-3),
                          "kernel_us_per_launch": {k: round(1e3 * v[0] / v[1], 2) for k, v in r["breakdown"].items()},
                          "kernel_ms_per_step": {k: round(v[0] / n, 4) for k, v in r["breakdown"].items()},
                          "shapes": r["hp"].shapes()}), flush=True)
        return

    fp64 = measure(0)
    progress(f"fp64 path {fp64['ms_step']:.3f} ms/step")
    mixed = measure(1)
    progress(f"mixed path {mixed['ms_step']:.3f} ms/step")

    # hprod residual on one state: the split hprod of both paths on the
    # initial state with the field at the pulse peak
    lf0 = td1c_amd.field_value(field, 0.0)
    lfp = td1c_amd.field_value(field, 0.5 * field.numcyc * td1c_amd.field_period(field))
    hps = {0: fp64["hp"], 1: mixed["hp"]}
    ks = {}
    for p in (0, 1):
        dw = hps[p].to_device(fp64["wfn"])
        hw = hps[p].htot(1.0, lfp, dw)
        torch.cuda.synchronize()
        ks[p] = hw.cpu().numpy().copy()
    residual = float(np.linalg.norm(ks[1] - ks[0]) / np.linalg.norm(ks[0]))
    progress(f"hprod residual {residual:.3e} (norms {np.linalg.norm(ks[0]):.6e}, {np.linalg.norm(ks[1]):.6e}; "
             f"precisions {fp64['prob'].precision}, {mixed['prob'].precision})")

    # dipole(t) and final ionisation probabilities, mixed against FP64
    gate = None
    if args.gate_steps > 0:
        every = max(1, args.gate_steps // 40)
        series, ipx = {}, {}
        for p in (0, 1):
            hp = hps[p]
            d = hp.to_device(fp64["wfn"])
            out, t = [], 0.0
            for _ in range(args.gate_steps // every):
                rk.propagate(hp, d, t, dt, every, field)
                t += every * dt
                out.append(hp.dipole(td1c_amd.field_value(field, t), d))
            series[p] = np.array(out)
            ipx[p] = hp.ipx(2, 20.0, td1c_amd.field_value(field, t), d)
            progress(f"dipole series of path {p} done")
        derr = np.abs(series[1] - series[0])
        gate = {"steps": args.gate_steps, "dt": dt, "sampled_every": every,
                "dipole_max_abs_err": float(derr.max()),
                "dipole_max_rel_err": float(derr.max() / max(np.abs(series[0]).max(), 1e-300)),
                "dipole_fp64_final": float(series[0][-1]),
                "ipx_fp64": [float(x) for x in ipx[0]], "ipx_mixed": [float(x) for x in ipx[1]],
                "ipx_max_abs_err": float(np.abs(np.array(ipx[1]) - np.array(ipx[0])).max()),
                "ipx_radius": 20.0}

    # the host-pointer entry point hprod_htot_ (wave function copied in and
    # the derivative copied out on every call, pageable host memory)
    host_entry = None
    if not args.no_extra:
        import ctypes

        hp = fp64["hp"]
        hp.bind()
        lib = td1c_amd.load_library()
        w0 = fp64["wfn"]
        orb = np.ascontiguousarray(w0[:hp.norb])
        cic = np.ascontiguousarray(w0[hp.norb:])
        hw, hc = np.zeros_like(orb), np.zeros_like(cic)
        dtc = ctypes.c_double(1.0)
        lfh = np.ascontiguousarray(lfp)
        ptrs = [a.ctypes.data_as(ctypes.c_void_p) for a in (lfh, orb, cic, hw, hc)]
        for _ in range(2):
            lib.hprod_htot_(ctypes.byref(dtc), *ptrs)
        nh, t0 = 10, time.perf_counter()
        for _ in range(nh):
            lib.hprod_htot_(ctypes.byref(dtc), *ptrs)
        host_entry = {"hprod_evals_per_s": nh / (time.perf_counter() - t0),
                      "bytes_over_pcie_per_call": 2 * (orb.nbytes + cic.nbytes),
                      "entry": "hprod_htot_ (host arrays in pageable memory, synchronous)"}
        progress(f"host entry {host_entry['hprod_evals_per_s']:.1f} hprod evals/s")

    # the 16-orbital mean-field workload: hprod rates and residual
    extra = None
    if not args.no_extra:
        rates, kk = {}, {}
        for p in (0, 1):
            prob5 = W.mf16_l63(precision=p)
            hp5 = td1c_amd.HProd(prob5)
            w5 = W.bound_wfn(prob5, 5007, hp5.lcic)
            hp5.set_wfn0(w5)
            d5 = hp5.to_device(w5)
            h5 = torch.empty_like(d5)
            lf5 = np.array([0.0, 0.0, 0.05, 0.0, 0.0, 0.05, 0.0, 0.0, 0.0])
            for _ in range(3):
                hp5.htot(1.0, lf5, d5, h5)
            torch.cuda.synchronize()
            n5 = 20
            t0 = time.perf_counter()
            for _ in range(n5):
                hp5.htot(1.0, lf5, d5, h5)
            torch.cuda.synchronize()
            rates[p] = n5 / (time.perf_counter() - t0)
            kk[p] = h5.cpu().numpy()
            hp5.close()
        progress("mean-field workload done")
        extra = {"workload": "mf16_l63: 16 orbitals (136 pair densities), 100x21 FEDVR (1999 pts) on [0,200], "
                             "lmax1 63/lmax2 126, one hprod per call",
                 "fp64_hprod_evals_per_s": rates[0], "mixed_hprod_evals_per_s": rates[1],
                 "hprod_residual": float(np.linalg.norm(kk[1] - kk[0]) / np.linalg.norm(kk[0]))}

    etd = None
    if args.etd_steps > 0:
        etd = {"fp64": measure_etdrb(fp64["hp"], fp64["wfn"], args.etd_steps, compare_steps=100),
               "mixed": measure_etdrb(mixed["hp"], fp64["wfn"], args.etd_steps)}
        etd["split2_ms_per_step_fp64"] = fp64["ms_step"]
        progress(f"etdrb {etd['fp64']['ms_per_step']:.3f} / {etd['mixed']['ms_per_step']:.3f} ms/step")
        etd["others_fp64"] = {}
        for nm, cls in (("etdrk4", td1c_amd.ETDRK4), ("lawrk", td1c_amd.LAWRK), ("expab", td1c_amd.EXPAB)):
            r = measure_etdrb(fp64["hp"], fp64["wfn"], args.etd_steps, compare_steps=100, scheme=cls)
            r.pop("event_ms_one_step", None)
            etd["others_fp64"][nm] = r
            progress(f"{nm} {r['ms_per_step']:.3f} ms/step")

    scan = None
    if args.scan_members > 0:
        scan = scan_batch(args.scan_members, dt, 10, args.scan_steps, args.scan_gate_steps)

    prob, hp = fp64["prob"], fp64["hp"]
    s = dict(nr=prob.nr, L1=prob.lmax1 + 1, L2=prob.lmax2 + 1, nlat=prob.lmax2 + 1, nfun=prob.nfun,
             ndyn=prob.nfun - prob.nfcore, npair=hp.npair, nfcore=prob.nfcore, lcic=hp.lcic,
             nactpair=prob.nact * (prob.nact + 1) // 2, ndvr=prob.ndvr)
    dname = fp64["dname"]
    dms, dcount = fp64["dtime"]
    avg_s = dms / dcount * 1e-3
    abytes = kernel_bytes(dname, s, 16)
    aflops = kernel_flops(dname, s)
    gbs = abytes / avg_s / 1e9 if abytes > 0 else 0.0
    tfs = aflops / avg_s / 1e12
    # the binding resource: the larger fraction of its peak
    if tfs / MFMA_PEAK["f64"] > gbs / HBM_PEAK_GBS:
        roof = {"bound": "mfma", "achieved": tfs, "peak": MFMA_PEAK["f64"], "unit": "TFLOP/s",
                "frac": tfs / MFMA_PEAK["f64"], "hbm_frac": gbs / HBM_PEAK_GBS}
    else:
        roof = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                "mfma_frac": tfs / MFMA_PEAK["f64"]}
    ms_step = fp64["ms_step"]
    step_bytes = 4 * hprod_bytes(s, 16) + 15 * (hp.wfn_size * 16)
    mstep_bytes = 4 * hprod_bytes(s, 8) + 15 * (hp.wfn_size * 16)

    def per_kernel(r):
        n = r["nprof"]
        bd = sorted(r["breakdown"].items(), key=lambda kv: -kv[1][0])
        return ({k: v[0] / n for k, v in bd}, {k: 1e3 * v[0] / v[1] for k, v in bd})

    kms, kus = per_kernel(fp64)
    mkms, mkus = per_kernel(mixed)
    result = {
        "metric": "hprod evals/s + propagations/s, mixed vs FP64 at dipole err <=1e-6; %HBM peak",
        "value": 4.0 / (ms_step * 1e-3),
        "unit": "hprod evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic (seeded bound-state-like orbitals, random-init)",
        "config": {"workload": (f"{args.workload}: split2 step (clh1itr half steps + RK4 of the split hprod + mask), "
                                f"{prob.nfe}x{prob.ndvr + 1} FEDVR ({prob.nr} pts) on [0,{prob.rmax:g}], "
                                f"lmax1 {prob.lmax1}/lmax2 {prob.lmax2}, Z={prob.znuc:g}, {prob.nfun} orbitals "
                                f"({prob.nfcore} frozen, {prob.nact} active), 800 nm sin2 4 cycles "
                                f"{field.fint:.2g} W/cm2 velocity gauge"),
                   "precision": "fp64", "split_type": 1, "rk_order": 4, "dt": dt,
                   "steps_per_propagation": nprop, "parallelism": "single GPU"},
        "propagations_per_s": 1.0 / (nprop * ms_step * 1e-3),
        "hbm_frac_step": step_bytes / (ms_step * 1e-3) / (HBM_PEAK_GBS * 1e9),
        "fp64": {"hprod_evals_per_s": 4.0 / (fp64["ms_step"] * 1e-3), "ms_per_step": fp64["ms_step"],
                 "propagations_per_s": 1.0 / (nprop * fp64["ms_step"] * 1e-3)},
        "mixed": {"hprod_evals_per_s": 4.0 / (mixed["ms_step"] * 1e-3), "ms_per_step": mixed["ms_step"],
                  "propagations_per_s": 1.0 / (nprop * mixed["ms_step"] * 1e-3),
                  "speedup_vs_fp64": fp64["ms_step"] / mixed["ms_step"],
                  "hbm_frac_step": mstep_bytes / (mixed["ms_step"] * 1e-3) / (HBM_PEAK_GBS * 1e9),
                  "hprod_residual_vs_fp64": residual,
                  "accuracy_vs_fp64": gate},
        "roofline": dict(roof, kernel=dname, traffic=recorded_traffic(args.workload, dname, "fp64"),
                         algorithmic_bytes_per_launch=abytes, flops_per_launch=aflops, avg_launch_us=avg_s * 1e6,
                         timing="HIP events around this kernel only, on its launch stream; the dominant "
                                "main-stream kernel of the FP64 step"),
        "kernel_ms_per_step": kms,
        "kernel_us_per_launch": kus,
        "mixed_kernel_ms_per_step": mkms,
        "norm_after": fp64["norm"],
        "mean_field_stress": extra,
        "host_entry": host_entry,
        "scan_batch": scan,
        "etdrb": etd,
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(fp64["prob"], fp64["wfn"], lf0, args.cpu_seconds, args.workload)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(strict(result), allow_nan=False), flush=True)


if __name__ == "__main__":
    main()
