"""Benchmark of the single-spin-flip annealing path on one MI355X.

Default workload (the metric's configuration): dense Sherrington-Kirkpatrick
N=4096, J ~ N(0, 1/N) rounded to FP32, h = 0, 2048 replicas, 500 sweeps,
geometric beta 0.1 -> 10 (500 levels, one sweep each).  A step is one full
anneal of all replicas from random initial states (uploaded before the timed
region).  `value` is flip attempts/s of the batched mixed-precision (FP32)
path; the plain per-replica FP64 path is timed on the same workload for the
ratio.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--config A..E]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # B/s, MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(cfg, sweeps=None, replicas=None):
    from tests import helpers as H

    if cfg == "E":
        n, R, L = 4096, replicas or 2048, sweeps or 500
        return dict(name="dense SK N=4096 gaussian fp32 J, 2048 replicas, 500 sweeps, "
                         "beta 0.1->10 geometric", n=n, R=R,
                    E=H.sk_gauss(n, 0xE005), beta=H.geometric_betas(0.1, 10.0, L),
                    mcs=np.ones(L, np.int64), fast="fp32", plain="fp64", tol_rel=1e-9)
    if cfg == "A":
        n, R, L = 256, replicas or 64, sweeps or 1000
        return dict(name="dense SK N=256 +-1 J, 64 reads, 1000 sweeps, beta 0.1->10",
                    n=n, R=R, E=H.sk_pm1(n, 0xA001), beta=H.geometric_betas(0.1, 10.0, L),
                    mcs=np.ones(L, np.int64), fast="int32", plain="fp64", tol_rel=0.0)
    raise SystemExit(f"unknown config {cfg}")


def prec_code(p):
    from openjij_amd import capi

    return {"fp64": capi.PREC_FP64, "fp32": capi.PREC_FP32, "int32": capi.PREC_INT32}[p]


def time_path(W, prec, steps, warmup, seed0, options=None):
    """Returns per-step wall seconds, per-step kernel ms, energies of the last step."""
    from openjij_amd import capi
    from tests import helpers as H

    ctx = capi.Context.dense(W["E"], W["R"], 1, 0.0, prec_code(prec))
    for k, v in (options or {}).items():
        ctx.set_option(k, v)
    walls, kms = [], []
    energies = None
    for k in range(warmup + steps):
        init = H.random_pm1(W["R"], W["n"], seed0 + k)
        ctx.set_spins(init)  # upload outside the timed region
        t0 = time.perf_counter()
        ctx.run(W["beta"], W["mcs"], seed0 + k, 0)  # returns after the kernel has finished
        t1 = time.perf_counter()
        info = ctx.info()
        if k >= warmup:
            walls.append(t1 - t0)
            kms.append(info["last_kernel_ms"])
            log(f"  {prec} step {k - warmup}: wall {t1 - t0:.4f} s, kernel "
                f"{info['last_kernel_ms']:.2f} ms")
    energies = ctx.energies()
    drift = ctx.drift()
    info = ctx.info()
    ctx.close()
    return np.array(walls), np.array(kms), energies, drift, info


def cpu_baseline(W, reads=6):
    """Reference restatement (FP64, xorshift + libstdc++ uniforms, sequential
    sweep) on `reads` replicas of the same instance and schedule, one thread."""
    from oracle import oracle as orc
    from tests import helpers as H

    Es = H.symmetric_ext(W["E"])
    init = H.random_pm1(reads, W["n"], 99)
    seeds = np.arange(reads, dtype=np.uint32) + 1
    t0 = time.perf_counter()
    orc.ref_sa_dense_batch(Es, init, seeds, W["beta"], W["mcs"], nthreads=1)
    dt = time.perf_counter() - t0
    attempts = reads * int(W["mcs"].sum()) * W["n"]
    return dict(value=attempts / dt, unit="flip attempts/s", cores=1, kind="port",
                sample=f"{reads} reads x {int(W['mcs'].sum())} sweeps x N={W['n']} of the "
                       f"same instance and schedule, 1 thread, {dt:.1f} s")


def tts99(energies, e_ref, tol, tau):
    ps = float(np.mean(energies <= e_ref + tol))
    if ps >= 1.0:
        return 0.0, ps
    if ps <= 0.0:
        return float("inf"), ps
    return float(tau * np.log(0.01) / np.log(1.0 - ps)), ps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="E")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--sweeps", type=int, default=None,
                    help="shorter schedule (profiling runs only; not a bench figure)")
    ap.add_argument("--replicas", type=int, default=None)
    ap.add_argument("--group", type=int, default=None, help="replicas per workgroup")
    ap.add_argument("--xcd-window", type=int, default=None)
    ap.add_argument("--refresh-every", type=int, default=None)
    ap.add_argument("--ring-rows", type=int, default=None)
    ap.add_argument("--round-rows", type=int, default=None)
    ap.add_argument("--colown", type=int, default=None)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; replicas would shard across GPUs "
              "with no communication (see DESIGN.md), but no multi-GPU figure is produced.",
              file=sys.stderr)
        sys.exit(2)

    from openjij_amd import capi

    if capi.device_count() < 1:
        raise SystemExit("bench.py: no HIP device visible")
    W = workload(args.config, args.sweeps, args.replicas)
    attempts = W["R"] * int(W["mcs"].sum()) * W["n"]
    log(f"workload: {W['name']} ({attempts:.3e} flip attempts per step)")

    opts = {}
    if args.group is not None:
        opts["group"] = args.group
    if args.xcd_window is not None:
        opts["xcd_window"] = args.xcd_window
    if args.refresh_every is not None:
        opts["refresh_every"] = args.refresh_every
    if args.ring_rows is not None:
        opts["ring_rows"] = args.ring_rows
    if args.round_rows is not None:
        opts["round_rows"] = args.round_rows
    if args.colown is not None:
        opts["colown"] = args.colown
    walls, kms, e_fast, drift, info = time_path(W, W["fast"], args.steps, args.warmup, 1000,
                                                opts)
    step_s = float(walls.mean())
    value = attempts / step_s
    kernel_ms = float(kms.mean())

    plain = None
    if not args.no_plain:
        pw, pk, e_plain, _, pinfo = time_path(W, W["plain"], 1, 0, 1000 + args.warmup)
        plain = dict(value=attempts / float(pw.mean()), kernel_ms=float(pk.mean()))
    else:
        e_plain = e_fast

    e_ref = float(min(e_fast.min(), e_plain.min()))
    tol = W["tol_rel"] * abs(e_ref)
    tau = step_s / W["R"]
    tts, ps = tts99(e_fast, e_ref, tol, tau)

    bJ = info["coupling_bytes"]
    b_alg = W["n"] * bJ  # coupling bytes per attempt (one row per attempt)
    achieved = attempts * b_alg / (kernel_ms * 1e-3)
    cpu = None if args.no_cpu else cpu_baseline(W)

    out = {
        "metric": "flip attempts/s + TTS99, batched mixed-prec vs plain FP64 GPU SA, equal residual E",
        "value": value,
        "unit": "flip attempts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if W["fast"] == "fp32" else ("int32" if W["fast"] == "int32" else "f64"),
        "data": "synthetic (seeded splitmix64 instance)",
        "config": {"workload": W["name"], "replicas": W["R"], "spins": W["n"],
                   "sweeps": int(W["mcs"].sum()), "path": W["fast"],
                   "kernel": info["kernel"], "refresh_every": info["refresh_every"],
                   "group": info["group"], "ring_rows": info["ring_rows"],
                   "lds_bytes": info["lds_bytes"], "lds_budget": info["lds_budget"]},
        "tts99_s": tts,
        "tts99_success_prob": ps,
        "plain_fp64": None if plain is None else {
            "value": plain["value"], "kernel_ms": plain["kernel_ms"],
            "speedup_of_value": value / plain["value"]},
        "energy": {"best": e_ref, "mean_fast": float(e_fast.mean()),
                   "mean_plain": float(e_plain.mean()), "max_field_drift": drift},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
