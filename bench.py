"""Benchmark of the single-spin-flip annealing path on one MI355X.

Default workload (the metric's configuration, E): dense Sherrington-Kirkpatrick
N=4096, J ~ N(0, 1/N) rounded to FP32, h = 0, 2048 replicas, 500 sweeps,
geometric beta 0.1 -> 10 (500 levels, one sweep each).  A step is one full
anneal of all replicas from random initial states (uploaded before the timed
region).  `value` is flip attempts/s of the batched mixed-precision (FP32)
path; the plain per-replica FP64 path is timed on the same workload for the
ratio, and the two final-energy ensembles are compared (two-sample KS).

Other workloads (parity-test shapes, measured the same way):
  A  dense SK N=256 +-1, 64 reads, 1000 sweeps (exact int32 path vs FP64)
  B  3D Edwards-Anderson L=16 +-1 (CSR), 1024 reads, 1000 levels x 2 sweeps
  C  SQA dense SK N=256 gaussian, M=32, gamma=5, 256 reads, 1000 sweeps
  D  4096 uneven QUBO instances (N 32..1024 log-uniform), 16 reads each

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--config A..E]
  python bench.py --calibrate-copy     # plain 256 MiB device copy (HBM counter scale)

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # B/s, MI355X HBM3E
CPU_THREADS = 16   # host-core share of one GPU on the measurement machine


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


# --------------------------------------------------------------------------
# workloads (seeded splitmix64 instances; identical inputs for GPU and CPU)
# --------------------------------------------------------------------------
def workload(cfg, sweeps=None, replicas=None):
    from tests import helpers as H

    if cfg == "E":
        n, R, L = 4096, replicas or 2048, sweeps or 500
        return dict(cfg=cfg, kind="dense", name=f"dense SK N={n} gaussian fp32 J, {R} replicas, "
                    f"{L} sweeps, beta 0.1->10 geometric", n=n, R=R, M=1, gamma=0.0,
                    E=H.sk_gauss(n, 0xE005), beta=H.geometric_betas(0.1, 10.0, L), s=None,
                    mcs=np.ones(L, np.int64), fast="fp32", plain="fp64", tol_rel=1e-9, deg=n,
                    cpu_reads=(3, 32))
    if cfg == "A":
        n, R, L = 256, replicas or 64, sweeps or 1000
        return dict(cfg=cfg, kind="dense", name=f"dense SK N={n} +-1 J, {R} reads, {L} sweeps, "
                    "beta 0.1->10 geometric", n=n, R=R, M=1, gamma=0.0, E=H.sk_pm1(n, 0xA001),
                    beta=H.geometric_betas(0.1, 10.0, L), s=None, mcs=np.ones(L, np.int64),
                    fast="int32", plain="fp64", tol_rel=0.0, deg=n, cpu_reads=(640, 2560))
    if cfg == "B":
        Lside, R, L = 16, replicas or 1024, sweeps or 1000
        n, rp, ci, va = H.ea3d_csr(Lside, 0xB002)
        return dict(cfg=cfg, kind="csr", name=f"3D EA L={Lside} +-1 (CSR, N={n}), {R} reads, "
                    f"{L} levels x 2 sweeps, beta 0.1->5 geometric", n=n, R=R, M=1, gamma=0.0,
                    csr=(rp, ci, va), beta=H.geometric_betas(0.1, 5.0, L), s=None,
                    mcs=np.full(L, 2, np.int64), fast="int32", plain="fp64", tol_rel=0.0, deg=6,
                    cpu_reads=(64, 512))
    if cfg == "C":
        n, R, L, M = 256, replicas or 256, sweeps or 1000, 32
        k = np.arange(L, dtype=np.float64)
        return dict(cfg=cfg, kind="sqa", name=f"SQA dense SK N={n} gaussian, M={M}, gamma=5, "
                    f"beta=10, {R} reads, {L} sweeps, s 0->0.998 linear", n=n, R=R, M=M,
                    gamma=5.0, E=H.sk_gauss(n, 0xC003), beta=np.full(L, 10.0),
                    s=(4.99 / 5.0) * k / (L - 1), mcs=np.ones(L, np.int64), fast="fp32",
                    plain="fp64", tol_rel=1e-9, deg=n + 2, cpu_reads=(8, 64))
    if cfg == "D":
        K, R, L = 4096, replicas or 16, sweeps or 1000
        ns = H.qubo_batch_sizes(K, 0xD004)
        Es, betas = [], []
        for k, n in enumerate(ns):
            Q = H.qubo_upper(int(n), 0xD004 + 1 + k)
            E, _ = H.qubo_to_scaled_ising(Q)
            S = np.abs(Q) + np.abs(np.triu(Q, 1)).T
            dE_bound = max(float(S.sum(axis=1).max()), 1.0)
            # beta in QUBO units, then / 4 for the scaled Ising energies
            betas.append(H.geometric_betas(np.log(2.0) / dE_bound, np.log(100.0), L) / 4.0)
            Es.append(E)
        return dict(cfg=cfg, kind="batch", name=f"{K} QUBO instances, N log-uniform 32..1024 "
                    f"(sum N = {int(ns.sum())}), int [-8,8] p=1/2, {R} reads each, {L} sweeps, "
                    "per-instance geometric beta ln2/dE_max -> ln100", n=int(ns.sum()), R=R,
                    M=1, gamma=0.0, Es=Es, ns=ns, betas=np.stack(betas), s=None,
                    mcs=np.ones(L, np.int64), fast="int32", plain="fp64", tol_rel=0.0,
                    deg=float((ns.astype(np.float64) ** 2).sum() / ns.sum()), cpu_reads=(64, 1024))
    raise SystemExit(f"unknown config {cfg}")


def prec_code(p):
    from openjij_amd import capi

    return {"fp64": capi.PREC_FP64, "fp32": capi.PREC_FP32, "int32": capi.PREC_INT32}[p]


def make_ctx(W, prec):
    from openjij_amd import capi

    if W["kind"] == "batch":
        return capi.BatchContext.create(W["Es"], W["R"], prec_code(prec))
    if W["kind"] == "csr":
        rp, ci, va = W["csr"]
        return capi.Context.csr(W["n"], rp, ci, va, W["R"], 1, 0.0, prec_code(prec))
    return capi.Context.dense(W["E"], W["R"], W["M"], W["gamma"], prec_code(prec))


def init_spins(W, seed):
    from tests import helpers as H

    if W["kind"] == "batch":
        flat = H.random_pm1(W["R"], W["n"], seed).ravel()  # split per problem below
        out, o = [], 0
        for n in W["ns"]:
            out.append(flat[o:o + W["R"] * n].reshape(W["R"], n))
            o += W["R"] * n
        return out
    s = H.random_pm1(W["R"] * W["M"], W["n"], seed)
    return s.reshape(W["R"], W["M"], W["n"]) if W["M"] > 1 else s


def time_path(W, prec, steps, warmup, seed0, options=None):
    """Per-step wall seconds, per-step kernel ms (HIP events on the context's
    stream around the sweep kernel), final energies per replica, drift, info."""
    ctx = make_ctx(W, prec)
    for k, v in (options or {}).items():
        ctx.set_option(k, v)
    walls, kms = [], []
    for k in range(warmup + steps):
        ctx.set_spins(init_spins(W, seed0 + k))  # upload outside the timed region
        t0 = time.perf_counter()
        if W["kind"] == "batch":
            ctx.run_batch(W["betas"], W["mcs"], seed0 + k, 0)
        else:
            ctx.run(W["beta"], W["mcs"], seed0 + k, 0, s=W["s"])  # returns after the kernel ends
        t1 = time.perf_counter()
        info = ctx.info()
        if k >= warmup:
            walls.append(t1 - t0)
            kms.append(info["last_kernel_ms"])
            log(f"  {prec} step {k - warmup}: wall {t1 - t0:.4f} s, kernel "
                f"{info['last_kernel_ms']:.2f} ms")
    e = ctx.energies()
    if W["kind"] == "batch":
        energies = e  # [K][R]
    else:
        energies = e.min(axis=1) if W["M"] > 1 else e.reshape(-1)
    drift = ctx.drift()
    info = ctx.info()
    ctx.close()
    return np.array(walls), np.array(kms), energies, drift, info


# --------------------------------------------------------------------------
# CPU baseline: the oracle's restatement of the reference path (FP64,
# xorshift + libstdc++ uniforms, reference sweep order, same schedule)
# --------------------------------------------------------------------------
def _cpu_run(W, reads, threads):
    from concurrent.futures import ThreadPoolExecutor

    from oracle import oracle as orc
    from tests import helpers as H

    seeds = np.arange(reads, dtype=np.uint32) + 1
    init = None if W["kind"] == "batch" else H.random_pm1(reads * W["M"], W["n"], 99)
    t0 = time.perf_counter()
    if W["kind"] == "batch":
        # `reads` instances spread over the size range, one read each
        K = len(W["Es"])
        pick = np.linspace(0, K - 1, reads).astype(int)

        def one(q):
            k = pick[q]
            n = W["Es"][k].shape[0] - 1
            orc.ref_sa_dense_batch(H.symmetric_ext(W["Es"][k]), H.random_pm1(1, n, 99 + q),
                                   seeds[q:q + 1], W["betas"][k], W["mcs"], nthreads=1)

        t0 = time.perf_counter()
        with ThreadPoolExecutor(threads) as ex:
            list(ex.map(one, range(reads)))
        dt = time.perf_counter() - t0
        att = sum(W["Es"][k].shape[0] - 1 for k in pick) * int(W["mcs"].sum())
        return att / dt, dt
    if W["kind"] == "dense":
        orc.ref_sa_dense_batch(H.symmetric_ext(W["E"]), init, seeds, W["beta"], W["mcs"],
                               nthreads=threads)
    elif W["kind"] == "csr":
        rp, ci, va = W["csr"]
        orc.ref_sa_csr_batch(rp, ci, va, W["n"], init, seeds, W["beta"], W["mcs"],
                             nthreads=threads)
    else:  # one SQA read per call; ctypes drops the GIL, so reads run in parallel
        Es = H.symmetric_ext(W["E"])
        init = init.reshape(reads, W["M"], W["n"])

        def one(r):
            orc.ref_sqa_dense(Es, W["gamma"], init[r], int(seeds[r]), W["beta"], W["s"],
                              W["mcs"])

        with ThreadPoolExecutor(threads) as ex:
            list(ex.map(one, range(reads)))
    dt = time.perf_counter() - t0
    return reads * W["M"] * int(W["mcs"].sum()) * W["n"] / dt, dt


def cpu_baseline(W):
    threads = max(1, min(CPU_THREADS, os.cpu_count() or 1))
    r1, rN = W["cpu_reads"]
    v1, dt1 = _cpu_run(W, r1, 1)
    log(f"  cpu 1 thread: {v1:.3e} attempts/s ({r1} reads, {dt1:.1f} s)")
    rN = max(rN, threads) // threads * threads
    vN, dtN = _cpu_run(W, rN, threads)
    log(f"  cpu {threads} threads: {vN:.3e} attempts/s ({rN} reads, {dtN:.1f} s)")
    sweeps = int(W["mcs"].sum())
    if W["kind"] == "batch":
        what = (f"{rN} of the instances (evenly spaced in the list), one read each, "
                f"{sweeps} sweeps, their own schedules")
    else:
        what = (f"{rN} reads x {sweeps} sweeps x N={W['n']}"
                + (f" x M={W['M']}" if W["M"] > 1 else "") + " of the same instance and schedule")
    return dict(value=vN, unit="flip attempts/s", cores=threads, kind="port",
                sample=f"{what}, {threads} threads (OpenMP/threads over reads), {dtN:.1f} s; "
                       f"CPU {cpu_model()}",
                one_thread=dict(value=v1, reads=r1, seconds=dt1))


def tts99(energies, e_ref, tol, tau):
    """The reference's TTS formula (openjij/utils/benchmark.py:205-223)."""
    ps = float(np.mean(energies <= e_ref + tol))
    if ps >= 1.0:
        return 0.0, ps
    if ps <= 0.0:
        return float("inf"), ps
    return float(tau * np.log(0.01) / np.log(1.0 - ps)), ps


def profiled_traffic(kernel_name, workload_name):
    """HBM bytes per run of `kernel_name` on `workload_name` (all of the run's
    launches), from a committed rocprofv3 counter summary under profiles/
    (FETCH_SIZE and WRITE_SIZE, scaled by the copy-kernel calibration).  None
    when no summary covers this kernel and workload."""
    import glob

    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_pmc.txt"))):
        kv = dict(re.findall(r"^([a-z_]+):\s*(.+?)\s*$", open(path).read(), re.M))
        if kv.get("kernel") == kernel_name and kv.get("workload") == workload_name:
            try:
                return float(kv["hbm_bytes_per_run"])
            except (KeyError, ValueError):
                return None
    return None


def calibrate_copy():
    import torch

    nbytes = 256 << 20
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ms = []
    for _ in range(5):
        e0.record(st)
        dst.copy_(src)
        e1.record(st)
        e1.synchronize()
        ms.append(e0.elapsed_time(e1))
    best = min(ms)
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes,
                      "best_ms": best, "GB_per_s_read_plus_write": 2 * nbytes / best / 1e6}),
          flush=True)


KERNEL_NAMES = {1: "k_dense_wave", 2: "k_sqa_wave", 3: "k_csr_level", 4: "k_dense_shared",
                5: "k_dense_colown", 6: "k_batch_wave", 7: "k_dense_lds"}


def kernel_name(info):
    if info["kernel"] == 3 and info.get("waves_per_replica", 1) > 1:
        return "k_csr_wg"  # several waves per replica
    return KERNEL_NAMES.get(info["kernel"], str(info["kernel"]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="E")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--sweeps", type=int, default=None,
                    help="shorter schedule (profiling runs only; not a bench figure)")
    ap.add_argument("--replicas", type=int, default=None)
    ap.add_argument("--group", type=int, default=None, help="replicas per workgroup")
    ap.add_argument("--xcd-window", type=int, default=None)
    ap.add_argument("--ring-rows", type=int, default=None)
    ap.add_argument("--round-rows", type=int, default=None)
    ap.add_argument("--colown", type=int, default=None)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; replicas would shard across GPUs "
              "with no communication during a run, but no multi-GPU figure is produced.",
              file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return

    from openjij_amd import capi

    if capi.device_count() < 1:
        raise SystemExit("bench.py: no HIP device visible")
    W = workload(args.config, args.sweeps, args.replicas)
    attempts = W["R"] * W["M"] * int(W["mcs"].sum()) * W["n"]
    log(f"workload: {W['name']} ({attempts:.3e} flip attempts per step)")

    opts = {}
    for key in ("group", "xcd_window", "ring_rows", "round_rows", "colown"):
        v = getattr(args, key)
        if v is not None:
            opts[key] = v
    walls, kms, e_fast, drift, info = time_path(W, W["fast"], args.steps, args.warmup, 1000,
                                                opts)
    step_s = float(walls.mean())
    value = attempts / step_s
    kernel_ms = float(kms.mean())

    plain = None
    if not args.no_plain:
        # same seed and initial state as the last timed step of the fast path:
        # with integer couplings the two paths must then end in identical states
        pw, pk, e_plain, _, pinfo = time_path(W, W["plain"], 1, 0,
                                              1000 + args.warmup + args.steps - 1)
        plain = dict(value=attempts / float(pw.mean()), kernel_ms=float(pk.mean()),
                     kernel=kernel_name(pinfo))
    else:
        e_plain = e_fast

    from tests import helpers as H

    same_as_plain = None if plain is None else bool(np.array_equal(e_fast, e_plain))
    ks_p = None
    if W["kind"] == "batch":
        # residual energy of every read against its instance's best (both paths)
        best = np.minimum(e_fast.min(axis=1), e_plain.min(axis=1))[:, None]
        r_fast, r_plain = (e_fast - best).ravel(), (e_plain - best).ravel()
        tts, ps = tts99(r_fast, 0.0, 0.0, step_s / (W["R"] * len(W["ns"])))
        if plain is not None:
            ks_p = H.ks_2samp_pvalue(r_fast, r_plain)
        e_ref = float("nan")
        e_fast, e_plain = r_fast, r_plain  # reported means are mean residuals
    else:
        e_ref = float(min(e_fast.min(), e_plain.min()))
        tol = W["tol_rel"] * abs(e_ref)
        tts, ps = tts99(e_fast, e_ref, tol, step_s / W["R"])
        if plain is not None:
            ks_p = H.ks_2samp_pvalue(e_fast, e_plain)

    kname = kernel_name(info)
    bJ = info["coupling_bytes"]
    # one coupling row (deg couplings, plus an int32 column index each for CSR)
    # consulted per attempt
    b_alg = W["deg"] * (bJ + (4 if W["kind"] == "csr" else 0))
    achieved = attempts * b_alg / (kernel_ms * 1e-3)
    default_run = (not opts and args.sweeps is None and args.replicas is None)
    traffic = profiled_traffic(kname, W["name"]) if default_run else None
    cpu = None if args.no_cpu else cpu_baseline(W)

    dtype = {"fp32": "f32", "int32": "int32", "fp64": "f64"}[W["fast"]]
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
        "dtype": dtype,
        "data": "synthetic (seeded splitmix64 instance)",
        "config": {"workload": W["name"], "replicas": W["R"], "spins": W["n"],
                   "problems": len(W["ns"]) if W["kind"] == "batch" else 1,
                   "trotter": W["M"], "sweeps": int(W["mcs"].sum()), "path": W["fast"],
                   "kernel": kname, "refresh_every": info["refresh_every"],
                   "group": info["group"], "ring_rows": info["ring_rows"],
                   "lds_bytes": info["lds_bytes"], "lds_budget": info["lds_budget"],
                   "kernel_ms": kernel_ms},
        "tts99_s": tts,
        "tts99_success_prob": ps,
        "plain_fp64": None if plain is None else {
            "value": plain["value"], "kernel_ms": plain["kernel_ms"], "kernel": plain["kernel"],
            "speedup_of_value": value / plain["value"]},
        "energy": {"best": e_ref if W["kind"] != "batch" else None,
                   "residual_vs_instance_best": W["kind"] == "batch",
                   "mean_fast": float(e_fast.mean()),
                   "mean_plain": float(e_plain.mean()), "ks_pvalue_fast_vs_plain": ks_p,
                   "final_energies_identical_to_plain": same_as_plain,
                   "max_field_drift": drift},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": traffic, "bytes_per_attempt": b_alg},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
