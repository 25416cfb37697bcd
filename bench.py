"""Benchmark of the single-spin-flip annealing path on one MI355X.

Headline workload (the metric's configuration, E): dense Sherrington-
Kirkpatrick N=4096, J ~ N(0, 1/N) rounded to FP32, h = 0, 2048 replicas, 500
sweeps, geometric beta 0.1 -> 10 (500 levels, one sweep each).  A step is one
full anneal of all replicas from random initial states (uploaded before the
timed region) on the batched mixed-precision path: FP32 fields refreshed on
the matrix cores every 16 sweeps (bf16 hi/lo couplings, FP32 accumulation),
then the FP64 energies of the final states and the FP64 drift check of the
carried fields, all inside the timed region.  `value` is flip attempts/s of
that step.  The plain per-replica FP64 path is timed on the same workload
(anneal + FP64 energies) for the ratio.

The same JSON line carries configs A-D (the other shapes of the contract),
measured the same way, under "configs":
  A  dense SK N=256 +-1, 64 reads, 1000 sweeps (exact int32 path vs FP64)
  B  3D Edwards-Anderson L=16 +-1 (CSR), 1024 reads, 1000 levels x 2 sweeps
  C  SQA dense SK N=256 gaussian, M=32, gamma=5, 256 reads, 1000 sweeps
  D  4096 uneven QUBO instances (N 32..1024 log-uniform), 16 reads each

Per config: TTS99 and success probability of both paths against one fixed
target (the best energy of a separate 4x-longer anneal of the fast path, plus
a stated tolerance), the roofline of the dominant kernel from flips and rows
counted on the GPU, and the CPU baseline (the oracle's restatement of the
reference algorithm, timed on this host).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--config A..E|T]
  (T, not part of the default run: SQA on the 3D EA L=16 CSR graph, 8 slices,
  1024 reads, on the sparse Trotter kernel; CPU baseline: the oracle's FP64
  replay of the same path)
  python bench.py --calibrate-copy     # plain 256 MiB device copy (HBM counter scale)

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import glob
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12       # B/s, MI355X HBM3E
VALU_PEAK = 78.6e12     # FP32 / INT32 / FP64 vector ops/s without packed FP32 (disabled here)
MFMA_BF16_PEAK = 2.5e15  # dense bf16 MFMA flop/s
# L2 -> CU read bandwidth: 16 channels x 128 B/clk per XCD, 8 XCDs, 2.4 GHz
L2_PEAK = 16 * 128 * 8 * 2.4e9
MFMA_F64_PEAK = 78.6e12  # FP64 MFMA flop/s
CPU_THREADS = 16        # host-core share of one GPU on the measurement machine
XCDS = 8                # L2 domains of the MI355X
METRIC = "flip attempts/s + TTS99, batched mixed-prec vs plain FP64 GPU SA, equal residual E"


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
                    mcs=np.ones(L, np.int64), fast="fp32", plain="fp64", refresh=16,
                    tol_rel=5e-3, integer=False, deg=n, cpu_reads=(3, 32))
    if cfg == "A":
        n, R, L = 256, replicas or 64, sweeps or 1000
        return dict(cfg=cfg, kind="dense", name=f"dense SK N={n} +-1 J, {R} reads, {L} sweeps, "
                    "beta 0.1->10 geometric", n=n, R=R, M=1, gamma=0.0, E=H.sk_pm1(n, 0xA001),
                    beta=H.geometric_betas(0.1, 10.0, L), s=None, mcs=np.ones(L, np.int64),
                    fast="int32", plain="fp64", refresh=0, tol_rel=0.0, integer=True, deg=n,
                    cpu_reads=(640, 2560))
    if cfg == "B":
        Lside, R, L = 16, replicas or 1024, sweeps or 1000
        n, rp, ci, va = H.ea3d_csr(Lside, 0xB002)
        return dict(cfg=cfg, kind="csr", name=f"3D EA L={Lside} +-1 (CSR, N={n}), {R} reads, "
                    f"{L} levels x 2 sweeps, beta 0.1->5 geometric", n=n, R=R, M=1, gamma=0.0,
                    csr=(rp, ci, va), beta=H.geometric_betas(0.1, 5.0, L), s=None,
                    mcs=np.full(L, 2, np.int64), fast="int32", plain="fp64", refresh=0,
                    tol_rel=5e-3, integer=True, deg=6, cpu_reads=(64, 512))
    if cfg == "C":
        n, R, L, M = 256, replicas or 256, sweeps or 1000, 32
        k = np.arange(L, dtype=np.float64)
        return dict(cfg=cfg, kind="sqa", name=f"SQA dense SK N={n} gaussian, M={M}, gamma=5, "
                    f"beta=10, {R} reads, {L} sweeps, s 0->0.998 linear", n=n, R=R, M=M,
                    gamma=5.0, E=H.sk_gauss(n, 0xC003), beta=np.full(L, 10.0),
                    s=(4.99 / 5.0) * k / (L - 1), mcs=np.ones(L, np.int64), fast="fp32",
                    plain="fp64", refresh=0, tol_rel=5e-3, integer=False, deg=n + 2,
                    cpu_reads=(8, 64))
    if cfg == "T":
        Lside, R, L, M = 16, replicas or 1024, sweeps or 500, 8
        n, rp, ci, va = H.ea3d_csr(Lside, 0xB002)
        k = np.arange(L, dtype=np.float64)
        return dict(cfg=cfg, kind="sqa_csr", name=f"SQA 3D EA L={Lside} +-1 (CSR, N={n}), M={M}, "
                    f"gamma=1, beta=5, {R} reads, {L} sweeps, s 0->0.998 linear", n=n, R=R, M=M,
                    gamma=1.0, csr=(rp, ci, va), beta=np.full(L, 5.0),
                    s=(4.99 / 5.0) * k / (L - 1), mcs=np.ones(L, np.int64), fast="fp32",
                    plain="fp64", refresh=0, tol_rel=1.5e-2, integer=False, deg=6 + 2,
                    cpu_reads=(4, 64))  # tol: the reads of 500 sweeps end 1.6% above the
                                        # 4x-longer anneal's best on average
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
                    mcs=np.ones(L, np.int64), fast="int32", plain="fp64", refresh=0,
                    tol_rel=0.0, integer=True,
                    deg=float((ns.astype(np.float64) ** 2).sum() / ns.sum()), cpu_reads=(64, 1024))
    raise SystemExit(f"unknown config {cfg}")


def prec_code(p):
    from openjij_amd import capi

    return {"fp64": capi.PREC_FP64, "fp32": capi.PREC_FP32, "int32": capi.PREC_INT32}[p]


def make_ctx(W, prec):
    from openjij_amd import capi

    if W["kind"] == "batch":
        return capi.BatchContext.create(W["Es"], W["R"], prec_code(prec))
    if W["kind"] in ("csr", "sqa_csr"):
        rp, ci, va = W["csr"]
        return capi.Context.csr(W["n"], rp, ci, va, W["R"], W["M"], W["gamma"], prec_code(prec))
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


def min_energies(W, e):
    """Per read: the energy (classical) or the minimum over Trotter slices."""
    if W["kind"] == "batch":
        return e  # [K][R]
    return e.min(axis=1) if W["M"] > 1 else e.reshape(-1)


def time_path(W, prec, steps, warmup, seed0, options=None, sweeps_scale=1):
    """One context, `warmup` + `steps` anneals.  A timed step is the anneal
    (ojmi_run: sweep kernels and any field refreshes) plus the FP64 energies
    of the final states and, for FP32 fields, the FP64 drift check.
    Returns per-step wall seconds, per-step info dicts, final energies of the
    last step, drift, and the context's info."""
    ctx = make_ctx(W, prec)
    for k, v in (options or {}).items():
        ctx.set_option(k, v)
    beta, mcs = W.get("beta"), W["mcs"] * sweeps_scale
    walls, infos = [], []
    drift, e = 0.0, None
    for k in range(warmup + steps):
        ctx.set_spins(init_spins(W, seed0 + k))  # upload outside the timed region
        t0 = time.perf_counter()
        if W["kind"] == "batch":
            ctx.run_batch(W["betas"], mcs, seed0 + k, 0)
        else:
            ctx.run(beta, mcs, seed0 + k, 0, s=W["s"])  # returns after the kernels end
        e = ctx.energies()                              # FP64, on the GPU
        if prec == "fp32" and W["M"] == 1 and W["kind"] != "batch":
            drift = ctx.drift()                         # FP64 drift check of the carried fields
        t1 = time.perf_counter()
        info = ctx.info()
        if k >= warmup:
            walls.append(t1 - t0)
            infos.append(info)
            log(f"  {prec} step {k - warmup}: wall {t1 - t0:.4f} s, anneal "
                f"{info['last_kernel_ms']:.2f} ms (refresh {info['last_refresh_ms']:.2f} ms in "
                f"{info['last_refreshes']}), flips {info['last_flips']:.4g}")
    spins = ctx.get_spins()
    final_info = ctx.info()
    ctx.close()
    return np.array(walls), infos, min_energies(W, e), drift, final_info, spins


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
    elif W["kind"] == "sqa_csr":
        # the oracle's FP64 replay of the path; its OpenMP loop over reads
        # takes the thread count of the calling thread (libgomp's setting)
        import ctypes

        rp, ci, va = W["csr"]
        a, B = orc.expand_transverse(W["beta"], W["s"], W["mcs"], W["gamma"], W["M"])
        ctypes.CDLL("libgomp.so.1").omp_set_num_threads(threads)
        t0 = time.perf_counter()
        orc.replay_sqa_csr("fp64", rp, ci, va, W["n"], init.reshape(reads, W["M"], W["n"]), 5, 0,
                           a, B)
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
    nproc = os.cpu_count() or 1
    threads = max(1, min(CPU_THREADS, nproc))
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
                       f"CPU {cpu_model()}, nproc {nproc}; the host's share of one GPU is "
                       f"{CPU_THREADS} cores",
                nproc=nproc, one_thread=dict(value=v1, reads=r1, seconds=dt1))


# --------------------------------------------------------------------------
# TTS99 against a fixed target
# --------------------------------------------------------------------------
def tts_block(W, e_fast, e_plain, e_long, tau_fast, tau_plain):
    """Success probability p_s = fraction of reads with E <= target and
    TTS99 = tau ln(0.01) / ln(1 - p_s) (openjij/utils/benchmark.py:205-223),
    tau = step time per read.  Target: the best energy of a separate 4x-longer
    anneal (per instance for batches), plus tol_rel * |E_best|; it depends on
    neither timed path's own samples."""
    from openjij_amd.utils.benchmark import time_to_solution

    if W["kind"] == "batch":
        best = e_long.min(axis=1)[:, None]
        target = best + W["tol_rel"] * np.abs(best)
        ps_f = float(np.mean(e_fast <= target))
        ps_p = None if e_plain is None else float(np.mean(e_plain <= target))
        tgt = {"per_instance_best_of_long_run": True, "tol_rel": W["tol_rel"]}
    else:
        best = float(e_long.min())
        target = best + W["tol_rel"] * abs(best)
        ps_f = float(np.mean(e_fast <= target))
        ps_p = None if e_plain is None else float(np.mean(e_plain <= target))
        tgt = {"best_of_long_run": best, "tol_rel": W["tol_rel"], "target": target}
    out = dict(target=tgt, fast=dict(p_s=ps_f, tts99_s=float(time_to_solution(ps_f, tau_fast, 0.99)),
                                     tau_s=tau_fast))
    if ps_p is not None:
        out["plain_fp64"] = dict(p_s=ps_p, tts99_s=float(time_to_solution(ps_p, tau_plain, 0.99)),
                                 tau_s=tau_plain)
    return out


# --------------------------------------------------------------------------
# roofline of the dominant kernel, from counts made on the GPU
# --------------------------------------------------------------------------
def profiled(kernel_name, workload_name):
    """HBM bytes per run of `kernel_name` on `workload_name` from a committed
    rocprofv3 counter summary under profiles/ (FETCH_SIZE + WRITE_SIZE scaled
    by the copy-kernel calibration), or None."""
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_pmc.txt"))):
        kv = dict(re.findall(r"^([a-z_]+):\s*(.+?)\s*$", open(path).read(), re.M))
        if kv.get("kernel") == kernel_name and kv.get("workload") == workload_name:
            try:
                return float(kv["hbm_bytes_per_run"])
            except (KeyError, ValueError):
                return None
    return None


KERNEL_NAMES = {1: "k_dense_wave", 2: "k_sqa_wave", 3: "k_csr_level", 4: "k_dense_shared",
                5: "k_dense_colown", 6: "k_batch_wave", 7: "k_dense_lds", 8: "k_batch_persist",
                9: "k_batch_wave", 10: "k_dense_mfma", 11: "k_sqa_csr"}


def kernel_name(info):
    if info["kernel"] == 3 and info.get("waves_per_replica", 1) > 1:
        return "k_csr_wg"  # several waves per replica
    if info["kernel"] == 1 and info.get("waves_per_replica", 1) > 1:
        return "k_dense_wave_pipe"  # threshold helper waves
    if info["kernel"] == 1 and info.get("lds_bytes", 0) > 0:
        return "k_dense_wave_lds"  # coupling rows in LDS
    return KERNEL_NAMES.get(info["kernel"], str(info["kernel"]))


def roofline(W, info, sweep_ms, launches, default_run):
    """Per launch of the sweep kernel.

    hbm: algorithmic bytes = what the kernel must bring on chip from HBM.
      Dense and Trotter kernels share coupling rows through each XCD's L2, so
      every row is fetched at least once per XCD per sweep it is flipped in:
      min(rows streamed, 8 x rows per sweep) x row bytes, plus the spins in and
      out; CSR and LDS-resident couplings are read once per launch.
    valu: the arithmetic the sequential algorithm does: one N-wide update
      (N fused multiply-adds) per accepted flip, counted on the GPU; CSR
      recomputes each field from its deg neighbours at every attempt.
    l2: bytes of coupling rows the workgroups stream (each row once per
      workgroup per 64-spin block it was flipped in, counted on the GPU; the
      matrix-core sweep streams all 64 rows of every block), against the L2
      read bandwidth: what binds the matrix-core sweep.
    """
    bJ = info["coupling_bytes"]
    T = int(W["mcs"].sum())
    flips, rows = info["last_flips"], info["last_rows"]
    spins_io = 2 * W["R"] * W["M"] * W["n"]
    if W["kind"] in ("csr", "sqa_csr"):
        ell = W["n"] * 16 * (4 + bJ)  # level-order ELL arrays, 16 slots wide
        alg = ell + spins_io
        flop = 2.0 * W["R"] * W["M"] * T * W["n"] * W["deg"]
        row_bytes = 0.0
    elif W["kind"] == "batch":
        nsq = float((W["ns"].astype(np.float64) ** 2).sum())
        alg = nsq * bJ + spins_io                       # every problem once
        flop = 2.0 * flips * W["deg"]                   # attempt-weighted mean row length
        row_bytes = float(rows) * W["deg"] * bJ
    else:
        row = W["n"] * bJ
        per_sweep = W["n"] * W["M"]                      # rows a sweep can flip (per slice)
        alg = min(float(rows), XCDS * per_sweep * T) * row + spins_io
        if W["n"] * W["n"] * bJ <= 4 << 20:              # fits one L2: read once per XCD
            alg = XCDS * W["n"] * row + spins_io
        flop = 2.0 * flips * W["n"]
        row_bytes = float(rows) * row
    t = sweep_ms * 1e-3 / launches                       # average launch duration
    a_hbm = alg / launches / t
    a_valu = flop / launches / t
    traffic = profiled(kernel_name(info), W["name"]) if default_run else None
    return {
        "bound": "hbm", "achieved": a_hbm / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
        "frac": a_hbm / HBM_PEAK,
        "traffic": None if traffic is None else traffic / launches,
        "alg_bytes_per_launch": alg / launches,
        "launch_ms": t * 1e3, "launches": launches,
        "binding": "l2" if info["kernel"] == 10 else "valu",
        "valu": {"achieved": a_valu / 1e12, "peak": VALU_PEAK / 1e12, "unit": "TFLOP/s",
                 "frac": a_valu / VALU_PEAK, "flops_per_launch": flop / launches,
                 "accepted_flips": flips},
        "l2_rows": {"rows_streamed": rows, "bytes_per_launch": row_bytes / launches,
                    "GB_per_s": row_bytes / launches / t / 1e9, "peak_GB_per_s": L2_PEAK / 1e9,
                    "frac": row_bytes / launches / t / L2_PEAK},
    }


def mfma_summary(W, info, refresh_ms_mean):
    """Matrix-core refresh: 2 (hi, lo) x 2 R N^2 flops per refresh."""
    if not info["last_refreshes"]:
        return None
    per = refresh_ms_mean / info["last_refreshes"] * 1e-3
    flop = 2.0 * 2.0 * W["R"] * W["n"] * W["n"]
    return {"bound": "mfma", "refreshes_per_step": info["last_refreshes"],
            "ms_per_refresh": per * 1e3, "flops_per_refresh": flop,
            "achieved": flop / per / 1e12, "peak": MFMA_BF16_PEAK / 1e12, "unit": "TFLOP/s",
            "frac": flop / per / MFMA_BF16_PEAK}


# --------------------------------------------------------------------------
def run_config(cfg, steps, warmup, with_cpu, with_plain, opts, sweeps=None, replicas=None,
               with_tts=True):
    from tests import helpers as H

    W = workload(cfg, sweeps, replicas)
    attempts = W["R"] * W["M"] * int(W["mcs"].sum()) * W["n"]
    log(f"[{cfg}] workload: {W['name']} ({attempts:.3e} flip attempts per step)")
    o = dict(opts or {})
    if W["refresh"] and "refresh_every" not in o:
        o["refresh_every"] = W["refresh"]
    walls, infos, e_fast, drift, info, spins = time_path(W, W["fast"], steps, warmup, 1000, o)
    step_s = float(walls.mean())
    value = attempts / step_s
    kms = np.array([i["last_kernel_ms"] for i in infos])
    rms = np.array([i["last_refresh_ms"] for i in infos])
    launches = max(1, info["last_refreshes"]) if W["refresh"] else 1
    sweep_ms = float((kms - rms).mean())

    plain, e_plain, p_step = None, None, None
    if with_plain:
        # same seed and initial state as the last timed step of the fast path:
        # with integer couplings the two paths must end in identical states
        # plain: FP64 fields and FP64 couplings (sparse systems keep their
        # FP64 values rather than int8 quadruples)
        popts = {"csr_pack8": 0} if W["kind"] in ("csr", "sqa_csr") else None
        pw, pinfos, e_plain, _, pinfo, _ = time_path(W, W["plain"], 1, 0,
                                                     1000 + warmup + steps - 1, popts)
        p_step = float(pw.mean())
        plain = dict(value=attempts / p_step, ms_per_step=p_step * 1e3,
                     kernel_ms=float(np.mean([i["last_kernel_ms"] for i in pinfos])),
                     kernel=kernel_name(pinfo), speedup_of_value=value * p_step / attempts)
    # the fixed TTS target: a separate 4x-longer anneal of the fast path
    tts = None
    if with_tts:
        _, _, e_long, _, _, _ = time_path(W, W["fast"], 1, 0, 7777, o, sweeps_scale=4)
        reads = W["R"] * (len(W["ns"]) if W["kind"] == "batch" else 1)
        tts = tts_block(W, e_fast, e_plain, e_long, step_s / reads,
                        None if p_step is None else p_step / reads)

    energy = {"mean_fast": float(np.mean(e_fast)),
              "mean_plain": None if e_plain is None else float(np.mean(e_plain)),
              "ks_pvalue_fast_vs_plain": None if e_plain is None else H.ks_2samp_pvalue(
                  np.ravel(e_fast), np.ravel(e_plain)),
              "final_energies_identical_to_plain": None if e_plain is None else bool(
                  np.array_equal(e_fast, e_plain)),
              "max_field_drift": drift}
    if W["kind"] == "dense" and W["M"] == 1:
        # FP64 energies reported by the GPU against the oracle's exact
        # (double-double) energies of the same states, on a subset of reads
        from oracle import oracle as orc

        sub = min(16, W["R"])
        ex = orc.energy_exact(W["E"], spins[:sub])
        energy["max_rel_error_vs_exact_fp64"] = float(
            np.max(np.abs(e_fast[:sub] - ex) / np.maximum(1.0, np.abs(ex))))
        energy["checked_reads"] = sub
    default_run = not opts and sweeps is None and replicas is None
    out = {
        "value": value, "unit": "flip attempts/s", "steps": steps, "warmup": warmup,
        "ms_per_step": step_s * 1e3,
        "dtype": {"fp32": "f32", "int32": "int32", "fp64": "f64"}[W["fast"]],
        "config": {"workload": W["name"], "replicas": W["R"], "spins": W["n"],
                   "problems": len(W["ns"]) if W["kind"] == "batch" else 1,
                   "trotter": W["M"], "sweeps": int(W["mcs"].sum()), "path": W["fast"],
                   "kernel": kernel_name(info), "refresh_every": info["refresh_every"],
                   "mfma_segments": info["mfma_segments"],
                   "group": info["group"], "lds_bytes": info["lds_bytes"],
                   "anneal_ms": float(kms.mean()), "refresh_ms": float(rms.mean()),
                   "sweep_kernel_ms": sweep_ms},
        "tts99": tts,
        "plain_fp64": plain,
        "energy": energy,
        "roofline": roofline(W, info, sweep_ms, launches, default_run),
        "mfma_refresh": mfma_summary(W, info, float(rms.mean())),
        "cpu_baseline": cpu_baseline(W) if with_cpu else None,
    }
    return out


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default=None,
                    help="one config only (A..E, or T: sparse Trotter); default: E as the "
                         "headline plus A-D")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-tts", action="store_true",
                    help="skip the 4x-longer anneal that fixes the TTS99 target")
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--sweeps", type=int, default=None,
                    help="shorter schedule (profiling runs only; not a bench figure)")
    ap.add_argument("--replicas", type=int, default=None)
    ap.add_argument("--group", type=int, default=None, help="replicas per workgroup")
    ap.add_argument("--xcd-window", type=int, default=None)
    ap.add_argument("--colown", type=int, default=None)
    ap.add_argument("--refresh-every", type=int, default=None)
    ap.add_argument("--option", action="append", default=[], metavar="KEY=VALUE",
                    help="any ojmi_set_option key for the head config's fast-path context")
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
    opts = {}
    for key in ("group", "xcd_window", "colown", "refresh_every"):
        v = getattr(args, key)
        if v is not None:
            opts[key] = v
    for kv in args.option:
        k, _, v = kv.partition("=")
        opts[k] = int(v)
    head_cfg = args.config or "E"
    head = run_config(head_cfg, args.steps, args.warmup, not args.no_cpu, not args.no_plain, opts,
                      args.sweeps, args.replicas, not args.no_tts)
    others = {}
    if args.config is None:
        sub_steps, sub_warm = min(args.steps, 5), 1
        for cfg in ("A", "B", "C", "D"):
            others[cfg] = run_config(cfg, sub_steps, sub_warm, not args.no_cpu,
                                     not args.no_plain, {}, with_tts=not args.no_tts)
    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "flip attempts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": head["dtype"],
        "data": "synthetic (seeded splitmix64 instances)",
        "config": head["config"],
        "tts99": head["tts99"],
        "plain_fp64": head["plain_fp64"],
        "energy": head["energy"],
        "roofline": head["roofline"],
        "mfma_refresh": head["mfma_refresh"],
        "cpu_baseline": head["cpu_baseline"],
    }
    if others:
        out["configs"] = others
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
