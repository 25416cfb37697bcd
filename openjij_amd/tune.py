"""Launch-shape autotuner of the engine (run on the GPU it tunes for).

For each bucket of problem shape it times every candidate launch shape the
engine offers on a short anneal of a representative synthetic instance,
keeps the fastest, and writes

  * ``openjij_amd/tuned_shapes.json``: every measurement and the choices;
  * ``openjij_amd/csrc/ojmi_tuned.h``: the chosen table, compiled into
    ``libojmi.so`` as the engine's default shapes (rebuild after tuning).

Buckets and candidates:
  * CSR graphs (3D Edwards-Anderson, degree 6): waves per replica 1 / 2 / 4 / 8,
    per field type and replica count;
  * dense classical FP32 / INT32 systems: one wave per replica, or G replicas
    per workgroup sharing coupling rows (column-owned fields, G = 4 / 8, with
    and without XCD pacing; rows staged in LDS, G = 8 / 16), per field type,
    slot count (64-spin slots) and replica count;
  * batches of uneven problems: one launch largest first, one launch per slot
    class, or one persistent launch over per-XCD queues;
  * large dense FP32 systems (2048 replicas; N = 2048, 4096): a geometric
    anneal run 16 sweeps at a time on the matrix-core sweep and on the
    column-owned kernel, each segment's acceptance rate counted on the GPU;
    the table keeps the acceptance rate at which the two cross (the
    per-segment kernel choice of ``mfma_sweep`` 2), or 0 when the matrix-core
    sweep never wins.

  python -m openjij_amd.tune [--quick]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FT = {"fp64": 0, "fp32": 1, "int32": 2}   # FieldType of ojmi_kernels.h


def _splitmix(seed, count):
    with np.errstate(over="ignore"):
        k = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def _pm1(shape, seed):
    w = _splitmix(seed, int(np.prod(shape)))
    return np.where((w >> np.uint64(63)) == 1, 1, -1).astype(np.int8).reshape(shape)


def _sk(n, seed, integer):
    w = _splitmix(seed, n * n)
    if integer:
        J = ((w >> np.uint64(61)).astype(np.int64) - 4).astype(float).reshape(n, n)
    else:
        u = ((w >> np.uint64(11)).astype(np.float64) + 0.5) * 2.0 ** -53
        J = (np.sqrt(-2 * np.log(u)) * np.cos(2 * np.pi * u[::-1])).astype(np.float32)
        J = J.astype(np.float64).reshape(n, n) / np.sqrt(n)
    E = np.zeros((n + 1, n + 1))
    E[:n, :n] = np.triu(J, 1)
    E[n, n] = 1.0
    return E


def _ea3d(L, seed):
    n = L ** 3
    w = _splitmix(seed, 3 * n)
    sign = np.where((w >> np.uint64(63)) == 1, 1.0, -1.0)
    rows = [dict() for _ in range(n + 1)]
    for i in range(n):
        x, y, z = i % L, (i // L) % L, i // (L * L)
        for d, j in enumerate([((x + 1) % L) + L * y + L * L * z, x + L * ((y + 1) % L) + L * L * z,
                               x + L * y + L * L * ((z + 1) % L)]):
            rows[i][j] = rows[i].get(j, 0.0) + sign[3 * i + d]
            rows[j][i] = rows[j].get(i, 0.0) + sign[3 * i + d]
    rows[n][n] = 1.0
    rp, col, val = [0], [], []
    for i in range(n + 1):
        for j in sorted(rows[i]):
            if rows[i][j] != 0:
                col.append(j)
                val.append(rows[i][j])
        rp.append(len(col))
    return n, np.array(rp, np.int64), np.array(col, np.int64), np.array(val)


def _geometric(bmin, bmax, L):
    r = (bmax / bmin) ** (1.0 / (L - 1))
    return bmin * r ** np.arange(L)


def _time(ctx, beta, mcs, spins, reps, run=None):
    best = float("inf")
    for k in range(reps):
        ctx.set_spins(spins)
        if run is None:
            ctx.run(beta, mcs, 1 + k, 0)
        else:
            run(ctx, k)
        best = min(best, ctx.info()["last_kernel_ms"])
    return best


def _pick(times, order, tol=0.02):
    """The first candidate in `order` (simplest shapes first) within `tol` of
    the fastest: measurement noise does not flip a table entry."""
    best = min(times.values())
    for c in order:
        if c in times and times[c] <= best * (1 + tol):
            return c
    return min(times, key=times.get)


def tune_csr(capi, quick):
    n, rp, ci, va = _ea3d(12 if quick else 16, 0xB0B)
    L = 60 if quick else 150
    beta, mcs = _geometric(0.1, 5.0, L), np.ones(L, np.int64)
    prec = {"int32": capi.PREC_INT32, "fp32": capi.PREC_FP32, "fp64": capi.PREC_FP64}
    rows, table = [], []
    for ft in ("int32", "fp32", "fp64"):
        rs = (256, 1024, 4096) if quick else (256, 512, 1024, 2048, 4096)
        for R in rs:
            spins = _pm1((R, n), R)
            times = {}
            with capi.Context.csr(n, rp, ci, va, R, precision=prec[ft]) as ctx:
                # (3 waves, option csr_waves 3, is no candidate: 150-sweep
                # runs here favoured it by 3% on this graph at R >= 1024, but
                # 2000-sweep anneals took 54.9 ms against 51.6 with 4)
                for W in (1, 2, 4, 8):
                    ctx.set_option("csr_waves", W)
                    if ctx.info()["waves_per_replica"] != W:
                        continue
                    times[W] = _time(ctx, beta, mcs, spins, 2)
            best = _pick(times, (4, 2, 1, 8))
            rows.append({"ftype": ft, "R": R, "ms": times, "best": best})
            print(f"csr {ft} R={R}: {times} -> {best}", flush=True)
            table.append((FT[ft], R, best))
    return rows, table


def tune_dense(capi, quick):
    L = 12 if quick else 30
    beta, mcs = _geometric(0.1, 8.0, L), np.ones(L, np.int64)
    cands = [(1, 1, 0), (4, 1, 2), (4, 1, 0), (8, 1, 2), (8, 0, 2), (16, 0, 2)]
    rows, table = [], []
    for ft, prec, integer in (("fp32", capi.PREC_FP32, False), ("int32", capi.PREC_INT32, True)):
        for n in ((1024, 4096) if quick else (512, 1024, 2048, 4096)):
            ns = 1 << max(0, (n + 63) // 64 - 1).bit_length()
            E = _sk(n, 0xDE + n, integer)
            for R in ((256, 2048) if quick else (64, 256, 2048)):
                spins = _pm1((R, n), n + R)
                times = {}
                with capi.Context.dense(E, R, 1, 0.0, prec) as ctx:
                    ctx.set_option("mfma_sweep", 0)  # the per-flip kernels' own shapes
                    for G, col, win in cands:
                        try:
                            ctx.set_option("group", G)
                            ctx.set_option("colown", col)
                            ctx.set_option("xcd_window", win)
                        except capi.OjmiError:
                            continue
                        key = f"G{G}/{'colown' if col else 'lds'}/w{win}" if G > 1 else "wave"
                        times[key] = (_time(ctx, beta, mcs, spins, 2), (G, col, win))
                order = ["wave", "G4/colown/w2", "G4/colown/w0", "G8/colown/w2", "G8/lds/w2",
                         "G16/lds/w2"]
                best = _pick({k: v[0] for k, v in times.items()}, order)
                G, col, win = times[best][1]
                rows.append({"ftype": ft, "n": n, "slots": ns, "R": R,
                             "ms": {k: v[0] for k, v in times.items()}, "best": best})
                print(f"dense {ft} n={n} R={R}: "
                      f"{ {k: round(v[0], 2) for k, v in times.items()} } -> {best}", flush=True)
                table.append((FT[ft], ns, R, G, col, win))
    return rows, table


def _qubo_ising(n, seed):
    """Integer QUBO, Q_ij in [-8, 8] at density 1/2, as x4-scaled Ising
    couplings (J_ij = Q_ij, h_i = 2 Q_ii + sum_j Q_ij), and the bound
    max_i sum_j |Q_ij| of a flip's energy change."""
    w = _splitmix(seed, 2 * n * n)
    vals = ((w[: n * n] >> np.uint64(32)) * np.uint64(17) >> np.uint64(32)).astype(np.int64) - 8
    keep = (w[n * n:] >> np.uint64(11)).astype(np.float64) * 2.0 ** -53 < 0.5
    Q = np.triu(np.where(keep, vals, 0).reshape(n, n)).astype(float)
    U = np.triu(Q, 1)
    E = np.zeros((n + 1, n + 1))
    E[:n, :n] = U
    E[:n, n] = 2 * np.diag(Q) + U.sum(axis=1) + U.sum(axis=0)
    E[n, n] = 1.0
    S = np.abs(Q) + np.abs(U).T
    return E, max(float(S.sum(axis=1).max()), 1.0)


def tune_batch(capi, quick):
    """Uneven QUBO batches: sizes log-uniform in [32, 1024], 16 reads, the
    per-instance geometric schedule ln2 / dE_bound -> ln 100 (in QUBO units)."""
    K, R, L = (512, 16, 100) if quick else (1024, 16, 200)
    w = _splitmix(0xBA7C, K)
    u = (w >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    ns = np.rint(32.0 * 32.0 ** u).astype(int)
    Es, betas = [], []
    for k, n in enumerate(ns):
        E, bound = _qubo_ising(int(n), 0xBA + k)
        Es.append(E)
        betas.append(_geometric(np.log(2.0) / bound, np.log(100.0), L) / 4.0)
    betas = np.stack(betas)
    mcs = np.ones(L, np.int64)
    spins = [_pm1((R, int(n)), 7 + k) for k, n in enumerate(ns)]
    times = {}
    with capi.BatchContext.create(Es, R, capi.PREC_INT32) as ctx:
        for mode in (0, 1, 2):
            ctx.set_option("batch_mode", mode)
            times[mode] = _time(ctx, None, mcs, spins, 2,
                                run=lambda c, k: c.run_batch(betas, mcs, 1 + k, 0))
    best = _pick(times, (0, 1, 2))
    print(f"batch modes {times} -> {best}", flush=True)
    return {"problems": K, "replicas": R, "sweeps": L, "ms": times, "best": best}, best


def tune_mfma(capi, quick):
    K, R = 16, 2048
    L = 160 if quick else 500
    beta = _geometric(0.1, 10.0, L)
    rows, table = [], []
    for n in ((4096,) if quick else (2048, 4096)):
        ns = 1 << max(0, (n + 63) // 64 - 1).bit_length()
        E = _sk(n, 0xAC + n, False)
        spins = _pm1((R, n), n + 7)
        seg = {}
        for mode in (0, 1):
            with capi.Context.dense(E, R, 1, 0.0, capi.PREC_FP32) as ctx:
                ctx.set_option("group", 4)
                ctx.set_option("colown", 1)
                ctx.set_option("refresh_every", K)
                ctx.set_option("mfma_sweep", mode)
                ctx.set_spins(spins)
                out = []
                for c in range(0, L, K):
                    b = beta[c:c + K]
                    ctx.run(b, np.ones(len(b), np.int64), 3, c)
                    i = ctx.info()
                    out.append((i["last_flips"] / (R * n * len(b)),
                                (i["last_kernel_ms"] - i["last_refresh_ms"]) / len(b)))
                seg[mode] = out
        # segments of equal acceptance (the same spins, seed and schedule up to
        # the kernels' rounding): the crossover is where the matrix-core sweep
        # stops winning as the acceptance falls
        pairs = [(a, t0, t1) for (a, t0), (_, t1) in zip(seg[0], seg[1])]
        wins = [a for a, t0, t1 in pairs if t1 < t0]
        if not wins:
            permille = 0
        else:
            lo = min(wins)
            below = [a for a, t0, t1 in pairs if a < lo]
            hi = max(below) if below else 0.0
            permille = max(1, int(round(500.0 * (lo + hi))))
        rows.append({"n": n, "slots": ns, "R": R,
                     "segments": [{"accept": a, "colown_ms_per_sweep": t0,
                                   "mfma_ms_per_sweep": t1} for a, t0, t1 in pairs],
                     "accept_permille": permille})
        print(f"mfma n={n}: crossover at {permille} per mille "
              f"(colown {sum(p[1] for p in pairs) * K:.1f} ms, "
              f"matrix cores {sum(p[2] for p in pairs) * K:.1f} ms per anneal)", flush=True)
        table.append((ns, permille))
    return rows, table


def write_header(path, device, csr, dense, batch_mode, mfma):
    lines = [
        "// ojmi_tuned.h -- launch shapes chosen by openjij_amd/tune.py (generated;",
        f"// measured on {device}).  Re-run the tuner on another GPU and rebuild.",
        "#pragma once",
        "#include <stdint.h>",
        "",
        "namespace ojmi {",
        "namespace tuned {",
        "",
        "// CSR graphs of degree <= 16: waves per replica, by field type (0 FP64,",
        "// 1 FP32, 2 INT32) and replica count (first entry with R <= r_max).",
        "struct CsrShape {",
        "  int ftype;",
        "  int64_t r_max;",
        "  int waves;",
        "};",
        "constexpr CsrShape kCsr[] = {",
    ]
    by_ft = {}
    for ft, R, W in csr:
        by_ft.setdefault(ft, []).append((R, W))
    for ft in sorted(by_ft):
        ents = sorted(by_ft[ft])
        for i, (R, W) in enumerate(ents):
            # bucket edges halfway (geometrically) between the measured R
            r_max = int(np.sqrt(R * ents[i + 1][0])) if i + 1 < len(ents) else (1 << 40)
            lines.append(f"    {{{ft}, {r_max}, {W}}},")
    lines += [
        "};",
        "",
        "// Dense classical FP32 / INT32 systems: replicas per workgroup (1: one wave",
        "// per replica), column-owned fields (1) or rows staged in LDS (0), XCD",
        "// pacing window, by field type, 64-spin slot count and replica count",
        "// (first entry with matching type and slots and R <= r_max).",
        "struct DenseShape {",
        "  int ftype;",
        "  int slots;",
        "  int64_t r_max;",
        "  int group;",
        "  int colown;",
        "  int xcd_window;",
        "};",
        "constexpr DenseShape kDense[] = {",
    ]
    by = {}
    for ft, ns, R, G, col, win in dense:
        by.setdefault((ft, ns), []).append((R, G, col, win))
    for key in sorted(by):
        ents = sorted(by[key])
        for i, (R, G, col, win) in enumerate(ents):
            r_max = int(np.sqrt(R * ents[i + 1][0])) if i + 1 < len(ents) else (1 << 40)
            lines.append(f"    {{{key[0]}, {key[1]}, {r_max}, {G}, {col}, {win}}},")
    lines += [
        "};",
        "",
        "// Batches of problems: default launch (0 fused largest first, 1 per slot",
        "// class, 2 persistent per-XCD queues).",
        f"constexpr int kBatchMode = {batch_mode};",
        "",
        "// Dense FP32 systems with replicas to fill the device: per 16-sweep",
        "// segment the matrix-core sweep while the acceptance rate is at least",
        "// accept_permille per mille, the column-owned kernel below (0: never the",
        "// matrix-core sweep), by 64-spin slot count.",
        "struct MfmaChoice {",
        "  int slots;",
        "  int accept_permille;",
        "};",
        "constexpr MfmaChoice kMfma[] = {",
    ]
    for ns, pm in sorted(mfma):
        lines.append(f"    {{{ns}, {pm}}},")
    lines += [
        "};",
        "",
        "}  // namespace tuned",
        "}  // namespace ojmi",
        "",
    ]
    with open(path, "w") as f:
        f.write("\n".join(lines))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--quick", action="store_true", help="fewer buckets and sweeps")
    ap.add_argument("--out", default=os.path.join(HERE, "csrc", "ojmi_tuned.h"))
    ap.add_argument("--json", default=os.path.join(HERE, "tuned_shapes.json"))
    args = ap.parse_args()
    from openjij_amd import capi

    if capi.device_count() < 1:
        raise SystemExit("tune: no HIP device visible")
    device = "HIP device 0"
    try:  # the device name, from the HIP runtime the engine links
        import ctypes

        try:
            hip = ctypes.CDLL("libamdhip64.so")
        except OSError:
            hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
        buf = ctypes.create_string_buffer(256)
        if hip.hipInit(0) == 0 and hip.hipDeviceGetName(buf, 256, 0) == 0:
            device = buf.value.decode() or device
    except OSError:  # pragma: no cover
        pass
    t0 = time.time()
    csr_rows, csr = tune_csr(capi, args.quick)
    dense_rows, dense = tune_dense(capi, args.quick)
    batch_row, batch_mode = tune_batch(capi, args.quick)
    mfma_rows, mfma = tune_mfma(capi, args.quick)
    with open(args.json, "w") as f:
        json.dump({"device": device, "seconds": time.time() - t0, "csr": csr_rows,
                   "dense": dense_rows, "batch": batch_row, "mfma": mfma_rows}, f, indent=1)
    write_header(args.out, device, csr, dense, batch_mode, mfma)
    print(f"wrote {args.out} and {args.json} ({time.time() - t0:.0f} s)", flush=True)


if __name__ == "__main__":
    main()
