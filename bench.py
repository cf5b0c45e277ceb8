"""Benchmark: time to a TRUE relative residual 1e-6 on one MI355X, learned
8x8 block-Jacobi preconditioned PCG against plain Jacobi-PCG, plus the SpMV's
HBM roofline, on the thermal-FEM workload (P1 FEM stiffness, 1100 x 1100
node grid, n = 1,205,604, ~8.4M nonzeros, log-normal conductivity sigma = 1
with 20 inclusions at 1e3 contrast, b ~ U[-1,1], x0 = 0, synthetic, seed
20263203).

A step is one complete time to solution: from the CSR matrix and b
resident in HBM to x in HBM with ||b - A x|| / ||b|| <= 1e-6 (FP64, on the
GPU).  Everything a method needs per matrix is inside the step:

  learned   (the value)  the matrix set up on the device (parallel
            strength aggregation into blocks of <= 8 rows, SELL layout),
            the solver, the learned 8x8 blocks (device features, BF16
            MFMA inference, FP16 factors) and the solve with the accurate
            iteration: FP32 r and z, matrix values as FP32 hi + 16-bit
            remainders, FP64-accumulated SpMV, FP64 p, A p and x, FP64
            reductions, residual replacement on the FP64 true residual, and
            the true-residual stagnation fallback to Jacobi armed;
  accurate  the same iteration with plain Jacobi (its own device set-up);
  refined   FP32 Jacobi-PCG inner solves restarted from FP64 outer residuals;
  plain     FP32 Jacobi-PCG stopped on its recursive residual (its true
            residual is reported: FP32 arithmetic does not reach 1e-6 on this
            matrix), in the three-kernel form (the reference's) and the
            fused form.
The network and the layouts (sort window, SpMV inner loop, workgroups) are
chosen beforehand on OTHER matrices of the family (training split, untimed),
as a deployment would; the solved matrix is held out.  The speedups are
ratios of these totals; setup and solve are also reported apart, with the
number of right-hand sides from which the learned method is ahead.

Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--profile]

--profile runs only the learned solve with a fixed layout (--sigma, --grid;
no autotuning), for kernel profilers.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "project-xplus_amd"))

METRIC = "time-to-1e-6 residual: learned-precond vs plain Jacobi-PCG; SpMV HBM GB/s"
HBM_PEAK_GBS = 8000.0
SEED = 20261203 + 1000 * 2
TUNE_SEED = 20266303          # the training-split matrix the layouts are tuned on
# Workgroup counts a lone SpMV is tuned over.  A lone SpMV sums no partials,
# so it can take far more workgroups than an iteration kernel: beyond the
# MALL, 16,384 workgroups of four wavefronts stream 5-10% faster than 1,024.
SPMV_GRIDS = (256, 512, 1024, 2048, 4096, 8192, 16384, 32768)


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def timed(X, fn, warmup: int, steps: int):
    for _ in range(warmup):
        fn()
    X.synchronize()
    t0 = time.perf_counter()
    out = None
    for _ in range(steps):
        out = fn()
    X.synchronize()
    return (time.perf_counter() - t0) * 1000.0 / max(steps, 1), out


def roof(kernel: str, alg_bytes: float, us: float, note: str) -> dict:
    achieved = alg_bytes / (us * 1e-6) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": kernel, "avg_us": round(us, 3),
            "alg_bytes": int(alg_bytes), "alg_bytes_formula": note}


def profiled_traffic() -> dict:
    """HBM bytes per launch measured by the rocprofv3 FETCH_SIZE / WRITE_SIZE
    passes committed under profiles/ (lines 'traffic_bytes_per_launch K B')."""
    out = {}
    for f in sorted((ROOT / "profiles").glob("*.md")):
        for line in f.read_text().splitlines():
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch":
                out[parts[1]] = (int(parts[2]), f"profiles/{f.name}")
    return out


def spmv_tuned(X, a, reps: int = 200):
    """The standalone FP32 SpMV on its own matrix (sigma 64), tuned for the
    SpMV alone: the fastest of inner loop x workgroup count by its own timing
    (the solvers' matrices are tuned for whole iterations, which can pick a
    layout that is slower for a lone SpMV)."""
    m = X.Matrix(a, sigma=64)
    best = None
    for v in (0, 1, 2):
        m.set_variant(v)
        for g in SPMV_GRIDS:
            m.set_grid(g)
            us = m.spmv_f32_time(20, reps)
            if best is None or us < best[0]:
                best = (us, v, g)
    m.set_variant(best[1])
    m.set_grid(best[2])
    return m, {"spmv_variant": best[1], "grid": best[2]}


def spmv_roofline(m, a, reps: int = 200) -> dict:
    """Standalone FP32 SELL SpMV (k_spmv_f32), HIP events on its stream."""
    us = m.spmv_f32_time(20, reps)
    r = roof("k_spmv_f32", 8 * a.nnz + 12 * a.n, us, "8 nnz + 12 n (FP32 value + int32 column; row ptr, x, y)")
    r["stream_bytes"] = m.info()["bytes_spmv"]
    # warm (above: the working set can stay in the 256 MB MALL between
    # launches) and cold (a 1 GiB copy before every launch: HBM)
    cold = m.spmv_f32_time_cold(20, 1 << 30)
    r["cold"] = {"avg_us": round(cold, 3), "achieved": round((8 * a.nnz + 12 * a.n) / (cold * 1e-6) / 1e9, 1),
                 "frac": round((8 * a.nnz + 12 * a.n) / (cold * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
                 "protocol": "1 GiB device copy before each of 20 launches"}
    prof = profiled_traffic()
    if r["kernel"] in prof:
        r["traffic"], r["traffic_source"] = prof[r["kernel"]]
    return r


def stream_reference(X) -> dict:
    """The project-written streaming kernels over 1 GiB (beyond the MALL):
    the best HBM rate a read / copy reaches on this GPU, for scale -- plain
    16-byte loads and non-temporal 16-byte loads four in flight per thread,
    over 1024 / 4096 / 16384 workgroups."""
    nb = 1 << 30
    grids = (1024, 4096, 16384)
    reads = {(mode, g): X.stream_time(nb, mode, g) for mode in (0, 2) for g in grids}
    copies = {(mode, g): X.stream_time(nb, mode, g) for mode in (1, 3) for g in grids}
    (mode, g), r = min(reads.items(), key=lambda kv: kv[1])
    (mc, gc), c = min(copies.items(), key=lambda kv: kv[1])
    plain = min(v for (m, _), v in reads.items() if m == 0)
    # a launch that moves (almost) nothing: 16 KiB read by 256 / 512 workgroups
    # of 1024 threads, the fixed cost every iteration kernel pays
    floor = {g: round(X.stream_time(1 << 14, 0, g, 200), 3) for g in (256, 512)}
    return {"read_gbs": round(nb / (r * 1e-6) / 1e9, 1), "copy_gbs": round(2 * nb / (c * 1e-6) / 1e9, 1),
            "read_frac": round(nb / (r * 1e-6) / 1e9 / HBM_PEAK_GBS, 4), "bytes": nb,
            "plain_read_gbs": round(nb / (plain * 1e-6) / 1e9, 1), "launch_floor_us": floor,
            "kernels": f"best read: {'k_stream_read_nt (non-temporal, 4 loads in flight)' if mode == 2 else 'k_stream_read'} "
                       f"at {g} workgroups; copy: "
                       f"{'k_stream_copy_nt (non-temporal stores)' if mc == 3 else 'k_stream_copy'} at {gc} workgroups; "
                       f"16-byte loads"}


def spmv_beyond_mall(X, mesh: int = 2400) -> dict:
    """The FP32 SpMV on a matrix of the same family too large for the 256 MB
    MALL (algorithmic bytes 391 MB at 2400 x 2400), so that warm launches
    stream the matrix from HBM: tuned over inner loop x workgroups, warm and
    cold (1 GiB copy before each launch)."""
    a = X.gen_fem2d(mesh, 1.0, 20, 1e3, 0.3, SEED + mesh)
    d = X.DeviceCsr(a)
    m = X.Matrix.from_device(d, sigma=64)
    alg = 8 * a.nnz + 12 * a.n
    best = None
    for v in (0, 1, 2):
        m.set_variant(v)
        for g in SPMV_GRIDS:
            m.set_grid(g)
            us = m.spmv_f32_time(10, 50)
            if best is None or us < best[0]:
                best = (us, v, g)
    m.set_variant(best[1])
    m.set_grid(best[2])
    cold = m.spmv_f32_time_cold(10, 1 << 30)
    m.close()
    d.close()
    out = {"mesh": mesh, "n": a.n, "nnz": a.nnz, "alg_bytes": alg, "layout": {"spmv_variant": best[1],
                                                                                 "grid": best[2]},
           "warm_us": round(best[0], 3), "warm_gbs": round(alg / (best[0] * 1e-6) / 1e9, 1),
           "warm_frac": round(alg / (best[0] * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
           "cold_us": round(cold, 3), "cold_gbs": round(alg / (cold * 1e-6) / 1e9, 1),
           "cold_frac": round(alg / (cold * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)}
    log(f"spmv beyond the MALL: {out}")
    return out


def bitmap_block_spmv(X, a) -> dict:
    """The same matrix in the reference's 4x4 bitmap-block layout (plain
    layout, 144-byte blocks + a 4-byte block column each): the direct FP64
    SpMV (k_bitmap_spmv, HIP events around each launch) against the bytes
    the format streams, the host packer, and the import into a device matrix
    (device unpack + rule-3 set-up)."""
    t0 = time.perf_counter()
    bm = X.bitmap_pack(a, 0)
    pack_s = time.perf_counter() - t0
    d = bm.to_device()
    x = X.DeviceArray.from_numpy(np.random.default_rng(1).standard_normal(a.n))
    y = X.DeviceArray(a.n, np.float64)
    us = d.time(x, y, 20, 200)
    nbr = (a.n + 3) // 4
    alg = bm.nblocks * (144 + 4) + 4 * (nbr + 1) + 16 * a.n
    X.synchronize()
    t0 = time.perf_counter()
    m = X.Matrix.from_bitmap(d, block_rows=8, blocking=3, keep_lo=True)
    X.synchronize()
    imp_ms = 1e3 * (time.perf_counter() - t0)
    m.close()
    for buf in (x, y, d):
        buf.close()
    out = {"kernel": "k_bitmap_spmv", "layout": "plain (window 0)", "nblocks": bm.nblocks,
           "entries_per_block": round(a.nnz / max(bm.nblocks, 1), 3), "avg_us": round(us, 3), "alg_bytes": alg,
           "alg_bytes_note": "148 B per block (SpmvBlock + block column) + 4 B per block row + x and y (16 B per row)",
           "achieved_gbs": round(alg / (us * 1e-6) / 1e9, 1),
           "frac": round(alg / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
           "pack_host_s": round(pack_s, 3), "import_setup_ms": round(imp_ms, 2)}
    log(f"bitmap-block SpMV: {out}")
    return out


def iteration_roofline(s, a, block: bool, iters: int = 200) -> dict:
    """The accurate iteration's two kernels timed with HIP events between
    launches on the solver stream; the dominant one is k_pcg_spmv_fused_acc."""
    even, odd = s.kernel_times_split(iters)
    us = (even + odd) / 2.0
    b_fused = 12 * a.nnz + 40 * a.n
    b_upd = (54 if block else 40) * a.n
    r = roof("k_pcg_spmv_fused_acc", b_fused, us[0],
             "12 nnz + 40 n (FP64 value + int32 column; row ptr, z, p and A p read and written in FP64)")
    r["update_kernel"] = roof("k_pcg_update_xrz_acc", b_upd, us[1],
                              ("54" if block else "40") + " n on average over an even and an odd iteration ("
                              + ("38 / 70" if block else "24 / 56") + " n: r r/w, A p FP64, "
                              + ("the 36 FP32 entries of an SPD 8x8 block, 18 B per row" if block else "1/diag")
                              + ", z; the odd one adds x r/w FP64 and two FP64 directions)")
    for name, t, per_n in (("even", even[1], 38 if block else 24), ("odd", odd[1], 70 if block else 56)):
        q = roof("k_pcg_update_xrz_acc", per_n * a.n, t, f"{per_n} n")
        r["update_kernel"][name] = {k: q[k] for k in ("avg_us", "alg_bytes", "achieved", "frac")}
    prof = profiled_traffic()
    for rr in (r, r["update_kernel"]):
        if rr["kernel"] in prof:
            rr["traffic"], rr["traffic_source"] = prof[rr["kernel"]]
    r["iteration_us"] = round(float(us[0] + us[1]), 3)
    r["iteration_gbs"] = round((b_fused + b_upd) / ((us[0] + us[1]) * 1e-6) / 1e9, 1)
    return r


def host_description() -> dict:
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model, "machine": os.uname().machine}


def cpu_golden_worker(mesh: int) -> int:
    """Child process: the oracle's FP64 golden Jacobi-PCG (one thread, -O2)
    on the same matrix, a full solve to relative residual 1e-6, timed around
    the solve only (matrix generation excluded).  Prints one JSON line."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as O
    import xplus_amd as X
    a = X.gen_fem2d(mesh, 1.0, 20, 1e3, 0.3, SEED)
    tau = (1e-6 * float(np.linalg.norm(a.b))) ** 2
    t0 = time.perf_counter()
    g = O.golden(a, tau=tau, trace_cap=1)
    dt = time.perf_counter() - t0
    res = O.residual_norm(a, g["x"]) / float(np.linalg.norm(a.b))
    print(json.dumps({"ms": dt * 1000.0, "iterations": g["iterations"], "converged": g["converged"],
                      "true_rel_residual": res}), flush=True)
    return 0


def start_cpu_golden(mesh: int):
    """Starts the golden run as a child process (before this process touches
    the GPU): it runs on one host core while the GPU work proceeds."""
    import subprocess
    return subprocess.Popen([sys.executable, str(Path(__file__).resolve()), "--cpu-golden-worker", "--mesh",
                             str(mesh)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)


def finish_cpu_golden(proc, wait_s: float) -> dict:
    try:
        out, err = proc.communicate(timeout=wait_s)
    except Exception:  # noqa: BLE001 -- a timeout or an I/O error: no measurement
        proc.kill()
        proc.communicate()
        return None
    if proc.returncode != 0:
        log(f"cpu golden failed: {err[-500:]}")
        return None
    r = json.loads(out.strip().splitlines()[-1])
    return {"value": round(r["ms"], 1), "unit": "ms (time to true relative residual 1e-6)", "cores": 1,
            "kind": "port", "iterations": r["iterations"], "true_rel_residual": r["true_rel_residual"],
            "sample": "one full solve of the same system (oracle FP64 golden Jacobi-PCG, "
                      "oracle/xplus_oracle.c, one thread, -O2), run on the host while the GPU methods ran",
            "host": host_description()}


def train_network(X, args):
    """Trains the block network on the training split of the FEM family
    (grids from --train-grids, seeds 20266203 + i; the solved matrix uses
    seed 20263203, so it is held out).  Untimed."""
    t0 = time.perf_counter()
    sets = None
    grids = [int(g) for g in args.train_grids.split(",") if g]
    for i, g in enumerate(grids):
        bs = X.BlockSet(X.gen_fem2d(g, 1.0, 20, 1e3, 0.3, 20266203 + i), blocking=args.blocking, lam=args.lam)
        # blocking 3: the training blocks come from the same device aggregation
        if sets is None:
            sets = bs
        else:
            sets.append(bs)
    feat_s = time.perf_counter() - t0
    net = X.Mlp(seed=20266203)
    t0 = time.perf_counter()
    hist = net.train(sets, args.train_steps, batch=4096, lr=2e-3, seed=1)
    train_s = time.perf_counter() - t0
    info = {"grids": grids, "blocks": len(sets), "blocking": args.blocking, "lambda": args.lam,
            "steps": args.train_steps, "batch": 4096,
            "loss_first": float(np.mean(hist[:10])), "loss_last": float(np.mean(hist[-50:])),
            "host_feature_s": round(feat_s, 2), "train_s": round(train_s, 2)}
    log(f"trained: {info}")
    return net, info


def accurate_solve(X, a, m, args, net=None, kernels: bool = False) -> dict:
    """The accurate iteration with residual replacement to true relative
    residual 1e-6 (learned blocks when net is given; their setup timed)."""
    s = X.Pcg(m, rel_tol=1e-6, accurate=True, stall_limit=0)
    s.set_rhs(a.b, a.x0)
    setup_ms = 0.0
    if net is not None:
        X.synchronize()
        t0 = time.perf_counter()
        s.set_mlp_preconditioner(net)
        X.synchronize()
        setup_ms = (time.perf_counter() - t0) * 1000.0
    ms, (rel, repl, iters) = timed(X, lambda: s.run_reliable(1e-6, 0.0), args.warmup, args.steps)
    x = s.x()
    check = float(np.linalg.norm(a.b - X.csr_spmv_f64(a, x)) / np.linalg.norm(a.b))
    out = {"setup_ms": round(setup_ms, 3), "solve_ms": round(ms, 3), "total_ms": round(setup_ms + ms, 3),
           "iterations": iters, "replacements": repl, "true_rel_residual": float(rel[0]),
           "recomputed_rel_residual": check, "device_ms": round(s.last_timing()[0], 3),
           "us_per_iteration": round(ms * 1000.0 / max(iters, 1), 3), "fallback": s.status()["fallback"]}
    if kernels:
        out["roofline"] = iteration_roofline(s, a, net is not None)
    s.close()
    if not (rel[0] <= 1e-6 and check <= 1.01e-6):
        raise RuntimeError(f"accurate solve missed 1e-6: {out}")
    return out


def learned_matrix(X, a, args, tune: bool, keep64: bool = False):
    t0 = time.perf_counter()
    ml = X.Matrix(a, sigma=args.sigma if not tune else 256, keep_lo=not keep64, keep64=keep64, block_rows=8,
                  grid=args.grid, blocking=args.blocking)
    log(f"learned matrix (blocking {args.blocking}) built in {time.perf_counter() - t0:.2f} s "
        f"({'device' if args.blocking == 3 else 'host'})")
    t = ml.autotune(fused=2, tune_sigma=not args.no_tune_sigma, iters=64) if tune else None
    return ml, t


def calibration_copies(X) -> None:
    """A copy of known size: 256 MiB read + 256 MiB written by the runtime's
    copy kernel, three times -- the scale for memory-traffic counters."""
    src, dst = X.DeviceArray(1 << 26, np.float32), X.DeviceArray(1 << 26, np.float32)
    for _ in range(3):
        dst.copy_from(src)
    src.close()
    dst.close()


def profile_run(X, a, args) -> int:
    calibration_copies(X)
    # the standalone SpMV on the Jacobi matrix, in the layout spmv_tuned picks
    # for this workload (sigma 64, inner loop 0, 512 workgroups)
    m = X.Matrix(a, sigma=64, grid=512)
    spmv = spmv_roofline(m, a)
    m.close()
    bitmap = bitmap_block_spmv(X, a)
    net, _ = train_network(X, args)
    ml, _ = learned_matrix(X, a, args, tune=False)
    out = accurate_solve(X, a, ml, args, net=net, kernels=True)
    out["matrix"] = ml.info()
    # the value's path: the same blocks as the smoother of the multilevel cycle
    s = X.Pcg(ml, rel_tol=1e-6, accurate=True, stall_limit=0)
    s.set_rhs(a.b, a.x0)
    s.set_mlp_preconditioner(net)
    s.set_multilevel(args.ml_coarse_max, args.ml_omega, args.ml_gamma, args.ml_coarse_blocking, 0, args.ml_theta)
    X.synchronize()
    t0 = time.perf_counter()
    rel, repl, iters = s.run_reliable(1e-6, 0.0)
    X.synchronize()
    out["multilevel"] = {"solve_ms": round((time.perf_counter() - t0) * 1000.0, 3), "iterations": iters,
                         "true_rel_residual": float(rel[0]), "hierarchy": s.multilevel_info(),
                         "vcycle_us": round(s.multilevel_time(50), 3)}
    s.close()
    out["spmv_roofline"] = spmv
    out["bitmap_block_spmv"] = bitmap
    print(json.dumps({"profile": "learned accurate solve", **out}), flush=True)
    return 0


def workload_c3(X, args) -> int:
    """FP64-refined solves of the FEM family at 1100 x 1100 and 1600 x 1600
    nodes to true relative residual 1e-10: the restarted FP32 refinement the
    configuration describes (inner FP32 Jacobi-PCG to 1e-5, FP64 outer
    residual and solution update), the accurate iteration with the final
    FP64 check, and the same with the learned 8x8 blocks (the network trained
    once, untimed; its device setup timed), each timed with everything
    resident.  The matrices keep exact FP64 values (keep64): for a 1e-10
    target the accurate iteration reads 8 bytes per entry either way (FP32
    hi + FP32 lo, ~48 bits, or FP64), and the ~48-bit values slow the last
    decades on the 1600 x 1600 matrix (110,677 iterations against 57,747)."""
    out = []
    net, tinfo = train_network(X, args)
    for mesh in (1100, 1600):
        a = X.gen_fem2d(mesh, 1.0, 20, 1e3, 0.3, 20261203 + 1000 * 3 + mesh)
        m = X.Matrix(a, sigma=256, keep64=True)
        m.autotune(fused=True, tune_sigma=False, iters=64)
        r = X.Pcg(m, rel_tol=1e-10, fused=True)
        r.set_rhs(a.b, a.x0)
        ref_ms, (rel, outer, inner) = timed(X, lambda: r.run_refined(1e-10, 1e-5, 60), 0, 1)
        r.close()
        m.autotune(fused=2, tune_sigma=False, iters=64)
        s = X.Pcg(m, rel_tol=1e-10, accurate=True)
        s.set_rhs(a.b, a.x0)
        acc_ms, (rel2, repl, iters) = timed(X, lambda: s.run_reliable(1e-10, 0.1), 0, 1)
        acc01_ms, (rel4, repl4, iters4) = timed(X, lambda: s.run_reliable(1e-10, 0.01), 0, 1)
        s.close()
        m.close()
        log(f"c3 {mesh}: refined {ref_ms:.1f} ms, accurate {acc_ms:.1f} ms ({iters} iterations), "
            f"delta 0.01: {acc01_ms:.1f} ms ({iters4})")
        # block solves capped at 3x the Jacobi count, so one that stalls ends;
        # replacement when r^T r dropped 1e4-fold (delta 0.01)
        ml, _ = learned_matrix(X, a, args, tune=True, keep64=True)
        blocks = {}
        for name in ("exact", "learned", "learned_multilevel"):
            q = X.Pcg(ml, rel_tol=1e-10, accurate=True, stall_limit=0, max_iters=3 * iters)
            q.set_rhs(a.b, a.x0)
            X.synchronize()
            t0 = time.perf_counter()
            if name.startswith("learned"):
                q.set_mlp_preconditioner(net)
            else:
                q.set_preconditioner(1)
            if name == "learned_multilevel":
                q.set_multilevel(args.ml_coarse_max, args.ml_omega, args.ml_gamma, args.ml_coarse_blocking, 0, args.ml_theta)
            X.synchronize()
            setup_ms = (time.perf_counter() - t0) * 1000.0
            b_ms, (rel3, repl3, iters3) = timed(X, lambda: q.run_reliable(1e-10, 0.01), 0, 1)
            blocks[name] = {"ms": round(setup_ms + b_ms, 1), "setup_ms": round(setup_ms, 2), "iterations": iters3,
                            "replacements": repl3, "true_rel_residual": float(rel3[0]),
                            "status": X.STATUS_NAMES[q.status()["status"]]}
            q.close()
            log(f"c3 {mesh}: {name} blocks {blocks[name]}")
        ml.close()
        row = {"mesh": mesh, "n": a.n, "nnz": a.nnz,
               "refined": {"ms": round(ref_ms, 1), "outer": outer, "inner_iterations": inner,
                           "true_rel_residual": float(rel[0])},
               "accurate": {"ms": round(acc_ms, 1), "iterations": iters, "replacements": repl,
                            "true_rel_residual": float(rel2[0]), "delta": 0.1},
               "accurate_delta_0.01": {"ms": round(acc01_ms, 1), "iterations": iters4, "replacements": repl4,
                                       "true_rel_residual": float(rel4[0])},
               "exact_blocks_accurate": blocks["exact"], "learned_accurate": blocks["learned"],
               "learned_multilevel_accurate": blocks["learned_multilevel"]}
        log(f"c3: {row}")  # one line per grid: the run prints as it goes
        out.append(row)
    print(json.dumps({"workload": "FP64-refined thermal-FEM to true relative residual 1e-10", "unit": "ms",
                      "training": tinfo, "results": out}), flush=True)
    return 0


def workload_c4(X, args) -> int:
    """256 independent 3D 7-point diffusion systems, edges k ~ U{16..64},
    log-normal coefficients sigma 1.5, solved together in one batched solve
    (one SELL matrix, per-system state, one graph) to relative residual 1e-6
    each: the FP32 iteration to the recursive residual and the accurate
    iteration to the true one; each with the workgroups of stopped systems
    marked idle between graph launches (the default) and with every
    workgroup launched ("_every_workgroup")."""
    rng = np.random.default_rng(20264203)
    ks = rng.integers(16, 65, 256)
    t0 = time.perf_counter()
    systems = [X.gen_diffusion3d(int(k), 1.5, 20261203 + 1000 * 4 + i) for i, k in enumerate(ks)]
    gen_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    m = X.Matrix(systems, sigma=256, keep_lo=True)
    build_s = time.perf_counter() - t0
    n, nnz = sum(a.n for a in systems), sum(a.nnz for a in systems)
    log(f"c4: {len(systems)} systems, n={n}, nnz={nnz}, generated in {gen_s:.1f} s, SELL built in {build_s:.1f} s")
    b = np.concatenate([a.b for a in systems])
    res = {"systems": len(systems), "n": n, "nnz": nnz, "k_min": int(ks.min()), "k_max": int(ks.max())}
    for name, kw in (("fp32", {"fused": True}), ("accurate", {"accurate": True})):
        m.autotune(fused=2 if name == "accurate" else True, tune_sigma=False, iters=32)
        for full in (True, False):
            s = X.Pcg(m, rel_tol=1e-6, full_work_map=full, **kw)
            s.set_rhs(b)
            if name == "fp32":
                ms, _ = timed(X, s.run, args.warmup, args.steps)
                rel = s.true_residual() / np.array([np.linalg.norm(a.b) for a in systems])
            else:
                ms, (rel, _, _) = timed(X, lambda: s.run_reliable(1e-6, 0.0), args.warmup, args.steps)
            its = [s.status(k)["iterations"] for k in range(len(systems))]
            sts = [s.status(k)["status"] for k in range(len(systems))]
            key = name if not full else name + "_every_workgroup"
            # algorithmic bytes of every system's iterations (the kernels'
            # per-iteration formulas, FP32: 8 nnz + 56 n; accurate: 12 nnz +
            # 94 n), over the whole batch's time
            per_nnz, per_n = (8, 56) if name == "fp32" else (12, 94)
            moved = sum(it * (per_nnz * a.nnz + per_n * a.n) for it, a in zip(its, systems))
            res[key] = {"ms": round(ms, 2), "iterations_min": min(its), "iterations_max": max(its),
                        "iterations_sum": int(sum(its)),
                        "converged": int(sum(1 for v in sts if v == X.CONVERGED)),
                        "max_true_rel_residual": float(np.max(rel)),
                        "alg_bytes": int(moved), "achieved_gbs": round(moved / (ms * 1e-3) / 1e9, 1),
                        "frac": round(moved / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                        "alg_bytes_formula": f"sum over systems of iterations x ({per_nnz} nnz + {per_n} n)"}
            log(f"c4 {key}: {res[key]}")
            s.close()
    print(json.dumps({"workload": "256 batched 3D 7-point systems to relative residual 1e-6", "unit": "ms",
                      **res}), flush=True)
    return 0


def workload_jacobi(X, args) -> int:
    """Plain Jacobi iteration (the reference's Cuper-jacobi-iteration
    contract: a fixed count of FP32 updates x = D^-1 (b - R x) from x0 = 0,
    one k_jacobi_sweep launch per update, no stop test)
    on a thermal2-sized 2D diffusion matrix (1108 x 1108 grid,
    n = 1,227,664, 5-point, log-normal conductivity sigma 1) with
    b = A x_ref (the dataset converter's smooth x_ref).  1000 updates per
    step, graph-captured, matrix and vectors resident."""
    if args.profile:
        calibration_copies(X)
    a = X.gen_diffusion2d(1108, 1108, 1.0, 20261203 + 1000 * 5)
    b = a.to_scipy() @ X.reference_solution(a.n)      # CSR-order FP64 row sums, as the converter
    m = X.Matrix(a)
    info = m.info()
    iters = 1000
    s = X.Jacobi(m, max_iters=iters)
    s.set_rhs(b)
    sweep = {}
    if args.grid_sweep:
        for g in SPMV_GRIDS:
            m.set_grid(g)
            sweep[g] = round(s.kernel_time(100), 3)
            log(f"jacobi grid {g}: {sweep[g]} us")
        m.set_grid(min(sweep, key=sweep.get) if args.grid == 0 else args.grid)
    elif args.grid:
        m.set_grid(args.grid)
    ms, _ = timed(X, s.run, args.warmup, args.steps)
    st = s.status()
    us = s.kernel_time(200)
    spmv_us = m.spmv_f32_time(20, 200)     # the plain FP32 SpMV of the same matrix, for scale
    off = a.nnz - a.n
    alg = (6 if info["idx16"] else 8) * off + 16 * a.n
    r = roof("k_jacobi_sweep", alg, us, ("6" if info["idx16"] else "8") + " (nnz - n) + 16 n (FP32 off-diagonal value + "
             "column; x, b, 1/D read, x_new written)")
    prof = profiled_traffic()
    if r["kernel"] in prof:
        r["traffic"], r["traffic_source"] = prof[r["kernel"]]
    cpu = None
    if not args.no_cpu and not args.profile:
        sys.path.insert(0, str(ROOT / "tests"))
        import oracle_lib as O
        k = 4
        t0 = time.perf_counter()
        O.jacobi(a, b=b, max_iters=k)
        dt = time.perf_counter() - t0
        k2 = max(k, min(400, int(k * 15.0 / max(dt, 1e-3))))
        t0 = time.perf_counter()
        O.jacobi(a, b=b, max_iters=k2)
        dt2 = time.perf_counter() - t0
        cpu = {"value": round(dt2 * 1000.0 / k2 * iters, 1), "unit": f"ms per {iters} updates (projected)",
               "cores": 1, "kind": "port",
               "sample": f"{k2} updates of the same system in {dt2:.2f} s (oracle FP32 Jacobi, one thread)"}
    print(json.dumps({"workload": f"Jacobi iteration, {iters} FP32 updates, 2D diffusion 1108 x 1108 "
                                  f"(n={a.n}, nnz={a.nnz})", "metric": "ms per 1000 Jacobi updates",
                      "value": round(ms, 2), "unit": "ms", "higher_is_better": False,
                      "updates_per_s": round(iters / (ms * 1e-3), 1), "iterations": st["iterations"],
                      "final_diff": st["diff"], "graph_launches": s.timing()[1], "grid": m.info()["grid"],
                      "grid_sweep_us": sweep, "spmv_f32_us_same_matrix": round(spmv_us, 3), "roofline": r,
                      "cpu_baseline": cpu, "dtype": "f32", "data": "synthetic"}), flush=True)
    return 0


def tune_layouts(X, args) -> dict:
    """Layouts chosen by the autotuner on a training-split matrix of the same
    family and size (seed TUNE_SEED), untimed: the sort window, SpMV inner
    loop and workgroup count of each method, passed to the held-out matrix's
    device set-up.  Returns {method: {sigma, spmv_variant, grid, ...}}."""
    t0 = time.perf_counter()
    t = X.gen_fem2d(args.mesh, 1.0, 20, 1e3, 0.3, TUNE_SEED)
    out = {}
    m = X.Matrix(t, sigma=256)
    out["plain_fused"] = m.autotune(fused=True, tune_sigma=not args.no_tune_sigma, iters=64)
    out["plain_unfused"] = m.autotune(fused=False, tune_sigma=False, iters=64)
    m.close()
    m = X.Matrix(t, sigma=256, keep_lo=True)
    out["accurate_jacobi"] = m.autotune(fused=2, tune_sigma=not args.no_tune_sigma, iters=64)
    m.close()
    m = X.Matrix(t, sigma=256, keep_lo=True, block_rows=8, blocking=args.blocking)
    out["learned"] = m.autotune(fused=2, tune_sigma=not args.no_tune_sigma, iters=64)
    m.close()
    log(f"layouts tuned on the training matrix (seed {TUNE_SEED}) in {time.perf_counter() - t0:.1f} s: {out}")
    return out


METHODS = {
    # name: (matrix options, solver options, run)
    "plain_unfused": ({"block_rows": 1}, {"fused": False}, "run"),
    "plain_fused": ({"block_rows": 1}, {"fused": True}, "run"),
    "refined_jacobi": ({"block_rows": 1, "keep_lo": True}, {"fused": None}, "refined"),
    "accurate_jacobi": ({"block_rows": 1, "keep_lo": True}, {"accurate": True}, "reliable"),
    "learned": ({"block_rows": 8, "keep_lo": True}, {"accurate": True}, "reliable"),
    # ~45 iterations: direct launches in chunks of 16 (no graph to capture
    # and instantiate per solve; the GPU, at ~170 us per iteration, stays
    # ahead of the host's launches)
    "learned_multilevel": ({"block_rows": 8, "keep_lo": True},
                           {"accurate": True, "graph_iters": 16, "use_graph": False}, "reliable"),
}


def time_to_solution(X, d, a, name, layout, args, net=None, fused_refined=True) -> dict:
    """W untimed + K timed complete solves of method `name`: set-up on the
    device from the resident CSR (matrix, solver, right-hand side,
    preconditioner), solve, solution into device memory.  Setup and solve
    are timed apart (a device synchronisation between them)."""
    mopts, popts, run = METHODS[name]
    mkw = dict(mopts)
    mkw.update({"sigma": layout["sigma"], "grid": layout["grid"], "spmv_variant": layout["spmv_variant"]})
    learned = name.startswith("learned")
    if learned:
        mkw["blocking"] = args.ml_blocking if name == "learned_multilevel" else args.blocking
    pkw = dict(popts)
    if pkw.get("fused", 0) is None:
        pkw["fused"] = fused_refined
    if learned:
        pkw["stall_limit"] = args.stall_limit
    xout = X.DeviceArray(a.n, np.float64)
    info = {}

    def one():
        X.synchronize()
        t0 = time.perf_counter()
        m = X.Matrix.from_device(d, **mkw)
        s = X.Pcg(m, rel_tol=1e-6, **pkw)
        s.set_rhs_device(d.b.ptr)
        if net is not None:
            s.set_mlp_preconditioner(net)
        if name == "learned_multilevel":
            s.set_multilevel(args.ml_coarse_max, args.ml_omega, args.ml_gamma, args.ml_coarse_blocking, 0, args.ml_theta)
        X.synchronize()
        t1 = time.perf_counter()
        if run == "run":
            s.run()
            r = None
        elif run == "refined":
            r = s.run_refined(1e-6, args.inner_rel, 50)
        else:
            r = s.run_reliable(1e-6, 0.0)
        s.x_device(xout.ptr)
        t2 = time.perf_counter()
        if name == "learned_multilevel":                 # after the timed region
            info["multilevel"] = s.multilevel_info()
            info["ml_rounds"] = s.multilevel_rounds()
            info["vcycle_us"] = s.multilevel_time(50)
        info["status"] = s.status()
        info["ret"] = r
        info["matrix"] = m.info()
        info["setup_stats"] = m.setup_stats()
        s.close()
        m.close()
        return (t1 - t0) * 1000.0, (t2 - t1) * 1000.0

    # warm-up steps run the whole path too; the first of a process also pays
    # the device allocations the library's caching allocator then reuses
    warm = [one() for _ in range(args.warmup)]
    tt = [one() for _ in range(args.steps)]
    x = xout.numpy()
    xout.close()
    setup = float(np.mean([v[0] for v in tt]))
    solve = float(np.mean([v[1] for v in tt]))
    st = info["status"]
    res = float(np.linalg.norm(a.b - X.csr_spmv_f64(a, x)) / np.linalg.norm(a.b))
    out = {"total_ms": round(setup + solve, 3), "setup_ms": round(setup, 3), "solve_ms": round(solve, 3),
           "warmup_total_ms": [round(v[0] + v[1], 3) for v in warm],
           "iterations": st["iterations"], "status": X.STATUS_NAMES[st["status"]], "true_rel_residual": res,
           "layout": {k: layout[k] for k in ("sigma", "spmv_variant", "grid")}}
    if run == "refined":
        out["outer_steps"], out["inner_iterations"] = info["ret"][1], info["ret"][2]
        out["iterations"] = info["ret"][2]
    if run == "reliable":
        out["replacements"] = info["ret"][1]
        out["iterations"] = info["ret"][2]
        out["fallback"] = st["fallback"]
    if learned:
        out["setup_stats"] = info["setup_stats"]
        out["matrix"] = info["matrix"]
    if "multilevel" in info:
        ml = info["multilevel"]
        out["multilevel"] = {"levels": ml["levels"], "n": ml["n"], "nnz": ml["nnz"],
                             "hierarchy_setup_ms": round(ml["setup_ms"], 3),
                             "phase_ms": [round(float(v), 3) for v in ml["phase_ms"]],
                             "coarse_max": args.ml_coarse_max, "omega": args.ml_omega, "gamma": args.ml_gamma,
                             "coarse_blocking": args.ml_coarse_blocking, "theta": args.ml_theta,
                             "aggregation_rounds": info.get("ml_rounds", []),
                             "vcycle_us": round(info["vcycle_us"], 3)}
    out["us_per_iteration"] = round(solve * 1000.0 / max(out["iterations"], 1), 3)
    log(f"{name}: {out}")
    return out


def break_even(learned: dict, base: dict):
    """Right-hand sides (each a full solve on the same matrix, the set-up
    done once) from which the learned method's total time is the lower."""
    ds = learned["setup_ms"] - base["setup_ms"]
    dt = base["solve_ms"] - learned["solve_ms"]
    if dt <= 0:
        return None
    return max(1, int(np.ceil(ds / dt))) if ds > 0 else 1


def workload_c2(X, args, cpu_proc) -> int:
    a = X.gen_fem2d(args.mesh, 1.0, 20, 1e3, 0.3, SEED)
    X.DeviceCsr(a).close()          # first-touch of the runtime's transfer path, untimed
    X.synchronize()
    t0 = time.perf_counter()
    d = X.DeviceCsr(a)              # CSR + b from host memory over PCIe
    X.synchronize()
    upload_ms = (time.perf_counter() - t0) * 1000.0
    log(f"FEM mesh {args.mesh}: n={a.n} nnz={a.nnz}, CSR and b resident in HBM (upload {upload_ms:.2f} ms)")
    net, tinfo = train_network(X, args)
    layouts = tune_layouts(X, args)
    res = {}
    fused_refined = layouts["plain_fused"]["us_per_iter"] < layouts["plain_unfused"]["us_per_iter"]
    for name in ("learned_multilevel", "learned", "accurate_jacobi", "plain_unfused", "plain_fused",
                 "refined_jacobi"):
        lay = layouts["plain_fused" if name == "refined_jacobi" and fused_refined else
                      "plain_unfused" if name == "refined_jacobi" else
                      "learned" if name == "learned_multilevel" else name]
        res[name] = time_to_solution(X, d, a, name, lay, args, net=net if name.startswith("learned") else None,
                                     fused_refined=fused_refined)
    learned, acc = res["learned_multilevel"], res["accurate_jacobi"]
    for name in ("learned_multilevel", "learned", "accurate_jacobi", "refined_jacobi"):
        if not res[name]["true_rel_residual"] <= 1.01e-6:
            raise RuntimeError(f"{name} missed 1e-6: {res[name]}")
    plain_form = min(("plain_unfused", "plain_fused"), key=lambda k: res[k]["total_ms"])
    plain = res[plain_form]

    # the learned iteration's kernels against the roofline, on its own instance
    ml = X.Matrix.from_device(d, sigma=layouts["learned"]["sigma"], grid=layouts["learned"]["grid"],
                              spmv_variant=layouts["learned"]["spmv_variant"], keep_lo=True, block_rows=8,
                              blocking=args.blocking)
    s = X.Pcg(ml, rel_tol=1e-6, accurate=True)
    s.set_rhs_device(d.b.ptr)
    s.set_mlp_preconditioner(net)
    iroof = iteration_roofline(s, a, True)
    s.close()
    # what the network does to the FP32 iteration's count (recursive 1e-6)
    it = {}
    for name in ("jacobi", "learned"):
        q = X.Pcg(ml, rel_tol=1e-6, fused=True, stall_limit=0)
        if name == "learned":
            q.set_mlp_preconditioner(net)
        q.set_rhs_device(d.b.ptr)
        q.run()
        it[name] = q.status()["iterations"]
        q.close()
    ml.close()
    log(f"FP32 iterations to recursive 1e-6: {it}")

    ms_, layout = spmv_tuned(X, a)
    spmv = spmv_roofline(ms_, a)
    spmv["layout"] = layout
    ms_.close()
    d.close()
    spmv["stream_reference"] = stream_reference(X)
    spmv["beyond_mall"] = spmv_beyond_mall(X)
    spmv["beyond_mall"]["frac_of_stream_read"] = round(spmv["beyond_mall"]["warm_gbs"] /
                                                       spmv["stream_reference"]["read_gbs"], 4)
    log(f"spmv: {spmv}")
    bitmap = bitmap_block_spmv(X, a)
    cpu = None
    if cpu_proc is not None:
        cpu = finish_cpu_golden(cpu_proc, 300.0)
        log(f"cpu baseline: {cpu}")

    value = learned["total_ms"]
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(value, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"thermal-FEM P1 {args.mesh}x{args.mesh} nodes, n={a.n}, nnz={a.nnz}, "
                               f"time from the CSR and b in HBM to x with true relative residual 1e-6",
                   "n": a.n, "nnz": a.nnz,
                   "method": "PCG preconditioned by a multilevel V-cycle whose level-0 smoother is the learned "
                             "8x8 block-Jacobi preconditioner: blocks = parallel strength aggregates of <= 8 rows "
                             "computed on the device, blocks from a BF16 MFMA MLP (which starts as, and on these "
                             "aggregates stays within 1% of, the exact block inverses); coarse levels = Galerkin "
                             "products over the aggregates, set up on the device with their own aggregates and "
                             "exact block factors, dense coarsest inverse; one symmetric FP32 V-cycle per "
                             "iteration; FP32 r/z and FP32 hi + 16-bit-remainder matrix values, FP64-accumulated "
                             "SpMV, FP64 p/Ap/x, FP64 reductions, residual replacement on the FP64 true residual, "
                             "true-residual stagnation fallback to Jacobi; value = device set-up (matrix, blocks, "
                             "hierarchy) + solve; 'learned' is the same without the coarse levels (the learned "
                             "block-Jacobi PCG alone)",
                   "network_training": tinfo, "layouts_tuned_on": f"training matrix seed {TUNE_SEED}",
                   "layouts": layouts},
        "true_rel_residual": learned["true_rel_residual"],
        "iterations": learned["iterations"],
        "learned_multilevel": learned,
        "learned": res["learned"],
        "speedup_learned_block_jacobi_vs_accurate_jacobi": round(acc["total_ms"] / res["learned"]["total_ms"], 3),
        "accurate_jacobi": acc,
        "refined_jacobi": res["refined_jacobi"],
        "plain_forms": {k: res[k] for k in ("plain_unfused", "plain_fused")},
        "plain_form": plain_form,
        "plain_ms": plain["total_ms"],
        "plain_true_rel_residual": plain["true_rel_residual"],
        "speedup_learned_vs_accurate_jacobi": round(acc["total_ms"] / value, 3),
        "speedup_learned_vs_plain": round(plain["total_ms"] / value, 3),
        "speedup_learned_vs_refined_jacobi": round(res["refined_jacobi"]["total_ms"] / value, 3),
        "speedup_note": "ratios of complete times to solution (set-up + solve); accurate Jacobi is the fastest "
                        "plain-Jacobi solve that reaches true 1e-6 (the primary baseline); plain FP32 Jacobi-PCG "
                        "stops on its recursive residual at a true residual far above 1e-6",
        "solve_only_speedup_vs_accurate_jacobi": round(acc["solve_ms"] / learned["solve_ms"], 3),
        "break_even_rhs_vs_accurate_jacobi": break_even(learned, acc),
        "fp32_iterations_to_recursive_1e-6": it,
        "spmv_hbm_gbs": spmv["achieved"],
        "spmv_roofline": spmv,
        "bitmap_block_spmv": bitmap,
        "roofline": iroof,
        "cpu_baseline": cpu,
        "host_upload_ms": round(upload_ms, 3),
        "value_with_host_upload_ms": round(value + upload_ms, 3),
        "host_upload_note": "CSR (row_ptr, col_idx, values) and b copied from pageable host memory; value starts "
                            "with them resident in HBM",
    }
    if cpu is not None:
        out["speedup_vs_cpu_golden"] = round(cpu["value"] / value, 1)
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=0)
    ap.add_argument("--sigma", type=int, default=4096, help="sort window of --profile (no autotuning there)")
    ap.add_argument("--mesh", type=int, default=1100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--inner-rel", type=float, default=1e-2)
    ap.add_argument("--no-tune-sigma", action="store_true")
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--train-grids", default="512,700,900,1100")
    ap.add_argument("--profile", action="store_true")
    ap.add_argument("--grid-sweep", action="store_true", help="jacobi: time the update kernel at several grids")
    ap.add_argument("--workload", default="c2", choices=["c2", "c3", "c4", "jacobi"],
                    help="c2: the benchmark (default); c3: FP64-refined 1100/1600 FEM to 1e-10; "
                         "c4: 256 batched 3D systems; jacobi: 1000 plain Jacobi updates")
    ap.add_argument("--blocking", type=int, default=3,
                    help="8-row blocks: 0 consecutive rows, 1 strength aggregates (row-order seeds), "
                         "2 strength-ordered aggregates (host), 3 parallel strength aggregates (device)")
    ap.add_argument("--stall-limit", type=int, default=4096,
                    help="learned solve: iterations between true-residual stagnation checks (0: off)")
    # off-block weight of the training loss: 0 makes the loss's minimiser the
    # exact block inverse the network starts from; any off-block weight
    # moves the blocks away from it and costs iterations (strength-ordered
    # aggregates, FP32 iteration: 0.3 +2%, 1.0 +10%; parallel aggregates,
    # accurate iteration: 0.05 +7%, 0.5 +14%)
    ap.add_argument("--lam", type=float, default=0.0)
    ap.add_argument("--ml-coarse-max", type=int, default=512, help="multilevel: rows of the dense coarsest level")
    ap.add_argument("--ml-omega", type=float, default=0.8, help="multilevel: smoother damping")
    ap.add_argument("--ml-gamma", type=float, default=1.6, help="multilevel: coarse-correction scale")
    ap.add_argument("--ml-theta", type=float, default=0.08, help="multilevel: strength threshold of the aggregation")
    ap.add_argument("--ml-blocking", type=int, default=4, choices=(3, 4),
                    help="multilevel: level 0's smoother blocks (device aggregation rule 3 or 4)")
    ap.add_argument("--ml-coarse-blocking", type=int, default=1, choices=(1, 3, 4),
                    help="multilevel: coarse smoother blocks (1: pieces of the level's aggregates; 3, 4: rules)")
    ap.add_argument("--cpu-golden-worker", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_golden_worker:
        return cpu_golden_worker(args.mesh)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(a system would shard as a row partition with one halo exchange and two scalar "
              "all-reduces per iteration; batches by system)", file=sys.stderr)
        return 2
    cpu_proc = None
    if args.workload == "c2" and not args.profile and not args.no_cpu:
        cpu_proc = start_cpu_golden(args.mesh)   # before this process touches the GPU

    import xplus_amd as X

    if X.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        if cpu_proc is not None:
            cpu_proc.kill()
        return 1
    if args.workload == "c3":
        return workload_c3(X, args)
    if args.workload == "c4":
        return workload_c4(X, args)
    if args.workload == "jacobi":
        return workload_jacobi(X, args)
    if args.profile:
        a = X.gen_fem2d(args.mesh, 1.0, 20, 1e3, 0.3, SEED)
        if args.grid <= 0:
            args.grid = 512
        return profile_run(X, a, args)
    try:
        return workload_c2(X, args, cpu_proc)
    finally:
        if cpu_proc is not None and cpu_proc.poll() is None:
            cpu_proc.kill()


if __name__ == "__main__":
    sys.exit(main())
