"""Benchmark: solved states/s of the 442@512 classic tablebase build on one
MI355X (plain per-layer GPU schedule), with the dominant kernel's HBM
roofline and the reference's own CPU builder timed on the host beside it.

One step = one complete build of pattern 442 at target 512 (304 layers:
forward enumeration + dedup, final layers, backward solve) with the seeds
resident and every solved layer kept in HBM (no file I/O in the timed
region).  value = sum of raw layer sizes / build wall time.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes
import importlib
import json
import os
import shutil
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "solved states/s for 442@512 build (batched vs plain); query QPS at p99 latency"
# single-kernel stats families and the kernel each times on this workload
# (the forward pass sorts unique layers: 4 keys per thread up to 512 Ki keys,
# 8 above, so each radix family holds both tile instances)
KERNEL_OF = {"radix_scatter": "egtb::k_radix_scatter<8, 4>", "radix_hist": "egtb::k_radix_hist<8, 4>",
             "solve": "egtb::k_solve<float>", "expand": "egtb::k_expand_hash", "extract": "egtb::k_set_extract"}
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_442_512_plain_build.md")


def profiled_traffic(kernel):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (already scaled by the copy calibration),
    or None when the summary has no row for it."""
    try:
        with open(TRAFFIC_PROFILE) as fh:
            for line in fh:
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cells) >= 5 and cells[0].strip("`") == kernel:
                    return round((float(cells[3]) + float(cells[4])) * 1e6, 0)
    except (OSError, ValueError):
        pass
    return None


def algorithmic_build_bytes(res):
    """Compulsory bytes of one build: forward Bf(i) = 8 N_{i-1} + 16 C_i + 8 N_i
    over steps 1..steps-2, backward Bb(i) = 12 N_i + 12 (N'_{i+1} + N'_{i+2})
    + 12 N'_i over layers 0..steps-3 (float32 rates)."""
    info = [res.layer_info(i) for i in range(res.steps)]
    n = [x[0] for x in info]
    live = [x[1] for x in info]
    fwd = sum(8 * n[i - 1] + 16 * (info[i][2] + info[i][3]) + 8 * n[i] for i in range(1, res.steps - 1))
    bwd = sum(12 * n[i] + 12 * (live[i + 1] + live[i + 2]) + 12 * live[i] for i in range(0, res.steps - 2))
    return fwd + bwd


def kernel_stats(native, names):
    L = native.lib()
    out = {}
    for nm in names:
        la, ms, by = ctypes.c_uint64(), ctypes.c_double(), ctypes.c_double()
        if L.egtb_stats_get(nm.encode(), ctypes.byref(la), ctypes.byref(ms), ctypes.byref(by)) == 0:
            out[nm] = dict(launches=int(la.value), ms=ms.value, bytes=by.value)
    return out


def host_cpu():
    """nproc, the CPU model and the cores this process may run on."""
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    return {"nproc": os.cpu_count() or 1, "affinity_cores": allowed, "model": model,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


TMPFS = "/dev/shm" if os.path.isdir("/dev/shm") else None


def cpu_baseline(pattern, target, budget_threads, n_states):
    """Time the reference's own classic builder (compiled from its sources
    into oracle/_ref) on the same workload; fall back to the oracle port on
    a smaller target if that library is absent.  budget_threads: the cores
    this run may use; the reference itself takes clamp(nproc, 4, 32)
    (engine_core/BookBuilder.py:72), which the caller reports beside it."""
    pkg = importlib.import_module("2048endgametablebase_amd")
    p = pkg.patterns.build_params(pattern, target)
    threads = max(4, min(32, budget_threads))
    work = tempfile.mkdtemp(prefix="egtb_cpu_", dir=TMPFS)
    try:
        from oracle.ref import ref

        if ref.available():
            t0 = time.perf_counter()
            ref.run_build(p["seeds"], p["masks"], p["shifts"], p["symm_mode"], target, p["steps"],
                          p["docheck_step"], os.path.join(work, f"{pattern}_{2**target}_"), is_free=p["is_free"],
                          threads=threads)
            dt = time.perf_counter() - t0
            # same layer sets as the GPU build (golden-checked), so the same count
            return dict(value=round(n_states / dt, 1), unit="states/s", cores=threads, kind="reference",
                        sample=f"full {pattern}@{2**target} build by the reference's run_pattern_build_cpp "
                               f"(g++ -Ofast, {threads} OpenMP threads, layer files on tmpfs): "
                               f"{n_states} states in {dt:.1f} s")
        from oracle import oracle as orc

        t_small = target - 1
        ps = pkg.patterns.build_params(pattern, t_small)
        t0 = time.perf_counter()
        ob = orc.OracleBuild(ps["seeds"], ps["masks"], ps["shifts"], ps["symm_mode"], t_small, ps["steps"],
                             ps["docheck_step"], threads=threads)
        dt = time.perf_counter() - t0
        n = sum(ob.raw(i).size for i in range(ps["steps"]))
        ob.close()
        return dict(value=round(n / dt, 1), unit="states/s", cores=threads, kind="port",
                    sample=f"oracle port, {pattern}@{2**t_small} in memory, {n} states in {dt:.1f} s")
    finally:
        shutil.rmtree(work, ignore_errors=True)


def cpu_baseline_advanced(pattern, target, threads):
    """The reference's advanced builder (compiled from its sources into
    oracle/_ref) on the same forward pass as the GPU leg, stopped after its
    raw layers are written to tmpfs; None when that library is absent."""
    pkg = importlib.import_module("2048endgametablebase_amd")
    from oracle.ref import ref

    if not ref.available():
        return None
    p = pkg.patterns.ad_params(pattern, target)
    work = tempfile.mkdtemp(prefix="egtb_cpu_ad_", dir=TMPFS)
    try:
        prefix = os.path.join(work, f"{pattern}_{2**target}_")
        t0 = time.perf_counter()
        ref.run_build_ad(p["seeds"], p["masks"], p["symm_mode"], p["n_free_32k"], p["fixed_32k_shifts"], target,
                         p["steps"], p["docheck_step"], prefix, is_free=p["is_free"],
                         small_tile_sum_limit=p["small_tile_sum_limit"], threads=threads)
        dt = time.perf_counter() - t0
        n = sum(os.path.getsize(prefix + str(i)) // 8 for i in range(p["steps"] - 1))
    finally:
        shutil.rmtree(work, ignore_errors=True)
    return dict(value=round(n / dt, 1), unit="states/s", cores=threads, kind="reference",
                sample=f"{pattern}@{2**target} advanced forward pass by the reference's run_pattern_build_ad_cpp "
                       f"(g++ -Ofast, {threads} OpenMP threads, raw layer files on tmpfs): {n} states in {dt:.1f} s")


def cpu_baseline_advanced_build(pattern, target, threads):
    """The reference's whole advanced build (forward pass and solver,
    run_pattern_build_ad_cpp compiled into oracle/_ref) with its files on
    tmpfs: rates written per second; None when that library is absent."""
    pkg = importlib.import_module("2048endgametablebase_amd")
    from oracle.ref import ref

    if not ref.available():
        return None
    p = pkg.patterns.ad_params(pattern, target)
    work = tempfile.mkdtemp(prefix="egtb_cpu_adb_", dir=TMPFS)
    try:
        prefix = os.path.join(work, f"{pattern}_{2**target}_")
        t0 = time.perf_counter()
        ref.run_build_ad(p["seeds"], p["masks"], p["symm_mode"], p["n_free_32k"], p["fixed_32k_shifts"], target,
                         p["steps"], p["docheck_step"], prefix, is_free=p["is_free"],
                         small_tile_sum_limit=p["small_tile_sum_limit"], threads=threads, stop_after_forward=False)
        dt = time.perf_counter() - t0
        n = 0
        for i in range(p["steps"] - 2):
            folder = f"{prefix}{i}b"
            n += sum(os.path.getsize(os.path.join(folder, f)) // 4 for f in os.listdir(folder) if f.endswith(".b"))
    finally:
        shutil.rmtree(work, ignore_errors=True)
    return dict(value=round(n / dt, 1), unit="kept rates/s", cores=threads, kind="reference",
                seconds=round(dt, 3),
                sample=f"{pattern}@{2**target} advanced build (forward pass and solver) by the reference's "
                       f"run_pattern_build_ad_cpp (g++ -Ofast, {threads} OpenMP threads, files on tmpfs): "
                       f"{n} float32 rates kept in {dt:.1f} s")


def advanced_build_leg(pkg, pattern, target, reps=2):
    """Best wall time of the whole GPU advanced build (forward pass and
    solver, float32 books in HBM and freed as the solve passes them, no
    files)."""
    adv = importlib.import_module("2048endgametablebase_amd.advanced")
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        r = adv.build_ad(pattern, target)
        dt = time.perf_counter() - t0
        info = r.solve_info()
        rec = {"fwd_ms": round(r.fwd_ms, 2), "solve_ms": round(info["solve_ms"], 2), "states": int(r.sizes.sum()),
               "rates": int(info["values"])}
        r.close()
        if best is None or dt < best[0]:
            best = (dt, rec)
    dt, rec = best
    return dict({"workload": f"{pattern}@{2**target} advanced build (forward pass and solver), books in HBM",
                 "ms": round(dt * 1000.0, 2)}, **rec, rates_per_s=round(rec["rates"] / dt, 1))


def advanced_leg(pkg, pattern, target, reps=3):
    """Best wall time of the GPU advanced (masked) forward pass, raw layers
    kept only while needed, no files."""
    adv = importlib.import_module("2048endgametablebase_amd.advanced")
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        r = adv.generate_ad(pattern, target)
        dt = time.perf_counter() - t0
        n, c, layers = int(r.sizes.sum()), int(r.candidates), int(r.sizes.size)
        r.close()
        best = dt if best is None else min(best, dt)
    return {"workload": f"{pattern}@{2**target} advanced (masked) forward pass, layers in HBM", "layers": layers,
            "states": n, "boards_before_dedup": c, "ms": round(best * 1000.0, 2),
            "states_per_s": round(n / best, 1)}


def serving_sweep(pkg, pattern, target, seconds):
    """Continuous-batching server on the device-resident table: Poisson
    arrivals at rates doubling from 10k QPS, 80% on-table / 20% fallback
    queries, until p99 reaches 5 ms or the server stops keeping up (achieved
    below 90% of offered)."""
    serving = importlib.import_module("2048endgametablebase_amd.serving")
    res = pkg.build_in_memory(pattern, target)
    table = pkg.Table.from_result(res, pattern, target)
    pool = serving.make_query_pool(res, pattern, n=1 << 20, off_table_frac=0.2, seed=0)
    rows = []
    with serving.Server(slots=4, max_batch=8192) as srv:
        tid = srv.attach(table)
        srv.loadgen(tid, pool, qps=20000, duration_s=0.3, seed=99)  # warm-up
        stop = None
        for qps in (10000, 25000, 50000, 100000, 200000, 400000, 800000, 1600000, 3200000, 6400000):
            r = srv.loadgen(tid, pool, qps=qps, duration_s=seconds if qps <= 800000 else seconds / 2, seed=qps)
            rows.append({"qps_offered": qps, "qps_achieved": round(r["qps_achieved"], 1),
                         "p50_ms": round(r["p50_ms"], 4), "p99_ms": round(r["p99_ms"], 4),
                         "p999_ms": round(r["p999_ms"], 4), "max_ms": round(r["max_ms"], 4),
                         "max_batch": int(r["max_batch"]),
                         "fallback_frac": round(r["fallbacks"] / max(r["completed"], 1), 4)})
            if r["p99_ms"] >= 5.0:
                stop = "p99 reached 5 ms"
                break
            if r["qps_achieved"] < 0.9 * qps:
                stop = "achieved below 90% of offered (the load generator or server saturated)"
                break
        st = srv.stats()
    ok = [r for r in rows if r["p99_ms"] < 5.0 and r["qps_achieved"] >= 0.9 * r["qps_offered"]]
    best = max(ok, key=lambda r: r["qps_achieved"]) if ok else None
    table.close()
    res.close()
    return {"workload": f"{pattern}@{2**target} table in HBM, 4096 games, Poisson arrivals, 20% off-table "
                        "(the expectimax player's depth-1 search, two moves deep, as the fallback)",
            "qps_at_p99_lt_5ms": best["qps_achieved"] if best else None,
            "p99_ms_at_that_qps": best["p99_ms"] if best else None,
            "sweep_stopped": stop or "end of the rate list",
            "sweep": rows, "batches": st["batches"], "max_batch": st["max_batch"]}


def io_inclusive(pkg, pattern, target, reps=2):
    """The same build through start_build's native call with its files: every
    <i>.book plus stats.txt written to tmpfs (where the CPU baseline writes
    too), float32 rates."""
    p = pkg.patterns.build_params(pattern, target)
    best = None
    for _ in range(reps):
        work = tempfile.mkdtemp(prefix="egtb_io_", dir=TMPFS)
        try:
            t0 = time.perf_counter()
            pkg.run_pattern_build(p["seeds"], p["masks"], p["shifts"], p["symm_mode"], target, p["steps"],
                                  p["docheck_step"], os.path.join(work, f"{pattern}_{2**target}_"),
                                  is_free=p["is_free"], success_rate_dtype="float32", write_files=True)
            dt = time.perf_counter() - t0
            nbytes = sum(os.path.getsize(os.path.join(work, f)) for f in os.listdir(work))
        finally:
            shutil.rmtree(work, ignore_errors=True)
        best = dt if best is None else min(best, dt)
    return {"ms": round(best * 1000.0, 2), "files_bytes": nbytes, "dir": TMPFS or tempfile.gettempdir(),
            "dtype": "float32"}


def fp64_max_error(pkg, pattern, target):
    """Largest |float32 rate - float64 rate| over every state of the build
    (the float64 build is the FP64 recomputation; both are bit-exact to the
    reference's builds of their type).  A state missing from one book
    (rate 0 there) counts with rate 0."""
    import numpy as np

    a = pkg.build_in_memory(pattern, target, success_rate_dtype="float32")
    b = pkg.build_in_memory(pattern, target, success_rate_dtype="float64")
    worst, n = 0.0, 0
    try:
        for i in range(a.steps):
            ka, ra = a.book(i)
            kb, rb = b.book(i)
            keys = np.union1d(ka, kb)
            va = np.zeros(keys.size)
            vb = np.zeros(keys.size)
            va[np.searchsorted(keys, ka)] = ra.astype(np.float64)
            vb[np.searchsorted(keys, kb)] = rb
            if keys.size:
                worst = max(worst, float(np.abs(va - vb).max()))
            n += keys.size
    finally:
        a.close()
        b.close()
    return {"max_abs_error": worst, "states_compared": n, "tolerance": 1e-6}


def build_leg(pkg, pattern, target, reps=1, **kw):
    """Wall time and states/s of one plain build configuration."""
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        r = pkg.build_in_memory(pattern, target, **kw)
        dt = time.perf_counter() - t0
        info = [r.layer_info(i) for i in range(r.steps)]
        mem = r.memory()
        f, b, _ = r.timing()
        r.close()
        if best is None or dt < best[0]:
            best = (dt, info, mem, f, b)
    dt, info, mem, f, b = best
    n = sum(x[0] for x in info)
    out = {"workload": f"{pattern}@{2**target}", "layers": len(info), "states": n,
           "live_records": sum(x[1] for x in info), "ms": round(dt * 1000.0, 2), "fwd_ms": round(f, 2),
           "bwd_ms": round(b, 2), "states_per_s": round(n / dt, 1)}
    if kw.get("host_tables"):
        out.update(peak_device_bytes=mem["peak_device_bytes"], streamed_bytes=mem["streamed_bytes"],
                   stream_GBps=round(mem["streamed_bytes"] / max(mem["stream_ms"], 1e-9) / 1e6, 2)
                   if mem.get("stream_ms") else None)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pattern", default="442")
    ap.add_argument("--target", type=int, default=9)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the I/O-inclusive, FP64-error, L3@256 and 444@2048 legs")
    ap.add_argument("--serve-seconds", type=float, default=1.5)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="copy 1 GiB with torch first (a kernel of known bytes for counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (a layer would be sharded by "
              "key range, with an all-to-all of candidates); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    pkg = importlib.import_module("2048endgametablebase_amd")
    native = pkg.native
    L = native.lib()
    torch.cuda.init()

    def one_build():
        return pkg.build_in_memory(args.pattern, args.target, success_rate_dtype="float32")

    if args.calibrate_copy:
        src = torch.empty(1 << 27, dtype=torch.int64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    for _ in range(args.warmup):
        one_build().close()
    torch.cuda.synchronize()

    # timed region: statistics off, so no event pairs are recorded
    L.egtb_stats_enable(0)
    totals = []
    fwd = bwd = 0.0
    n_states = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    first = None
    for _ in range(args.steps):
        res = one_build()
        f, b, tt = res.timing()
        fwd += f
        bwd += b
        totals.append(tt)
        if first is None:
            first = res  # counted after the timed region
        else:
            res.close()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    layers = first.steps
    info = [first.layer_info(i) for i in range(first.steps)]
    n_states = sum(x[0] for x in info)
    live = sum(x[1] for x in info)
    cands = sum(x[2] + x[3] for x in info)
    first.close()
    ms_per_step = wall * 1000.0 / args.steps
    value = n_states / (wall / args.steps)

    # instrumented builds: HIP events around each launch group on the
    # library's stream (the stream the kernels run on), per kernel family
    stat_builds = 2
    L.egtb_stats_reset()
    L.egtb_stats_enable(1)
    for _ in range(stat_builds):
        r = one_build()
        alg_bytes = algorithmic_build_bytes(r)
        r.close()
    torch.cuda.synchronize()
    L.egtb_stats_enable(0)
    stats = kernel_stats(native, ["expand_count", "expand", "extract", "sort", "unique", "solve", "compact",
                                  "hash", "radix_hist", "radix_scatter"])
    single = {k: v for k, v in stats.items() if k in KERNEL_OF and v["launches"]}
    roofline = None
    if single:
        name, st = max(single.items(), key=lambda kv: kv[1]["ms"])
        avg_ms = st["ms"] / st["launches"]
        per_launch = st["bytes"] / st["launches"]
        achieved = per_launch / (avg_ms / 1000.0) / 1e9 if avg_ms > 0 else 0.0
        roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": profiled_traffic(KERNEL_OF[name]),
                    "kernel": KERNEL_OF[name], "family": name, "alg_bytes_per_launch": round(per_launch, 0),
                    "avg_launch_ms": round(avg_ms, 5), "launches_per_build": st["launches"] // stat_builds,
                    "build": {"alg_bytes": int(alg_bytes),
                              "achieved_GBps": round(alg_bytes / (ms_per_step / 1000.0) / 1e9, 2),
                              "frac": round(alg_bytes / (ms_per_step / 1000.0) / 1e9 / HBM_PEAK_GBS, 5)}}

    batched = None
    if not args.no_batched:
        bm = importlib.import_module("2048endgametablebase_amd.batched")
        bm.build_batch(bm.SMALL_JOBS).close()  # warm-up
        batched = {"workload_24_jobs": [f"{p}@{2**t}" for p, t in bm.DEFAULT_JOBS],
                   "uneven_24_jobs": bm.compare_with_plain(bm.DEFAULT_JOBS, reps=2),
                   "uneven_24_jobs_planned": bm.compare_with_plain(bm.DEFAULT_JOBS, reps=2, plan="auto"),
                   "small_24_jobs_planned": bm.compare_with_plain(bm.SMALL_JOBS, reps=3, plan="auto"),
                   "small_24_jobs": bm.compare_with_plain(bm.SMALL_JOBS, reps=3),
                   "tiny_24_jobs": bm.compare_with_plain(bm.TINY_JOBS, reps=3)}

    serving = None
    if not args.no_serving:
        serving = serving_sweep(pkg, args.pattern, args.target, args.serve_seconds)

    extra = None
    if not args.no_extra:
        io = io_inclusive(pkg, args.pattern, args.target)
        io["states_per_s"] = round(n_states / (io["ms"] / 1000.0), 1)
        extra = {"io_inclusive": io,
                 "fp64": fp64_max_error(pkg, args.pattern, args.target),
                 "C0_L3_256": build_leg(pkg, "L3", 8, reps=3),
                 "C4_444_2048_in_hbm": build_leg(pkg, "444", 11),
                 "C4_444_2048_host_tables": build_leg(pkg, "444", 11, host_tables=True),
                 "advanced_442_256": advanced_leg(pkg, "442", 8),
                 "advanced_LL_64": advanced_leg(pkg, "LL", 6),
                 "advanced_4442_64": advanced_leg(pkg, "4442", 6, reps=2),
                 "advanced_build_442_256": advanced_build_leg(pkg, "442", 8),
                 "advanced_build_LL_64": advanced_build_leg(pkg, "LL", 6),
                 "advanced_build_free8_64": advanced_build_leg(pkg, "free8", 6)}

    cpu = None
    hostinfo = host_cpu()
    if not args.no_cpu:
        # the box allots 16 cores to one GPU (OMP_NUM_THREADS); the
        # reference's own rule is clamp(nproc, 4, 32)
        share = int(os.environ.get("OMP_NUM_THREADS") or min(hostinfo["affinity_cores"], 16))
        try:
            cpu = cpu_baseline(args.pattern, args.target, share, n_states)
            cpu["host"] = hostinfo
            cpu["reference_thread_rule"] = max(4, min(32, hostinfo["nproc"]))
            cpu["note"] = (f"timed at {cpu.get('cores')} threads, every core this run may use "
                           f"(the box allots {share} to one GPU); the reference's clamp(nproc,4,32) would "
                           f"take {cpu['reference_thread_rule']} on this {hostinfo['nproc']}-thread host")
        except Exception as exc:  # reported, never fatal for the GPU number
            cpu = {"error": str(exc)}
        if extra is not None:
            try:
                ad = cpu_baseline_advanced("442", 8, max(4, min(32, share)))
                if ad is not None:
                    gpu = extra["advanced_442_256"]["states_per_s"]
                    ad["gpu_states_per_s"] = gpu
                    ad["gpu_over_cpu"] = round(gpu / ad["value"], 1)
                    cpu["advanced_forward"] = ad
            except Exception as exc:
                cpu["advanced_forward"] = {"error": str(exc)}
            try:
                adb = cpu_baseline_advanced_build("442", 8, max(4, min(32, share)))
                if adb is not None:
                    gpu_s = extra["advanced_build_442_256"]["ms"] / 1000.0
                    adb["gpu_seconds"] = round(gpu_s, 4)
                    adb["gpu_over_cpu"] = round(adb["seconds"] / gpu_s, 1)  # same build, wall time ratio
                    cpu["advanced_build"] = adb
            except Exception as exc:
                cpu["advanced_build"] = {"error": str(exc)}

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "states/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 keys / f32 rates (f64 accumulate)",
        "data": "synthetic: fully determined by pattern 442, target 512, seeds of the catalogue",
        "config": {"workload": f"{args.pattern}@{2**args.target} classic build, float32 rates, p4=0.1",
                   "layers": layers, "states_per_build": n_states,
                   "live_records": live, "candidates": cands, "fwd_ms": round(fwd / args.steps, 2),
                   "bwd_ms": round(bwd / args.steps, 2),
                   "schedule": "plain (per-layer launches)",
                   "kernels": {k: {"ms_per_build": round(v["ms"] / stat_builds, 3),
                                   "launches_per_build": v["launches"] // stat_builds} for k, v in stats.items()}},
        "roofline": roofline,
        "batched": batched,
        "serving": serving,
        "legs": extra,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
