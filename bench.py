"""Benchmark: circuit layers/s of the batched MIPT engine at L=20 x 4096
trajectories (16 p values in [0, 0.3] x 256, seeds 1..4096, T=40 layers,
closed ring, Haar U(4), complex64 state), next to the plain
one-trajectory-per-launch GPU path and the FP64 CPU oracle, with S(L/2)/TMI
errors against the oracle on a replayed subset.

A step = one full run of the batch: draw every gate, flag and uniform on
device, then evolve all 4096 trajectories through all 40 layers (gate sweeps,
Born probabilities, collapse and renormalisation), states resident in HBM.
Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
from concurrent.futures import ThreadPoolExecutor
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "circuit layers/s at L=20 x 4096 traj, batched vs plain; S/TMI err vs FP64"
HBM_PEAK_GBS = 8000.0
FP32_MATRIX_PEAK_TFLOPS = 157.3
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "pass_kernel_L20x112_traffic.md")


def traffic_ratio():
    """HBM-side bytes per algorithmic byte of the gate sweep, from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (calibrated against a plain copy)."""
    try:
        with open(TRAFFIC_PROFILE) as fh:
            for line in fh:
                if line.startswith("traffic_per_algorithmic_byte"):
                    return float(line.split("=")[1])
    except OSError:
        pass
    return None


def hip_runtime():
    h = ctypes.CDLL("libamdhip64.so")
    h.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    h.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    h.hipEventSynchronize.argtypes = [ctypes.c_void_p]
    h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
    h.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    return h


def host_cpu():
    """nproc, the cores this process may use, and the CPU model name."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "usable_cores": len(os.sched_getaffinity(0)), "model": model}


def linspace(first, last, count):
    # cpp/include/mipt/types.hpp:23-38
    if count == 1:
        return [first]
    step = (last - first) / (count - 1)
    v = [first + step * i for i in range(count)]
    v[-1] = last
    return v


def _oracle_evolve(args):
    """CPU FP64 oracle run of one trajectory of the workload (cpu_baseline leg).
    Runs in a thread: the oracle's C calls release the GIL."""
    L, layers, p, seed = args
    from oracle import oracle as O
    psi, out, p1, (g, f, u) = O.simulate(L, layers, p, seed, True, "philox", 0)
    return psi, out, p1, u, f


def _oracle_observables(args):
    psi, L = args
    from oracle import oracle as O
    return O.half_chain_entropy(psi, L), O.tmi(psi, L)


T_START = time.perf_counter()


def log(msg: str) -> None:
    print(f"[bench {time.perf_counter() - T_START:7.1f}s] {msg}", file=sys.stderr, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=2, help="untimed steps (the second captures the HIP graph)")
    ap.add_argument("--L", type=int, default=20)
    ap.add_argument("--layers", type=int, default=40)
    ap.add_argument("--per-p", type=int, default=256)
    ap.add_argument("--np", type=int, default=16)
    ap.add_argument("--plain", type=int, default=8, help="trajectories timed on the plain path")
    ap.add_argument("--replay", type=int, default=16,
                    help="trajectories checked against the FP64 oracle (one per host thread, up to 16)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the oracle leg (profiling runs)")
    ap.add_argument("--tile-bits", type=int, default=0)
    ap.add_argument("--no-autotune", action="store_true", help="keep the default tile configuration")
    ap.add_argument("--no-graph", action="store_true", help="launch every kernel eagerly (no HIP graph replay)")
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--bond-arith", type=int, default=0,
                    help="1 VALU, 2 VALU + matrix cores, 3 matrix cores (0: the autotuner's choice)")
    ap.add_argument("--streams", type=int, default=2, help="tiled groups in flight during the timed steps")
    ap.add_argument("--profile-every", type=int, default=1,
                    help="time every n-th gate sweep with HIP events (1: all, so the mean weighs every launch kind)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 256 MiB plain copy (traffic-counter calibration)")
    ap.add_argument("--rdm-L", type=int, default=0,
                    help="also time the half-chain RDM entropy at this L (Bell-pair state, S = L/2 exactly)")
    ap.add_argument("--rdm-reps", type=int, default=3)
    ap.add_argument("--no-decoder", action="store_true", help="skip the learned probe decoder leg")
    ap.add_argument("--no-uneven", action="store_true", help="skip the uneven-length / 2D-grid leg")
    ap.add_argument("--no-observables", action="store_true", help="skip the S/TMI readout timing")
    ap.add_argument("--no-c1", action="store_true", help="skip the L=12 open-chain (CPU-path) workload")
    ap.add_argument("--no-c4", action="store_true", help="skip the L=26 x 64 workload")
    ap.add_argument("--obs-traj", type=int, default=256)
    ap.add_argument("--uneven-per", type=int, default=512, help="trajectories per (L, p) in the uneven leg")
    ap.add_argument("--decoder-train", type=int, default=200_000)
    ap.add_argument("--decoder-test", type=int, default=20_000)
    ap.add_argument("--decoder-epochs", type=int, default=16)
    args = ap.parse_args(argv)

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU; trajectories shard across GPUs without "
              "communication until a final per-p merge, which is not built here", file=sys.stderr)
        return 2

    import mipt_amd as M

    if args.copy_calibration:
        nbytes = M.stream_copy_calibration(25)
        log(f"calibration copy of {nbytes} bytes done")
    L, layers = args.L, args.layers
    ps = linspace(0.0, 0.3, args.np)
    specs = []
    for pi, p in enumerate(ps):
        for r in range(args.per_p):
            specs.append(M.Spec(L, layers, p, len(specs) + 1, True))
    ntraj = len(specs)
    log(f"setting up {ntraj} trajectories (L={L}, T={layers})")
    b = M.Batch(specs, tile_bits=args.tile_bits, threads=args.threads, bond_arith=args.bond_arith)
    b.set_streams(args.streams)
    tune = None
    if not args.no_autotune and args.tile_bits == 0 and args.bond_arith == 0:
        b.generate()
        best, plan, ms = b.autotune(4, with_plan=True)
        names = ["13b/256t valu", "13b/256t valu+mfma", "13b/256t mfma", "14b/256t", "13b/128t", "12b/64t"]
        plans = ["plan 1 (register-slot bonds first)", "plan 2 (spread windows)", "plan 3 (both)"]
        tune = {"choice": names[best] if best >= 0 else None, "plan_variant": plan,
                "probe_ms": {n: round(float(m), 2) for n, m in zip(names + plans, ms)}}
        log(f"autotune: {tune}")
    if not args.no_graph:
        b.set_graph(True)
    hip = hip_runtime()
    stream = ctypes.c_void_p(b.stream)
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    hip.hipEventCreate(ctypes.byref(e0))
    hip.hipEventCreate(ctypes.byref(e1))

    host_issue = []

    def step():
        b.reset()
        b.generate()
        t = time.perf_counter()
        b.evolve()
        host_issue.append(time.perf_counter() - t)

    for w in range(args.warmup):
        step()
        b.synchronize()
        log(f"warmup step {w + 1}/{args.warmup} done")
    b.set_profiling(0 if not args.no_graph else args.profile_every)
    b.synchronize()
    hip.hipEventRecord(e0, stream)
    t_wall = time.perf_counter()
    for k in range(args.steps):
        step()
        log(f"timed step {k + 1}/{args.steps} issued")
    hip.hipEventRecord(e1, stream)
    hip.hipEventSynchronize(e1)
    b.synchronize()
    t_wall = time.perf_counter() - t_wall
    ms = ctypes.c_float()
    hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
    ms_per_step = ms.value / args.steps
    layer_trajs = ntraj * layers
    value = layer_trajs / (ms_per_step / 1e3)
    # per-kernel timing: one more eager step outside the timed region, every
    # group on the batch stream (so no other group's kernels overlap a sweep)
    # and every (profile_every-th) sweep bracketed by HIP events on that stream
    b.set_graph(False)
    b.set_streams(1)
    b.set_profiling(args.profile_every)
    hip.hipEventRecord(e0, stream)
    step()
    hip.hipEventRecord(e1, stream)
    hip.hipEventSynchronize(e1)
    b.synchronize()
    ms1 = ctypes.c_float()
    hip.hipEventElapsedTime(ctypes.byref(ms1), e0, e1)
    single_stream_ms = ms1.value
    b.set_streams(args.streams)
    log(f"profiled single-stream eager step done ({single_stream_ms:.1f} ms)")
    st = b.pass_stats()
    achieved = st["bytes_avg"] / (st["avg_ms"] * 1e-3) / 1e9 if st["avg_ms"] > 0 else 0.0
    # all sweeps of a step together: algorithmic bytes of the step over its wall time
    aggregate = 16.0 * (1 << L) * layer_trajs / (ms_per_step * 1e-3) / 1e9
    ratio = traffic_ratio() if L == 20 else None

    log(f"timed steps done: {ms_per_step:.1f} ms/step, {value:.0f} layers/s; host issue of evolve "
        f"{1e3 * sum(host_issue[-args.steps:]) / args.steps:.1f} ms/step")
    # plain one-trajectory-per-launch path on the same trajectories' device draws
    plain_lps = plain_host_lps = None
    if args.plain > 0:
        # one C-ABI call per trajectory (blocked matrices staged once, the
        # per-layer histogram round trip kept); the layer-by-layer host loop
        # over the same engine is timed beside it
        eng = M.PlainEngine(False)
        idx = [int(i * ntraj / args.plain) for i in range(args.plain)]
        data = [b.export_layers(i) for i in idx]
        M.run_plain_trajectory_native(eng, L, *data[0], True, 4, 0, 10)  # warm
        t = time.perf_counter()
        for g, f, u in data:
            M.run_plain_trajectory_native(eng, L, g, f, u, True, 4, 0, 10)
        dt = time.perf_counter() - t
        plain_lps = len(data) * layers / dt
        t = time.perf_counter()
        for g, f, u in data:
            M.run_plain_trajectory(eng, L, g, f, u, True, 4, 0, 10)
        plain_host_lps = len(data) * layers / (time.perf_counter() - t)
        eng.close()
        log(f"plain path: {plain_lps:.0f} layers/s one call per trajectory, {plain_host_lps:.0f} layers/s host loop")

    result = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "layers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (device Philox draws; seeds 1..%d)" % ntraj,
        "config": {"workload": f"batched p-sweep: L={L} ring, T={layers}, {args.np} p in [0,0.3] x {args.per_p} "
                               f"= {ntraj} trajectories, Haar U(4), complex64 state",
                   "L": L, "layers": layers, "trajectories": ntraj, "parallelism": "none (1 GPU)"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None if ratio is None else round(ratio * st["bytes_avg"]),
                     "kernel": "pass_kernel (gate sweep)", "timed_launches": st["timed"],
                     "avg_launch_ms": round(st["avg_ms"], 4), "algorithmic_bytes_per_launch": round(st["bytes_avg"]),
                     "streams_in_flight": 1, "launches_per_step": st["launches"],
                     "single_stream_step_ms": round(single_stream_ms, 3),
                     "sweep_time_share": round(st["launches"] * st["avg_ms"] / single_stream_ms, 4)
                     if single_stream_ms > 0 else None,
                     "aggregate_sweep_GBps": round(aggregate, 1),
                     "aggregate_frac": round(aggregate / HBM_PEAK_GBS, 4),
                     # the same launches as arithmetic: a dense 4x4 complex bond is 16 real
                     # multiply-adds per amplitude, L/2 bonds per layer of the ring, so
                     # 2 * (L/2) FLOP per algorithmic byte (16 B per amplitude and layer)
                     "arith": {"flop_per_launch": round(st["bytes_avg"] * L),
                               "achieved_TFLOPs": round(achieved * L / 1e3, 2),
                               "peak_TFLOPs": FP32_MATRIX_PEAK_TFLOPS,
                               "frac": round(achieved * L / 1e3 / FP32_MATRIX_PEAK_TFLOPS, 4)}},
        "autotune": tune,
        "graph": not args.no_graph,
        "plain_layers_per_s": None if plain_lps is None else round(plain_lps, 1),
        "plain": None if plain_lps is None else {
            "trajectories": args.plain, "block_targets": 4,
            "driver": "one C-ABI call per trajectory (mipt_hip_engine_run_trajectory)",
            "host_loop_layers_per_s": round(plain_host_lps, 1)},
        "batched_vs_plain": None if plain_lps is None else round(value / plain_lps, 2),
    }

    # ---- oracle leg: CPU baseline + S/TMI/outcome parity on a replayed subset
    if not args.no_cpu and args.replay > 0:
        idx = [int(i * ntraj / args.replay) for i in range(args.replay)]
        cpu = host_cpu()
        cores = max(1, min(len(idx), cpu["usable_cores"], 16))
        jobs = [(L, layers, specs[i].p, specs[i].seed) for i in idx]
        from oracle import oracle as O
        O.lib()
        log(f"oracle leg: {len(jobs)} trajectories on {cores} threads")
        t = time.perf_counter()
        with ThreadPoolExecutor(cores) as pool:
            res = list(pool.map(_oracle_evolve, jobs))
        wall = time.perf_counter() - t
        cpu_lps = len(idx) * layers / wall
        log(f"oracle evolution done in {wall:.1f} s ({cpu_lps:.2f} layers/s)")
        with ThreadPoolExecutor(cores) as pool:
            obs_all = list(pool.map(_oracle_observables, [(r[0], L) for r in res]))
        log("oracle observables done")
        s_err = tmi_err = amp_err = 0.0
        mism = flagged_n = 0
        for i, (psi, out, p1, u, f), obs in zip(idx, res, obs_all):
            got, flg, _, _ = b.outcomes(i)
            near = np.abs(u - p1) < 1e-6
            mism += int(np.sum((got != out) & (f == 1) & ~near))
            flagged_n += int(np.sum(flg))
            ob = b.observables(i)
            s_err = max(s_err, abs(ob["s_half"] - obs[0]))
            tmi_err = max(tmi_err, abs(ob["tmi"] - obs[1]))
            if not flg.any():
                amp_err = max(amp_err, float(np.abs(b.get_state(i).astype(np.complex128) - psi).max()))
        result["cpu_baseline"] = {"value": round(cpu_lps, 2), "unit": "layers/s", "cores": cores, "kind": "port",
                                  "sample": f"{len(idx)} trajectories of the workload (L={L}, T={layers}) evolved "
                                            f"in complex128 by the oracle, one host thread per trajectory",
                                  "nproc": cpu["nproc"], "usable_cores": cpu["usable_cores"], "cpu_model": cpu["model"]}
        result["parity"] = {"replayed": len(idx), "outcome_mismatches": mism, "flagged_draws": flagged_n,
                            "s_half_err_max": s_err, "tmi_err_max": tmi_err, "amp_err_max": amp_err}
        result["batched_vs_cpu"] = round(value / cpu_lps, 1)
    else:
        result["cpu_baseline"] = None
    b.close()  # the legs below need the device memory
    if args.rdm_L:
        result["rdm_gemm"] = rdm_timing(M, args.rdm_L, args.rdm_reps)
    if not args.no_cpu and not args.no_c1:
        try:
            result["c1"] = c1_leg(M)
        except Exception as e:
            result["c1"] = {"error": repr(e)}
            log(f"C1 leg failed: {e!r}")
    if not args.no_c4:
        try:
            result["c4"] = c4_leg(M, hip)
        except Exception as e:
            result["c4"] = {"error": repr(e)}
            log(f"C4 leg failed: {e!r}")
    if not args.no_observables:
        try:
            result["observables"] = observables_leg(M, L, layers, ps, args.obs_traj)
        except Exception as e:
            result["observables"] = {"error": repr(e)}
            log(f"observables leg failed: {e!r}")
    if not args.no_uneven:
        try:
            result["uneven"] = uneven_leg(M, hip, args.uneven_per)
        except Exception as e:
            result["uneven"] = {"error": repr(e)}
            log(f"uneven leg failed: {e!r}")
    if not args.no_decoder:
        try:
            result["decoder"] = decoder_leg(args.decoder_train, args.decoder_test, args.decoder_epochs)
        except Exception as e:  # the headline line must still print
            result["decoder"] = {"error": repr(e)}
            log(f"decoder leg failed: {e!r}")
    print(json.dumps(result))
    return 0


def _oracle_c1(args):
    n, layers, p, seed = args
    from oracle import oracle as O
    psi, out, p1, (g, f, u) = O.simulate(n, layers, p, seed, False, "philox", 0)
    return psi, out, p1, f, u


def c1_leg(M):
    """The reference's CPU-path workload: open chain L=12 from |0...0>, T=24,
    p=0.16, 200 trajectories with seeds trajectory_seed(0, i).  The FP64
    oracle runs all 200 on the host's cores (the CPU baseline on this
    workload); the GPU runs them in the batched engine (complex64, one launch)
    and on the plain engine in complex128; outcomes, S(6) and the TMI of
    3-site regions are checked against the oracle for every trajectory."""
    from oracle import oracle as O
    n, layers, p, count = 12, 24, 0.16, 200
    seeds = [int(O.lib().orc_trajectory_seed(0, i)) for i in range(count)]
    cores = max(1, min(len(os.sched_getaffinity(0)), 16))
    t = time.perf_counter()
    with ThreadPoolExecutor(cores) as pool:
        ref = list(pool.map(_oracle_c1, [(n, layers, p, sd) for sd in seeds]))
    cpu_s = time.perf_counter() - t
    specs = [M.Spec(n, layers, p, sd, False) for sd in seeds]
    b = M.Batch(specs)
    b.generate()
    b.evolve()
    b.synchronize()  # warm
    b.reset()
    t = time.perf_counter()
    b.generate()
    b.evolve()
    b.synchronize()
    gpu_s = time.perf_counter() - t
    eng = M.PlainEngine(True)
    draws = [b.export_layers(i) for i in range(16)]
    M.run_plain_trajectory_native(eng, n, *draws[0], False, 4, 0, 10)  # warm
    t = time.perf_counter()
    for g, f, u in draws:
        M.run_plain_trajectory_native(eng, n, g, f, u, False, 4, 0, 10)
    plain_s = time.perf_counter() - t
    eng.close()
    sh, tm = b.observables_all(3, 1)
    mism = flagged = 0
    s_err = tmi_err = 0.0
    for i, (psi, out, p1, f, u) in enumerate(ref):
        got, flg, _, _ = b.outcomes(i)
        near = np.abs(u - p1) < 1e-6
        mism += int(np.sum((got != out) & (f == 1) & ~near))
        flagged += int(flg.sum())
        if not flg.any():
            s_err = max(s_err, abs(sh[i] - O.half_chain_entropy(psi, n)))
            tmi_err = max(tmi_err, abs(tm[i, 0] - O.tmi(psi, n, 3, 0)))
    b.close()
    log(f"C1: oracle {count * layers / cpu_s:.0f} layers/s on {cores} threads, batched GPU "
        f"{count * layers / gpu_s:.0f} layers/s, plain complex128 {16 * layers / plain_s:.0f} layers/s")
    return {"workload": "L=12 open chain, T=24, p=0.16, 200 trajectories (seeds trajectory_seed(0, i))",
            "cpu_oracle_layers_per_s": round(count * layers / cpu_s, 1), "cpu_threads": cores,
            "gpu_batched_layers_per_s": round(count * layers / gpu_s, 1),
            "gpu_plain_complex128_layers_per_s": round(16 * layers / plain_s, 1),
            "parity": {"trajectories": count, "outcome_mismatches": mism, "flagged_draws": flagged,
                       "s_half_err_max": s_err, "tmi_err_max": tmi_err}}


def observables_leg(M, L, layers, ps, n):
    """S(L/2) and TMI (cycle 0) of n trajectories of the benchmark's shape,
    RDMs built on the GPU and their eigenvalues solved in batches of one size
    (mipt_hip_batch_observables_all), against one entropy call at a time."""
    specs = [M.Spec(L, layers, ps[i % len(ps)], 70_000_000 + i, True) for i in range(n)]
    b = M.Batch(specs)
    b.generate()
    b.evolve()
    b.synchronize()
    b.observables_all(0, 1)  # warm: workspaces, solver handle
    t = time.perf_counter()
    sh, tm = b.observables_all(0, 1)
    dt = time.perf_counter() - t
    k = min(8, n)
    t = time.perf_counter()
    for i in range(k):
        b.observables(i)
    dt1 = (time.perf_counter() - t) / k
    b.close()
    log(f"observables: batched {1e3 * dt / n:.2f} ms/trajectory, one at a time {1e3 * dt1:.2f} ms/trajectory")
    return {"trajectories": n, "batched_ms_per_trajectory": round(1e3 * dt / n, 3),
            "single_ms_per_trajectory": round(1e3 * dt1, 3), "s_half_mean": round(float(np.mean(sh)), 6),
            "tmi_mean": round(float(np.mean(tm)), 6)}


def _timed_evolve(M, hip, specs):
    """Draw and evolve one batch; returns (GPU ms from events on the batch
    stream, layer-trajectories)."""
    b = M.Batch(specs)
    stream = ctypes.c_void_p(b.stream)
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    hip.hipEventCreate(ctypes.byref(e0))
    hip.hipEventCreate(ctypes.byref(e1))
    hip.hipEventRecord(e0, stream)
    b.generate()
    b.evolve()
    hip.hipEventRecord(e1, stream)
    hip.hipEventSynchronize(e1)
    b.synchronize()
    ms = ctypes.c_float()
    hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
    b.outcomes(0)  # raises if a trajectory hit a zero-weight branch
    work = sum(s.layers for s in specs)
    b.close()
    return ms.value, work


def uneven_leg(M, hip, per):
    """Uneven-size workload in ONE batch: chains L in {14,...,24} with T = 2L
    and p in {0.10, 0.16, 0.22}, `per` trajectories per (L, p), interleaved in
    the spec list (one evolve call; the engine groups same-length trajectories
    into tile sweeps and keeps two groups in flight); the state memory is
    checked against the device's free memory first and reported.  Plus the
    4 x 5 grid (L=20, four bond colours, open) at p = 0.12, 1024 trajectories,
    T = 40.  Throughput in layer-trajectories/s."""
    out = {"workload": f"one batch: L in 14..24 interleaved, T=2L, p in (0.10,0.16,0.22) x {per}; "
                       f"4x5 grid p=0.12 x 1024, T=40"}
    specs = []
    seed = 50_000_000
    for r in range(per):
        for p in (0.10, 0.16, 0.22):
            for L in range(14, 26, 2):
                seed += 1
                specs.append(M.Spec(L, 2 * L, p, seed, True))
    state_bytes = sum(8 << s.L for s in specs)
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total))
    out["state_GiB"] = round(state_bytes / 2**30, 2)
    out["device_free_GiB"] = round(free.value / 2**30, 2)
    if state_bytes > 0.97 * free.value:
        # keep every length and the interleaving; scale the count per (L, p)
        scale = 0.97 * free.value / state_bytes
        keep = max(1, int(per * scale))
        specs = [sp for i, sp in enumerate(specs) if (i // 18) < keep]
        out["per_L_p_run"] = keep
        out["state_GiB"] = round(sum(8 << s.L for s in specs) / 2**30, 2)
        log(f"uneven leg: {state_bytes / 2**30:.1f} GiB of states exceeds free memory, running {keep} per (L, p)")
    else:
        out["per_L_p_run"] = per
    log(f"uneven leg: {len(specs)} trajectories in one batch, {out['state_GiB']} GiB of states")
    ms, work = _timed_evolve(M, hip, specs)
    out["trajectories"] = len(specs)
    out["chains_layers_per_s"] = round(work / (ms * 1e-3), 1)
    log(f"uneven leg: mixed lengths {work / (ms * 1e-3):.0f} layers/s")
    specs = [M.Spec(20, 40, 0.12, 60_000_000 + i, False, grid_x=5) for i in range(1024)]
    ms, work = _timed_evolve(M, hip, specs)
    out["grid_4x5_layers_per_s"] = round(work / (ms * 1e-3), 1)
    log(f"uneven leg: 4x5 grid {work / (ms * 1e-3):.0f} layers/s")
    return out


def c4_leg(M, hip, n=64, L=26, layers=52, p=0.16):
    """The large single-size workload: L=26 ring (512 MiB complex64 states),
    T=52, p=0.16, 64 trajectories.  Evolution throughput (layer-trajectories/s
    over HIP events on the batch stream), the gate sweeps' per-launch
    algorithmic HBM GB/s on one stream, and the 8192 x 8192 half-chain RDM
    Gram of one trajectory: kernel time and algorithmic FLOPs (lower triangle)
    against the FP64 (default) and FP32 MFMA peaks."""
    specs = [M.Spec(L, layers, p, 90_000_000 + i, True) for i in range(n)]
    b = M.Batch(specs)
    stream = ctypes.c_void_p(b.stream)
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    hip.hipEventCreate(ctypes.byref(e0))
    hip.hipEventCreate(ctypes.byref(e1))
    b.generate()
    b.synchronize()
    hip.hipEventRecord(e0, stream)
    b.evolve()
    hip.hipEventRecord(e1, stream)
    hip.hipEventSynchronize(e1)
    b.synchronize()
    ms = ctypes.c_float()
    hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
    lps = n * layers / (ms.value * 1e-3)
    log(f"C4: {n} x L={L} x {layers} layers in {ms.value:.0f} ms, {lps:.0f} layers/s")
    # per-launch sweep timing on one stream (an extra run, outside the timed one)
    b.reset()
    b.set_streams(1)
    b.set_profiling(1)
    b.evolve()
    b.synchronize()
    st = b.pass_stats()
    sweep_gbs = st["bytes_avg"] / (st["avg_ms"] * 1e-3) / 1e9 if st["avg_ms"] > 0 else 0.0
    out = {"workload": f"L={L} ring, T={layers}, p={p}, {n} trajectories, complex64 (512 MiB states)",
           "layers_per_s": round(lps, 1), "evolve_ms": round(ms.value, 1),
           "sweep": {"achieved_GBps": round(sweep_gbs, 1), "peak_GBps": HBM_PEAK_GBS,
                     "frac": round(sweep_gbs / HBM_PEAK_GBS, 4), "avg_launch_ms": round(st["avg_ms"], 4),
                     "timed_launches": st["timed"], "algorithmic_bytes_per_launch": round(st["bytes_avg"])}}
    # half-chain Gram (8192 x 8192 from an 8192 x 8192 reshape), FP64 then FP32 MFMA
    ptr = b.state_ptr(0)
    kept = list(range(L // 2))
    grams = {}
    for bits, peak in ((64, 78.6e12), (32, 157.3e12)):
        M.set_gram_precision(bits)
        M.set_gram_profiling(True)
        s_half = M.entropy(ptr, L, kept)
        g = M.gram_stats()
        M.set_gram_profiling(False)
        tf = g["flops_avg"] / (g["avg_ms"] * 1e-3) if g["avg_ms"] > 0 else 0.0
        grams[f"fp{bits}_mfma"] = {"kernel_ms": round(g["avg_ms"], 3), "flop": g["flops_avg"],
                                   "achieved_TFLOPs": round(tf / 1e12, 2), "peak_TFLOPs": peak / 1e12,
                                   "frac": round(tf / peak, 4), "s_half_bits": s_half}
        log(f"C4: half-chain Gram FP{bits} MFMA {g['avg_ms']:.1f} ms, {tf / 1e12:.1f} TFLOP/s, S(13)={s_half:.6f}")
    M.set_gram_precision(64)
    out["rdm_gram"] = grams
    b.close()
    return out


def decoder_leg(n_train, n_test, epochs, batch=1024, lr=1e-3):
    """Learned probe decoder workload: L=16, one shared Haar realisation (seed
    7), 32 unmeasured scrambling layers, probe attached to site 8, 32 measured
    layers at p ~ U[0.05, 0.30]; n_train training and n_test held-out
    trajectories (disjoint seeds); BF16-MFMA training on the GPU; mean
    |S_pred - S_exact| per p bin against the p-bin mean entropy."""
    from mipt_amd import decoder as dec
    log(f"decoder leg: simulating {n_train} + {n_test} trajectories")
    train = dec.generate_probe_dataset(np.arange(1, n_train + 1))
    test = dec.generate_probe_dataset(np.arange(10_000_001, 10_000_001 + n_test))
    sim_tps = (n_train + n_test) / (train.seconds + test.seconds)
    log(f"decoder leg: simulation {sim_tps:.0f} trajectories/s; training {epochs} epochs")
    d = dec.Decoder(32, 16, batch, seed=5)
    d.set_dataset(np.concatenate([train.records, test.records]), np.concatenate([train.bloch, test.bloch]))
    rng = np.random.default_rng(0)
    steps = n_train // batch
    t = time.perf_counter()
    loss = None
    for e in range(epochs):
        loss = d.train(rng.permutation(n_train)[: steps * batch], batch, lr=dec.cosine_lr(lr, e, epochs))
        log(f"decoder epoch {e + 1}/{epochs}: mean batch loss {loss:.4f}")
    train_s = time.perf_counter() - t
    d.predict(n_train, min(n_test, batch))  # warm
    t = time.perf_counter()
    pred = d.predict(n_train, n_test)
    inf_s = time.perf_counter() - t
    s_exact = dec.entropy_from_bloch(test.bloch)
    err = np.abs(dec.entropy_from_bloch(pred) - s_exact)
    base, bins, edges = dec.p_bin_baseline(train, test)
    err_b = np.abs(base - s_exact)
    per_bin = [{"p": [round(float(edges[k]), 3), round(float(edges[k + 1]), 3)], "n": int((bins == k).sum()),
                "decoder": round(float(err[bins == k].mean()), 4), "p_bin_mean": round(float(err_b[bins == k].mean()), 4)}
               for k in range(len(edges) - 1)]
    d.close()
    return {"workload": "L=16 ring, shared Haar gates (seed 7), 32 scrambling layers, probe at site 8, "
                        "32 measured layers, p ~ U[0.05,0.30]",
            "train": n_train, "held_out": n_test, "epochs": epochs, "batch": batch, "lr": lr, "lr_schedule": "half cosine per epoch",
            "final_train_loss": round(loss, 4), "train_seconds": round(train_s, 2),
            "mean_abs_dS_nats": round(float(err.mean()), 4), "baseline_mean_abs_dS_nats": round(float(err_b.mean()), 4),
            "per_p_bin": per_bin, "inferences_per_s": round(n_test / inf_s, 1),
            "simulation_trajectories_per_s": round(sim_tps, 1)}


def rdm_timing(M, n, reps):
    """Half-chain entropy of a Bell-pair state (pairs (i, i + n/2)), whose RDM
    is 2^(n/2) square with k = 2^(n/2): gather + Gram (8 m^2 k FLOPs) + FP64
    eigensolve.  The Gram kernel's own duration comes from rocprofv3."""
    h = n // 2
    idx = np.zeros(1, np.int64)
    for i in range(h):
        idx = np.concatenate([idx, idx | (1 << i) | (1 << (i + h))])
    psi = np.zeros(1 << n, np.complex64)
    psi[idx] = 1.0 / math.sqrt(len(idx))
    eng = M.PlainEngine(False)
    eng.reset(n)
    eng.set_from_host(psi)
    del psi
    kept = list(range(h))
    s = M.entropy(eng.data, n, kept)  # warm: workspaces, solver handle
    times = []
    for r in range(reps):
        t = time.perf_counter()
        s = M.entropy(eng.data, n, kept)
        times.append(time.perf_counter() - t)
        log(f"rdm L={n} rep {r + 1}/{reps}: {times[-1] * 1e3:.1f} ms, S={s:.9f}")
    eng.close()
    m = 1 << h
    return {"L": n, "m": m, "k": m, "gram_flop": 8.0 * m * m * m, "entropy_call_ms": round(1e3 * min(times), 2),
            "s_half": s, "s_exact": float(h), "s_err": abs(s - h)}


if __name__ == "__main__":
    sys.exit(main())
