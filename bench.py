"""Benchmark: ECG pair energy+gradient evaluations on one MI355X.

Workload: non-BO LiH (n = 5, nst = 24, Stark field on), K = 2048 basis
functions (2,098,176 unordered pairs), parameters seed 1; one step = one full
FP64 energy + analytic gradient evaluation (outprod) with x resident in HBM.
Prints one JSON line.

    python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--K 2048]
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)


def cpu_baseline(sysd, K, x, rows, threads):
    """Oracle (C, OpenMP) on a bounded sample of the same evaluation: `rows`
    gradient rows (rows x K ordered pairs x nst terms of gradel) and `rows`
    element rows of the lower triangle, extrapolated by the exact pair counts."""
    import numpy as np
    from oracle.oracle import Oracle
    o = Oracle(sysd)
    iv = x[:K * sysd.blk]
    step = max(1, K // rows)
    r0 = [i * step for i in range(rows)]
    # gradient rows are timed as one contiguous block (cost per row is
    # independent of the row index: every row spans all K kets)
    t0 = time.perf_counter()
    o.grada_rows(x, 1.0, 1.0, K // 2, rows, nthreads=threads)
    tg = time.perf_counter() - t0
    # element rows: row i costs i+1 pairs; sample rows spread over [0, K)
    t0 = time.perf_counter()
    npairs_e = 0
    for r in r0:
        o.matel_rows(iv, K, r, 1, nthreads=threads)
        npairs_e += r + 1
    te = time.perf_counter() - t0
    full = tg * K / rows + te * (K * (K + 1) / 2) / npairs_e
    return {
        "value": (K * (K + 1) / 2) / full,
        "unit": "pairs/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle/necg_oracle.c (C restatement of the reference NumPy path), "
                   f"{rows} gradient rows ({rows * K} ordered pairs x {sysd.nst} terms) + "
                   f"{rows} element rows ({npairs_e} pairs) of the K={K} LiH evaluation, "
                   f"{tg + te:.1f} s, extrapolated to one full evaluation "
                   f"({full:.0f} s)"),
        "seconds_sampled": tg + te,
        "est_seconds_per_eval": full,
    }


def host_cpu():
    """The host's core count and model, and the CPU share a run may use."""
    import subprocess
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    nproc = os.cpu_count() or 1
    share = int(os.environ.get("OMP_NUM_THREADS") or nproc)
    return {"nproc": nproc, "model": model, "share": max(1, min(share, nproc))}


def reference_structure_baseline(sysd, K, x, E, norm, share):
    """The NumPy restatement of the reference's per-pair loops
    (oracle/kron_ref.py: one matel/gradel call per (term, bra, ket) in the
    reference's Kronecker operator form, the make_grad GRADA loop) on a
    bounded sample of the same evaluation: one process (1 core,
    single-threaded BLAS), then `share` processes with rows dealt
    rank-cyclically as the Fortran code deals them; both extrapolated."""
    from oracle.kron_ref import time_sample
    out = {}
    for procs in (1, share):
        r = time_sample(sysd, x, K, E, norm, procs=procs, rows_per_proc=2, kets=48)
        out[f"cores_{procs}"] = {
            "value": (K * (K + 1) / 2) / r["est_seconds_per_eval"], "unit": "pairs/s",
            "cores": procs, "kind": "port",
            "sample": (f"oracle/kron_ref.py, {r['gradient_blocks']} ordered pairs x "
                       f"{sysd.nst} terms of gradel + {r['element_pairs']} pairs of matel + "
                       f"one GRADA row, {r['wall_s']:.1f} s wall on {procs} process(es), "
                       f"extrapolated to one full K={K} evaluation "
                       f"({r['est_seconds_per_eval']:.0f} s)"),
            "est_seconds_per_eval": r["est_seconds_per_eval"],
            "s_per_gradient_block": r["s_per_gradient_block"],
            "s_per_element_pair": r["s_per_element_pair"]}
    return out


def batched_leg(reps=5):
    """64 independent systems (He/H2/LiH, K log-uniform in [32, 1024], seeds
    100..163) evaluated together (autotuned tiles, HIP-graph replay) against
    the plain per-system path (one outprod per system, back to back)."""
    import torch
    from nbo_correlatedgaussians_amd.batch import Batch, BatchGraph, autotune, mixed_systems
    from nbo_correlatedgaussians_amd.engine import Engine
    systems, Ks, xs = mixed_systems()
    b = Batch(systems, Ks, "cuda:0")
    xa = b.pack(xs)
    best, _ = autotune(b, xa, reps=2)
    G = BatchGraph(b, xa)
    for _ in range(2):
        G.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        G.replay()
    torch.cuda.synchronize()
    tb = (time.perf_counter() - t0) / reps
    b.check("batched leg")
    engs = [Engine(s, "cuda:0") for s in systems]
    xds = [torch.as_tensor(x, device="cuda:0") for x in xs]
    for e, x in zip(engs, xds):
        e.outprod_t(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        for e, x in zip(engs, xds):
            e.outprod_t(x)
    torch.cuda.synchronize()
    tp = (time.perf_counter() - t0) / reps
    pairs = b.pairs
    return {"workload": "64 systems He/H2/LiH, K log-uniform [32,1024], seeds 100..163; "
                        "energy+gradient of all",
            "pairs": pairs, "ms": tb * 1e3, "pairs_per_s": pairs / tb,
            "plain_ms": tp * 1e3, "plain_pairs_per_s": pairs / tp, "speedup": tp / tb,
            "tiling": {"item_flops": best[0], "grid_per_cu": best[1]}}


def served_leg(nreq, rate):
    """32 resident systems (He/H2/LiH, K log-uniform in [32, 4096], seeds
    200..231); nreq requests with Poisson arrivals at `rate`/s, 70 % single-
    function replacement trials, 30 % full energy+gradient evaluations."""
    from nbo_correlatedgaussians_amd.batch import mixed_systems
    from nbo_correlatedgaussians_amd.serve import Server, latency_stats, poisson_trace
    systems, Ks, xs = mixed_systems(count=32, seed0=200, kmax=4096)
    srv = Server(systems, xs, "cuda:0", tune=True)
    st = latency_stats(srv.run(poisson_trace(systems, srv.Ks, nreq, rate=rate, seed=7)))
    sat = latency_stats(srv.run(poisson_trace(systems, srv.Ks, nreq, rate=1e9, seed=7)))
    tpl, ppl = srv.run_plain(poisson_trace(systems, srv.Ks, nreq, rate=1e9, seed=7))
    out = {"workload": f"32 resident systems He/H2/LiH, K log-uniform [32,4096], seeds "
                       f"200..231; {nreq} requests, Poisson arrivals at {rate:g}/s, 70% "
                       f"single-function replacement trials, 30% full energy+gradient",
           "rate_per_s": rate}
    out.update({k: v for k, v in st.items() if k != "requests"})
    out["saturated_pairs_per_s"] = sat["pairs_per_s"]
    out["saturated_p99_ms"] = sat["p99_ms"]
    out["plain_pairs_per_s"] = ppl / tpl
    out["speedup_vs_plain"] = sat["pairs_per_s"] / (ppl / tpl)
    out["tiling"] = {"item_flops": srv.batch.tiling[0]}
    return out


def outprod_leg(eng, xd, K, reps=5):
    """The same evaluation through the scipy drop-in route: one outprod call
    (fused energy + gradient, E and g written to device buffers)."""
    import torch
    f = torch.empty(1, dtype=torch.float64, device=xd.device)
    g = torch.empty_like(xd)
    eng.outprod_t(xd, f, g)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        eng.outprod_t(xd, f, g)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    eng.check("outprod")
    xh = xd.cpu().numpy()
    t0 = time.perf_counter()
    for _ in range(reps):
        eng.outprod(xh)          # host x in, host (E, g) out: PCIe both ways
    th = (time.perf_counter() - t0) / reps
    # energy alone (no gradient: the lower-triangle element pass + quotient)
    eng.energy_t(xd)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        eng.energy_t(xd)
    torch.cuda.synchronize()
    te = (time.perf_counter() - t0) / reps
    eng.check("energy")
    pairs = K * (K + 1) // 2
    return {"workload": f"LiH K={K}: outprod(x) -> (E, g)", "ms": dt * 1e3,
            "pairs_per_s": pairs / dt, "host_buffers_ms": th * 1e3,
            "host_buffers_pairs_per_s": pairs / th, "energy": float(f.item()),
            "energy_only_ms": te * 1e3}


def small_lih_input():
    """LiH K = 12 as the reference's input generator makes it from seed 0
    (input_file_generation/generate_lih_input.py; restated byte for byte by
    generated_lih_input), with the file's C = 1/K used as is."""
    import nbo_correlatedgaussians_amd as necg
    _, s, iv, C = necg.generated_lih_input(12, 0)
    return s, 12, necg.pack_x(iv, C)


def optimizer_leg(maxiter=20):
    """LiH K = 12 from the generator (seed 0): one outprod, a 20-iteration
    truncated-Newton run (LMQN constants through scipy TNC) and a 20-iteration
    L-BFGS-B run (python/necg.py:159-171), both driven by the GPU outprod.
    The same runs on the host are part of the CPU baseline."""
    from nbo_correlatedgaussians_amd.driver import optimize
    from nbo_correlatedgaussians_amd.engine import Engine
    s, K, x0 = small_lih_input()
    eng = Engine(s, "cuda:0")
    E0 = eng.outprod(x0)[0]
    out = {"workload": f"LiH K={K} made by the input generator from seed 0: one outprod, "
                       f"{maxiter} iterations of truncated Newton (scipy TNC with the LMQN "
                       f"constants) and of L-BFGS-B",
           "maxiter": maxiter, "E0": E0}
    for method, key in (("TN", "tn"), ("L-BFGS-B", "lbfgsb")):
        t0 = time.perf_counter()
        r = optimize(eng.outprod, x0, method=method, maxiter=maxiter, siv=K * s.blk, M=K)
        out[key] = {"seconds": time.perf_counter() - t0, "nit": r.nit, "nfev": r.nfev,
                    "E_final": r.fun, "energies": r.energies}
    return out


def cpu_optimizer(opt):
    """The optimizer leg's runs with the oracle (one thread) as the objective.
    L-BFGS-B trajectories are compared iteration by iteration.  Truncated
    Newton forms Hessian-vector products from gradient differences over a
    ~1e-8 step: on this input, seeded gradient noise of 1e-15 relative (a few
    ulp) moves the oracle-driven run's energies by O(1) from its first or
    second iteration on (tests/test_driver_cpu.py), so two runs whose
    gradients differ at the 1e-13 level are not comparable beyond E0 and,
    at best, the first iteration."""
    from nbo_correlatedgaussians_amd.driver import optimize
    from oracle.oracle import Oracle
    s, K, x0 = small_lih_input()
    o = Oracle(s)
    out = {"cores": 1, "E0": o.outprod(x0, 1)[0]}
    out["E0_dE_vs_gpu"] = abs(out["E0"] - opt["E0"])
    for method, key in (("TN", "tn"), ("L-BFGS-B", "lbfgsb")):
        t0 = time.perf_counter()
        r = optimize(lambda x: o.outprod(x, 1), x0, method=method, maxiter=opt["maxiter"],
                     siv=K * s.blk, M=K)
        out[key] = {"seconds": time.perf_counter() - t0, "nit": r.nit, "nfev": r.nfev,
                    "E_final": r.fun, "energies": r.energies}
    g = opt["lbfgsb"]["energies"]
    c = out["lbfgsb"]["energies"]
    n = min(len(g), len(c))
    out["lbfgsb_max_dE_vs_gpu"] = max(abs(a - b) for a, b in zip(g[:n], c[:n])) if n else None
    out["lbfgsb_iterations_compared"] = n
    return out


def per_ordering_leg(eng, xd, K, reps=3):
    """The same outprod with the per-ordering pair pass (one evaluation per
    ordered pair, bra derivatives only, as the reference's make_grad loop):
    the gain of the two-sided form on this workload."""
    import torch
    f = torch.empty(1, dtype=torch.float64, device=xd.device)
    g = torch.empty_like(xd)
    eng.set_pair_mode(2)
    try:
        eng.outprod_t(xd, f, g)
        eng.timing(1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            eng.outprod_t(xd, f, g)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1000.0 / reps
        pair_ms, launches = eng.timing(2)
        eng.timing(0)
        eng.check("per-ordering leg")
    finally:
        eng.set_pair_mode(0)
    return {"workload": f"LiH K={K}: outprod(x) with one pair evaluation per ordered pair",
            "kernel": "pair_kernel<5,MODE_SPLIT>", "ms": ms,
            "pair_kernel_ms": pair_ms / max(launches, 1),
            "pairs_per_s": K * (K + 1) / 2 / (ms / 1000.0), "energy": float(f.item())}


def properties_leg(eng, xd, K, reps=5):
    """The driver's dipole pass (both sides symmetrised: nst^2 term pairs per
    basis pair) on the headline basis (LiH K = 2048, seed 1)."""
    import torch
    out = eng.dipole_t(xd)
    eng.check("dipole warm-up")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        eng.dipole_t(xd, out)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    eng.check("dipole")
    o = out.cpu().numpy()
    eng.set_dipole_mode(2)                   # the full nst^2 double sum, for comparison
    eng.dipole_t(xd, out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        eng.dipole_t(xd, out)
    torch.cuda.synchronize()
    dt_full = (time.perf_counter() - t0) / reps
    eng.set_dipole_mode(0)
    eng.check("dipole (full double sum)")
    of = out.cpu().numpy()
    from nbo_correlatedgaussians_amd.flops import dipole_flops_per_eval
    nst = eng.system.nst
    pairs = K * (K + 1) // 2
    fl = dipole_flops_per_eval(eng.system.n, nst, K)
    return {"workload": f"dipole pass, LiH K={K}: {pairs} pairs; symmetry-reduced ({nst} term "
                        f"pairs per basis pair) and the full {nst}^2 double sum",
            "ms": dt * 1e3, "pairs_per_s": pairs / dt,
            "full_ms": dt_full * 1e3, "full_term_pairs_per_s": pairs * nst * nst / dt_full,
            "full_model_flop": fl, "full_achieved_tflops": fl / dt_full / 1e12,
            "full_frac_fp64_peak": fl / dt_full / 1e12 / PEAK_FP64_TFLOPS,
            "max_diff_reduced_vs_full": float(abs(o - of).max() / abs(of).max()),
            "dipole_au": [float(v) for v in o[-3:]]}


def fp64_fma_rate(iters=200000):
    """The FP64 FMA rate this GPU sustains (necg_fp64_fma_rate: eight
    independent FMA chains per lane, two waves per SIMD), in TFLOP/s."""
    import ctypes as ct
    from nbo_correlatedgaussians_amd import _lib
    lib = _lib.load()
    r = ct.c_double(0.0)
    _lib.raise_for(lib.necg_fp64_fma_rate(0, iters, ct.byref(r)), "necg_fp64_fma_rate")
    return r.value


def traffic_probe(K, seed, pair_mode=0):
    """For the rocprofv3 counter passes: a 256 MiB device copy (the
    calibration of FETCH_SIZE / WRITE_SIZE) followed by two outprod calls
    (pair_mode 2: with the per-ordering pair pass)."""
    import torch
    import nbo_correlatedgaussians_amd as necg
    from nbo_correlatedgaussians_amd.engine import Engine
    a = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(2):
        b.copy_(a)
    torch.cuda.synchronize()
    s = necg.lih_system()
    iv, C = necg.generate_parameters(s, K, seed)
    eng = Engine(s, "cuda:0")
    xd = torch.as_tensor(necg.pack_x(iv, C), device="cuda:0")
    eng.set_pair_mode(pair_mode)
    for _ in range(2):
        eng.outprod_t(xd)
    torch.cuda.synchronize()
    eng.check("traffic probe")
    print("traffic probe done", flush=True)


def batch_probe():
    """For rocprofv3 passes over the batched pair kernel: the 64-system mix of
    the batched leg, autotuned, then two evaluations with the chosen tiling."""
    import torch
    from nbo_correlatedgaussians_amd.batch import Batch, autotune, mixed_systems
    systems, Ks, xs = mixed_systems()
    b = Batch(systems, Ks, "cuda:0")
    xa = b.pack(xs)
    best, _ = autotune(b, xa, reps=1)
    xd = torch.as_tensor(xa, device="cuda:0")
    torch.cuda.synchronize()
    for _ in range(2):
        b.outprod_t(xd)
    torch.cuda.synchronize()
    b.check("batch probe")
    print(f"batch probe done (item_flops {best[0]:g})", flush=True)


def _recorded(fname, key):
    """A figure recorded by a counter pass in profiles/<fname>, or None."""
    path = os.path.join(ROOT, "profiles", fname)
    try:
        with open(path, encoding="utf-8") as fh:
            for line in fh:
                if line.startswith(key + ":"):
                    return float(line.split(":")[1])
    except OSError:
        pass
    return None


def recorded_traffic():
    """HBM bytes per pair-kernel launch (FETCH_SIZE/WRITE_SIZE passes)."""
    return _recorded("pair2_kernel_lih_k2048_traffic.md", "traffic_bytes_per_launch")


def recorded_executed_flop():
    """Executed FP64 FLOP per pair-kernel launch (SQ_INSTS_VALU_*_F64 passes)."""
    return _recorded("pair2_kernel_lih_k2048_fp64_counters.md", "executed_flop_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--K", type=int, default=2048)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-rows", type=int, default=48)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-props", action="store_true")
    ap.add_argument("--no-opt", action="store_true")
    ap.add_argument("--path", choices=("autograd", "outprod"), default="autograd",
                    help="timed step: E = ecg_energy(x); E.backward(), or outprod(x)")
    ap.add_argument("--serve-requests", type=int, default=10000)
    ap.add_argument("--serve-rate", type=float, default=1000.0)
    ap.add_argument("--traffic-probe", action="store_true")
    ap.add_argument("--pair-mode", type=int, default=0,
                    help="with --traffic-probe: 0 the library's choice, 2 the per-ordering pass")
    ap.add_argument("--batch-probe", action="store_true")
    args = ap.parse_args()
    if args.traffic_probe:
        traffic_probe(args.K, args.seed, args.pair_mode)
        return
    if args.batch_probe:
        batch_probe()
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: this build measures one GPU only (asked for {args.gpus} GPUs, "
              f"WORLD_SIZE={world}); the pair loop would shard as contiguous pair ranges "
              f"with one all-reduce per evaluation, which is not built here.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import nbo_correlatedgaussians_amd as necg
    from nbo_correlatedgaussians_amd.engine import Engine
    from nbo_correlatedgaussians_amd.flops import (flops_per_eval, flops_per_eval_v0,
                                                   mass_is_rank1, pairs_per_eval,
                                                   two_sided_term_flops)

    sysd = necg.lih_system()
    K = args.K
    iv, C = necg.generate_parameters(sysd, K, args.seed)
    x = necg.pack_x(iv, C)
    eng = Engine(sysd, "cuda:0")
    xd = torch.as_tensor(x, device="cuda:0")
    f = torch.empty(1, dtype=torch.float64, device="cuda:0")
    g = torch.empty(x.size, dtype=torch.float64, device="cuda:0")
    from nbo_correlatedgaussians_amd.autograd import ecg_energy
    xr = xd.clone().requires_grad_(True)

    def step():
        if args.path == "autograd":      # forward: pair kernel + quotient; backward: adjoint
            xr.grad = None
            E = ecg_energy(xr, eng)
            E.backward()
            return E
        eng.outprod_t(xd, f, g)
        return f

    for _ in range(args.warmup):
        E = step()
    eng.check("warmup")
    print(f"bench: warmup done, E = {E.item():.12e}", file=sys.stderr, flush=True)

    eng.timing(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        E = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    eng.check("timed steps")
    ms_pair, launches = eng.timing(2)
    eng.timing(0)
    ms_step = (t1 - t0) * 1000.0 / args.steps
    pairs = pairs_per_eval(K)
    value = pairs / (ms_step / 1000.0)
    two_sided = eng.pair_form()[1]
    rank1 = mass_is_rank1(sysd.mass.tolist())
    fl = flops_per_eval(sysd.n, sysd.nst, K, rank1)
    fl_v0 = flops_per_eval_v0(sysd.n, sysd.nst, K)
    pair_ms = ms_pair / max(launches, 1)
    achieved = fl / (pair_ms / 1000.0) / 1e12
    kt, ts, w = eng.tiling()
    ex_fl = recorded_executed_flop() if (K == 2048 and args.seed == 1 and two_sided) else None
    fma_peak = fp64_fma_rate()
    out = {
        "metric": "ECG pair energy+grad evals/s (FP64, % of peak); served p50/p99 latency",
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"LiH n=5 nst=24 K={K} seed={args.seed}: one energy+analytic "
                               f"gradient per step, "
                               + ("E = ecg_energy(x) (torch.autograd.Function) + E.backward()"
                                  if args.path == "autograd" else "outprod(x)"),
                   "K": K, "pairs": pairs, "parallelism": "single GPU"},
        "roofline": {"bound": "fp64_valu",
                     "engine": "FP64 vector ALU, not MFMA and not HBM (the n <= 6 pair "
                               "algebra is too small and irregular for matrix cores; "
                               "MI355X FP64 vector and matrix peaks are both 78.6 TFLOP/s)",
                     "kernel": "pair2_kernel<5>" if two_sided else "pair_kernel<5,MODE_SPLIT>",
                     "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS,
                     "traffic": (recorded_traffic()
                                 if (K == 2048 and args.seed == 1 and two_sided) else None),
                     "traffic_source": "profiles/pair2_kernel_lih_k2048_traffic.md",
                     "flop_model": (f"two-sided n-space op count, "
                                    f"{two_sided_term_flops(sysd.n, rank1=rank1)} FLOP per "
                                    f"(pair, term) (flops.two_sided_term"
                                    f"{', diagonal-plus-constant mass' if rank1 else ''})"),
                     "flop_per_launch": fl, "pair_kernel_ms": pair_ms,
                     "v0_model_flop_per_launch": fl_v0,
                     "v0_model_tflops": fl_v0 / (pair_ms / 1000.0) / 1e12,
                     "executed_flop_per_launch": ex_fl,
                     "executed_frac": (ex_fl / (pair_ms / 1000.0) / 1e12 / PEAK_FP64_TFLOPS
                                       if ex_fl else None),
                     "executed_source": "profiles/pair2_kernel_lih_k2048_fp64_counters.md",
                     "fma_probe_tflops": fma_peak,
                     "frac_of_fma_probe": achieved / fma_peak if fma_peak else None,
                     "tiling": {"ket_tile": kt, "term_split": ts, "waves": w}},
        "energy": float(E.item()),
    }
    if args.path == "autograd":
        out["outprod"] = outprod_leg(eng, xd, K)
        print(f"bench: outprod leg done", file=sys.stderr, flush=True)
    if two_sided:
        out["per_ordering"] = per_ordering_leg(eng, xd, K)
        print(f"bench: per-ordering leg done", file=sys.stderr, flush=True)
    if not args.no_props:
        out["properties"] = properties_leg(eng, xd, K)
        print(f"bench: properties leg done", file=sys.stderr, flush=True)
    if not args.no_opt:
        out["optimizer"] = optimizer_leg()
        print(f"bench: optimizer leg done", file=sys.stderr, flush=True)
    if not args.no_batch:
        out["batched"] = batched_leg()
        print(f"bench: batched leg done", file=sys.stderr, flush=True)
    if not args.no_serve:
        out["served"] = served_leg(args.serve_requests, args.serve_rate)
        print(f"bench: served leg done", file=sys.stderr, flush=True)
    if not args.no_cpu:
        host = host_cpu()
        threads = host["share"]          # the box's CPU share for this GPU
        out["cpu_baseline"] = cpu_baseline(sysd, K, x, args.cpu_rows, threads)
        out["cpu_baseline"]["host"] = host
        print(f"bench: oracle baseline done", file=sys.stderr, flush=True)
        Eh = float(E.item())          # norm does not change the work timed
        out["cpu_baseline"]["numpy_reference_structure"] = reference_structure_baseline(
            sysd, K, x, Eh, 1.0, threads)
        print(f"bench: reference-structure baseline done", file=sys.stderr, flush=True)
        if "optimizer" in out:
            out["cpu_baseline"]["optimizer"] = cpu_optimizer(out["optimizer"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
