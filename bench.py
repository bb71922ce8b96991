#!/usr/bin/env python3
"""Benchmark of the mesh-propagation hot path on one MI355X.

Workload: Clements MZI mesh, N=256 modes, L=256 layers, ideal 50:50 splitters, phases
U[0, 2pi) (float32), B=4096 complex64 unit-norm CN(0,1) input vectors.  One step is one
fwd+grad pass of the loss J = sum |U x - y|^2 with y = U_target x (U_target a second
random-phase mesh, precomputed), as one library call (MESHPROP_OP_FWDGRAD): one prologue
for the forward rows, the inverse rows and dT/dp; fwdprop; the adjoint with dJ/du_out =
2 (u_out - y) formed in registers; the fixed-order dp reduction writing dp.  --path separate
runs fwdprop and backdiff as two calls (two prologues, dp zeroed and accumulated) instead.
Inputs are resident in HBM before timing.

Prints one JSON line.  value = MZI-vector updates/s = B * sum(lens) / step time (an update
is one active crossing applied to one vector through the forward and the adjoint pass,
128 FLOP).  The roofline object is for the dominant kernel (the backdiff propagation,
96 FLOP per update), timed with HIP events on the stream it runs on.  cpu_baseline times the
compiled reference CPU engine (c/dot.cpp, built into oracle/_ref by oracle/Makefile) on a
bounded slice of the same workload.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MZI-vector updates/s fwd+grad, N=256 L=256 B=4096; p50/p99 request latency"
N, L, B = 256, 256, 4096
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def host_cpu():
    """The host's stated cores: os.cpu_count(), the cores this process may run on (affinity) and
    the lscpu model name."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    model = None
    try:
        import subprocess
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in txt.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except Exception:
        pass
    return {"os_cpu_count": os.cpu_count(), "affinity_cores": aff, "model": model}


def cpu_baseline(p, u, y, seconds_target=12.0):
    """Reference c/dot.cpp (complex128, as mesh.py:98-113 calls it) on a slice of the batch:
    mode 0 (forward) + mode 2 (VJP), timed with one OpenMP thread per core this process may use
    (the reported value) and with one thread.  Falls back to the NumPy oracle ("port") if the
    compiled reference is absent."""
    from oracle import meshprop_oracle as O
    g = O.clements(N)
    T = O.slot_T(g, "mzi", p.astype(np.float64))
    host = host_cpu()
    full = max(1, host["affinity_cores"])
    nact = int(np.sum(g.lens))
    try:
        from oracle import ref_engine
        if not ref_engine.available():
            raise FileNotFoundError(ref_engine.LIB_PATH)
        kind = "reference"

        def one(bs, threads):
            x = u[:bs].astype(np.complex128)
            t0 = time.perf_counter()
            out = ref_engine.run(g, T, x, 0, threads=threads)
            ref_engine.run(g, T, out, 2, dv=2 * (out - y[:bs]), threads=threads)
            return time.perf_counter() - t0
    except Exception:
        kind = "port"
        full = 1

        def one(bs, threads):
            x = u[:bs].astype(np.complex128)
            t0 = time.perf_counter()
            out = O.fwd(g, T, x)
            O.vjp_inverse(g, T, out, 2 * (out - y[:bs]))
            return time.perf_counter() - t0

    def timed(threads, target):
        bs = 16
        t = one(bs, threads)
        # grow the slice toward ~target seconds of CPU work, within the batch, then repeat it
        while t < target / 4 and bs < B:
            bs = min(B, bs * 4)
            t = one(bs, threads)
        times = [t]
        while sum(times) < target and len(times) < 7:
            times.append(one(bs, threads))
        t = float(np.median(times))
        return bs * nact / t, bs, t, len(times)

    value, bs, t, n = timed(full, seconds_target)
    one_thread = timed(1, seconds_target / 2) if full > 1 else (value, bs, t, n)
    out = {"value": value, "unit": "MZI-vector updates/s", "cores": full, "kind": kind,
           "single_thread_value": one_thread[0], "host": host}
    if kind == "reference":
        out["sample"] = (f"c/dot.cpp complex128 fwd (mode 0) + VJP (mode 2), same N=256 L=256 Clements mesh, "
                         f"{bs} of the {B} vectors, {full} OpenMP threads (one per core this process may use), "
                         f"median of {n} runs of {t:.2f} s; single thread: {one_thread[1]} vectors, "
                         f"{one_thread[2]:.2f} s per run.  Timing only: at more than one thread the reference "
                         f"VJP races on its dT and dph accumulators (c/dot.cpp:135, 279-285), so its gradient "
                         f"is not used as a result here")
    else:
        out["sample"] = f"NumPy FP64 oracle fwd + inverse-propagation VJP, {bs} vectors, {t:.2f} s per run"
    return out


def load_traffic(name):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 --pmc summary."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def bench_batched(args):
    """Heterogeneous mix (64 meshes, half Clements / half butterfly, N in {32..512}, batch
    log-uniform in [1, 2048]), one fwd+grad job per mesh: the grouped persistent launch
    (autotuned tiles, replayed hipGraph) against the per-mesh plain launches of the drop-in path,
    issued one mesh after another (as host calls, and recorded into a graph)."""
    import torch
    from meshes_amd import workloads as WL
    from meshes_amd.autotune import tune_group
    from meshes_amd.batched import MeshGroup, PlainRunner

    specs = WL.mix_specs(seed=args.seed)
    jobs = WL.mix_jobs(specs, seed=args.seed + 1)
    upd = WL.updates(jobs)
    plain = PlainRunner(jobs)
    t_plain = WL.time_ms(plain.run, args.steps, args.warmup)
    ref = [(j.y.clone(), j.dp.clone()) for j in jobs]
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        plain.run()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        plain.run()
    t_plain_graph = WL.time_ms(g.replay, args.steps, args.warmup)
    del g
    cfg, tuned_ms, sweep = tune_group(jobs, reps=3)
    grp = MeshGroup(jobs)
    grp.capture()
    t0 = time.perf_counter()
    t_grp = WL.time_ms(grp.run, args.steps, args.warmup)
    wall = (time.perf_counter() - t0) / (args.steps + args.warmup)
    torch.cuda.synchronize()
    err_y = max(float(torch.linalg.norm(j.y - y) / torch.linalg.norm(y)) for j, (y, _) in zip(jobs, ref))
    err_dp = max(float(torch.linalg.norm(j.dp - d) / torch.linalg.norm(d)) for j, (_, d) in zip(jobs, ref))
    ms = float(np.median(t_grp))
    out = {
        "metric": "MZI-vector updates/s, heterogeneous 64-mesh fwd+grad mix, grouped persistent launch",
        "value": upd / (ms * 1e-3), "unit": "MZI-vector updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32 (complex64 fields, float32 phases)",
        "data": "synthetic: seeded phases U[0,2pi), CN(0,1) unit-norm inputs and targets",
        "config": {"workload": "64 meshes: 32 Clements (depth N) + 32 butterfly (log2 N), N uniform in "
                               "{32,64,128,256,512}, batch log-uniform in [1,2048], fwd+grad each",
                   "seed": args.seed, "meshes": [list(t) for t in specs], "updates_per_step": upd,
                   "tiles": {str(k): v for k, v in cfg.items()}},
        "plain_ms": float(np.median(t_plain)), "plain_graph_ms": float(np.median(t_plain_graph)),
        "speedup_vs_plain": float(np.median(t_plain)) / ms,
        "speedup_vs_plain_graph": float(np.median(t_plain_graph)) / ms,
        "latency_ms": {"p50": ms, "p99": float(np.percentile(t_grp, 99))},
        "group_wall_ms": wall * 1e3,
        "max_rel_l2_vs_plain": {"y": err_y, "dp": err_dp},
        "autotune": {"best_ms": tuned_ms, "trials": len(sweep)},
    }
    print(json.dumps(out), flush=True)


def bench_serving(args):
    """Poisson request stream over 256 resident Clements meshes (N log-uniform in [16, 1024],
    depth N), 70 % forward / 30 % gradient requests, lognormal batches (median 64, sigma 1.5):
    the continuous-batching server against sequential plain launches, per arrival rate."""
    import torch
    from meshes_amd.serving import MeshServer, poisson_requests, resident_meshes

    rates = [float(r) for r in args.rates.split(",")]
    meshes = resident_meshes(256, seed=args.seed)
    server = MeshServer(meshes, hbm_budget=args.hbm_gib << 30, slots=3)
    warm = poisson_requests(meshes, rate=20000.0, n=200, seed=args.seed + 99)
    server.serve(warm)
    server.serve_sequential(warm[:20])
    sweep = []
    for k, rate in enumerate(rates):
        n = int(min(max(rate * args.window, 150), args.max_requests))
        reqs = poisson_requests(meshes, rate=rate, n=n, seed=args.seed + k)
        _, st = server.serve(reqs)
        seq_reqs = poisson_requests(meshes, rate=rate, n=n, seed=args.seed + k)
        _, sq = server.serve_sequential(seq_reqs)
        sweep.append({"rate_req_s": rate, "served": st, "sequential": sq})
        print(json.dumps({"rate": rate, "served_p50_ms": st["p50_ms"], "served_p99_ms": st["p99_ms"],
                          "seq_p50_ms": sq["p50_ms"], "seq_p99_ms": sq["p99_ms"]}), file=sys.stderr, flush=True)
        del reqs, seq_reqs
        torch.cuda.empty_cache()
    top = sweep[-1]
    out = {
        "metric": "served request throughput and p50/p99 latency, Poisson fwd/grad stream over 256 resident meshes",
        "value": top["served"]["throughput_req_s"], "unit": "requests/s", "n_gpus": 1,
        "steps": sum(s["served"]["requests"] for s in sweep), "warmup": len(warm), "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 (complex64 fields, float32 phases)",
        "data": "synthetic: seeded phases, Poisson arrivals, lognormal batches of CN(0,1) unit vectors on device",
        "config": {"workload": "256 Clements MZI meshes, N log-uniform in [16,1024], depth N; 70/30 fwd/grad; "
                               "batch lognormal median 64 sigma 1.5", "rates": rates, "slots": 3,
                   "hbm_budget_GiB": args.hbm_gib},
        "latency_ms": {"p50": top["served"]["p50_ms"], "p99": top["served"]["p99_ms"]},
        "sweep": sweep,
    }
    print(json.dumps(out), flush=True)


def bench_training(args):
    """Differentiable training through the autograd adjoint: imperfect lossy Clements mesh
    N = L = 512 (alpha, beta ~ N(0, 0.02 rad), 0.1 dB per crossing), B = 8192, Adam on the
    phases toward a Haar-random target, checkpoints every K in {8, 16, 32} layers.

    A path through 512 crossings at 0.1 dB each loses 51 dB, so |U x| is ~3e-3 of |x| whatever
    the phases and sum |U x - t|^2 hardly moves.  The default loss is therefore scale-invariant:
    J = 1 - F^2 with the batch fidelity F = |<Y, T>| / (|Y| |T|) (Frobenius inner product of the
    outputs Y = U X and the targets T); F is reported every 20 steps for either loss."""
    import torch
    from scipy.stats import unitary_group
    from meshes_amd import MeshNetworkGPU, geometry
    from meshes_amd.autograd import MeshLayer

    Nt, Lt, Bt = 512, 512, 8192
    rng = np.random.RandomState(6)
    lens, shifts = geometry.clements(Nt)
    p = rng.uniform(0, 2 * np.pi, (Lt, Nt // 2, 2)).astype(np.float32)
    s = np.zeros((Lt, Nt // 2, 14), np.float32)
    s[..., :2] = 0.02 * rng.standard_normal((Lt, Nt // 2, 2))
    s[..., 10:] = np.float32(np.log(10 ** (-0.1 / 20)))  # two stages per arm: 0.1 dB per crossing
    U = unitary_group.rvs(Nt, random_state=6)
    x = torch.as_tensor(geometry.unit_cn(rng, Bt, Nt).astype(np.complex64)).cuda()
    t = x @ torch.as_tensor(U.T.astype(np.complex64)).cuda()
    nact = int(np.sum(lens))
    runs = []
    for K in (8, 16, 32):
        layer = MeshLayer(MeshNetworkGPU(Nt, Lt, lens, shifts, p, s, "mzi_generic", is_phase=False), ckpt_every=K)
        opt = torch.optim.Adam(layer.parameters(), lr=args.lr)
        losses, fids, ev = [], [], []
        tn = torch.linalg.norm(t)
        for step in range(args.train_steps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            opt.zero_grad()
            y = layer(x)
            fid = torch.abs(torch.sum(y.conj() * t)) / (torch.linalg.norm(y) * tn)
            if args.train_loss == "fidelity":
                loss = 1 - fid ** 2
            else:
                loss = torch.sum(torch.abs(y - t) ** 2) / Bt
            loss.backward()
            opt.step()
            b.record()
            losses.append(loss.detach())
            fids.append(fid.detach())
            ev.append((a, b))
        torch.cuda.synchronize()
        ms = np.array([a.elapsed_time(b) for a, b in ev[3:]])
        lv = [float(v) for v in losses]
        fv = [float(v) for v in fids]
        runs.append({"K": K, "ms_per_step_p50": float(np.median(ms)), "ms_per_step_p99": float(np.percentile(ms, 99)),
                     "loss_first": lv[0], "loss_last": lv[-1], "loss_every_20": lv[::20],
                     "fidelity_first": fv[0], "fidelity_last": fv[-1], "fidelity_every_20": fv[::20],
                     "updates_per_s": Bt * nact / (np.median(ms) * 1e-3)})
        print(json.dumps(runs[-1]), file=sys.stderr, flush=True)
    best = min(runs, key=lambda r: r["ms_per_step_p50"])
    out = {
        "metric": "Adam training step (fwd with checkpoints + checkpointed adjoint + optimizer), lossy N=512 L=512 B=8192",
        "value": best["updates_per_s"], "unit": "MZI-vector updates/s", "n_gpus": 1, "steps": args.train_steps,
        "warmup": 3, "ms_per_step": best["ms_per_step_p50"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32 (complex64 fields, float32 phases)",
        "data": "synthetic: seeded phases and splitter errors, CN(0,1) unit inputs, Haar target (scipy, seed 6)",
        "config": {"workload": "imperfect lossy Clements MZI mesh (mzi_generic), N=512 L=512 B=8192, Adam lr "
                               f"{args.lr}, checkpoints every K layers", "K": [8, 16, 32], "loss": args.train_loss,
                   "fidelity": "|<Y, T>| / (|Y| |T|), Frobenius over the batch"},
        "runs": runs,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--tile-fwd", type=str, default="")
    ap.add_argument("--tile-bwd", type=str, default="")
    ap.add_argument("--path", choices=["fwdgrad", "separate"], default="fwdgrad",
                    help="fwd+grad as one library call (MESHPROP_OP_FWDGRAD), or fwdprop then backdiff calls")
    ap.add_argument("--workload", choices=["fwdgrad", "batched", "serving", "training"], default="fwdgrad",
                    help="fwdgrad: the single-mesh step (default); batched: the heterogeneous grouped mix; "
                         "serving: the continuous-batching request loop; training: Adam through the autograd "
                         "adjoint on the lossy N=512 mesh")
    ap.add_argument("--train-steps", type=int, default=200)
    ap.add_argument("--lr", type=float, default=0.003)
    ap.add_argument("--train-loss", choices=["fidelity", "l2"], default="fidelity",
                    help="training: 1 - F^2 with the scale-invariant batch fidelity F, or sum |U x - t|^2 / B")
    ap.add_argument("--rates", type=str, default="100,1000,5000,20000")
    ap.add_argument("--window", type=float, default=1.5, help="seconds of arrivals per rate (serving)")
    ap.add_argument("--max-requests", type=int, default=3000)
    ap.add_argument("--hbm-gib", type=int, default=24)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--probe-copy", type=int, default=0,
                    help="only copy a buffer of this many MiB a few times (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.probe_copy:
        import torch
        a = torch.empty(args.probe_copy << 20, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"probe_copy_MiB": args.probe_copy, "copies": 5}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the workload runs on a single MI355X; sharding the batch across "
              "GPUs (with one all-reduce of dp per step) is not run by this benchmark.", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batched":
        bench_batched(args)
        return
    if args.workload == "serving":
        bench_serving(args)
        return
    if args.workload == "training":
        bench_training(args)
        return
    import torch
    import meshes_amd._lib as Lb
    from meshes_amd import MeshNetworkGPU, geometry

    dev = torch.device("cuda", 0)
    lens, shifts = geometry.clements(N)
    p = np.random.RandomState(2).uniform(0, 2 * np.pi, (L, N // 2, 2)).astype(np.float32)
    pt = np.random.RandomState(3).uniform(0, 2 * np.pi, (L, N // 2, 2)).astype(np.float32)
    x = geometry.unit_cn(np.random.RandomState(4), B, N).astype(np.complex64)
    tile_f = tuple(int(t) for t in args.tile_fwd.split(",")) if args.tile_fwd else None
    tile_b = tuple(int(t) for t in args.tile_bwd.split(",")) if args.tile_bwd else None
    mesh = MeshNetworkGPU(N, L, lens, shifts, p, np.zeros(0), "mzi", is_phase=False)
    target = MeshNetworkGPU(N, L, lens, shifts, pt, np.zeros(0), "mzi", is_phase=False)
    xb = torch.as_tensor(x).to(dev).contiguous()          # (B, N): Fortran (N, B) buffer
    yb = target.dot(xb.T).T.contiguous()
    uout = torch.empty_like(xb)
    uin = torch.empty_like(xb)
    dp = torch.zeros(mesh.n_phase, dtype=torch.float32, device=dev)
    geo = (mesh.lens, mesh.shifts)
    ldp = N
    onecall = args.path == "fwdgrad"
    if onecall:
        ws_g_n = Lb.workspace_size(Lb.OP_FWDGRAD, "mzi", N, L, B, False, tile_b)
        ws_g = torch.empty(ws_g_n, dtype=torch.uint8, device=dev)
    else:
        ws_f_n = Lb.workspace_size(Lb.OP_FWDPROP, "mzi", N, L, B, False, tile_f)
        ws_b_n = Lb.workspace_size(Lb.OP_BACKDIFF, "mzi", N, L, B, False, tile_b)
        ws_f = torch.empty(ws_f_n, dtype=torch.uint8, device=dev)
        ws_b = torch.empty(ws_b_n, dtype=torch.uint8, device=dev)

    def fwd():
        Lb.run(Lb.OP_FWDPROP, "mzi", N, L, B, False, *geo, mesh.p_phase, None, ldp, None, 0, xb, None, uout, None,
               N, tile=tile_f, workspace=ws_f, ws_bytes=ws_f_n)

    def bwd(flags):
        Lb.run(Lb.OP_BACKDIFF, "mzi", N, L, B, False, *geo, mesh.p_phase, dp, ldp, None, 0, uout, None, uin, None,
               N, target=yb, tile=tile_b, workspace=ws_b, ws_bytes=ws_b_n, flags=flags)

    def fwdgrad(flags):
        Lb.run(Lb.OP_FWDGRAD, "mzi", N, L, B, False, *geo, mesh.p_phase, dp, ldp, None, 0, xb, None, uout, None,
               N, target=yb, tile=tile_b, workspace=ws_g, ws_bytes=ws_g_n, flags=flags)

    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(6)] for _ in range(args.steps)]

    def step(e=None, full=False):
        # e[2]..e[3] bracket the dominant kernel (the backdiff propagation); with full, e[0], e[1]
        # and e[4] also split the step into its stages (the instrumented pass after the timed one:
        # each event record costs a few microseconds of device time).
        if onecall:
            # The stages of one fwd+grad call, issued one by one so that each can be timed.
            if e and full:
                e[0].record()
            fwdgrad(Lb.ONLY_PROLOGUE)
            fwdgrad(Lb.SKIP_PROLOGUE | Lb.ONLY_FORWARD)
            if e:
                if full:
                    e[1].record()
                e[2].record()
            fwdgrad(Lb.SKIP_PROLOGUE | Lb.SKIP_FORWARD | Lb.SKIP_REDUCE)
            if e:
                e[3].record()
            fwdgrad(Lb.ONLY_REDUCE | Lb.DP_ASSIGN)  # dp = gradient: no separate zeroing of dp
            if e and full:
                e[4].record()
            return
        if e and full:
            e[0].record()
        dp.zero_()
        fwd()
        if e and full:
            e[1].record()
        bwd(Lb.ONLY_PROLOGUE)
        if e:
            e[2].record()
        bwd(Lb.SKIP_PROLOGUE | Lb.SKIP_REDUCE)
        if e:
            e[3].record()
        bwd(Lb.ONLY_REDUCE)
        if e and full:
            e[4].record()

    # Warmup: the device clocks reach their steady level only after some tens of steps.
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(ev[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kb_ms = np.array([ev[k][2].elapsed_time(ev[k][3]) for k in range(args.steps)])
    # Instrumented pass (not timed): per-step latency and the split of a step into its stages.
    for k in range(args.steps):
        step(ev[k], full=True)
    torch.cuda.synchronize()
    step_ms = np.array([ev[k][0].elapsed_time(ev[k][4]) for k in range(args.steps)])
    pre_ms = np.array([ev[k][0].elapsed_time(ev[k][1]) for k in range(args.steps)])

    nact = int(np.sum(lens))
    updates = B * nact
    ms_step = wall * 1e3 / args.steps
    kb_avg = float(np.mean(kb_ms))
    flop_k = 96.0 * updates
    kname = "k_bwd_nn<complex64, 4, 2, 8, 4> (backdiff propagation, NN-specialised adjoint)"
    traffic = load_traffic("k_bwd_c2_traffic.json")
    if onecall:
        kernels_ms = {"prologue_and_fwdprop": float(np.mean(pre_ms)), "backdiff_propagation": kb_avg}
    else:
        kernels_ms = {"fwdprop_total": float(np.mean(pre_ms)), "backdiff_propagation": kb_avg}
    achieved = flop_k / (kb_avg * 1e-3) / 1e12
    # correctness of the measured configuration: reconstructed input vs x, gradient finite
    rec_err = None if onecall else float(torch.linalg.norm(uin - xb) / torch.linalg.norm(xb))
    out = {
        "metric": METRIC,
        "value": updates / (ms_step * 1e-3),
        "unit": "MZI-vector updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (complex64 fields, float32 phases)",
        "data": "synthetic: seeded U[0,2pi) phases, CN(0,1) unit-norm inputs, target y = U_target x",
        "config": {"workload": "Clements MZI mesh fwd+grad (fwdprop + backdiff with fused L2 loss gradient), "
                               "N=256 L=256 B=4096, ideal splitters, complex64",
                   "N": N, "L": L, "B": B, "active_crossings": nact, "updates_per_step": updates,
                   "flop_per_update": 128, "path": args.path, "tile_fwd": tile_f, "tile_bwd": tile_b},
        "latency_ms": {"p50": float(np.percentile(step_ms, 50)), "p99": float(np.percentile(step_ms, 99)),
                       "what": "per-step fwd+grad latency (HIP events, instrumented pass after the timed one)"},
        "kernels_ms": kernels_ms,
        "gflops_per_s_fwd_plus_grad": 128.0 * updates / (ms_step * 1e-3) / 1e9,
        "reconstruction_rel_err": rec_err,
        "roofline": {"bound": "valu", "compute": "fp32 VALU, non-packed FMA (2x2 complex rotations: no MFMA shape)",
                     "kernel": kname,
                     "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": traffic,
                     "flop_per_launch": flop_k, "avg_launch_ms": kb_avg},
    }
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(p, x, yb.cpu().numpy(), args.cpu_seconds)
        except Exception as e:  # the GPU figure stands without it
            out["cpu_baseline"] = {"value": None, "error": repr(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
