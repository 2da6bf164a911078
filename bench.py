#!/usr/bin/env python3
"""Benchmark of the patchV1 simulation step on one MI355X.

Workload: the full-patch plain forward -- 64 blocks x 1024 = 65,536 V1
neurons (E AdEx, I LIF), 16,384 LGN cells, 30 LGN + 800 E + 200 I synapses
per neuron, 256x256-px drifting grating at 125 Hz, dt = 0.125 ms.  One
"step" is one dt of the whole path (LGN stage + V1 stages) with all inputs
resident in HBM.  Synthetic, seeded network and stimulus.

Prints one JSON line: neuron-steps/s over the timed steps, the roofline of
the dominant kernel (the fused V1 step) from HIP events over the timed
region, and a CPU baseline (the oracle, timed on a bounded sample).
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import re
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
HBM_PEAK_GBS = 8000.0


def load_package():
    name = "visual_cortex_amd"
    if name in sys.modules:
        return sys.modules[name]
    d = ROOT / "visual-cortex_amd"
    spec = importlib.util.spec_from_file_location(name, d / "__init__.py",
                                                  submodule_search_locations=[str(d)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def v1_bytes_per_neuron_step(spec, rho: float) -> float:
    """Algorithmic HBM bytes per V1 neuron-step of the fused V1 kernel:
    13 persistent state floats read + written, tonicDep read, sInfo written
    (4*(2*13 + 2) = 112), m LGN synapses (weight + index, 8 B each), the LGN
    sInfo table amortised over the neurons (4*nLGN/nV1), and the event-driven
    recurrent term 8*rho*K (K = outgoing synapses per neuron).  The
    background drive is generated in-kernel (0 B)."""
    K = spec.KE + spec.KI
    return 112.0 + 8.0 * spec.m + 4.0 * spec.nLGN / spec.nV1 + 8.0 * rho * K


def pmc_traffic(kernel: str, workload: str):
    """Physical HBM bytes per launch of `kernel` on `workload`, from the
    committed rocprofv3 counter summary profiles/<kernel>_<workload>_pmc.md
    (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by what the counter
    reads for a plain device copy of known size); None without one."""
    path = ROOT / "profiles" / f"{kernel}_{workload}_pmc.md"
    try:
        text = path.read_text()
    except OSError:
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)\s*$", text, re.M)
    return float(m.group(1)) if m else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16000)
    ap.add_argument("--warmup", type=int, default=640)
    ap.add_argument("--cpu-steps", type=int, default=8000,
                    help="steps of the CPU-baseline sample (0 = skip)")
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--reps", type=int, default=5, help="c4: timed forward/backward repetitions")
    ap.add_argument("--trials", type=int, default=32, help="c3: number of trials")
    ap.add_argument("--delay", type=float, default=0.0,
                    help="c1/c2 with axonal delays (noDelay 0) at this speedOfThought, mm/ms")
    ap.add_argument("--far", type=int, default=0,
                    help="with --delay: this many far (long-range) inputs per neuron")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run a plain 256 MiB device copy (x4) for HBM counter calibration")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} requested): the "
              "multi-GPU sharding of this path (by V1 blocks, or by trial) is not built.",
              file=sys.stderr)
        sys.exit(2)

    pkg = load_package()
    import numpy as np
    import torch
    from visual_cortex_amd import network
    from visual_cortex_amd.sim import GpuSim

    if args.config == "c3":
        return bench_batched(args)
    if args.config == "c4":
        return bench_lgn_adjoint(args)
    if args.config == "c5":
        return bench_v1_adjoint(args)
    spec = network.spec_c2() if args.config == "c2" else network.spec_c1()
    if args.delay > 0:
        spec = network.replace(spec, noDelay=0, speedOfThought=args.delay, nFar=args.far)
    elif args.far:
        print("--far needs --delay (far connections are delayed by distance)", file=sys.stderr)
        sys.exit(2)
    K, W = args.steps, args.warmup
    spec = network.replace(spec, nt=K + W) if hasattr(network, "replace") else spec
    t0 = time.time()
    net = network.build_network(spec)
    log(f"[bench] network {args.config}: nV1={spec.nV1} nLGN={spec.nLGN} "
        f"rec synapses={net.n_rec_synapses()} built in {time.time() - t0:.1f}s")

    sim = GpuSim(net, record=False)
    sim.run(W)
    torch.cuda.synchronize()
    log(f"[bench] warmup {W} steps done")

    out = torch.empty(K, spec.nV1, dtype=torch.float32, device="cuda")
    events = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.run(K, out=out, events=events)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    log(f"[bench] timed {K} steps in {elapsed:.3f}s")

    sim.check()
    spikes = int((out >= 1.0).sum().item())
    rho = spikes / float(K * spec.nV1)
    # HIP events on the V1 stream bracket every V1 launch (a persistent launch
    # covers one frame period, a one-step launch one step)
    ev_ms = sum(a.elapsed_time(b) for a, b, n in events)
    ev_n = sum(n for a, b, n in events)
    steps_per_launch = ev_n / len(events) if sim.persistent else 1.0
    v1_step_s = ev_ms / 1e3 / ev_n
    bpn = v1_bytes_per_neuron_step(spec, rho)
    bytes_per_launch = bpn * spec.nV1 * steps_per_launch
    launch_s = v1_step_s * steps_per_launch
    achieved = bytes_per_launch / launch_s / 1e9
    value = spec.nV1 * K / elapsed
    isE = (np.arange(spec.nV1) % spec.blockSize) < spec.nE
    per_n = (out >= 1.0).sum(0).cpu().numpy()
    T = K * spec.dt / 1000.0
    log(f"[bench] rates: E {per_n[isE].mean() / T:.2f} Hz, I {per_n[~isE].mean() / T:.2f} Hz; "
        f"rho={rho:.3e}; V1 {v1_step_s * 1e6:.2f} us/step ({launch_s * 1e6:.1f} us per launch of "
        f"{steps_per_launch:.0f} steps); {bpn:.1f} B/neuron-step")
    del out

    if args.copy_calib:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(4):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] copy calibration: 4 x {src.numel() * 4} bytes")
        del src, dst

    cpu = None
    if args.cpu_steps > 0:
        # the oracle's delayed gather scans the dense matrix: a short sample
        cpu = cpu_baseline(pkg, net, sim, min(args.cpu_steps, 48) if sim.delay else args.cpu_steps)

    kernel = "k_v1_multi" if sim.persistent else "k_v1_step"
    if sim.delay:
        kernel = "k_v1_step+k_rec_delay" + ("+k_far_mark+k_far_eval" if sim.far_on else "")
    line = {
        "metric": "V1 neuron-steps/s at 64 blocks dt=0.125ms",
        "value": value,
        "unit": "neuron-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": elapsed * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded network, drifting-grating stimulus)",
        "config": {"workload": f"plain forward, {'full patch' if args.config == 'c2' else 'small'}: "
                               f"{spec.nblock} blocks x {spec.blockSize} neurons, {spec.nLGN} LGN, "
                               f"m={spec.m}, K={spec.KE}+{spec.KI}, {spec.width}x{spec.height} px, "
                               f"dt={spec.dt} ms" + (f", axonal delays at {spec.speedOfThought} "
                                                      f"mm/ms (ring of {sim.depth} steps)"
                                                      if sim.delay else "")
                               + (f", {spec.nFar} far inputs per neuron" if sim.delay and
                                  sim.far_on else ""),
                   "nV1": spec.nV1, "nLGN": spec.nLGN, "dt_ms": spec.dt,
                   "spikes_per_neuron_step": rho},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (pmc_traffic(kernel, "full_patch")
                                 if args.config == "c2" and not sim.delay else None),
                     "kernel": kernel,
                     "steps_per_launch": steps_per_launch,
                     "bytes_per_launch": bytes_per_launch, "launch_us": launch_s * 1e6,
                     "step_us": v1_step_s * 1e6},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def bench_batched(args):
    """Batched trials: `--trials` independent trials (8-96 blocks, 8
    orientations, distinct seeds, the small plain-forward distributions),
    500 ms each, in one launch per step, against the same trials run one
    after the other (trials/s, and the ratio)."""
    load_package()
    import torch
    from visual_cortex_amd import batch
    from visual_cortex_amd.sim import GpuSim
    K, W = args.steps, args.warmup
    specs = batch.c3_specs(ntrial=args.trials, nt=K + W)
    t0 = time.time()
    nets = [batch.build_network_fast(s, device="cuda") for s in specs]
    bnet = batch.batch_networks(nets)
    log(f"[bench] {len(nets)} trials, blocks {[s.nblock for s in specs]}, nV1={bnet.spec.nV1}, "
        f"built in {time.time() - t0:.1f}s")
    t0 = time.time()
    sim = GpuSim(bnet, record=False, tune=True)
    log(f"[bench] autotuned one-step launch: stepWG {sim.shape[0]}, gatherTile {sim.shape[1]} "
        f"({time.time() - t0:.1f}s)")
    sim.run(W)
    torch.cuda.synchronize()
    events = []
    t0 = time.perf_counter()
    sim.run(K, events=events)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    launch = sum(a.elapsed_time(b) for a, b, n in events) / 1e3 / sum(n for a, b, n in events)
    log(f"[bench] batched: {K} steps in {tb:.3f}s ({tb / K * 1e6:.1f} us/step, V1 launch {launch * 1e6:.1f} us)")
    sim_shape = sim.shape
    del sim
    ts = 0.0
    for n in nets:
        s1 = GpuSim(n, record=False)
        s1.run(W)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s1.run(K)
        torch.cuda.synchronize()
        ts += time.perf_counter() - t0
        del s1
    log(f"[bench] serial: {len(nets)} trials x {K} steps in {ts:.3f}s")
    sp = bnet.spec
    line = {
        "metric": "batched trials/s",
        "value": len(nets) * (K * sp.dt / 500.0) / tb,
        "unit": "trials/s (500 ms trials)",
        "n_gpus": 1, "steps": K, "warmup": W, "ms_per_step": tb * 1e3 / K,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded networks, drifting gratings at 8 orientations)",
        "config": {"workload": f"{len(nets)} batched trials, 8-96 blocks x 1024 neurons, "
                               f"400 E + 100 I partners, 20 LGN inputs, dt={sp.dt} ms",
                   "nV1_total": sp.nV1, "blocks": [s.nblock for s in specs],
                   "stepWG": sim_shape[0], "gatherTile": sim_shape[1]},
        "serial_trials_per_s": len(nets) * (K * sp.dt / 500.0) / ts,
        "speedup_vs_serial": ts / tb,
        "neuron_steps_per_s": sp.nV1 * K / tb,
    }
    print(json.dumps(line), flush=True)


def bench_lgn_adjoint(args):
    """Differentiation of the LGN stage: 16,384 cells, 500 ms (4000 steps,
    63 frames) of 256x256 white-noise frames (mean 0.5, sigma 0.3), loss =
    mean squared error of the rates against the rates under a drifting
    grating.  Times the forward (frame responses + rates) and the backward
    (rate adjoint + frame adjoint + temporal-parameter adjoint) with HIP
    events on the stream they run on; reports backward / forward."""
    load_package()
    import torch
    from visual_cortex_amd import network as N
    from visual_cortex_amd.lgn_grad import LgnStage
    spec = N.spec_c2()
    T = 4000
    t0 = time.time()
    lgn = N.lgn_layer(spec)
    st = LgnStage(spec, lgn, T)
    frames = torch.as_tensor(N.white_noise_frames(spec, st.nF, seed=11), device="cuda")
    with torch.no_grad():
        tgt = st.rates(torch.as_tensor(N.grating_frames(spec, st.nF), device="cuda"), st.theta0)
    theta = st.theta0.clone()
    log(f"[bench] LGN adjoint: nLGN={spec.nLGN} {spec.width}x{spec.height} T={T} nF={st.nF}, "
        f"set up in {time.time() - t0:.1f}s")
    stream = torch.cuda.current_stream()

    def once(timing=None):
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        e[0].record(stream)
        fr, R, th = st.forward(frames, theta)
        e[1].record(stream)
        gFr = 2.0 * (fr - tgt) / fr.numel()
        dF, dTh = st.backward(frames, R, th, gFr)
        e[2].record(stream)
        if timing is not None:
            timing.append(e)
        return dF, dTh

    once()
    torch.cuda.synchronize()
    tim = []
    t0 = time.perf_counter()
    for _ in range(args.reps):
        once(tim)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / args.reps
    fwd = sum(a.elapsed_time(b) for a, b, c in tim) / len(tim) / 1e3
    bwd = sum(b.elapsed_time(c) for a, b, c in tim) / len(tim) / 1e3
    log(f"[bench] forward {fwd * 1e3:.2f} ms, backward {bwd * 1e3:.2f} ms, wall/rep {wall * 1e3:.2f} ms")
    line = {
        "metric": "LGN adjoint/fwd time",
        "value": bwd / fwd,
        "unit": "backward time / forward time",
        "n_gpus": 1, "steps": T, "warmup": 1, "ms_per_step": (fwd + bwd) * 1e3 / T,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (white-noise frames sigma 0.3, grating-driven target rates)",
        "config": {"workload": f"LGN stage differentiation: {spec.nLGN} cells, {st.nF} frames of "
                               f"{spec.width}x{spec.height} px, {T} steps, 6 parameters per cell",
                   "reps": args.reps},
        "forward_ms": fwd * 1e3, "backward_ms": bwd * 1e3,
    }
    print(json.dumps(line), flush=True)


def bench_v1_adjoint(args):
    """Differentiation of the V1 dynamics: 16 blocks = 16,384 conductance-LIF
    neurons, 20 LGN + 400 E + 100 I synapses per neuron, 250 ms (2000
    steps), checkpoints every 100 steps; loss = squared error of smoothed
    spike counts in 25 ms bins against a run with perturbed recurrent
    weights.  Times the forward (2000 steps, checkpoints kept) and the
    backward (segment recompute with snapshots + reverse sweep) with HIP
    events on their stream; reports backward / forward."""
    load_package()
    import torch
    from visual_cortex_amd import batch, network as N
    from visual_cortex_amd.v1_grad import V1Run, soft_bin_counts, soft_count_loss
    spec = N.spec_c5()
    T = spec.nt
    t0 = time.time()
    net = batch.build_network_fast(spec, device="cuda")
    run = V1Run(net, T, ckpt=100)
    sL, w = run.sim.sLGN.clone(), run.sim.wrec.clone()
    g = torch.Generator(device="cuda").manual_seed(3)
    with torch.no_grad():
        wp = w * torch.exp(0.2 * torch.randn(w.shape, generator=g, device="cuda"))
        tgt = soft_bin_counts(run.forward(sL, wp), spec.dt, 200, 1.0)
    log(f"[bench] V1 adjoint: nV1={spec.nV1} synapses={w.numel()} + {sL.numel()} T={T}, "
        f"set up in {time.time() - t0:.1f}s")
    stream = torch.cuda.current_stream()

    def once(timing=None):
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        e[0].record(stream)
        sInfo = run.forward(sL, w)
        e[1].record(stream)
        x = sInfo.detach().requires_grad_(True)
        loss = soft_count_loss(x, tgt, spec.dt, 200, 1.0)
        gOut, = torch.autograd.grad(loss, x)
        e[2].record(stream)
        dsl, dw = run.backward(gOut)
        e2 = torch.cuda.Event(enable_timing=True)
        e2.record(stream)
        if timing is not None:
            timing.append((e[0], e[1], e[2], e2))
        return sInfo, dsl, dw

    sInfo, _, _ = once()
    torch.cuda.synchronize()
    rho = float((sInfo >= 1).float().mean())
    tim = []
    for _ in range(args.reps):
        once(tim)
    torch.cuda.synchronize()
    fwd = sum(a.elapsed_time(b) for a, b, c, d in tim) / len(tim) / 1e3
    bwd = sum(c.elapsed_time(d) for a, b, c, d in tim) / len(tim) / 1e3
    log(f"[bench] forward {fwd * 1e3:.2f} ms ({spec.nV1 * T / fwd:.3e} neuron-steps/s), "
        f"backward {bwd * 1e3:.2f} ms; spikes/neuron-step {rho:.2e} "
        f"({rho / (spec.dt / 1000):.1f} Hz)")
    line = {
        "metric": "V1 adjoint/fwd time",
        "value": bwd / fwd,
        "unit": "backward time / forward time",
        "n_gpus": 1, "steps": T, "warmup": 1, "ms_per_step": (fwd + bwd) * 1e3 / T,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded LIF network, drifting grating, perturbed-weight target)",
        "config": {"workload": f"V1 spike-time adjoint: {spec.nblock} blocks x {spec.blockSize} "
                               f"LIF neurons, {spec.m} LGN + {spec.KE}+{spec.KI} recurrent "
                               f"synapses, {T} steps, checkpoint every 100",
                   "reps": args.reps, "weights": int(w.numel() + sL.numel())},
        "forward_ms": fwd * 1e3, "backward_ms": bwd * 1e3, "spikes_per_neuron_step": rho,
    }
    print(json.dumps(line), flush=True)


def cpu_baseline(pkg, net, sim, nsteps):
    """The oracle (C restatement of the reference's step equations, OpenMP)
    on the first `nsteps` steps of the same network from the same initial
    state, LGN tables taken from the GPU so that only the step loop is timed."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle
    tables = dict(TW=sim.TW.cpu().numpy(), SW=sim.SW.cpu().numpy(), SCx=sim.SCx.cpu().numpy(),
                  SCy=sim.SCy.cpu().numpy(), maxConvol=sim.maxConvol.cpu().numpy())
    t0 = time.time()
    o = oracle.OracleSim(net, pkg, tables=tables)
    log(f"[bench] cpu baseline set-up {time.time() - t0:.1f}s (dense conMat)")
    t0 = time.perf_counter()
    o.run(nsteps, record=False)
    el = time.perf_counter() - t0
    sp = net.spec
    log(f"[bench] cpu baseline {nsteps} steps in {el:.2f}s")
    return {"value": sp.nV1 * nsteps / el, "unit": "neuron-steps/s", "cores": o.o.num_threads(),
            "kind": "port",
            "sample": f"first {nsteps} steps ({nsteps * sp.dt:g} ms) of the same workload, "
                      f"same network and initial state, OpenMP over neurons/cells"}


if __name__ == "__main__":
    main()
