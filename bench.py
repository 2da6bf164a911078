#!/usr/bin/env python3
"""Benchmark of the patchV1 simulation step on one MI355X.

Headline workload: the full-patch plain forward -- 64 blocks x 1024 = 65,536
V1 neurons (E AdEx, I LIF), 16,384 LGN cells, 30 LGN + 800 E + 200 I synapses
per neuron, a 2 s drifting grating at 256x256 px and 125 Hz, dt = 0.125 ms.
One "step" is one dt of the whole path: the LGN stage's share of the step
(the temporal sum and spike draw of the step, and 1/64 of the spatial
response of the next frame) and the V1 stages, with every input resident in
HBM.  Synthetic, seeded network and stimulus.

Before the timed steps the network runs a fixed 200 ms pre-roll (whatever
--warmup is) so that the timed steps see its steady activity; then --warmup
untimed steps; then --steps timed steps, whose LGN work is issued inside the
timed region.

Prints one JSON line: neuron-steps/s over the timed steps; the roofline of
the dominant kernel (the persistent V1 kernel) from HIP events on its stream
over the timed region; a CPU baseline (the oracle, timed on a bounded
sample); and, in the same line, the other two figures of the metric --
batched trials/s (32 trials of 500 ms against the same trials run one after
the other) and backward/forward time of the two adjoints -- plus the LGN
stage's FLOP/s.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import platform
import re
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
METRIC = "V1 neuron-steps/s at 64 blocks dt=0.125ms; batched trials/s; adjoint/fwd time"
PREROLL_STEPS = 1600          # 200 ms at dt = 0.125 ms


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


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def v1_bytes_per_neuron_step(spec, rho: float) -> float:
    """Algorithmic HBM bytes per V1 neuron-step of the fused V1 kernel:
    13 persistent state floats read + written, tonicDep read, sInfo written
    (4*(2*13 + 2) = 112), m LGN synapses (weight + index, 8 B each), the LGN
    sInfo table amortised over the neurons (4*nLGN/nV1), and the event-driven
    recurrent term 8*rho*K (K = outgoing synapses per neuron).  The
    background drive is generated in-kernel (0 B)."""
    K = spec.KE + spec.KI
    return 112.0 + 8.0 * spec.m + 4.0 * spec.nLGN / spec.nV1 + 8.0 * rho * K


def pmc_traffic(name: str, steps_per_launch: float):
    """Physical HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 counter summary profiles/<name>_pmc.md (FETCH_SIZE and
    WRITE_SIZE of the timed dispatch, each scaled by what the counter reads
    for a plain device copy of known size), when that summary was taken on
    launches of the same length; None otherwise."""
    path = ROOT / "profiles" / f"{name}_pmc.md"
    try:
        text = path.read_text()
    except OSError:
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)\s*$", text, re.M)
    n = re.search(r"^steps_per_launch:\s*([0-9.]+)\s*$", text, re.M)
    if not m or not n or abs(float(n.group(1)) - steps_per_launch) > 1e-9:
        return None
    return float(m.group(1))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1280)
    ap.add_argument("--warmup", type=int, default=64)
    ap.add_argument("--cpu-steps", type=int, default=8000,
                    help="steps of the CPU-baseline sample (0 = skip)")
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--no-extras", action="store_true",
                    help="c1/c2: skip the batched-trials, adjoint and LGN-stage figures")
    ap.add_argument("--reps", type=int, default=3, help="c4/c5: timed forward/backward repetitions")
    ap.add_argument("--trials", type=int, default=32, help="c3: number of trials")
    ap.add_argument("--trial-steps", type=int, default=4000, help="c3: steps per trial (500 ms)")
    ap.add_argument("--delay", type=float, default=0.0,
                    help="c1/c2 with axonal delays (noDelay 0) at this speedOfThought, mm/ms")
    ap.add_argument("--far", type=int, default=0,
                    help="this many far (long-range) inputs per neuron (delayed by distance: "
                         "with --delay, or alone with undelayed near connections)")
    ap.add_argument("--learn", type=float, default=0.0,
                    help="c1/c2 with the feed-forward plasticity (learning 3) at this A_LGN")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run a plain 256 MiB device copy (x4) for HBM counter calibration")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} requested): the "
              "multi-GPU sharding of this path (by V1 blocks, or by trial) is not built; "
              "every figure is for one MI355X.", file=sys.stderr)
        sys.exit(2)

    load_package()
    if args.config == "c3":
        print(json.dumps(bench_batched(args)), flush=True)
        return
    if args.config == "c4":
        print(json.dumps(bench_lgn_adjoint(args)), flush=True)
        return
    if args.config == "c5":
        print(json.dumps(bench_v1_adjoint(args)), flush=True)
        return
    line = bench_plain(args)
    if args.config == "c2" and not args.no_extras and not args.delay and not args.far and not args.learn:
        import torch
        torch.cuda.empty_cache()
        line["batched_trials"] = bench_batched(args)
        torch.cuda.empty_cache()
        line["lgn_adjoint"] = bench_lgn_adjoint(args)
        torch.cuda.empty_cache()
        line["v1_adjoint"] = bench_v1_adjoint(args)
    print(json.dumps(line), flush=True)


def bench_plain(args) -> dict:
    pkg = load_package()
    import numpy as np
    import torch
    from visual_cortex_amd import network
    from visual_cortex_amd.sim import GpuSim

    spec = network.spec_c2() if args.config == "c2" else network.spec_c1()
    if args.delay > 0:
        spec = network.replace(spec, noDelay=0, speedOfThought=args.delay, nFar=args.far)
    elif args.far:
        spec = network.replace(spec, noDelay=1, noFarDelay=0, speedOfThought=1.0, nFar=args.far)
    if args.learn > 0:
        # the reference's evoked0.cfg time constants, a 5 Hz target rate
        spec = network.replace(spec, learning=3, A_LGN=args.learn, tauLTP=16.8, tauLTD=33.7,
                               tauTrip=114.0, r_LTD=1.0, tauAvg=(1000.0, 1000.0),
                               targetFR=(5.0, 5.0), gmaxLGN=100.0, gminLGN=0.0)
    K, W = args.steps, args.warmup
    # the configuration's whole stimulus (2 s / 1 s of frames), whatever --steps is
    spec = network.replace(spec, nt=max(spec.nt, PREROLL_STEPS + W + K))
    t0 = time.time()
    net = network.build_network(spec)
    log(f"[bench] network {args.config}: nV1={spec.nV1} nLGN={spec.nLGN} "
        f"rec synapses={net.n_rec_synapses()} frames={net.frames.shape[0]} "
        f"built in {time.time() - t0:.1f}s")

    sim = GpuSim(net, record=False)
    pre = torch.empty(PREROLL_STEPS, spec.nV1, dtype=torch.float32, device="cuda")
    sim.run(PREROLL_STEPS, out=pre)
    torch.cuda.synchronize()
    rho_pre = float((pre[-800:] >= 1.0).float().mean())
    del pre
    sim.run(W)
    torch.cuda.synchronize()
    n_v1_launches = sim.n_v1_launches
    log(f"[bench] pre-roll {PREROLL_STEPS} steps (last 100 ms: {rho_pre:.3e} spikes per "
        f"neuron-step) + warmup {W} steps done; {n_v1_launches} V1 launches so far")

    out = torch.empty(K, spec.nV1, dtype=torch.float32, device="cuda")
    events = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.run(K, out=out, events=events)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    sim.check()
    log(f"[bench] timed {K} steps (steps {PREROLL_STEPS + W}..{PREROLL_STEPS + W + K - 1}) in "
        f"{elapsed * 1e3:.3f} ms; V1 launches {n_v1_launches + 1}..{sim.n_v1_launches} of the run "
        f"(1-based)")

    spikes = int((out >= 1.0).sum().item())
    rho = spikes / float(K * spec.nV1)
    # HIP events on the V1 stream bracket every V1 launch (a persistent launch
    # covers the rest of a frame period, a one-step launch one step)
    ev_ms = sum(a.elapsed_time(b) for a, b, n in events)
    ev_n = sum(n for a, b, n in events)
    steps_per_launch = ev_n / len(events) if sim.persistent else 1.0
    v1_step_s = ev_ms / 1e3 / ev_n
    bpn = v1_bytes_per_neuron_step(spec, rho)
    bytes_per_launch = bpn * spec.nV1 * steps_per_launch
    launch_s = v1_step_s * steps_per_launch
    achieved = bytes_per_launch / launch_s / 1e9
    valid = rho > 0.0
    value = spec.nV1 * K / elapsed
    isE = (np.arange(spec.nV1) % spec.blockSize) < spec.nE
    per_n = (out >= 1.0).sum(0).cpu().numpy()
    T = K * spec.dt / 1000.0
    log(f"[bench] rates: E {per_n[isE].mean() / T:.2f} Hz, I {per_n[~isE].mean() / T:.2f} Hz; "
        f"rho={rho:.3e}; V1 {v1_step_s * 1e6:.2f} us/step ({launch_s * 1e6:.1f} us per launch of "
        f"{steps_per_launch:.1f} steps); {bpn:.1f} B/neuron-step; whole path "
        f"{elapsed / K * 1e6:.2f} us/step")
    if not valid:
        log("[bench] no V1 spike in the timed steps: the figure is not a measurement of the "
            "workload (value null)")
    del out

    stage_a = None
    if not args.no_extras:
        stage_a = lgn_stage_flops(sim, spec)
        log(f"[bench] LGN stage alone: {stage_a['us_per_period']:.1f} us per frame period, "
            f"{stage_a['tflops']:.3f} TFLOP/s")

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
        cpu = cpu_baseline(pkg, net, sim, min(args.cpu_steps, 48) if (sim.delay or sim.far_on)
                           else args.cpu_steps)

    kernel = "k_v1_multi" if sim.persistent else "k_v1_step"
    if sim.delay:
        kernel = "k_v1_step+k_rec_delay" + ("+k_far_mark+k_far_eval" if sim.far_on else "")
    elif sim.far_only:
        kernel = "k_v1_step+k_v1_step(apply)+k_far_mark+k_far_eval"
    if sim.learn:
        kernel += "+k_v1_learn_ff"
    workload = (f"plain forward, {'full patch' if args.config == 'c2' else 'small'}: "
                f"{spec.nblock} blocks x {spec.blockSize} neurons, {spec.nLGN} LGN, m={spec.m}, "
                f"K={spec.KE}+{spec.KI}, {spec.width}x{spec.height} px, "
                f"{net.frames.shape[0]} frames at {spec.frameRate} Hz, dt={spec.dt} ms, "
                f"timed after a {PREROLL_STEPS}-step pre-roll")
    if sim.delay:
        workload += f", axonal delays at {spec.speedOfThought} mm/ms (ring of {sim.depth} steps)"
    if sim.far_on:
        workload += f", {spec.nFar} far inputs per neuron"
    if sim.learn:
        workload += f", feed-forward plasticity (learning 3, A_LGN {spec.A_LGN})"
    traffic = None
    if args.config == "c2" and not sim.delay and not sim.far_on and sim.persistent:
        traffic = pmc_traffic("k_v1_multi_full_patch", steps_per_launch)
    line = {
        "metric": METRIC,
        "value": value if valid else None,
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
        "config": {"workload": workload, "nV1": spec.nV1, "nLGN": spec.nLGN, "dt_ms": spec.dt,
                   "preroll_steps": PREROLL_STEPS, "spikes_per_neuron_step": rho,
                   "valid": valid},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": kernel, "steps_per_launch": steps_per_launch,
                     "bytes_per_neuron_step": bpn,
                     "bytes_per_launch": bytes_per_launch, "launch_us": launch_s * 1e6,
                     "step_us": v1_step_s * 1e6},
        "cpu_baseline": cpu,
    }
    if stage_a is not None:
        line["lgn_stage"] = stage_a
    return line


def lgn_stage_flops(sim, spec) -> dict:
    """The LGN stage alone on copies of its state, for two frame periods,
    timed with HIP events on its stream: FLOP/s of the separable formulation
    (per frame nRegion*nSample*13 per cell for the spatial response, per step
    nRegion*nKernelSample*2 per cell for the temporal sum) and the equivalent
    rate of the reference's formulation, which recomputes the spatial sums of
    every frame in the window every step (SURVEY's nRegion*(F_w*nSample*13 +
    nKernelSample*2) per cell-step)."""
    import ctypes as C
    import torch
    from visual_cortex_amd import abi
    from visual_cortex_amd.abi import check, ptr
    npf = sim.tm["ntPerFrame"]
    nL, nS, nKS = spec.nLGN, spec.nSpatialSample1D ** 2, spec.nKernelSample
    st = sim.s_lgn
    lTR, lNL = sim.lTR.clone(), sim.lNL.clone()
    ls = abi.LgnState.from_buffer_copy(sim.ls)
    ls.leftTimeRate, ls.lastNegLogRand = ptr(lTR), ptr(lNL)
    sI = torch.empty(npf, nL, dtype=torch.float32, device=sim.dev)
    fr = torch.empty_like(sI)
    R = sim.Rring.clone()
    k0 = sim.it // npf + 4
    torch.cuda.synchronize()
    reps = 2
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for k in range(k0, k0 + reps):
        frame, ctr = sim._frame_of(k + 1)
        check(sim.lib.vc_lgn_frame_response_cells(
            sim._hs(st), ptr(frame), spec.width, spec.height, nL, 0, nL, spec.nSpatialSample1D,
            ptr(sim.SCx), ptr(sim.SCy), ptr(sim.SW), ptr(sim.coneType), ctr,
            sim.trialFrameStride, ptr(R[(k + 1) % sim.ringF])), "vc_lgn_frame_response_cells")
        check(sim.lib.vc_lgn_steps(sim._hs(st), C.byref(sim.lp), C.byref(ls), ptr(R), k * npf,
                                   npf, ptr(sI), ptr(fr), None), "vc_lgn_steps")
    e1.record(st)
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) / 1e3 / reps
    flops = 2 * nL * nS * 13 + npf * 2 * nL * nKS * 2
    tm = sim.tm
    Fw = (tm["nRetrace"] + npf - 1) // npf + 1
    ref_flops = npf * 2 * nL * (Fw * nS * 13 + nKS * 2)
    return {"us_per_period": t * 1e6, "flops_per_period": flops, "tflops": flops / t / 1e12,
            "frac_fp32_peak": flops / t / 1e12 / FP32_PEAK_TFLOPS,
            "reference_formulation_flops_per_period": ref_flops,
            "reference_equivalent_tflops": ref_flops / t / 1e12,
            "note": f"{nL} cells x 2 regions, {nS} samples, {nKS} taps, {npf} steps per frame; "
                    "spike draw included in the time, not in the FLOPs"}


def bench_batched(args) -> dict:
    """Batched trials: `--trials` independent trials (8-96 blocks, 8
    orientations, distinct seeds, the small plain-forward distributions) of
    --trial-steps steps (500 ms) each in one launch per step, against the
    same trials run one after the other (trials/s, and the ratio).  Both are
    timed over whole trials from step 0, with their LGN stages."""
    load_package()
    import torch
    from visual_cortex_amd import batch
    from visual_cortex_amd.sim import GpuSim
    K = args.trial_steps
    specs = batch.c3_specs(ntrial=args.trials, nt=K)
    t0 = time.time()
    nets = []
    for i, s in enumerate(specs):
        nets.append(batch.build_network_fast(s, device="cuda"))
        if i % 8 == 7:
            log(f"[bench] built {i + 1} of {len(specs)} trial networks ({time.time() - t0:.1f}s)")
    bnet = batch.batch_networks(nets)
    log(f"[bench] {len(nets)} trials, blocks {[s.nblock for s in specs]}, nV1={bnet.spec.nV1}, "
        f"built in {time.time() - t0:.1f}s")
    t0 = time.time()
    sim = GpuSim(bnet, record=False, tune=True)
    log(f"[bench] autotuned one-step launch: stepWG {sim.shape[0]}, gatherTile {sim.shape[1]} "
        f"({time.time() - t0:.1f}s)")
    # a fresh simulator (the tuner issued the first frame period's LGN stage):
    # the shape comes from the tuner's cache, the whole trial is timed
    del sim
    sim = GpuSim(bnet, record=False, tune=True)
    events = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.run(K, events=events)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    sim.check()
    launch = sum(a.elapsed_time(b) for a, b, n in events) / 1e3 / sum(n for a, b, n in events)
    log(f"[bench] batched: {K} steps in {tb:.3f}s ({tb / K * 1e6:.1f} us/step, V1 launch "
        f"{launch * 1e6:.1f} us)")
    sim_shape = sim.shape
    del sim
    ts = 0.0
    for i, n in enumerate(nets):
        if i % 8 == 0:
            log(f"[bench] serial trials {i + 1}..{min(i + 8, len(nets))}")
        s1 = GpuSim(n, record=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s1.run(K)
        torch.cuda.synchronize()
        ts += time.perf_counter() - t0
        s1.check()
        del s1
    log(f"[bench] serial: {len(nets)} trials x {K} steps in {ts:.3f}s")
    sp = bnet.spec
    return {
        "metric": "batched trials/s",
        "value": len(nets) / tb,
        "unit": f"trials/s ({K * sp.dt:g} ms trials)",
        "steps": K, "ms_per_step": tb * 1e3 / K, "higher_is_better": True,
        "config": {"workload": f"{len(nets)} batched trials, 8-96 blocks x 1024 neurons, "
                               f"400 E + 100 I partners, 20 LGN inputs, 8 orientations, "
                               f"dt={sp.dt} ms, {K} steps each",
                   "nV1_total": sp.nV1, "blocks": [s.nblock for s in specs],
                   "stepWG": sim_shape[0], "gatherTile": sim_shape[1]},
        "serial_trials_per_s": len(nets) / ts,
        "speedup_vs_serial": ts / tb,
        "neuron_steps_per_s": sp.nV1 * K / tb,
        "v1_launch_us": launch * 1e6,
    }


def bench_lgn_adjoint(args) -> dict:
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
    for _ in range(args.reps):
        once(tim)
    torch.cuda.synchronize()
    fwd = sum(a.elapsed_time(b) for a, b, c in tim) / len(tim) / 1e3
    bwd = sum(b.elapsed_time(c) for a, b, c in tim) / len(tim) / 1e3
    log(f"[bench] LGN forward {fwd * 1e3:.2f} ms, backward {bwd * 1e3:.2f} ms")
    return {
        "metric": "LGN adjoint/fwd time",
        "value": bwd / fwd,
        "unit": "backward time / forward time",
        "steps": T, "higher_is_better": False,
        "config": {"workload": f"LGN stage differentiation: {spec.nLGN} cells, {st.nF} frames of "
                               f"{spec.width}x{spec.height} px, {T} steps, 6 parameters per cell",
                   "reps": args.reps},
        "forward_ms": fwd * 1e3, "backward_ms": bwd * 1e3,
    }


def bench_v1_adjoint(args) -> dict:
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
    log(f"[bench] V1 forward {fwd * 1e3:.2f} ms ({spec.nV1 * T / fwd:.3e} neuron-steps/s), "
        f"backward {bwd * 1e3:.2f} ms; spikes/neuron-step {rho:.2e} "
        f"({rho / (spec.dt / 1000):.1f} Hz)")
    return {
        "metric": "V1 adjoint/fwd time",
        "value": bwd / fwd,
        "unit": "backward time / forward time",
        "steps": T, "higher_is_better": False,
        "config": {"workload": f"V1 spike-time adjoint: {spec.nblock} blocks x {spec.blockSize} "
                               f"LIF neurons, {spec.m} LGN + {spec.KE}+{spec.KI} recurrent "
                               f"synapses, {T} steps, checkpoint every 100",
                   "reps": args.reps, "weights": int(w.numel() + sL.numel())},
        "forward_ms": fwd * 1e3, "backward_ms": bwd * 1e3, "spikes_per_neuron_step": rho,
    }


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
            "kind": "port", "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "sample": f"first {nsteps} steps ({nsteps * sp.dt:g} ms) of the same workload, "
                      f"same network and initial state, OpenMP over neurons/cells"}


if __name__ == "__main__":
    main()
