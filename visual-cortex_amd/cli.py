"""`patch_fast -c <cfg> [-s seed] [-p print]` for the MI355X simulation step
(src/patch.cu:172-177).  Runs nt steps of the network the cfg's input files
describe and writes rawData / LGN_fr / LGN_sp like the reference when the cfg
asks for them (rawData, saveLGN_fr, getLGN_sp; hWrite for the h traces).

    python visual-cortex_amd/cli.py -c patchV1.cfg [-s seed]
"""
from __future__ import annotations

import argparse
import importlib.util
import sys
import time
from pathlib import Path


def _pkg():
    name = "visual_cortex_amd"
    if name in sys.modules:
        return sys.modules[name]
    d = Path(__file__).resolve().parent
    spec = importlib.util.spec_from_file_location(name, d / "__init__.py",
                                                  submodule_search_locations=[str(d)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="patch_fast")
    ap.add_argument("-c", "--cfg", required=True)
    ap.add_argument("-s", "--seed", type=int)
    ap.add_argument("-p", "--print", dest="print_log", default="false")
    a = ap.parse_args(argv)
    _pkg()
    from visual_cortex_amd import config, loader, snapshot
    from visual_cortex_amd.sim import GpuSim
    cmd = {} if a.seed is None else {"seed": a.seed}
    cfg = config.Config.from_file(a.cfg, cmd)
    t0 = time.time()
    net = loader.load_network(cfg)
    sp = net.spec
    print(f"{sp.nV1} V1 neurons in {sp.nblock} blocks, {sp.nLGN} LGN cells, "
          f"loaded in {time.time() - t0:.1f}s", flush=True)
    sim = GpuSim(net, record=False)
    snap = config.snapshot_options(cfg)
    if snap["restore"]:
        h = snapshot.restore(sim, snap["restore"], as_init=snap["as_init"])
        print(f"restored {snap['restore']} (step {h['it0']}"
              f"{', as the initial condition' if snap['as_init'] else ''})", flush=True)
    outputs = dict(rawData=cfg.bool("rawData", True), lgn_fr=cfg.bool("saveLGN_fr", True),
                   lgn_sp=cfg.bool("getLGN_sp", False), hWrite=cfg.bool("hWrite", True))
    paths = dict(rawData=cfg.path("fRawData"), lgn_fr=cfg.path("fLGN_fr"),
                 lgn_sp=cfg.path("fLGN_sp"))
    npf = sim.tm["ntPerFrame"]
    last = {"path": None}

    def take_snapshot(g, i):
        # every snapshotInterval steps and at the last step (src/patch.cu:6375),
        # at frame-period boundaries: the interval is rounded up to them
        every = -(-max(snap["interval"], 1) // npf) * npf
        if g.it % npf != 0 or not ((i + 1) % every == 0 or i == sp.nt - 1):
            return
        path = f"snapShot_{int(time.time())}{snap['output_suffix']}"
        snapshot.save(g, path)
        if last["path"] and last["path"] != path and snap["delete_previous"]:
            Path(last["path"]).unlink(missing_ok=True)
        last["path"] = path
        print(f"saved a snapshot in {path}", flush=True)

    # the strengths of a plastic run every sampleInterval_LGN_V1 steps
    # (learnData_FF = 1, src/patch.cu:6196-6200, 6936-6938)
    slgn_w = None
    if sim.learn and sp.learnData_FF == 1 and not cfg.bool("minimal", False):
        from visual_cortex_amd import formats
        slgn_w = formats.SLGNWriter(cfg.path("f_sLGN"), sp, int(net.sLGN.shape[0]))

    def after_step(g, i):
        if slgn_w is not None and (i + 1) % max(sp.sampleInterval_LGN_V1, 1) == 0:
            slgn_w.write(g.sLGN.cpu().numpy())
        take_snapshot(g, i)

    t0 = time.time()
    sim.run_with_outputs(sp.nt, paths, outputs, on_step=after_step)
    if slgn_w is not None:
        slgn_w.close()
    print(f"simulated {sp.nt} steps ({sp.nt * sp.dt:g} ms) in {time.time() - t0:.1f}s", flush=True)
    if not cfg.bool("minimal", False):          # src/patch.cu:6952-7028
        from visual_cortex_amd import formats
        from visual_cortex_amd.network import view_geometry
        x0, y0, nvd = view_geometry(sp)
        sr = cfg.floats("sRatioV1", [1.0])
        formats.write_patchv1_cfg(cfg.path("fPatchV1_cfg"), sp, sp.nLGN,
                                  cfg.path("fStimulus"), (nvd, x0, y0, 1.0, 1.0),
                                  sRatioV1=sr * 4 if len(sr) == 1 else sr)
    return 0


if __name__ == "__main__":
    sys.exit(main())
