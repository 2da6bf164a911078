"""ctypes driver of oracle/libvc_oracle.so -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, by __graft_entry__.smoke() (as the checker) and by the
CPU-baseline leg of bench.py.  Never by the product.

OracleSim runs the reference's step sequence on the CPU: frame response when
a frame enters, the LGN steps of a frame period, then for every step the V1
update kernel followed by recal_G over the dense V1_conMat
(src/patch.cu:5966-6546).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "libvc_oracle.so"


def _fp(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class Oracle:
    def __init__(self, abi_mod):
        if not LIB.exists():
            raise RuntimeError(f"{LIB} not built (make -C oracle)")
        self.abi = abi_mod
        L = C.CDLL(str(LIB))
        vp, u32, f32, i32 = C.c_void_p, C.c_uint32, C.c_float, C.c_int
        P = C.POINTER
        L.orc_v1_step.argtypes = [P(abi_mod.V1Params)] + [vp] * 14 + [u32, vp, vp, vp, vp, vp, vp, vp,
                                                                      u32, vp, vp, i32]
        L.orc_v1_step.restype = i32
        L.orc_lgn_frame_response.argtypes = [vp, u32, u32, u32, u32, vp, vp, vp, vp, vp,
                                             C.c_uint64, vp]
        L.orc_lgn_steps.argtypes = [P(abi_mod.LgnParams)] + [vp] * 11 + [u32, u32, vp, vp, vp]
        L.orc_lgn_init.argtypes = [P(abi_mod.LgnParams), vp, vp]
        L.orc_lgn_precompute.argtypes = [P(abi_mod.LgnSetup), vp, vp, vp, vp, vp]
        L.orc_expf.argtypes = [f32]; L.orc_expf.restype = f32
        L.orc_logf.argtypes = [f32]; L.orc_logf.restype = f32
        L.orc_philox.argtypes = [P(C.c_uint32 * 4), u32, u32, P(C.c_uint32 * 4)]
        L.orc_cond_single.argtypes = [P(f32), P(f32), f32, f32, f32, f32]
        L.orc_cond_decay.argtypes = [P(f32), P(f32), f32, f32, f32]
        L.orc_block_reduce_export.argtypes = [vp, i32, i32]; L.orc_block_reduce_export.restype = f32
        L.orc_bilinear_export.argtypes = [vp, i32, i32, f32, f32]; L.orc_bilinear_export.restype = f32
        L.orc_num_threads.restype = i32
        self.L = L

    # -- scalar helpers --------------------------------------------------
    def expf(self, x):
        return np.float32(self.L.orc_expf(float(x)))

    def logf(self, x):
        return np.float32(self.L.orc_logf(float(x)))

    def philox(self, ctr, k0, k1):
        c = (C.c_uint32 * 4)(*ctr)
        o = (C.c_uint32 * 4)()
        self.L.orc_philox(C.byref(c), k0, k1, C.byref(o))
        return list(o)

    def cond_single(self, g, h, f, ddt, rise, decay):
        G, H = C.c_float(g), C.c_float(h)
        self.L.orc_cond_single(C.byref(G), C.byref(H), f, ddt, rise, decay)
        return np.float32(G.value), np.float32(H.value)

    def cond_decay(self, g, h, ddt, rise, decay):
        G, H = C.c_float(g), C.c_float(h)
        self.L.orc_cond_decay(C.byref(G), C.byref(H), ddt, rise, decay)
        return np.float32(G.value), np.float32(H.value)

    def block_reduce(self, x, nThreads):
        x = np.ascontiguousarray(x, dtype=np.float32)
        return np.float32(self.L.orc_block_reduce_export(_fp(x), x.size, nThreads))

    def bilinear(self, img, x, y):
        img = np.ascontiguousarray(img, dtype=np.float32)
        H, W = img.shape
        return np.float32(self.L.orc_bilinear_export(_fp(img), W, H, x, y))

    def num_threads(self):
        return self.L.orc_num_threads()


class OracleSim:
    """The oracle's simulation of a Network (numpy state on the host)."""

    def __init__(self, net, pkg, tables: dict | None = None, state: dict | None = None):
        from importlib import import_module
        self.abi = import_module(pkg.__name__ + ".abi")
        params = import_module(pkg.__name__ + ".params")
        network = import_module(pkg.__name__ + ".network")
        self.o = Oracle(self.abi)
        self.net = net
        sp = net.spec
        self.tm = network.timing(sp)
        self.p = params.v1_params_struct(net)
        st = params.initial_state(net) if state is None else state
        self.st = {k: np.ascontiguousarray(v, dtype=np.float32).copy() for k, v in st.items()}
        self.tonic = np.ascontiguousarray(net.tonicDep, dtype=np.float32)
        self.conMat = np.ascontiguousarray(net.dense_conmat())
        self.nbr = np.ascontiguousarray(net.nbr, dtype=np.uint32)
        self.nNear = np.ascontiguousarray(net.nNear, dtype=np.uint32)
        self.nLGN_of = np.ascontiguousarray(net.nLGN_of, dtype=np.uint32)
        self.sLGN = np.ascontiguousarray(net.sLGN, dtype=np.float32)
        self.idLGN = np.ascontiguousarray(net.idLGN, dtype=np.uint32)
        bs = sp.blockSize
        self.ipre = np.zeros(sp.nV1, np.uint32)
        self.npre = np.zeros(2 * sp.nblock, np.uint32)
        ng = len(sp.riseTimeE) * sp.nE + len(sp.riseTimeI) * sp.nI
        self.og = np.zeros(sp.nblock * ng, np.float32)
        self.oh = np.zeros(sp.nblock * ng, np.float32)
        # LGN tables
        L = net.lgn
        self.lgn = {k: np.ascontiguousarray(L[k]) for k in L if isinstance(L[k], np.ndarray)}
        if tables is None:
            tables = self.precompute()
        self.tables = {k: np.ascontiguousarray(v, dtype=np.float32) for k, v in tables.items()}
        self.lp = params.lgn_params_struct(net)
        self.batch = getattr(net, "batch", None)
        if self.batch is not None:                  # host copies of the trial tables
            B = self.batch
            self.bt = {k: np.ascontiguousarray(getattr(B, k), dtype=np.uint32)
                       for k in ("blockTrial", "trialV1Off", "cellTrial", "trialLgnOff", "trialSeed")}
            self.p.blockTrial = self.bt["blockTrial"].ctypes.data
            self.p.trialV1Off = self.bt["trialV1Off"].ctypes.data
            self.p.trialSeed = self.bt["trialSeed"].ctypes.data
            self.lp.cellTrial = self.bt["cellTrial"].ctypes.data
            self.lp.trialLgnOff = self.bt["trialLgnOff"].ctypes.data
            self.lp.trialSeed = self.bt["trialSeed"].ctypes.data
        self.lTR = np.zeros(sp.nLGN, np.float32)
        self.lNL = np.zeros(sp.nLGN, np.float32)
        self.o.L.orc_lgn_init(C.byref(self.lp), _fp(self.lTR), _fp(self.lNL))
        self.Rring = np.zeros((self.tm["maxFrame"], sp.nLGN, 2), np.float32)
        init = np.empty((3, sp.height, sp.width), np.float32)
        for c in range(3):
            init[c] = net.init_lms[c]
        self.Rring[0] = self.frame_response(init)
        self.Rring[1:] = self.Rring[0]
        self.lgn_sInfo = np.zeros((self.tm["ntPerFrame"], sp.nLGN), np.float32)
        self.lgn_fr = np.zeros((self.tm["ntPerFrame"], sp.nLGN), np.float32)
        self.it = 0
        self.sInfo_hist = []
        self.lgn_hist = []
        self.fr_hist = []

    def precompute(self):
        from importlib import import_module
        network = import_module(self.net.spec.__class__.__module__)
        sp = self.net.spec
        x0, y0, nvd = network.view_geometry(sp)
        q = self.abi.LgnSetup()
        q.nLGN, q.nSample1D, q.nKernelSample = sp.nLGN, sp.nSpatialSample1D, sp.nKernelSample
        q.kernelSampleInterval = self.tm["kernelSampleInterval"]
        q.iKernelSampleT0 = self.tm["iKernelSampleT0"]
        q.maxSample1D = network.max_sample_1d(self.net)
        q.dt, q.nsig, q.normViewDistance, q.x0, q.y0 = sp.dt, sp.nsig, nvd, x0, y0
        keep = []
        for k in ("polar", "ecc", "rw", "rh", "orient", "k", "tauR", "tauD", "delay", "ratio",
                  "nR", "nD", "covariant"):
            a = np.ascontiguousarray(self.lgn[k], dtype=np.float32)
            keep.append(a)
            setattr(q, k, a.ctypes.data)
        nS = sp.nSpatialSample1D ** 2
        TW = np.zeros((sp.nLGN, 2, sp.nKernelSample), np.float32)
        SW = np.zeros(nS, np.float32)
        SCx = np.zeros((sp.nLGN, 2, nS), np.float32)
        SCy = np.zeros((sp.nLGN, 2, nS), np.float32)
        mc = np.zeros(sp.nLGN, np.float32)
        self.o.L.orc_lgn_precompute(C.byref(q), _fp(TW), _fp(SW), _fp(SCx), _fp(SCy), _fp(mc))
        return dict(TW=TW, SW=SW, SCx=SCx, SCy=SCy, maxConvol=mc)

    def frame_response(self, frame, per_trial: bool = False):
        """frame: [3][H][W], or with per_trial [ntrial][3][H][W] (batched)."""
        sp = self.net.spec
        frame = np.ascontiguousarray(frame, dtype=np.float32)
        R = np.zeros((sp.nLGN, 2), np.float32)
        t = self.tables
        ct = _fp(self.bt["cellTrial"]) if per_trial else None
        stride = int(frame[0].size) if per_trial else 0
        self.o.L.orc_lgn_frame_response(_fp(frame), sp.width, sp.height, sp.nLGN,
                                        sp.nSpatialSample1D, _fp(t["SCx"]), _fp(t["SCy"]),
                                        _fp(t["SW"]), _fp(self.lgn["coneType"]), ct, stride, _fp(R))
        return R

    def run(self, nsteps: int, record: bool = True):
        sp, tm = self.net.spec, self.tm
        npf = tm["ntPerFrame"]
        frames = self.net.frames
        t = self.tables
        kC = np.ascontiguousarray(self.lgn["k"][0])
        kS = np.ascontiguousarray(self.lgn["k"][1])
        st = self.st
        for _ in range(nsteps):
            it = self.it
            k_in = it % npf
            if k_in == 0:
                fidx = it // npf
                if self.batch is not None:                   # frames [ntrial][nF][3][H][W]
                    self.Rring[fidx % tm["maxFrame"]] = self.frame_response(
                        frames[:, fidx % frames.shape[1]], per_trial=True)
                else:
                    self.Rring[fidx % tm["maxFrame"]] = self.frame_response(
                        frames[fidx % frames.shape[0]])
                self.o.L.orc_lgn_steps(C.byref(self.lp), _fp(t["TW"]), _fp(kC), _fp(kS),
                                       _fp(t["maxConvol"]), _fp(self.lgn["c50"]),
                                       _fp(self.lgn["sharpness"]), _fp(self.lgn["logA"]),
                                       _fp(self.lgn["logB"]), _fp(self.lTR), _fp(self.lNL),
                                       _fp(self.Rring), it, npf, _fp(self.lgn_sInfo),
                                       _fp(self.lgn_fr), None)
                if record:
                    self.lgn_hist.append(self.lgn_sInfo.copy())
                    self.fr_hist.append(self.lgn_fr.copy())
            out = np.zeros(sp.nV1, np.float32)
            lgn_row = np.ascontiguousarray(self.lgn_sInfo[k_in])
            rc = self.o.L.orc_v1_step(C.byref(self.p), _fp(st["v"]), _fp(st["w"]), _fp(st["tBack"]),
                                      _fp(st["gFF"]), _fp(st["hFF"]), _fp(st["gE"]), _fp(st["hE"]),
                                      _fp(st["gI"]), _fp(st["hI"]), _fp(self.tonic),
                                      _fp(self.nLGN_of), _fp(self.sLGN), _fp(self.idLGN), _fp(lgn_row),
                                      self.net.maxNear, _fp(self.nNear), _fp(self.nbr),
                                      _fp(self.conMat), _fp(self.ipre), _fp(self.npre), _fp(self.og),
                                      _fp(self.oh), it, _fp(out), None, 1)
            if rc != 0:
                raise RuntimeError(f"orc_v1_step returned {rc}")
            if record:
                self.sInfo_hist.append(out)
            self.it += 1

    def sInfo(self):
        return np.stack(self.sInfo_hist)

    def lgn_sinfo(self):
        return np.concatenate(self.lgn_hist)
