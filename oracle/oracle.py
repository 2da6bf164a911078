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
                                                                      u32, vp, vp, i32, vp, vp, vp]
        L.orc_v1_step.restype = i32
        L.orc_v1_rec_delay.argtypes = [P(abi_mod.V1Params)] + [vp] * 4 + [u32] + [vp] * 4 + \
            [f32, vp, u32, u32]
        L.orc_v1_rec_delay.restype = i32
        L.orc_v1_rec_far.argtypes = [P(abi_mod.V1Params)] + [vp] * 8 + [f32, vp, u32, u32]
        L.orc_v1_rec_far.restype = i32
        L.orc_v1_gap.argtypes = [P(abi_mod.V1Params), u32, vp, vp, vp, vp, f32, vp, u32, u32, vp]
        L.orc_v1_gap.restype = i32
        L.orc_v1_rec_fast.argtypes = [P(abi_mod.V1Params)] + [vp] * 4 + [u32] + [vp] * 7 + [u32]
        L.orc_v1_rec_fast.restype = i32
        L.orc_v1_rec_dense.argtypes = [P(abi_mod.V1Params)] + [vp] * 4 + [u32] + [vp] * 4 + [u32]
        L.orc_v1_rec_dense.restype = i32
        L.orc_set_multi_all.argtypes = [i32]
        L.orc_v1_initial_spikes.argtypes = [P(abi_mod.V1Params), vp, P(abi_mod.V1SpInit), vp, vp,
                                            vp, vp, u32, u32]
        L.orc_v1_initial_spikes.restype = i32
        L.orc_lgn_frame_response.argtypes = [vp, u32, u32, u32, u32, vp, vp, vp, vp, vp,
                                             C.c_uint64, vp]
        L.orc_lgn_steps.argtypes = [P(abi_mod.LgnParams)] + [vp] * 11 + [u32, u32, vp, vp, vp]
        L.orc_lgn_steps_learn.argtypes = [P(abi_mod.LgnParams)] + [vp] * 11 + \
            [u32, u32, vp, vp, vp, f32, vp, vp, vp]
        L.orc_v1_learn_ff.argtypes = [P(abi_mod.V1Params), P(abi_mod.LearnFFParams)] + [vp] * 12
        L.orc_lgn_init.argtypes = [P(abi_mod.LgnParams), vp, vp]
        L.orc_lgn_precompute.argtypes = [P(abi_mod.LgnSetup), vp, vp, vp, vp, vp]
        L.orc_expf.argtypes = [f32]; L.orc_expf.restype = f32
        L.orc_logf.argtypes = [f32]; L.orc_logf.restype = f32
        L.orc_philox.argtypes = [P(C.c_uint32 * 4), u32, u32, P(C.c_uint32 * 4)]
        L.orc_cos2pi_vec.argtypes = [u32, vp, vp]
        L.orc_normal_vec.argtypes = [u32, vp, vp, vp]
        L.orc_draw_normal.argtypes = [u32] * 7
        L.orc_draw_normal.restype = f32
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
        self.sLGN = np.array(net.sLGN, dtype=np.float32, order="C")   # a copy: plasticity writes it
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
        # axonal delays (noDelay 0): ring of the last `depth` step outputs
        self.delay = getattr(sp, "noDelay", 1) == 0
        if self.delay:
            self.delayMat = np.ascontiguousarray(net.dense_delaymat())
            nnz = net.n_rec_synapses_csr()
            dmax = float(net.rec_distances()[:nnz].max()) if nnz else 0.0
            self.depth = network.train_depth(dmax, sp.speedOfThought, sp.dt)
            # delayed far connections (recal_G_vec + sum_G) read the same ring
            self.far_on = net.n_far() > 0 and net.far_delayed()
            if self.far_on:
                self.far = [np.ascontiguousarray(x) for x in net.far]
                fmax = float(self.far[3].max()) if self.far[3].size else 0.0
                self.depth = max(self.depth,
                                 network.train_depth(fmax, sp.speedOfThought, sp.dt))
            self.train = np.zeros((self.depth, sp.nV1), np.float32)
        # delayed far connections with undelayed near ones: a ring for them alone
        self.far_only = (not self.delay) and net.n_far() > 0 and net.far_delayed()
        if self.far_only:
            self.far = [np.ascontiguousarray(x) for x in net.far]
            fmax = float(self.far[3].max()) if self.far[3].size else 0.0
            self.depth = network.train_depth(fmax, sp.speedOfThought, sp.dt)
            self.train = np.zeros((self.depth, sp.nV1), np.float32)
        # gap junctions between inhibitory neurons: gapS from the strengths
        # (every inhibitory neuron, whether or not InhGap is on), the gap sums
        # when InhGap is on, first from the initial voltages (the pre-loop
        # recal_G, src/patch.cu:5687-5770)
        self.gapS = net.gap_strength_sums() if net.has_gap() else None
        self.gap = None
        if net.gap_on():
            self.gapMat = np.ascontiguousarray(net.dense_gapmat())
            self.gapDist = np.ascontiguousarray(net.dense_gapdist()) if self.delay else None
            self.gap = np.zeros(sp.nblock * sp.nI, np.float32)
            if self.delay:
                lag = net.gap_lags(sp.speedOfThought, sp.dt)
                need = int(lag.max()) + 1 if lag.size else 1
                if need > self.depth:
                    self.depth = need
                    self.train = np.zeros((self.depth, sp.nV1), np.float32)
        # the Ornstein-Uhlenbeck noise state (src/patch.cu:4086-4088: zeros)
        self.lastNoise = np.zeros(sp.nV1, np.float32)
        if any(x > 0 for x in sp.synFail) and (self.delay or self.far_only):
            raise NotImplementedError("synaptic failure between V1 neurons with conduction "
                                      "delays is not built")
        if self.delay or self.far_only:             # rand_spInit: every slot holds v
            self.train[:] = self.st["v"][None, :]
        init_row = self.st["v"][None, :].copy()
        if state is None and params.has_initial_spikes(sp):
            init_row = self._initial_spikes(params)
        if self.gap is not None:
            self._gap(self.train if self.delay else init_row, self.depth - 1 if self.delay else 0)
        self.lTR = np.zeros(sp.nLGN, np.float32)
        self.lNL = np.zeros(sp.nLGN, np.float32)
        self.o.L.orc_lgn_init(C.byref(self.lp), _fp(self.lTR), _fp(self.lNL))
        # a period of ntPerFrame steps holds denorm frames; its steps reach back
        # over maxFrame - 1 frames before the first of them
        self.ringF = self.tm["maxFrame"] - 1 + self.tm["denorm"]
        self.lp.maxFrame = self.ringF
        self.Rring = np.zeros((self.ringF, sp.nLGN, 2), np.float32)
        init = np.empty((3, sp.height, sp.width), np.float32)
        for c in range(3):
            init[c] = net.init_lms[c]
        self.Rring[0] = self.frame_response(init)
        self.Rring[1:] = self.Rring[0]
        self.lgn_sInfo = np.zeros((self.tm["ntPerFrame"], sp.nLGN), np.float32)
        self.lgn_fr = np.zeros((self.tm["ntPerFrame"], sp.nLGN), np.float32)
        # plasticity (learning 3): the LGN trace and its per-step values, the
        # post traces and the homeostatic totals (zero / the initial sums)
        self.learn = getattr(sp, "learning", 0) == 3
        if self.learn:
            self.learn_p = params.learn_ff_struct(net)
            self.xVar = np.zeros(sp.nLGN, np.float32)
            self.xStart = np.zeros_like(self.lgn_sInfo)
            self.xSpike = np.zeros_like(self.lgn_sInfo)
            self.lvar = {k: np.zeros(sp.nV1, np.float32) for k in ("vLTD", "vTrip", "vAvg")}
            self.totalFF = self.totalFFinf = None
            if sp.applyHomeo:
                self.totalFF, self.totalFFinf = params.learn_ff_totals(net)
        self.it = 0
        self.sInfo_hist = []
        self.lgn_hist = []
        self.fr_hist = []
        self.convol_hist = []

    def _initial_spikes(self, params):
        """rand_spInit (orc_v1_initial_spikes) over the reference's trainDepth
        steps before the first, then the pre-loop recal_G of the step before
        the first (src/patch.cu:5687-5770): the delayed gather at its ring
        slot, or without delays the dense gather of its spikes (their
        failure draws keyed by step -1).  Returns that step's outputs."""
        net, sp, st = self.net, self.net.spec, self.st
        if self.batch is not None or self.far_only or (self.delay and self.far_on):
            raise NotImplementedError("initial spikes with batched trials or far connections")
        D = params.reference_train_depth(net)
        q = params.spinit_struct(net, D)
        if self.delay:
            rc = self.o.L.orc_v1_initial_spikes(C.byref(self.p), _fp(self.nLGN_of), C.byref(q),
                                                _fp(st["v"]), _fp(st["w"]), _fp(st["tBack"]),
                                                _fp(self.train), self.depth, 0)
            if rc != 0:
                raise RuntimeError(f"orc_v1_initial_spikes returned {rc}")
            rc = self.o.L.orc_v1_rec_delay(C.byref(self.p), _fp(st["gE"]), _fp(st["hE"]),
                                           _fp(st["gI"]), _fp(st["hI"]), self.net.maxNear,
                                           _fp(self.nNear), _fp(self.nbr), _fp(self.conMat),
                                           _fp(self.delayMat), sp.speedOfThought,
                                           _fp(self.train), self.depth, self.depth - 1)
            if rc != 0:
                raise RuntimeError(f"orc_v1_rec_delay returned {rc}")
            return self.train[self.depth - 1][None, :].copy()
        ring = np.zeros((D, sp.nV1), np.float32)
        rc = self.o.L.orc_v1_initial_spikes(C.byref(self.p), _fp(self.nLGN_of), C.byref(q),
                                            _fp(st["v"]), _fp(st["w"]), _fp(st["tBack"]),
                                            _fp(ring), D, 0)
        if rc != 0:
            raise RuntimeError(f"orc_v1_initial_spikes returned {rc}")
        row = np.ascontiguousarray(ring[D - 1])
        rc = self.o.L.orc_v1_rec_dense(C.byref(self.p), _fp(st["gE"]), _fp(st["hE"]),
                                       _fp(st["gI"]), _fp(st["hI"]), self.net.maxNear,
                                       _fp(self.nNear), _fp(self.nbr), _fp(self.conMat), _fp(row),
                                       0xffffffff)
        if rc != 0:
            raise RuntimeError(f"orc_v1_rec_dense returned {rc}")
        self.init_ring = ring
        return row[None, :].copy()

    def _gap(self, train, cur):
        """The gap sums from the step outputs in ring slot `cur` of train."""
        sp = self.net.spec
        train = np.ascontiguousarray(train, dtype=np.float32)
        rc = self.o.L.orc_v1_gap(C.byref(self.p), self.net.maxNear, _fp(self.nNear),
                                 _fp(self.nbr), _fp(self.gapMat), _fp(self.gapDist),
                                 sp.speedOfThought, _fp(train), train.shape[0], cur,
                                 _fp(self.gap))
        if rc != 0:
            raise RuntimeError(f"orc_v1_gap returned {rc}")

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

    def run(self, nsteps: int, record: bool = True, lgn_only: bool = False):
        """nsteps steps.  lgn_only: the LGN stage alone (no V1 step); the
        linear filter output (convol, before the rectifier) of every step is
        kept in convol_hist, one array per frame period."""
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
                den = tm["denorm"]
                for fidx in range((it // npf) * den, (it // npf + 1) * den):
                    if self.batch is not None:               # frames [ntrial][nF][3][H][W]
                        self.Rring[fidx % self.ringF] = self.frame_response(
                            frames[:, fidx % frames.shape[1]], per_trial=True)
                    else:
                        self.Rring[fidx % self.ringF] = self.frame_response(
                            frames[fidx % frames.shape[0]])
                cv = np.zeros((npf, sp.nLGN), np.float32) if lgn_only else None
                lrn = self.learn
                self.o.L.orc_lgn_steps_learn(
                    C.byref(self.lp), _fp(t["TW"]), _fp(kC), _fp(kS), _fp(t["maxConvol"]),
                    _fp(self.lgn["c50"]), _fp(self.lgn["sharpness"]), _fp(self.lgn["logA"]),
                    _fp(self.lgn["logB"]), _fp(self.lTR), _fp(self.lNL), _fp(self.Rring), it,
                    npf, _fp(self.lgn_sInfo), _fp(self.lgn_fr), _fp(cv),
                    self.learn_p.tauLTP if lrn else 0.0, _fp(self.xVar) if lrn else None,
                    _fp(self.xStart) if lrn else None, _fp(self.xSpike) if lrn else None)
                if record:
                    self.lgn_hist.append(self.lgn_sInfo.copy())
                    self.fr_hist.append(self.lgn_fr.copy())
                if lgn_only:
                    self.convol_hist.append(cv)
            if lgn_only:
                self.it += 1
                continue
            out = np.zeros(sp.nV1, np.float32)
            lgn_row = np.ascontiguousarray(self.lgn_sInfo[k_in])
            rc = self.o.L.orc_v1_step(C.byref(self.p), _fp(st["v"]), _fp(st["w"]), _fp(st["tBack"]),
                                      _fp(st["gFF"]), _fp(st["hFF"]), _fp(st["gE"]), _fp(st["hE"]),
                                      _fp(st["gI"]), _fp(st["hI"]), _fp(self.tonic),
                                      _fp(self.nLGN_of), _fp(self.sLGN), _fp(self.idLGN), _fp(lgn_row),
                                      self.net.maxNear, _fp(self.nNear), _fp(self.nbr),
                                      _fp(self.conMat), _fp(self.ipre), _fp(self.npre), _fp(self.og),
                                      _fp(self.oh), it, _fp(out), None, 0 if self.delay else 1,
                                      _fp(self.gapS), _fp(self.gap), _fp(self.lastNoise))
            if rc != 0:
                raise RuntimeError(f"orc_v1_step returned {rc}")
            if self.delay:                  # recal_G_mat after compute_V, same step
                slot = it % self.depth
                self.train[slot] = out
                rc = self.o.L.orc_v1_rec_delay(C.byref(self.p), _fp(st["gE"]), _fp(st["hE"]),
                                               _fp(st["gI"]), _fp(st["hI"]), self.net.maxNear,
                                               _fp(self.nNear), _fp(self.nbr), _fp(self.conMat),
                                               _fp(self.delayMat), sp.speedOfThought,
                                               _fp(self.train), self.depth, slot)
                if rc != 0:
                    raise RuntimeError(f"orc_v1_rec_delay returned {rc}")
            if (self.delay and self.far_on) or self.far_only:   # recal_G_vec + sum_G
                slot = it % self.depth
                self.train[slot] = out
                off, pre, w, dist = self.far
                rc = self.o.L.orc_v1_rec_far(C.byref(self.p), _fp(st["gE"]), _fp(st["hE"]),
                                             _fp(st["gI"]), _fp(st["hI"]), _fp(off), _fp(pre),
                                             _fp(w), _fp(dist), sp.speedOfThought,
                                             _fp(self.train), self.depth, slot)
                if rc != 0:
                    raise RuntimeError(f"orc_v1_rec_far returned {rc}")
            if self.gap is not None:        # recal_G's gap sums, for the next step
                if self.delay:
                    self._gap(self.train, it % self.depth)
                else:
                    self._gap(out[None, :], 0)
            if self.learn:                  # the step's plasticity (end of compute_V)
                lv = self.lvar
                rc = self.o.L.orc_v1_learn_ff(
                    C.byref(self.p), C.byref(self.learn_p), _fp(self.nLGN_of), _fp(self.sLGN),
                    _fp(self.idLGN), _fp(lv["vLTD"]), _fp(lv["vTrip"]), _fp(lv["vAvg"]),
                    _fp(self.totalFF), _fp(self.totalFFinf), _fp(out), _fp(lgn_row),
                    _fp(np.ascontiguousarray(self.xStart[k_in])),
                    _fp(np.ascontiguousarray(self.xSpike[k_in])))
                if rc != 0:
                    raise RuntimeError(f"orc_v1_learn_ff returned {rc}")
            if record:
                self.sInfo_hist.append(out)
            self.it += 1

    def sInfo(self):
        return np.stack(self.sInfo_hist)

    def lgn_sinfo(self):
        return np.concatenate(self.lgn_hist)


# ---------------------------------------------------------------------------
# float64 LGN rate loss and gradient (orc_lgn64_*, oracle/vc_oracle_grad.c)

class _OrcLgn64(C.Structure):
    _fields_ = ([(n, C.c_uint32) for n in ("nLGN", "nS", "nKS", "interval", "iT0", "nRetrace",
                                          "npf", "nF", "T", "W", "H")]
                + [(n, C.c_float) for n in ("dt", "convolRatio", "frRatio")]
                + [(n, C.c_void_p) for n in ("SCx", "SCy", "SW", "coneType", "delay", "ratio",
                                            "nR", "nD", "maxConvol", "c50", "sharpness",
                                            "logA", "logB", "initFrame")])


def lgn_tables(o: Oracle, spec, lgn: dict, network_mod) -> dict:
    """orc_lgn_precompute for an LGN layer given as (spec, lgn dict)."""
    from types import SimpleNamespace
    tm = network_mod.timing(spec)
    x0, y0, nvd = network_mod.view_geometry(spec)
    q = o.abi.LgnSetup()
    q.nLGN, q.nSample1D, q.nKernelSample = spec.nLGN, spec.nSpatialSample1D, spec.nKernelSample
    q.kernelSampleInterval, q.iKernelSampleT0 = tm["kernelSampleInterval"], tm["iKernelSampleT0"]
    q.maxSample1D = network_mod.max_sample_1d(SimpleNamespace(spec=spec, lgn=lgn))
    q.dt, q.nsig, q.normViewDistance, q.x0, q.y0 = spec.dt, spec.nsig, nvd, x0, y0
    keep = []
    for k in ("polar", "ecc", "rw", "rh", "orient", "k", "tauR", "tauD", "delay", "ratio",
              "nR", "nD", "covariant"):
        a = np.ascontiguousarray(lgn[k], dtype=np.float32)
        keep.append(a)
        setattr(q, k, a.ctypes.data)
    nS = spec.nSpatialSample1D ** 2
    TW = np.zeros((spec.nLGN, 2, spec.nKernelSample), np.float32)
    SW = np.zeros(nS, np.float32)
    SCx = np.zeros((spec.nLGN, 2, nS), np.float32)
    SCy = np.zeros((spec.nLGN, 2, nS), np.float32)
    mc = np.zeros(spec.nLGN, np.float32)
    o.L.orc_lgn_precompute(C.byref(q), _fp(TW), _fp(SW), _fp(SCx), _fp(SCy), _fp(mc))
    return dict(TW=TW, SW=SW, SCx=SCx, SCy=SCy, maxConvol=mc)


class LgnGrad64:
    """The LGN rates of `nsteps` steps as a float64 function of (frames,
    theta), theta[c] = (kC, kS, tauR_C, tauD_C, tauR_S, tauD_S), with the
    loss mean((fr - target)^2) and its gradient."""

    def __init__(self, o: Oracle, spec, lgn: dict, tables: dict, nsteps: int, network_mod,
                 init_lms=(0.5, 0.5, 0.5)):
        tm = network_mod.timing(spec)
        npf = tm["ntPerFrame"]
        self.nF = -(-nsteps // npf)
        self.spec, self.T = spec, nsteps
        self.L = o.L
        vp = C.c_void_p
        self.L.orc_lgn64_loss.argtypes = [C.POINTER(_OrcLgn64), vp, vp, vp, vp, C.c_int,
                                          C.POINTER(C.c_long)]
        self.L.orc_lgn64_loss.restype = C.c_double
        self.L.orc_lgn64_grad.argtypes = [C.POINTER(_OrcLgn64), vp, vp, vp, vp, vp]
        self.L.orc_lgn64_rates.argtypes = [C.POINTER(_OrcLgn64), vp, vp, vp]
        f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
        self.keep = dict(
            SCx=f32(tables["SCx"]), SCy=f32(tables["SCy"]), SW=f32(tables["SW"]),
            maxConvol=f32(tables["maxConvol"]),
            coneType=np.ascontiguousarray(lgn["coneType"], dtype=np.uint32),
            **{k: f32(lgn[k]) for k in ("delay", "ratio", "nR", "nD", "c50", "sharpness",
                                        "logA", "logB")})
        init = np.empty((3, spec.height, spec.width), np.float64)
        for c in range(3):
            init[c] = np.float32(init_lms[c])
        self.keep["initFrame"] = init
        m = _OrcLgn64()
        m.nLGN, m.nS, m.nKS = spec.nLGN, spec.nSpatialSample1D ** 2, spec.nKernelSample
        m.interval, m.iT0 = tm["kernelSampleInterval"], tm["iKernelSampleT0"]
        m.nRetrace, m.npf, m.nF, m.T = tm["nRetrace"], npf, self.nF, nsteps
        m.W, m.H = spec.width, spec.height
        m.dt, m.convolRatio, m.frRatio = spec.dt, spec.convolRatio, spec.frRatioLGN
        for k, a in self.keep.items():
            setattr(m, k, a.ctypes.data)
        self.m = m
        self.theta0 = np.stack([lgn["k"][0], lgn["k"][1], lgn["tauR"][0], lgn["tauD"][0],
                                lgn["tauR"][1], lgn["tauD"][1]], 1).astype(np.float32)

    def _frames(self, frames):
        f = np.ascontiguousarray(frames[:self.nF], dtype=np.float64)
        assert f.shape[0] == self.nF, "need ceil(T / ntPerFrame) frames"
        return f

    def loss(self, frames, theta, target, cells=None):
        f = self._frames(frames)
        th = np.ascontiguousarray(theta, dtype=np.float64)
        tg = np.ascontiguousarray(target, dtype=np.float32)
        cl = None if cells is None else np.ascontiguousarray(cells, dtype=np.int32)
        pat = C.c_long(0)
        v = self.L.orc_lgn64_loss(C.byref(self.m), _fp(f), _fp(th), _fp(tg), _fp(cl),
                                  0 if cl is None else cl.size, C.byref(pat))
        return v, pat.value

    def rates(self, frames, theta):
        f = self._frames(frames)
        th = np.ascontiguousarray(theta, dtype=np.float64)
        fr = np.zeros((self.T, self.spec.nLGN), np.float64)
        self.L.orc_lgn64_rates(C.byref(self.m), _fp(f), _fp(th), _fp(fr))
        return fr

    def grad(self, frames, theta, target):
        f = self._frames(frames)
        th = np.ascontiguousarray(theta, dtype=np.float64)
        tg = np.ascontiguousarray(target, dtype=np.float32)
        df = np.zeros_like(f)
        dth = np.zeros_like(th)
        self.L.orc_lgn64_grad(C.byref(self.m), _fp(f), _fp(th), _fp(tg), _fp(df), _fp(dth))
        return df, dth

    def cells_touching(self, px, py, ch):
        """Cells any of whose samples read pixel (px, py) of channel ch."""
        sp = self.spec
        W, H = sp.width, sp.height
        xt = self.keep["SCx"].astype(np.float64) * W - 0.5
        yt = self.keep["SCy"].astype(np.float64) * H - 0.5
        i0 = np.floor(xt); j0 = np.floor(yt)
        hit = ((i0 == px) | (i0 + 1 == px)) & ((j0 == py) | (j0 + 1 == py))
        cone = self.keep["coneType"]                      # [2][nLGN]
        uses = {0: (0, 3, 4), 1: (1, 3, 4), 2: (2, 4)}[ch]
        ok = np.isin(cone, uses).any(0)
        return np.nonzero(hit.any(axis=(1, 2)) & ok)[0].astype(np.int32)


# ---------------------------------------------------------------------------
# float64 V1 LIF forward and tape gradient (orc_v1_lif64, oracle/vc_oracle_v1grad.c)

class V1Grad64:
    """The V1 steps of a Network (conductance LIF) in float64 as a function of
    its weights, with the reverse-mode tape gradient.  lgn [T][nLGN] is the
    LGN output of every step (an input); init is the state in snapshot rows."""

    def __init__(self, o: Oracle, net, params_mod, lgn: np.ndarray, init: np.ndarray):
        self.net, self.L = net, o.L
        sp = net.spec
        self.p = params_mod.v1_params_struct(net)
        vp, u32 = C.c_void_p, C.c_uint32
        self.L.orc_v1_lif64.argtypes = ([C.POINTER(o.abi.V1Params), u32, vp, vp, vp, u32, vp, vp,
                                         vp, vp, u32, u32, vp, vp, vp, vp, vp, vp, vp, vp,
                                         vp, vp, vp, vp])
        self.L.orc_v1_lif64.restype = C.c_int
        self.lgn = np.ascontiguousarray(lgn, dtype=np.float32)
        self.T = self.lgn.shape[0]
        self.init = np.ascontiguousarray(init, dtype=np.float32)
        self.tonic = np.ascontiguousarray(net.tonicDep, dtype=np.float32)
        self.wstate = np.zeros(sp.nV1, np.float32)
        self.nLGN_of = np.ascontiguousarray(net.nLGN_of, dtype=np.uint32)
        self.idLGN = np.ascontiguousarray(net.idLGN, dtype=np.uint32)
        self.nNear = np.ascontiguousarray(net.nNear, dtype=np.uint32)
        self.nbr = np.ascontiguousarray(net.nbr, dtype=np.uint32)
        self.conMat = np.ascontiguousarray(net.dense_conmat(), dtype=np.float64)
        self.sLGN = np.ascontiguousarray(net.sLGN, dtype=np.float64)

    def run(self, sLGN=None, conMat=None, pattern=None, gOut=None):
        """sInfo [T][nV1] (float64); with gOut also (dsLGN, dConMat)."""
        sp = self.net.spec
        sL = self.sLGN if sLGN is None else np.ascontiguousarray(sLGN, dtype=np.float64)
        cm = self.conMat if conMat is None else np.ascontiguousarray(conMat, dtype=np.float64)
        pat = None if pattern is None else np.ascontiguousarray(pattern, dtype=np.uint8)
        out = np.zeros((self.T, sp.nV1), np.float64)
        g = None if gOut is None else np.ascontiguousarray(gOut, dtype=np.float64)
        dsl = np.zeros_like(sL) if g is not None else None
        dcm = np.zeros_like(cm) if g is not None else None
        rc = self.L.orc_v1_lif64(C.byref(self.p), self.T, _fp(self.init), _fp(self.tonic),
                                 _fp(self.wstate), sL.shape[0], _fp(self.nLGN_of), _fp(self.idLGN),
                                 _fp(sL), _fp(self.lgn), self.lgn.shape[1], self.net.maxNear,
                                 _fp(self.nNear), _fp(self.nbr), _fp(cm), _fp(pat), _fp(out),
                                 _fp(g), _fp(dsl), _fp(dcm), None, None, None, None)
        if rc != 0:
            raise RuntimeError(f"orc_v1_lif64 returned {rc}")
        return out if g is None else (out, dsl, dcm)

    def tangent(self, pattern, uConMat=None, uSLGN=None, anchor=None):
        """Forward-mode derivative of every output [T][nV1] along the weight
        direction (uConMat dense like conMat, uSLGN like sLGN; None = 0) at
        the fixed spike pattern; with anchor [T][VC_SNAP_ROWS][nV1] (float32)
        the primal state follows that trajectory step by step."""
        sp = self.net.spec
        pat = np.ascontiguousarray(pattern, dtype=np.uint8)
        out = np.zeros((self.T, sp.nV1), np.float64)
        dOut = np.zeros((self.T, sp.nV1), np.float64)
        uc = None if uConMat is None else np.ascontiguousarray(uConMat, dtype=np.float64)
        us = None if uSLGN is None else np.ascontiguousarray(uSLGN, dtype=np.float64)
        an = None if anchor is None else np.ascontiguousarray(anchor, dtype=np.float32)
        if an is not None and an.shape != (self.T, 14, sp.nV1):
            raise ValueError(f"anchor must be [{self.T}][14][{sp.nV1}]")
        rc = self.L.orc_v1_lif64(C.byref(self.p), self.T, _fp(self.init), _fp(self.tonic),
                                 _fp(self.wstate), self.sLGN.shape[0], _fp(self.nLGN_of),
                                 _fp(self.idLGN), _fp(self.sLGN), _fp(self.lgn),
                                 self.lgn.shape[1], self.net.maxNear, _fp(self.nNear),
                                 _fp(self.nbr), _fp(self.conMat), _fp(pat), _fp(out), None, None,
                                 None, _fp(us), _fp(uc), _fp(an), _fp(dOut))
        if rc != 0:
            raise RuntimeError(f"orc_v1_lif64 returned {rc}")
        return dOut


def soft_bin_counts64(sInfo, dt, bin_steps, tau_s):
    """Spike counts per 25 ms-style bin, each spike weighted by a logistic-
    smoothed bin indicator sigma((t-lo)/tau_s) - sigma((t-hi)/tau_s) over its
    own bin and the two next to it; float64 twin of the torch loss.  Returns
    (counts [nBins][nV1], d counts / d sInfo as a function of the residual)."""
    T, N = sInfo.shape
    nb = T // bin_steps
    counts = np.zeros((nb, N))
    it, n = np.nonzero(sInfo >= 1.0)
    t = it * dt + (sInfo[it, n] - 1.0) * dt
    b0 = it // bin_steps
    parts = []
    for db in (-1, 0, 1):
        b = b0 + db
        ok = (b >= 0) & (b < nb)
        lo, hi = b * bin_steps * dt, (b + 1) * bin_steps * dt
        s_lo = 1.0 / (1.0 + np.exp(-(t - lo) / tau_s))
        s_hi = 1.0 / (1.0 + np.exp(-(t - hi) / tau_s))
        k = s_lo - s_hi
        dk = (s_lo * (1 - s_lo) - s_hi * (1 - s_hi)) / tau_s       # d k / d t
        np.add.at(counts, (b[ok], n[ok]), k[ok])
        parts.append((b, ok, dk))
    return counts, (it, n, parts)


def soft_count_loss64(sInfo, target, dt, bin_steps, tau_s):
    """mean((counts - target)^2) and d loss / d sInfo [T][nV1] (float64)."""
    counts, (it, n, parts) = soft_bin_counts64(sInfo, dt, bin_steps, tau_s)
    r = 2.0 * (counts - target) / counts.size
    g = np.zeros_like(sInfo)
    for b, ok, dk in parts:
        gg = np.zeros(it.size)
        gg[ok] = r[b[ok], n[ok]] * dk[ok] * dt                    # dt/dsInfo = dt
        np.add.at(g, (it, n), gg)
    return float(np.mean((counts - target) ** 2)), g
