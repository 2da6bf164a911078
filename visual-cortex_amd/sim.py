"""Python driver of the MI355X simulation step: owns every buffer as a torch
tensor on the GPU and calls libvc_hip.so through its C ABI.

The step order follows the reference's host loop (src/patch.cu:5966-6546):
a new stimulus frame enters every ntPerFrame steps; the LGN stage of a run of
steps only depends on the stimulus, so it is issued on its own stream for
exactly the steps of each V1 launch, one launch ahead (vc_lgn_steps), and the
spatial response of the frame that enters at the next period boundary is
computed in slices over the steps of the period before it
(vc_lgn_frame_response_cells), so every step carries the same share of the
LGN work; each V1 step (vc_v1_step) applies the previous step's recurrent
increments, updates every neuron and compacts its spikes.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import abi
from .abi import check, ptr
from .network import Network, max_sample_1d, timing, view_geometry
from .params import (has_initial_spikes, initial_state, lgn_params_struct,
                     reference_train_depth, spinit_struct, v1_params_struct)


class GpuSim:
    """One simulation of one network on one GPU."""

    def __init__(self, net: Network, device: str = "cuda", record: bool = True,
                 state: dict | None = None, ff_events: bool | None = None,
                 persistent: bool = True, dense_weights: bool | None = None,
                 tune: bool = False, noalloc: bool = False):
        self.lib = abi.load_library()
        self.net = net
        sp = net.spec
        self.dev = torch.device(device)
        self.stream = torch.cuda.current_stream(self.dev)
        self.tm = timing(sp)
        d = self.dev
        T = lambda a, dt=None: torch.as_tensor(np.ascontiguousarray(a), device=d) if dt is None \
            else torch.as_tensor(np.ascontiguousarray(a).astype(dt), device=d)
        # ---- V1 parameters and state
        self.p = v1_params_struct(net)
        st = initial_state(net) if state is None else state
        self.state = {k: T(v) for k, v in st.items()}
        self.tonic = T(net.tonicDep, np.float32)
        # gap junctions between inhibitory neurons: gapS (every inhibitory
        # neuron with a junction, as the reference's fast update kernel takes
        # it whether or not InhGap is on) and, with InhGap, the gap sums
        mI = sp.nblock * sp.nI
        self.gap_on = net.gap_on()
        self.gapS = T(net.gap_strength_sums()) if net.has_gap() else None
        self.gap = torch.zeros(mI, dtype=torch.float32, device=d) if self.gap_on else None
        if (self.gapS is not None or self.gap_on) and getattr(net, "batch", None) is not None:
            raise NotImplementedError("gap junctions with batched trials are not built")
        # the Ornstein-Uhlenbeck noise state, zeros at the start
        # (src/patch.cu:4086-4088)
        self.lastNoise = torch.zeros(sp.nV1, dtype=torch.float32, device=d)
        if any(x > 0 for x in sp.synFail) and (sp.noDelay == 0 or net.n_far() > 0):
            raise NotImplementedError("synaptic failure between V1 neurons (synFail) with "
                                      "conduction delays or far connections is not built")
        self.s = abi.V1State(*(ptr(self.state[k]) for k in
                               ("v", "w", "tBack", "gFF", "hFF", "gE", "hE", "gI", "hI")),
                             ptr(self.tonic), ptr(self.gapS), ptr(self.gap), ptr(self.lastNoise))
        # ---- FF synapses
        self.nLGN_of = T(net.nLGN_of.astype(np.uint32).view(np.int32))
        self.sLGN = T(net.sLGN, np.float32)
        self.idLGN = T(net.idLGN.astype(np.uint32).view(np.int32))
        self.ff = abi.FFConn(ptr(self.nLGN_of), ptr(self.sLGN), ptr(self.idLGN), net.sLGN.shape[0])
        # ---- event-driven LGN delivery (default when <= 32 inputs per neuron)
        if ff_events is None:
            ff_events = net.sLGN.shape[0] <= 32
        self.ev = None
        if ff_events:
            off, tgt = net.ff_reverse()
            self.revOff = T(off.view(np.int32))
            self.revTgt = T(tgt.view(np.int32) if tgt.size else np.zeros(1, np.int32))
            self.ffMask2 = [torch.zeros(self.tm["ntPerFrame"], sp.nV1, dtype=torch.int32, device=d)
                            for _ in range(2)]
            self.ev2 = [abi.FFEvents(ptr(self.revOff), ptr(self.revTgt), ptr(m), sp.nV1)
                        for m in self.ffMask2]
            self.ev = self.ev2[0]
        # ---- recurrent synapses (masked CSR)
        rm, ro, w = net.masked_csr_torch(d)
        self.rowMask = rm.contiguous()
        self.rowOff = ro.to(torch.int32).contiguous()          # u32 offsets, reinterpreted
        self.wrec = w.float().contiguous() if w.numel() else torch.zeros(1, device=d)
        # ---- batched trials: per-trial Philox keys and local ids
        self.batch = getattr(net, "batch", None)
        if self.batch is not None:
            B = self.batch
            self.bt = {k: T(np.ascontiguousarray(getattr(B, k)).astype(np.uint32).view(np.int32))
                       for k in ("blockTrial", "trialV1Off", "cellTrial", "trialLgnOff", "trialSeed")}
            self.p.blockTrial = ptr(self.bt["blockTrial"])
            self.p.trialV1Off = ptr(self.bt["trialV1Off"])
            self.p.trialSeed = ptr(self.bt["trialSeed"])
        self.nNear = T(net.nNear.astype(np.uint32).view(np.int32))
        self.nbr = T(net.nbr.astype(np.uint32).view(np.int32))
        # optional dense V1_conMat copy of the strengths (one gather round trip
        # fewer, but random rows of a multi-GiB matrix): off by default, the
        # masked CSR measured faster at the full-patch size
        self.wDense = self._dense_weights() if dense_weights else None
        self.rec = abi.RecConn(net.maxNear, ptr(self.nNear), ptr(self.nbr), ptr(self.rowMask),
                               ptr(self.rowOff), ptr(self.wrec), ptr(self.wDense))
        self.gap_lag = None
        if self.gap_on:
            if net.gap_far is not None and int(net.gap_far[0][-1]) > 0:
                raise NotImplementedError("far gap junctions (V1_gapVec) with InhGap are not "
                                          "built")
            mg, cnt, P, G = net.gap_ell()
            self.gap_t = [T(cnt.view(np.int32)), T(P.view(np.int32)), T(G)]
            self.rec.maxGap = mg
            self.rec.nGap, self.rec.gapPre, self.rec.gapG = (ptr(x) for x in self.gap_t)
            if getattr(sp, "noDelay", 1) == 0:
                self.gap_lag = T(net.gap_lags(sp.speedOfThought, sp.dt).view(np.int32))
        # ---- axonal delays (noDelay 0): spike-train ring of the last `depth`
        # step outputs and spike masks; every step is one vc_v1_step (no
        # undelayed increments) followed by vc_v1_rec_delay
        self.delay = getattr(sp, "noDelay", 1) == 0
        # far connections: delayed ones read the same ring (vc_v1_rec_far after
        # vc_v1_rec_delay); with noFarDelay 1 the reference's recal_G_vec adds
        # nothing (src/coredynamics.cu:219-220), so they are not run
        self.far_on = net.n_far() > 0 and net.far_delayed()
        # delayed far connections with undelayed near ones (noFarDelay 0, noDelay 1):
        # a ring for the far input only; each step is vc_v1_step, the step's own
        # recurrent increments (vc_v1_rec_apply, recal_G_mat_nd) and vc_v1_rec_far
        self.far_only = self.far_on and not self.delay
        if self.delay:
            self._setup_delay()
        elif self.far_only:
            self.speed = float(sp.speedOfThought)
            self._setup_far(1)
        # ---- plasticity of the LGN->V1 strengths (learning 3): one step at a
        # time, each vc_v1_step followed by vc_v1_learn_ff
        self.learn = getattr(sp, "learning", 0) == 3
        if getattr(sp, "learning", 0) not in (0, 3):
            raise NotImplementedError(f"learning = {sp.learning}: only learning = 3 is built")
        if self.learn:
            self._setup_learn()
            persistent = False
        # ---- spike buffers (double buffered)
        nT = sp.nV1 // 64
        self.masks = [torch.zeros(nT, dtype=torch.int64, device=d) for _ in range(2)]
        self.og = [torch.zeros(abi.VC_MAX_NG, sp.nV1, dtype=torch.float32, device=d) for _ in range(2)]
        self.oh = [torch.zeros(abi.VC_MAX_NG, sp.nV1, dtype=torch.float32, device=d) for _ in range(2)]
        self.spk = [abi.Spikes(ptr(self.masks[i]), ptr(self.og[i]), ptr(self.oh[i])) for i in range(2)]
        self.spk2 = (abi.Spikes * 2)(*self.spk)
        if self.gapS is not None or self.gap_on:
            persistent = False          # gap junctions need every step's voltages (vc_v1_run)
        if self.delay or self.far_only:
            zero = torch.zeros(nT, dtype=torch.int64, device=d)
            self.dog = torch.empty(abi.VC_MAX_NG, sp.nV1, dtype=torch.float32, device=d)
            self.doh = torch.empty(abi.VC_MAX_NG, sp.nV1, dtype=torch.float32, device=d)
            self._keep = zero
            self.spk_none = abi.Spikes(ptr(zero), ptr(self.dog), ptr(self.doh))
            self.ring_spk = [abi.Spikes(ptr(self.trainMask[k]), ptr(self.dog), ptr(self.doh))
                             for k in range(self.depth)]
            persistent = False
        # persistent multi-step launches (per-block step counters, time-out flag)
        self.persistent = persistent
        nbytes = int(self.lib.vc_v1_persistent_scratch_bytes(C.byref(self.p)))
        self.scratch = torch.zeros(nbytes // 4, dtype=torch.int32, device=d)
        self.epoch = 0                   # of the last persistent launch (1..65535, cycling)
        self.n_v1_launches = 0           # V1 step kernels dispatched by run()
        # persistent launches through the allocation-free entry point (its
        # records in the library's static array instead of self.scratch)
        self.noalloc = noalloc
        self.status = torch.zeros(1, dtype=torch.int32, device=d)
        # ---- streams (default priority) and the LGN pipeline state
        self.s_v1 = torch.cuda.current_stream(d)
        self.s_lgn = torch.cuda.Stream(device=d)
        self.lgn_evq = []                # (end step, event, on V1 stream) of issued LGN chunks
        self.slice_q = []                # frame shares queued by chunks on the V1 stream
        self.lgn_tail = None             # event after the LGN stream's latest work
        self.v1_done = [None, None]      # event: V1 steps reading buffer i finished
        self.lgn_to = 0                  # the LGN stage is issued for steps < lgn_to
        # frames < fr_frame have their spatial response in the ring; frame
        # fr_frame has its first fr_cells cells
        self.fr_frame, self.fr_cells = 0, 0
        # ---- LGN
        self._setup_lgn()
        self.it = 0
        self.cur = 0                 # index of the spike buffer holding the last step's spikes
        self.record = record
        self.sInfo_hist: list[torch.Tensor] = []
        self.lgn_hist: list[torch.Tensor] = []
        self.sInfo_step = torch.zeros(sp.nV1, dtype=torch.float32, device=d)
        self.sInfo_last = self.sInfo_step          # V1 output of the last step run (a view)
        # rand_spInit's initial spikes and the pre-loop recal_G of them (a run
        # that starts from a given state has none, as a restored one)
        self.init_row = self.state["v"]
        if state is None and has_initial_spikes(sp):
            self._initial_spikes()
        if self.gap_on:              # the pre-loop recal_G's gap sums (initial voltages)
            if self.delay:
                self._gap_after(self.train, (self.slot_off + self.depth - 1) % self.depth)
            else:
                self._gap_after(self.init_row, 0)
        self.s_lgn.wait_stream(self.s_v1)          # set-up work precedes the pipeline
        # launch shape of the one-step kernel (runs whose grid cannot be
        # resident at once): timed per workload when asked, else the default
        self.shape = None
        if tune:
            from .tune import tune_step
            self.shape = tune_step(self)

    def _setup_learn(self):
        """The plastic state: the post traces (zero at the start, as the
        reference's cudaMemset of learnVar), the homeostatic totals, the LGN
        cells' trace and the rule's constants."""
        from .params import learn_ff_struct, learn_ff_totals
        net, sp, d = self.net, self.net.spec, self.dev
        z = dict(dtype=torch.float32, device=d)
        self.lvar = {k: torch.zeros(sp.nV1, **z) for k in ("vLTD", "vTrip", "vAvg")}
        self.totalFF = self.totalFFinf = None
        if sp.applyHomeo:
            tot, inf = learn_ff_totals(net)
            self.totalFF = torch.as_tensor(tot, device=d)
            self.totalFFinf = torch.as_tensor(inf, device=d)
        self.learn_p = learn_ff_struct(net)
        self.learn_s = abi.LearnFFState(ptr(self.sLGN), ptr(self.lvar["vLTD"]),
                                        ptr(self.lvar["vTrip"]), ptr(self.lvar["vAvg"]),
                                        ptr(self.totalFF), ptr(self.totalFFinf))
        self.xVar = torch.zeros(sp.nLGN, **z)         # the LGN cells' LTP trace

    def _learn_step(self, buf: int, k_in: int, sInfo: torch.Tensor):
        """The plasticity of the step whose V1 outputs are `sInfo`, LGN row
        k_in of buffer buf (after that step's update, as the reference learns
        at the end of compute_V)."""
        nL = self.net.spec.nLGN
        o = k_in * nL * 4
        check(self.lib.vc_v1_learn_ff(self._hs(self.s_v1), C.byref(self.p), C.byref(self.learn_p),
                                      C.byref(self.ff), C.byref(self.learn_s), ptr(sInfo),
                                      ptr(self.lgn_sInfo2[buf]) + o, ptr(self.xStart2[buf]) + o,
                                      ptr(self.xSpike2[buf]) + o,
                                      ptr(self.lbits2[buf]) + k_in * self.nWords * 8, nL),
              "vc_v1_learn_ff")

    def learn_state_numpy(self) -> dict:
        """The plastic state on the host: sLGN [maxLGN][nV1], the post traces,
        the totals (with homeostasis) and the LGN trace at the end of the
        issued LGN steps."""
        torch.cuda.synchronize()
        out = {k: v.cpu().numpy() for k, v in self.lvar.items()}
        out["sLGN"] = self.sLGN.cpu().numpy()
        out["xVar"] = self.xVar.cpu().numpy()
        if self.totalFF is not None:
            out["totalFF"] = self.totalFF.cpu().numpy()
        return out

    def _initial_spikes(self):
        """rand_spInit (vc_v1_initial_spikes, src/coredynamics.cu:5-146) over
        the reference's trainDepth steps before the first, then the pre-loop
        recal_G (src/patch.cu:5687-5770), which gathers the step before the
        first: with axonal delays vc_v1_rec_delay at that step's ring slot;
        without, its spikes become the spike buffer the first step gathers."""
        net, sp, d = self.net, self.net.spec, self.dev
        if self.batch is not None:
            raise NotImplementedError("initial spikes (spE0/spI0) with batched trials are not "
                                      "built")
        if net.n_far() > 0 and net.far_delayed():
            raise NotImplementedError("initial spikes (spE0/spI0) with far connections are not "
                                      "built")
        D = reference_train_depth(net)
        q = spinit_struct(net, D)
        h = self._hs(self.s_v1)
        if self.delay:
            if D > self.depth:
                raise ValueError(f"spike-train ring of {self.depth} steps < trainDepth {D}")
            check(self.lib.vc_v1_initial_spikes(h, C.byref(self.p), C.byref(self.s),
                                                ptr(self.nLGN_of), C.byref(q), ptr(self.train),
                                                ptr(self.trainMask), self.depth, self.slot_off),
                  "vc_v1_initial_spikes")
            prev = (self.slot_off + self.depth - 1) % self.depth
            check(self.lib.vc_v1_rec_delay(h, C.byref(self.p), C.byref(self.s), C.byref(self.rec),
                                           ptr(self.syn), ptr(self.train), ptr(self.trainMask),
                                           self.depth, prev, self.window, ptr(self.status)),
                  "vc_v1_rec_delay")
        else:
            ring = torch.empty(D, sp.nV1, dtype=torch.float32, device=d)
            check(self.lib.vc_v1_initial_spikes(h, C.byref(self.p), C.byref(self.s),
                                                ptr(self.nLGN_of), C.byref(q), ptr(ring), None,
                                                D, 0), "vc_v1_initial_spikes")
            self.init_row = ring[D - 1]
            check(self.lib.vc_spikes_compact(h, C.byref(self.p), ptr(self.init_row),
                                             C.byref(self.spk[self.cur])), "vc_spikes_compact")
            self._init_ring = ring

    def _setup_delay(self):
        from .network import train_depth
        net, sp, d = self.net, self.net.spec, self.dev
        if net.pos is None:
            raise ValueError("axonal delays need the V1 positions (Network.pos)")
        nnz = net.n_rec_synapses_csr()
        if getattr(net, "rec_dist", None) is not None and net.csr is None:
            # distances read from the reference's V1_delayMat file
            self.dist = torch.as_tensor(net.rec_distances(), device=d)
        else:
            pos = torch.as_tensor(np.ascontiguousarray(net.pos, dtype=np.float64), device=d)
            self.dist = torch.zeros(max(nnz, 1), dtype=torch.float32, device=d)
            rec0 = abi.RecConn(net.maxNear, ptr(self.nNear), ptr(self.nbr), ptr(self.rowMask),
                               ptr(self.rowOff), ptr(self.wrec), None)
            check(self.lib.vc_v1_rec_distances(None, C.byref(self.p), C.byref(rec0), ptr(pos),
                                               ptr(self.dist)), "vc_v1_rec_distances")
        dmax = float(self.dist[:nnz].max().item()) if nnz else 0.0
        self.speed = float(sp.speedOfThought)
        # {w, rem, it2post, 0} per synapse (16-byte records)
        self.syn = torch.zeros(max(nnz, 1), 4, dtype=torch.int32, device=d)
        check(self.lib.vc_v1_rec_delay_table(None, nnz, ptr(self.wrec), ptr(self.dist),
                                             self.speed, sp.dt, ptr(self.syn)),
              "vc_v1_rec_delay_table")
        self.depth = train_depth(dmax, self.speed, sp.dt)
        self._setup_far(self.depth)

    def _setup_far(self, depth_near: int):
        """The far lists on the device (when delayed far input is on) and the
        spike-train ring, as long as the near and the far delays need."""
        from .network import train_depth
        net, sp, d = self.net, self.net.spec, self.dev
        self.depth = depth_near
        self.window = depth_near     # slots the near delays reach
        if self.far_on:
            mf, cnt, P, W, Dm = net.far_ell()
            self.far_t = [torch.as_tensor(np.ascontiguousarray(x), device=d)
                          for x in (cnt, P, W, Dm, net.far_pre_mask(), *net.far_out_index())]
            self.far_t.append(torch.zeros(sp.nV1 * ((mf + 63) // 64), dtype=torch.int64,
                                          device=d))
            self.far = abi.FarConn(mf, *(ptr(x) for x in self.far_t))
            fmax = float(net.far[3].max()) if net.far[3].size else 0.0
            self.depth = max(self.depth, train_depth(fmax, self.speed, sp.dt))
        if self.gap_lag is not None:         # the junctions' delays read the ring too
            self.depth = max(self.depth, int(self.gap_lag.max().item()) + 1)
        # every slot starts with the initial voltages (rand_spInit without
        # initial spikes, src/coredynamics.cu:119-133)
        self.train = self.state["v"][None, :].repeat(self.depth, 1).contiguous()
        self.trainMask = torch.zeros(self.depth, sp.nV1 // 64, dtype=torch.int64, device=d)
        self.slot_off = 0            # ring slot of step it is (it + slot_off) % depth

    def _gap_after(self, train: torch.Tensor, cur: int):
        """The gap sums after a step (recal_G's), from the step outputs in
        slot `cur` of `train` ([depth, nV1], or one row [nV1]), the junctions'
        delays applied when the ring is the delay path's."""
        depth = train.shape[0] if train.dim() == 2 else 1
        lag = ptr(self.gap_lag) if (self.gap_lag is not None and depth > 1) else None
        check(self.lib.vc_v1_gap(self._hs(self.s_v1), C.byref(self.p), C.byref(self.s),
                                 C.byref(self.rec), ptr(train), depth, cur, lag), "vc_v1_gap")

    def _delay_step(self, buf: int, k_in: int, it: int, sInfo=None, dep=None, gprev=None):
        """Step `it` with axonal delays: the update writes ring slot `slot`,
        then the delayed recurrent input arriving during the step is added
        (the reference's compute_V, then recal_G_mat).  Returns the slot."""
        sp = self.net.spec
        slot = (it + self.slot_off) % self.depth
        h = self._hs(self.s_v1)
        mptr = ptr(self.ffMask2[buf]) + k_in * sp.nV1 * 4 if self.ev is not None else None
        check(self.lib.vc_v1_step(h, C.byref(self.p), C.byref(self.s), C.byref(self.ff),
                                  C.byref(self.rec), C.byref(self.spk_none),
                                  C.byref(self.ring_spk[slot]),
                                  ptr(self.lgn_sInfo2[buf]) + k_in * sp.nLGN * 4, mptr, it,
                                  ptr(self.train[slot]), ptr(dep), ptr(gprev)), "vc_v1_step")
        check(self.lib.vc_v1_rec_delay(h, C.byref(self.p), C.byref(self.s), C.byref(self.rec),
                                       ptr(self.syn), ptr(self.train), ptr(self.trainMask),
                                       self.depth, slot, self.window, ptr(self.status)),
              "vc_v1_rec_delay")
        if self.far_on:
            check(self.lib.vc_v1_rec_far(h, C.byref(self.p), C.byref(self.s), C.byref(self.far),
                                         ptr(self.train), ptr(self.trainMask), self.depth,
                                         slot, self.speed, ptr(self.status)), "vc_v1_rec_far")
        if self.gap_on:
            self._gap_after(self.train, slot)
        if sInfo is not None:
            sInfo.copy_(self.train[slot])
        return slot

    def _far_only_step(self, buf: int, k_in: int, it: int, sInfo=None, dep=None, gprev=None):
        """Step `it` with undelayed near and delayed far connections: the update
        writes ring slot `slot`, the step's own recurrent increments are added
        (compute_V, then recal_G_mat_nd), then the far input (recal_G_vec +
        sum_G).  Returns the slot."""
        sp = self.net.spec
        slot = (it + self.slot_off) % self.depth
        h = self._hs(self.s_v1)
        mptr = ptr(self.ffMask2[buf]) + k_in * sp.nV1 * 4 if self.ev is not None else None
        check(self.lib.vc_v1_step(h, C.byref(self.p), C.byref(self.s), C.byref(self.ff),
                                  C.byref(self.rec), C.byref(self.spk_none),
                                  C.byref(self.ring_spk[slot]),
                                  ptr(self.lgn_sInfo2[buf]) + k_in * sp.nLGN * 4, mptr, it,
                                  ptr(self.train[slot]), ptr(dep), ptr(gprev)), "vc_v1_step")
        check(self.lib.vc_v1_rec_apply_at(h, C.byref(self.p), C.byref(self.s), C.byref(self.rec),
                                          C.byref(self.ring_spk[slot]), it), "vc_v1_rec_apply_at")
        check(self.lib.vc_v1_rec_far(h, C.byref(self.p), C.byref(self.s), C.byref(self.far),
                                     ptr(self.train), ptr(self.trainMask), self.depth, slot,
                                     self.speed, ptr(self.status)), "vc_v1_rec_far")
        if self.gap_on:
            self._gap_after(self.train[slot], 0)
        if sInfo is not None:
            sInfo.copy_(self.train[slot])
        return slot

    def _dense_weights(self) -> torch.Tensor:
        """The masked CSR expanded on the device to the dense layout
        [nblock][maxNear][bs pre][bs post] (flat)."""
        rm, ro, w = self.rowMask, self.rowOff.to(torch.int64) & 0xffffffff, self.wrec
        nS = rm.numel()
        dense = torch.zeros(nS * 64, dtype=torch.float32, device=self.dev)
        lanes = torch.arange(64, device=self.dev, dtype=torch.int64)
        ch = 1 << 20
        for a in range(0, nS, ch):
            b = min(nS, a + ch)
            bits = ((rm[a:b, None] >> lanes[None, :]) & 1).bool()
            seg, lane = torch.nonzero(bits, as_tuple=True)          # CSR order
            if seg.numel() == 0:
                continue
            o0 = int(ro[a])
            dense[(seg + a) * 64 + lane] = w[o0:o0 + seg.numel()]
        return dense

    # ------------------------------------------------------------------ LGN
    def _setup_lgn(self):
        net, sp, d = self.net, self.net.spec, self.dev
        L = net.lgn
        self.lgn_t = {k: torch.as_tensor(np.ascontiguousarray(L[k]), device=d)
                      for k in ("polar", "ecc", "rw", "rh", "orient", "k", "tauR", "tauD",
                                "delay", "ratio", "nR", "nD", "covariant", "c50", "sharpness",
                                "logA", "logB")}
        self.coneType = torch.as_tensor(L["coneType"].view(np.int32), device=d)
        x0, y0, nvd = view_geometry(sp)
        tm = self.tm
        q = abi.LgnSetup()
        q.nLGN, q.nSample1D, q.nKernelSample = sp.nLGN, sp.nSpatialSample1D, sp.nKernelSample
        q.kernelSampleInterval, q.iKernelSampleT0 = tm["kernelSampleInterval"], tm["iKernelSampleT0"]
        q.maxSample1D = max_sample_1d(net)
        q.dt, q.nsig, q.normViewDistance, q.x0, q.y0 = sp.dt, sp.nsig, nvd, x0, y0
        for k in ("polar", "ecc", "rw", "rh", "orient", "k", "tauR", "tauD", "delay", "ratio",
                  "nR", "nD", "covariant"):
            setattr(q, k, ptr(self.lgn_t[k]))
        self.lgn_setup = q
        nS = sp.nSpatialSample1D ** 2
        f32 = dict(dtype=torch.float32, device=d)
        self.TW = torch.zeros(sp.nLGN, 2, sp.nKernelSample, **f32)
        self.SW = torch.zeros(nS, **f32)
        self.SCx = torch.zeros(sp.nLGN, 2, nS, **f32)
        self.SCy = torch.zeros(sp.nLGN, 2, nS, **f32)
        self.maxConvol = torch.zeros(sp.nLGN, **f32)
        check(self.lib.vc_lgn_precompute(None, C.byref(q), ptr(self.TW), ptr(self.SW),
                                         ptr(self.SCx), ptr(self.SCy), ptr(self.maxConvol)),
              "vc_lgn_precompute")
        # the frames' spatial responses regenerate the sample coordinates from
        # the per-region geometry (64 bytes per cell instead of 16 per sample)
        self.geom = torch.zeros(sp.nLGN, 2, abi.VC_LGN_GEOM, **f32)
        check(self.lib.vc_lgn_sample_geometry(None, C.byref(q), ptr(self.geom)),
              "vc_lgn_sample_geometry")
        self.view = (float(x0), float(y0), float(nvd))
        self.kC = self.lgn_t["k"][0].contiguous()
        self.kS = self.lgn_t["k"][1].contiguous()
        self.lTR = torch.zeros(sp.nLGN, **f32)
        self.lNL = torch.zeros(sp.nLGN, **f32)
        # (leftTimeRate, lastNegLogRand) at the start of the last two issued
        # frame periods: the LGN stage runs ahead of V1, so a snapshot at a
        # period boundary takes the LGN state from here
        self.lgn_at = [torch.zeros(2, sp.nLGN, **f32) for _ in range(2)]
        self.lp = lgn_params_struct(net)
        if getattr(net, "batch", None) is not None:
            self.lp.cellTrial = ptr(self.bt["cellTrial"])
            self.lp.trialLgnOff = ptr(self.bt["trialLgnOff"])
            self.lp.trialSeed = ptr(self.bt["trialSeed"])
        check(self.lib.vc_lgn_init(None, C.byref(self.lp), ptr(self.lTR), ptr(self.lNL)), "vc_lgn_init")
        self.ls = abi.LgnState(ptr(self.TW), ptr(self.kC), ptr(self.kS), ptr(self.maxConvol),
                               ptr(self.lgn_t["c50"]), ptr(self.lgn_t["sharpness"]),
                               ptr(self.lgn_t["logA"]), ptr(self.lgn_t["logB"]),
                               ptr(self.lTR), ptr(self.lNL))
        # stimulus resident in HBM; pre-stimulus frame of the initial luminance
        self.frames = torch.as_tensor(net.frames, device=d)
        if getattr(net, "batch", None) is not None:      # [ntrial][nF][3][H][W]
            self.nF = self.frames.shape[1]
            self.trialFrameStride = int(self.frames[0].numel())
        else:
            self.nF = self.frames.shape[0]
            self.trialFrameStride = 0
        init = np.empty((3, sp.height, sp.width), np.float32)
        for c in range(3):
            init[c] = net.init_lms[c]
        self.init_frame = torch.as_tensor(init, device=d)
        # a period of ntPerFrame steps holds denorm frames; its steps reach back
        # over maxFrame - 1 frames before the first of them, and the next
        # period's denorm frames are built while they still run
        self.ringF = tm["maxFrame"] - 1 + 2 * tm["denorm"]
        self.lp.maxFrame = self.ringF
        self.Rring = torch.zeros(self.ringF, sp.nLGN, 2, **f32)
        self._frame_response(self.init_frame, 0)          # on the current stream
        for s in range(1, self.ringF):
            self.Rring[s].copy_(self.Rring[0])
        # double buffers: the LGN stage of frame period k+1 runs on its own
        # stream while the V1 steps of period k run
        self.lgn_sInfo2 = [torch.zeros(tm["ntPerFrame"], sp.nLGN, **f32) for _ in range(2)]
        if self.learn:                  # the LGN trace at each step's start and spike
            self.xStart2 = [torch.zeros(tm["ntPerFrame"], sp.nLGN, **f32) for _ in range(2)]
            self.xSpike2 = [torch.zeros(tm["ntPerFrame"], sp.nLGN, **f32) for _ in range(2)]
            self.nWords = (sp.nLGN + 63) // 64       # the LGN spike bitmap of each step
            self.lbits2 = [torch.zeros(tm["ntPerFrame"], self.nWords, dtype=torch.int64,
                                       device=self.dev) for _ in range(2)]
        self.lgn_fr2 = [torch.zeros(tm["ntPerFrame"], sp.nLGN, **f32) for _ in range(2)]
        self.lgn_sInfo = self.lgn_sInfo2[0]
        self.lgn_fr = self.lgn_fr2[0]

    def _frame_response(self, frame: torch.Tensor, slot: int):
        sp = self.net.spec
        check(self.lib.vc_lgn_frame_response_geom(None, ptr(frame), sp.width, sp.height,
                                                  sp.nLGN, 0, sp.nLGN, sp.nSpatialSample1D,
                                                  ptr(self.geom), *self.view, ptr(self.SW),
                                                  ptr(self.coneType), None, 0,
                                                  ptr(self.Rring[slot])),
              "vc_lgn_frame_response_geom")

    # ------------------------------------------------------------------ steps
    def _hs(self, stream) -> C.c_void_p:
        return C.c_void_p(stream.cuda_stream)

    def _frame_of(self, fidx: int):
        """(frame tensor, cellTrial pointer) of stimulus frame fidx (the file
        loops at its end, as the reference's reader does)."""
        if self.trialFrameStride:
            return self.frames[0, fidx % self.nF], ptr(self.bt["cellTrial"])
        return self.frames[fidx % self.nF], None

    def _frame_cells(self, stream, fidx: int, c0: int, c1: int):
        """Spatial responses of cells c0 .. c1-1 to frame fidx, into its ring slot."""
        if c1 <= c0:
            return
        sp = self.net.spec
        frame, ctr = self._frame_of(fidx)
        check(self.lib.vc_lgn_frame_response_geom(
            self._hs(stream), ptr(frame), sp.width, sp.height, sp.nLGN, c0, c1,
            sp.nSpatialSample1D, ptr(self.geom), *self.view, ptr(self.SW), ptr(self.coneType),
            ctr, self.trialFrameStride, ptr(self.Rring[fidx % self.ringF])),
            "vc_lgn_frame_response_geom")

    def _lgn_chunk(self, a: int, b: int, on_v1: bool = False):
        """Issue the LGN stage of steps a .. b-1 (inside one frame period k, a =
        lgn_to): the rest of frame k's spatial response if it is not complete,
        the temporal part and the spike draw of the steps into rows a..b-1
        (mod ntPerFrame) of buffer k % 2, and the share (b - a)/ntPerFrame of
        frame k+1's spatial response.  On the LGN stream, or with on_v1 on the
        V1 stream (for the V1 launch that follows at once: it then needs no
        wait on another stream, which can hold the host until the awaited work
        is done); the frame share is then queued for the LGN stream
        (_lgn_flush_slices, after the V1 launch)."""
        npf = self.tm["ntPerFrame"]
        sp = self.net.spec
        nL = sp.nLGN
        k, a_in = divmod(a, npf)
        b_in = b - k * npf
        if a != self.lgn_to or not 0 < b_in <= npf or b <= a:
            raise ValueError(f"LGN chunk {a}..{b} out of order (issued up to {self.lgn_to})")
        st = self.s_v1 if on_v1 else self.s_lgn
        if on_v1:
            # earlier LGN-stream work (frame shares): waited for only when it is
            # still running (a wait on another stream's event costs the host
            # tens of microseconds)
            if self.lgn_tail is not None and not self.lgn_tail.query():
                st.wait_event(self.lgn_tail)
            self.lgn_tail = None
        else:
            self._lgn_flush_slices()
        buf = k % 2
        if a_in == 0:
            if self.v1_done[buf] is not None and not on_v1:
                st.wait_event(self.v1_done[buf])       # V1 of period k-2 released it
            with torch.cuda.stream(st):
                self.lgn_at[buf][0].copy_(self.lTR)
                self.lgn_at[buf][1].copy_(self.lNL)
        D = self.tm["denorm"]
        for f, c0, c1 in self._frame_pieces((k + 1) * D * nL):   # period k's frames whole
            self._frame_cells(st, f, c0, c1)
        ev = None
        if self.ev is not None:
            e2 = self.ev2[buf]
            ev = C.byref(abi.FFEvents(e2.revOff, e2.revTgt, e2.mask + a_in * sp.nV1 * 4, e2.nV1))
        if self.learn:
            check(self.lib.vc_lgn_steps_learn(self._hs(st), C.byref(self.lp), C.byref(self.ls),
                                              ptr(self.Rring), a, b - a,
                                              ptr(self.lgn_sInfo2[buf]) + a_in * nL * 4,
                                              ptr(self.lgn_fr2[buf]) + a_in * nL * 4, ev,
                                              self.learn_p.tauLTP, ptr(self.xVar),
                                              ptr(self.xStart2[buf]) + a_in * nL * 4,
                                              ptr(self.xSpike2[buf]) + a_in * nL * 4,
                                              ptr(self.lbits2[buf]) + a_in * self.nWords * 8),
                  "vc_lgn_steps_learn")
        else:
            check(self.lib.vc_lgn_steps(self._hs(st), C.byref(self.lp), C.byref(self.ls),
                                        ptr(self.Rring), a, b - a,
                                        ptr(self.lgn_sInfo2[buf]) + a_in * nL * 4,
                                        ptr(self.lgn_fr2[buf]) + a_in * nL * 4, ev),
                  "vc_lgn_steps")
        # the V1 steps wait for this point only: the slice of the next frame
        # below is read by the next period's LGN steps, on the LGN stream
        e = torch.cuda.Event()
        e.record(st)
        self.lgn_evq.append((b, e, on_v1))
        self.lgn_to = b
        # the share b_in / npf of the next period's frames
        for f, c0, c1 in self._frame_pieces((k + 1) * D * nL + b_in * D * nL // npf):
            if on_v1:
                self.slice_q.append((f, c0, c1, e))
            else:
                self._frame_cells(st, f, c0, c1)
        if not on_v1:
            self._lgn_mark_tail()

    def _frame_pieces(self, end: int):
        """The (frame, c0, c1) pieces of spatial-response work from the current
        position (frame fr_frame, its first fr_cells cells done) up to the
        position `end` = frame * nLGN + cells, in order; advances the position."""
        nL = self.net.spec.nLGN
        out = []
        while self.fr_frame * nL + self.fr_cells < end:
            f, c0 = self.fr_frame, self.fr_cells
            c1 = min(nL, end - f * nL)
            out.append((f, c0, c1))
            if c1 == nL:
                self.fr_frame, self.fr_cells = f + 1, 0
            else:
                self.fr_cells = c1
        return out

    def _lgn_mark_tail(self):
        """An event after the LGN stream's latest work (see _lgn_chunk)."""
        self.lgn_tail = torch.cuda.Event()
        self.lgn_tail.record(self.s_lgn)

    def _lgn_flush_slices(self):
        """Issue the frame shares queued by chunks on the V1 stream, on the LGN
        stream after those chunks."""
        if not self.slice_q:
            return
        for f, c0, c1, e in self.slice_q:
            self.s_lgn.wait_event(e)
            self._frame_cells(self.s_lgn, f, c0, c1)
        self.slice_q = []
        self._lgn_mark_tail()

    def _lgn_period(self, k: int):
        """Issue the LGN stage of the whole frame period k (it must be next)."""
        npf = self.tm["ntPerFrame"]
        self._lgn_chunk(k * npf, (k + 1) * npf)

    def _lgn_issue_to(self, end: int, on_v1: bool = False):
        """Issue the LGN stage up to step `end` in chunks ending at period
        boundaries (on the V1 stream with on_v1, see _lgn_chunk)."""
        npf = self.tm["ntPerFrame"]
        while self.lgn_to < end:
            a = self.lgn_to
            self._lgn_chunk(a, min(end, (a // npf + 1) * npf), on_v1)

    def _lgn_wait(self, stream, end: int):
        """Make `stream` wait for the LGN chunk that covers step end-1 (nothing
        to wait for when that chunk was issued on the V1 stream itself)."""
        while self.lgn_evq and self.lgn_evq[0][0] < end:
            self.lgn_evq.pop(0)
        if not self.lgn_evq:
            raise RuntimeError(f"LGN stage of step {end - 1} not issued")
        b, e, on_v1 = self.lgn_evq[0]
        if not (on_v1 and stream is self.s_v1):
            stream.wait_event(e)

    def run(self, nsteps: int, out: torch.Tensor | None = None, events: list | None = None):
        """Advance `nsteps` steps.  V1 sInfo of step i goes to out[i] if `out`
        ([nsteps, nV1] float32 on the device) is given, else to the history
        when recording, else to a scratch row.  If `events` is a list, a
        (start, end, n) triple of timing events bracketing each run of n V1
        step launches is appended (on the V1 stream).  The LGN stage of each V1
        launch's steps is issued on its own stream one launch ahead, within
        the call: the call issues the LGN work of exactly its own steps (when
        recording, of whole frame periods, which the LGN history keeps)."""
        sp = self.net.spec
        npf = self.tm["ntPerFrame"]
        if out is None and self.record:
            out = torch.empty(nsteps, sp.nV1, dtype=torch.float32, device=self.dev)
            self.sInfo_hist.append(out)
        end_all = self.it + nsteps

        def lgn_end(a):                # LGN steps a V1 launch starting at a needs
            e = min(end_all, (a // npf + 1) * npf)
            return (a // npf + 1) * npf if self.record else e

        done = 0
        sv = self.s_v1
        while done < nsteps:
            it = self.it
            k = it // npf
            k_in = it % npf
            n = min(npf - k_in, nsteps - done)
            # this launch's LGN steps, if not issued ahead: on the V1 stream
            self._lgn_issue_to(lgn_end(it), on_v1=True)
            buf = k % 2
            self._lgn_wait(sv, it + n if not self.record else (k + 1) * npf)
            if k_in == 0 and self.record:
                self.lgn_hist.append(self.lgn_sInfo2[buf].clone())
            if out is not None:
                optr, ostride = ptr(out) + done * sp.nV1 * 4, sp.nV1
                self.sInfo_last = out[done + min(npf - k_in, nsteps - done) - 1]
            else:
                optr, ostride = ptr(self.sInfo_step), 0
                self.sInfo_last = self.sInfo_step
            mptr = ptr(self.ffMask2[buf]) + k_in * sp.nV1 * 4 if self.ev is not None else None
            if events is not None:
                e0 = torch.cuda.Event(enable_timing=True)
                e0.record(sv)
            if self.learn:
                # one step at a time, each followed by its plasticity
                for i in range(n):
                    row = out[done + i] if out is not None else None
                    if self.delay:
                        slot = self._delay_step(buf, k_in + i, it + i, sInfo=row)
                        srow = self.train[slot]
                    elif self.far_only:
                        slot = self._far_only_step(buf, k_in + i, it + i, sInfo=row)
                        srow = self.train[slot]
                    else:
                        srow = row if row is not None else self.sInfo_step
                        m1 = mptr + i * sp.nV1 * 4 if mptr is not None else None
                        check(self.lib.vc_v1_run(
                            self._hs(sv), C.byref(self.p), C.byref(self.s), C.byref(self.ff),
                            C.byref(self.rec), self.spk2, self.cur,
                            ptr(self.lgn_sInfo2[buf]) + (k_in + i) * sp.nLGN * 4, sp.nLGN, m1,
                            sp.nV1, it + i, 1, ptr(srow), 0), "vc_v1_run")
                        self.cur ^= 1
                    self._learn_step(buf, k_in + i, srow)
                self.sInfo_last = srow
                self.n_v1_launches += n
            elif self.delay and self.gap_on:
                # one step at a time: each step's gap sums follow it
                for i in range(n):
                    row = out[done + i] if out is not None else None
                    slot = self._delay_step(buf, k_in + i, it + i, sInfo=row)
                self.sInfo_last = out[done + n - 1] if out is not None else self.train[slot]
            elif self.delay:
                # the n steps in one native call (vc_v1_delay_run)
                slot0 = (it + self.slot_off) % self.depth
                check(self.lib.vc_v1_delay_run(
                    self._hs(sv), C.byref(self.p), C.byref(self.s), C.byref(self.ff),
                    C.byref(self.rec), ptr(self.syn), C.byref(self.far) if self.far_on else None,
                    ptr(self.train), ptr(self.trainMask), self.depth, self.window, slot0,
                    ptr(self.lgn_sInfo2[buf]) + k_in * sp.nLGN * 4, sp.nLGN, mptr, sp.nV1,
                    it, n, self.speed, ptr(self._keep), ptr(self.dog), ptr(self.doh),
                    optr if out is not None else None, sp.nV1, ptr(self.status)),
                    "vc_v1_delay_run")
                slot = (it + n - 1 + self.slot_off) % self.depth
                self.sInfo_last = out[done + n - 1] if out is not None else self.train[slot]
            elif self.far_only and not self.gap_on:
                # the n steps in one native call, undelayed near connections (syn NULL)
                slot0 = (it + self.slot_off) % self.depth
                check(self.lib.vc_v1_delay_run(
                    self._hs(sv), C.byref(self.p), C.byref(self.s), C.byref(self.ff),
                    C.byref(self.rec), None, C.byref(self.far), ptr(self.train),
                    ptr(self.trainMask), self.depth, self.window, slot0,
                    ptr(self.lgn_sInfo2[buf]) + k_in * sp.nLGN * 4, sp.nLGN, mptr, sp.nV1,
                    it, n, self.speed, ptr(self._keep), ptr(self.dog), ptr(self.doh),
                    optr if out is not None else None, sp.nV1, ptr(self.status)),
                    "vc_v1_delay_run")
                slot = (it + n - 1 + self.slot_off) % self.depth
                self.sInfo_last = out[done + n - 1] if out is not None else self.train[slot]
                self.n_v1_launches += n
            elif self.far_only:
                for i in range(n):
                    row = out[done + i] if out is not None else None
                    slot = self._far_only_step(buf, k_in + i, it + i, sInfo=row)
                self.sInfo_last = out[done + n - 1] if out is not None else self.train[slot]
            if self.delay or self.far_only or self.learn:
                if events is not None:
                    e1 = torch.cuda.Event(enable_timing=True)
                    e1.record(sv)
                    events.append((e0, e1, n))
                self.it += n
                done += n
                if self.it % npf == 0:
                    e = torch.cuda.Event()
                    e.record(sv)
                    self.v1_done[buf] = e
                self._lgn_ahead(end_all, lgn_end)
                continue
            args = (self._hs(sv), C.byref(self.p), C.byref(self.s), C.byref(self.ff),
                    C.byref(self.rec), self.spk2, self.cur,
                    ptr(self.lgn_sInfo2[buf]) + k_in * sp.nLGN * 4, sp.nLGN, mptr, sp.nV1, it, n,
                    optr, ostride)
            rc = abi.VC_E_UNSUPPORTED
            if self.persistent and self.noalloc:
                rc = self.lib.vc_v1_run_persistent_noalloc(*args, ptr(self.status))
                if rc == abi.VC_E_UNSUPPORTED:          # grid cannot be co-resident
                    self.persistent = False
                else:
                    check(rc, "vc_v1_run_persistent_noalloc")
            elif self.persistent:
                self.epoch = self.epoch % 65535 + 1
                rc = self.lib.vc_v1_run_persistent(*args, ptr(self.scratch), self.epoch,
                                                   ptr(self.status))
                if rc == abi.VC_E_UNSUPPORTED:          # grid cannot be co-resident
                    self.persistent = False
                else:
                    check(rc, "vc_v1_run_persistent")
            if not self.persistent:
                check(self.lib.vc_v1_run(*args), "vc_v1_run")
            self.n_v1_launches += 1 if self.persistent else n
            if events is not None:
                e1 = torch.cuda.Event(enable_timing=True)
                e1.record(sv)
                events.append((e0, e1, n))
            self.cur ^= (n & 1)
            self.it += n
            done += n
            if self.it % npf == 0:
                e = torch.cuda.Event()
                e.record(sv)
                self.v1_done[buf] = e
            self._lgn_ahead(end_all, lgn_end)

    def _lgn_ahead(self, end_all: int, lgn_end):
        """After a V1 launch: the frame shares its chunk queued, then the LGN
        stage of the next launch of the call, on the LGN stream (it overlaps
        this launch)."""
        self._lgn_flush_slices()
        if self.it < end_all:
            self._lgn_issue_to(lgn_end(self.it))

    def run_with_outputs(self, nsteps: int, paths: dict, outputs: dict, on_step=None):
        """Run `nsteps` steps one launch at a time and write the reference's
        output files: rawData (sInfo, dep, w, v, gFF, hFF, gE, gI, hE, hI, gap
        per step, src/patch.cu:4688-4716, 6139-6164, src/stats.cu:38-97),
        LGN_fr and LGN_sp.  The cortical g/h of step t are known after the
        recurrent increments of its spikes, i.e. at step t+1 (or finalize).
        on_step(self, i), if given, runs after step i (0-based) of the call,
        e.g. to write a snapshot."""
        from . import formats as F
        sp = self.net.spec
        nV1, npf = sp.nV1, self.tm["ntPerFrame"]
        ngFF, ngE, ngI = self.p.ngFF, self.p.ngE, self.p.ngI
        mI = sp.nblock * sp.nI
        hW = bool(outputs.get("hWrite", True))
        raw = F.RawDataWriter(paths["rawData"], sp.dt, nsteps, nV1, sp.iModel, mI, hW, ngFF, ngE,
                              ngI) if outputs.get("rawData") else None
        fr_w = F.StepWriter(paths["lgn_fr"], nsteps, sp.nLGN) if outputs.get("lgn_fr") else None
        sp_w = F.StepWriter(paths["lgn_sp"]) if outputs.get("lgn_sp") else None
        dev = self.dev
        sInfo = torch.empty(nV1, dtype=torch.float32, device=dev)
        dep = torch.empty(nV1, dtype=torch.float32, device=dev)
        gprev = torch.empty(2 * (ngE + ngI), nV1, dtype=torch.float32, device=dev)
        held = None
        st = self.state

        def g_parts(g):
            return g[:ngE], g[ngE:ngE + ngI], g[ngE + ngI:2 * ngE + ngI], g[2 * ngE + ngI:]

        for i in range(nsteps):
            it = self.it
            k, k_in = divmod(it, npf)
            self._lgn_issue_to((k + 2) * npf)           # this period and the next
            buf = k % 2
            self._lgn_wait(self.s_v1, (k + 1) * npf)
            if k_in == 0:
                n_here = min(npf, nsteps - i)
                if fr_w:
                    fr_w.write(self.lgn_fr2[buf][:n_here].cpu().numpy())
                if sp_w:
                    sp_w.write(self.lgn_sInfo2[buf][:n_here].cpu().numpy())
            if self.delay:
                self._delay_step(buf, k_in, it, sInfo=sInfo, dep=dep, gprev=gprev)
            elif self.far_only:
                self._far_only_step(buf, k_in, it, sInfo=sInfo, dep=dep, gprev=gprev)
            else:
                nxt = 1 - self.cur
                mptr = ptr(self.ffMask2[buf]) + k_in * nV1 * 4 if self.ev is not None else None
                check(self.lib.vc_v1_step(self._hs(self.s_v1), C.byref(self.p), C.byref(self.s),
                                          C.byref(self.ff), C.byref(self.rec),
                                          C.byref(self.spk[self.cur]), C.byref(self.spk[nxt]),
                                          ptr(self.lgn_sInfo2[buf]) + k_in * sp.nLGN * 4, mptr,
                                          it, ptr(sInfo), ptr(dep), ptr(gprev)), "vc_v1_step")
                if self.gap_on:
                    self._gap_after(sInfo, 0)
                self.cur = nxt
            if self.learn:
                self._learn_step(buf, k_in, sInfo)
            self.it += 1
            self.sInfo_last = sInfo
            if raw:
                if held is not None:
                    gE, gI, hE, hI = (x.cpu().numpy() for x in g_parts(gprev))
                    raw.write_step(*held[:6], gE, gI, hE, hI, gap=held[6])
                # the gap sums after the step (recal_G's), zero without InhGap
                held = (sInfo.cpu().numpy(), dep.cpu().numpy(), st["w"].cpu().numpy(),
                        st["v"].cpu().numpy(), st["gFF"][:ngFF].cpu().numpy(),
                        st["hFF"][:ngFF].cpu().numpy(),
                        self.gap.cpu().numpy() if self.gap is not None else None)
            if self.it % npf == 0:
                e = torch.cuda.Event()
                e.record(self.s_v1)
                self.v1_done[buf] = e
            if on_step is not None:
                on_step(self, i)
        if raw:
            self.finalize()
            torch.cuda.synchronize()
            raw.write_step(*held[:6], st["gE"].cpu().numpy(), st["gI"].cpu().numpy(),
                           st["hE"].cpu().numpy(), st["hI"].cpu().numpy(), gap=held[6])
            raw.close()
        for w in (fr_w, sp_w):
            if w:
                w.close()

    def seek(self, it0: int):
        """Restart the step pipeline at step it0 (a multiple of ntPerFrame):
        no pending spikes, the frames' spatial responses that the temporal
        kernels of the period reach back to recomputed from the resident
        stimulus, and the LGN stage due from period it0 / ntPerFrame.  The V1
        and LGN state are left as they are (snapshot.restore sets them)."""
        tm, sp = self.tm, self.net.spec
        npf, maxF = tm["ntPerFrame"], tm["maxFrame"]
        if it0 % npf != 0:
            raise ValueError(f"a run can restart at a frame period boundary only "
                             f"(step {it0}, {npf} steps per frame)")
        torch.cuda.synchronize()
        K = it0 // npf
        KF = K * tm["denorm"]                    # frames entered before step it0
        self._frame_response(self.init_frame, 0)
        for slot in range(1, self.ringF):
            self.Rring[slot].copy_(self.Rring[0])
        cur = torch.cuda.current_stream(self.dev)
        for fidx in range(max(0, KF - maxF + 1), KF):
            self._frame_cells(cur, fidx, 0, sp.nLGN)
        if self.ev is not None:
            for m in self.ffMask2:
                m.zero_()
        for i in range(2):
            self.masks[i].zero_()
            self.og[i].zero_()
            self.oh[i].zero_()
        self.cur = 0
        self.it = it0
        self.lgn_to = it0
        self.fr_frame, self.fr_cells = KF, 0
        self.lgn_evq = []
        self.slice_q = []
        self.lgn_tail = None
        self.v1_done = [None, None]
        torch.cuda.synchronize()
        self.s_lgn.wait_stream(self.s_v1)

    def finalize(self):
        """Apply the last step's pending recurrent increments (the reference
        applies them inside the same step, in recal_G).  With axonal delays,
        or with far connections, nothing is pending: each step ends with its
        recurrent input."""
        if self.delay or self.far_only:
            return
        # (the pending spikes are those of step it - 1, which keys their
        # synaptic-failure draws)
        check(self.lib.vc_v1_rec_apply_at(None, C.byref(self.p), C.byref(self.s),
                                          C.byref(self.rec), C.byref(self.spk[self.cur]),
                                          (self.it - 1) & 0xffffffff), "vc_v1_rec_apply_at")
        # the increments are now in the state: clear the pending list
        check(self.lib.vc_spikes_clear(None, self.net.spec.nV1, C.byref(self.spk[self.cur])),
              "vc_spikes_clear")

    def check(self):
        """Wait for the device and raise if a persistent launch timed out
        waiting for a neighbour block (its results are then invalid)."""
        torch.cuda.synchronize()
        if int(self.status.item()) != 0:
            if self.delay or self.far_only:
                if int(self.status.item()) & 2:
                    raise abi.VCError("a far connection's index entry is out of range")
                raise abi.VCError("a synapse's delay does not fit the spike-train ring")
            raise abi.VCError("persistent V1 launch timed out waiting for neighbour blocks")

    def sInfo(self) -> np.ndarray:
        self.check()
        return torch.cat(self.sInfo_hist).cpu().numpy()

    def lgn_sinfo(self) -> np.ndarray:
        self.check()
        return torch.cat(self.lgn_hist).cpu().numpy()

    def state_numpy(self) -> dict:
        self.check()
        return {k: v.cpu().numpy() for k, v in self.state.items()}
