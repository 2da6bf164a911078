"""The reference simulator's configuration interface: `patch_fast -c <cfg>
[-s seed] [-p print]` with a Boost program_options config file
(src/patch.cu:172-405).

Semantics kept from Boost program_options as the reference uses it:
  * `key = value` lines, `#` starts a comment, blank lines ignored,
    `[section]` prefixes the following keys with `section.`;
  * vector options take comma-separated values and a repeated key appends
    (e.g. two `pFF = a, b` lines give [a, b, a, b]; src/patchV1.cfg:127-128);
  * the command line is stored first, so it wins over the file;
  * file options are prefixed with inputFolder / resourceFolder /
    outputFolder unless they start with `!` (src/patch.cu:407-586), and the
    suffix options become `-<suffix>.bin` (src/patch.cu:587-605).
"""
from __future__ import annotations

from dataclasses import replace
from pathlib import Path

# file option -> (default name, folder option, suffix option)
FILES = {
    "fStimulus": ("stimulus.bin", "resourceFolder", None),
    "fV1_allpos": ("V1_allpos", "resourceFolder", "res_suffix"),
    "fLGN_vpos": ("LGN_vpos", "resourceFolder", "res_suffix"),
    "fConnectome_cfg": ("connectome_cfg", "inputFolder", "conV1_suffix"),
    "fV1_conMat": ("V1_conMat", "inputFolder", "conV1_suffix"),
    "fV1_delayMat": ("V1_delayMat", "inputFolder", "conV1_suffix"),   # src/patch.cu:352, 467-470
    "fNeighborBlock": ("neighborBlock", "inputFolder", "conV1_suffix"),
    "fV1_vec": ("V1_vec", "inputFolder", "conV1_suffix"),             # src/patch.cu:355, 482-486
    "fV1_gapMat": ("V1_gapMat", "inputFolder", "conV1_suffix"),       # src/patch.cu:353, 472-476
    "fV1_gapVec": ("V1_gapVec", "inputFolder", "conV1_suffix"),       # src/patch.cu:356, 487-491
    "fLGN_V1_ID": ("LGN_V1_idList", "inputFolder", "conLGN_suffix"),
    "fLGN_V1_s": ("LGN_V1_sList", "inputFolder", "conLGN_suffix"),
    "fLGN": ("LGN", "outputFolder", "output_suffix"),
    "fRawData": ("rawData", "outputFolder", "output_suffix"),
    "fLGN_fr": ("LGN_fr", "outputFolder", "output_suffix"),
    "fLGN_sp": ("LGN_sp", "outputFolder", "output_suffix"),
    "f_sLGN": ("sLGN", "outputFolder", "output_suffix"),                   # src/patch.cu:367, 4755
    "fLGN_surfaceID": ("LGN_surfaceID", "resourceFolder", "res_suffix"),   # src/patch.cu:2573
    "fConStats": ("conStats", "inputFolder", "conV1_suffix"),              # src/patch.cu:354, 477
    "fV1_feature": ("V1_feature", "resourceFolder", "res_suffix"),         # src/patch.cu:350, 422
    "fPatchV1_cfg": ("patchV1_cfg", "outputFolder", "output_suffix"),      # src/patch.cu:362, 507
}

# options that hold vectors (everything else is a scalar; the last value wins)
VECTOR_KEYS = {
    "sRatioLGN", "sRatioV1", "synPerConFF", "synPerCon", "gapRatio", "riseTimeFF", "riseTimeE",
    "riseTimeI", "decayTimeFF", "decayTimeE", "decayTimeI", "v0", "w0", "gFF0", "gE0", "gI0",
    "pFF", "pE", "pI", "vR", "vThres", "gL", "C", "tRef", "vT", "deltaT", "tau_w", "a", "b",
    "nTypeHierarchy", "spE0", "spI0", "noisyDep", "tonicDep", "minTonicRatio", "synFailFF",
    "synFail", "boostOri", "iOri", "bgStrength", "pBG",
    # plasticity (src/patch.cu:239-258)
    "A_LGN", "r_LTD", "A_V1", "A_Q", "tauQ", "tauLTP", "tauLTD", "tauTrip", "tauAvg",
    "targetFR", "gmaxLGN", "gmaxE", "gmaxQ", "gminLGN", "gminE", "gminQ",
}


def parse_cfg_text(text: str, into: dict | None = None) -> dict:
    """Parse cfg text into {key: [str, ...]} (repeated keys append)."""
    out = {} if into is None else into
    section = ""
    for raw in text.splitlines():
        line = raw.split("#", 1)[0].strip()
        if not line:
            continue
        if line.startswith("[") and line.endswith("]"):
            section = line[1:-1].strip() + "."
            continue
        if "=" not in line:
            raise ValueError(f"cfg line without '=': {raw!r}")
        k, v = line.split("=", 1)
        k = section + k.strip()
        vals = [x.strip() for x in v.split(",")] if k in VECTOR_KEYS else [v.strip()]
        out.setdefault(k, []).extend(x for x in vals if x != "")
    return out


class Config:
    """Command-line options first (they win), then the cfg file."""

    def __init__(self, cfg_text: str = "", cmdline: dict | None = None):
        self.file = parse_cfg_text(cfg_text)
        self.cmd = {k: [str(v)] for k, v in (cmdline or {}).items()}

    @classmethod
    def from_file(cls, path, cmdline: dict | None = None):
        return cls(Path(path).read_text(), cmdline)

    def has(self, key):
        return key in self.cmd or key in self.file

    def raw(self, key):
        if key in self.cmd:
            return self.cmd[key]
        return self.file.get(key)

    def str(self, key, default=None):
        r = self.raw(key)
        return default if r is None else r[-1]

    def float(self, key, default=None):
        r = self.raw(key)
        return default if r is None else float(r[-1])

    def int(self, key, default=None):
        r = self.raw(key)
        return default if r is None else int(float(r[-1]))

    def bool(self, key, default=None):
        r = self.raw(key)
        if r is None:
            return default
        s = r[-1].strip().lower()
        if s in ("1", "true", "yes", "on"):
            return True
        if s in ("0", "false", "no", "off"):
            return False
        raise ValueError(f"{key}: not a bool: {r[-1]!r}")

    def floats(self, key, default=None):
        r = self.raw(key)
        return default if r is None else [float(x) for x in r]

    def path(self, key) -> str:
        """Resolved path of a file option (folder prefix, '!' rule, suffix)."""
        name, folder_key, suffix_key = FILES[key]
        name = self.str(key, name)
        if name.startswith("!"):
            base = name[1:]
        else:
            base = self.str(folder_key, "") + name
        if suffix_key is None:
            return base
        suf = self.str(suffix_key, "")
        return base + ("-" + suf if suf else "") + ".bin"


def snapshot_options(cfg: Config) -> dict:
    """Snapshot / restore options of the reference (src/patch.cu:194, 306-307,
    340, 439-446, 611-619): snapshotInterval (steps), delPrevSnapshot, the
    file to restore from (fSnapshot under inputFolder, '!' for a literal
    path, then '_<snapshot_suffix>.bin') and asInit (use it as the initial
    condition rather than resume).  A resume needs snapshot_suffix equal to
    output_suffix, as in the reference."""
    restore = None
    if cfg.has("fSnapshot"):
        name = cfg.str("fSnapshot")
        base = name[1:] if name.startswith("!") else cfg.str("inputFolder", "") + name
        restore = base + "_" + cfg.str("snapshot_suffix", "") + ".bin"
    as_init = cfg.bool("asInit", True)
    if restore and not as_init and cfg.str("snapshot_suffix", "") != cfg.str("output_suffix", ""):
        raise ValueError("resuming from a snapshot needs snapshot_suffix == output_suffix")
    suf = cfg.str("output_suffix", "")
    return dict(interval=cfg.int("snapshotInterval", 10000),
                delete_previous=cfg.bool("delPrevSnapshot", True), restore=restore,
                as_init=as_init, output_suffix=("-" + suf) if suf else "")


def _pairs(vals, n_types, per):
    """A flat [type][per] list -> tuple of per-type tuples."""
    if vals is None:
        return None
    if len(vals) < n_types * per:
        raise ValueError(f"expected {n_types}x{per} values, got {len(vals)}")
    return tuple(tuple(vals[t * per:(t + 1) * per]) for t in range(n_types))


def learning_from_config(cfg: Config, dt: float) -> dict:
    """The plasticity keys (src/patch.cu:219-258, 289, 1092-1330).  learning 3
    (only the FF_E triplet rule) is built; 1, 2 and 4 (cortical E->E and
    I->E rules) are refused.  The reference's checks that apply to learning 3
    are made: tauAvg and targetFR of size 1 or 2, tauQ/A_Q/gmaxQ/gminQ of one
    size, tauLTP/tauLTD/r_LTD of tauTrip's size (one rule, max_nLearnTypeFF_E
    = 1), gmaxLGN/gminLGN of size 1 (or the number of FF rules)."""
    mode = cfg.int("learning", 0)
    if mode == 0:
        return {}
    if mode != 3:
        raise NotImplementedError(f"learning = {mode}: only learning = 3 (the feed-forward "
                                  "rule onto excitatory neurons) is built")
    kw = {"learning": 3}

    def need(key, sizes=None):
        v = cfg.floats(key)
        if v is None or (sizes is not None and len(v) not in sizes):
            raise ValueError(f"{key} is not given" if v is None else
                             f"{key} has {len(v)} values, expected one of {sorted(sizes)}")
        return v
    for key in ("tauAvg", "targetFR"):
        v = need(key, {1, 2})
        kw[key] = (v[0], v[-1])
    q = [cfg.floats(k) or [] for k in ("tauQ", "A_Q", "gmaxQ", "gminQ")]
    if len({len(x) for x in q}) > 1:
        raise ValueError(f"size of tauQ: {len(q[0])} should be the same as A_Q: {len(q[1])} "
                         "(and gmaxQ, gminQ)")
    if cfg.int("nLearnTypeFF_E", 1) != 1:
        raise NotImplementedError("nLearnTypeFF_E != 1 (max_nLearnTypeFF_E = 1)")
    trip = need("tauTrip")
    for key in ("tauLTP", "tauLTD", "r_LTD"):
        if len(need(key)) != len(trip):
            raise ValueError(f"the size of {key} should be the same as tauTrip: {len(trip)}")
    kw.update(tauLTP=cfg.floats("tauLTP")[0], tauLTD=cfg.floats("tauLTD")[0],
              tauTrip=trip[0], r_LTD=cfg.floats("r_LTD")[0], A_LGN=need("A_LGN")[0])
    for key in ("gmaxLGN", "gminLGN"):
        kw[key] = need(key, {1})[0]
    kw["applyHomeo"] = cfg.int("applyHomeo", 0)
    if kw["applyHomeo"] not in (0, 1, 2):
        raise ValueError(f"applyHomeo = {kw['applyHomeo']} (0, 1 or 2)")
    kw["symmetricHomeo"] = 1 if cfg.bool("symmetricHomeo", True) else 0
    kw["decayTimeHomeoFF"] = cfg.float("decayTimeHomeoFF", 250.0)
    kw["FF_InfRatio"] = cfg.float("FF_InfRatio", 1.0)
    kw["rebound"] = cfg.int("rebound", 0)
    kw["learnData_FF"] = cfg.int("learnData_FF", 1)
    if kw["learnData_FF"] > 1:
        raise NotImplementedError("learnData_FF = 2 (the per-step learning variables) is not "
                                  "built; 0 or 1 (the strengths every sampleInterval_LGN_V1 "
                                  "steps)")
    if cfg.bool("store_dsLGN", False):
        raise NotImplementedError("store_dsLGN is not built")
    kw["sampleInterval_LGN_V1"] = cfg.int("sampleInterval_LGN_V1", 100)
    return kw


def spec_from_config(cfg: Config, base=None):
    """ModelSpec from the cfg keys of src/patch.cu:180-374 (keys the model
    uses); everything not given keeps the base spec's value."""
    from . import network
    sp = base if base is not None else network.ModelSpec()
    kw = {}
    for key in ("dt", "nsig", "tau", "convolRatio", "frRatioLGN"):
        if cfg.has(key):
            kw[key] = cfg.float(key)
    for key in ("nt", "nKernelSample", "nSpatialSample1D", "frameRate", "iModel"):
        if cfg.has(key):
            kw[key] = cfg.int(key)
    if cfg.has("seed"):
        kw["seed"] = cfg.int("seed")
    for key in ("vR", "vThres", "gL", "C", "tRef", "vT", "deltaT", "tau_w", "a", "b",
                "tonicDep", "sRatioLGN", "bgStrength"):
        v = cfg.floats(key)
        if v is not None:
            kw[key] = tuple(v[:2]) if len(v) >= 2 else (v[0], v[0])
    for key in ("riseTimeFF", "decayTimeFF", "riseTimeE", "decayTimeE", "riseTimeI", "decayTimeI"):
        v = cfg.floats(key)
        if v is not None:
            kw[key] = tuple(v)
    ngFF = len(kw.get("riseTimeFF", sp.riseTimeFF))
    ngE = len(kw.get("riseTimeE", sp.riseTimeE))
    ngI = len(kw.get("riseTimeI", sp.riseTimeI))
    for key, ng in (("pFF", ngFF), ("pE", ngE), ("pI", ngI), ("pBG", ngFF)):
        v = cfg.floats(key)
        if v is not None:
            kw[key] = _pairs(v, 2, ng)
    if cfg.has("bgRate"):
        kw["bgRate"] = cfg.float("bgRate")
    # noise and synaptic failure (src/patch.cu:206-207, 273-280, 647-661,
    # 860-903): per-type lists take one value for every type; synFail and
    # synPerCon are [pre type][post type], negative failure rates become 0
    for key in ("noisyDep", "synFailFF", "synPerConFF"):
        v = cfg.floats(key)
        if v is not None:
            if len(v) not in (1, 2):
                raise ValueError(f"{key} needs 1 or nType = 2 values, got {len(v)}")
            kw[key] = (v[0], v[0]) if len(v) == 1 else tuple(v)
    for key in ("synFail", "synPerCon"):
        v = cfg.floats(key)
        if v is not None:
            if len(v) != 4:
                raise ValueError(f"{key} needs nType x nType = 4 values, got {len(v)}")
            kw[key] = tuple(max(x, 0.0) for x in v) if key == "synFail" else tuple(v)
    if any(x < 0 for x in kw.get("noisyDep", ())):
        raise ValueError("noisyDep must be >= 0")
    if cfg.has("noiseOnTonic"):
        kw["noiseOnTonic"] = 1 if cfg.bool("noiseOnTonic") else 0
    if cfg.has("tau_noise"):
        kw["tau_noise"] = cfg.float("tau_noise")
    # initial spikes: [simple, complex] rates per type (Hz), src/patch.cu:270-272, 811-837
    for key in ("spE0", "spI0"):
        v = cfg.floats(key)
        if v is not None:
            if len(v) != 2:
                raise ValueError(f"the size of {key} is {len(v)} != 1 x 2")
            kw[key] = tuple(v)
    if cfg.has("SCsplit"):
        kw["SCsplit"] = cfg.int("SCsplit")
    # the initial-state draws (src/patch.cu:225-229, 311, 4405-4570); the
    # reference requires gFF0, gE0 and gI0 (and w0 with AdEx), so a cfg that
    # gives them gets the reference's initial state
    for key, per in (("w0", 2), ("gFF0", 4 * ngFF), ("gE0", 4 * ngE), ("gI0", 4 * ngI)):
        v = cfg.floats(key)
        if v is not None:
            if len(v) != 2 * per:
                raise ValueError(f"size of {key} is {len(v)}, not nType x {per}")
            kw["init_" + key] = tuple(v)
    if any(k in kw for k in ("init_gFF0", "init_gE0", "init_gI0")):
        for key in ("init_gFF0", "init_gE0", "init_gI0"):
            if key not in kw:
                raise ValueError(f"{key[5:]} is required with the other initial distributions")
        if kw.get("iModel", sp.iModel) == 1 and "init_w0" not in kw:
            raise ValueError("w0 is required for the AdEx model (iModel 1)")
    if cfg.has("use_v0"):
        kw["use_v0"] = 1 if cfg.bool("use_v0") else 0
    if cfg.has("noDelay"):
        kw["noDelay"] = cfg.int("noDelay")
    if cfg.has("speedOfThought"):
        kw["speedOfThought"] = cfg.float("speedOfThought")
    # far connections: noFarDelay defaults to 1, and 1 with noDelay 0 is refused,
    # as the reference does (src/patch.cu:237, 839-842)
    kw["noFarDelay"] = cfg.int("noFarDelay", 1)
    if kw["noFarDelay"] and not kw.get("noDelay", sp.noDelay):
        raise ValueError("noFarDelay = 1 with noDelay = 0: delay or no delay is the question")
    # gap junctions between inhibitory neurons (src/patch.cu:238); their
    # strengths come from V1_gapMat / V1_gapVec, scaled by gapRatio (loader)
    kw["InhGap"] = 1 if cfg.bool("InhGap", False) else 0
    kw.update(learning_from_config(cfg, kw.get("dt", sp.dt)))
    if cfg.has("flat_retina") and not cfg.bool("flat_retina"):
        raise NotImplementedError("only flat_retina = true is built")
    if cfg.bool("virtual_LGN", False):
        raise NotImplementedError("virtual_LGN is not built")
    return replace(sp, **kw)
