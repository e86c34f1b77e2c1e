"""Learned diagonal bands for the PairHMM forward pass.

The band model is a 36 -> 64 -> 64 -> 2 ReLU MLP in FP32 whose forward,
backward and Adam kernels run on the GPU (csrc/band_kernels.hip).  Features
are computed on the device per pair (csrc/band_kernels.hip:band_features_kernel).
Training labels are the narrowest diagonal span holding 99.999 % of each
pair's posterior mass, from an FP64 forward-backward pass (bv_band_labels).

Targets: T0 = (label centre - anchor) / 16, T1 = log2(label width).  At run
time (csrc/band_kernels.hip:band_from_prediction_kernel) a pair whose
predicted width plus margins fits gets the banded kernel's 64 diagonals
centred on the predicted centre; the banded kernel judges every band by its
edge-zone touch mass and sends rejected pairs to the full matrix.
A trained model is saved as an .npz with a format version, the generator
workload/seeds it was trained on, the epoch count and a SHA-256 of the
parameters; `train()` can resume from such a file.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import math
import time
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import pairhmm, synth

DEFAULT_MODEL = Path(__file__).resolve().parent / "models" / "band_mlp.npz"
# the same parameters in the library's own file format (bv_band_model_read),
# which the C++ drop-in loads without Python
DEFAULT_MODEL_BIN = DEFAULT_MODEL.with_suffix(".bin")
MODEL_MAGIC = b"BVBANDM\0"
N_FEAT = 36
N_PARAMS = 6658
FORMAT_VERSION = 1
LABEL_MASS = 0.99999
# weighting of the two outputs in the loss (centre error matters more)
LOSS_W = (1.0, 0.25)

_bound = False


def _lib():
    global _bound
    lib = pairhmm.library()
    if not _bound:
        vp, i64 = C.c_void_p, C.c_int64
        lib.bv_band_model_create.argtypes = [C.POINTER(vp), C.c_uint64]
        lib.bv_band_model_destroy.argtypes = [vp]
        lib.bv_band_model_destroy.restype = None
        lib.bv_band_model_get_params.argtypes = [vp, vp]
        lib.bv_band_model_set_params.argtypes = [vp, vp]
        lib.bv_band_model_predict.argtypes = [vp, vp, i64, vp]
        lib.bv_band_model_loss_grad.argtypes = [vp, vp, vp, i64, C.c_float, C.c_float, vp, vp]
        lib.bv_band_model_train.argtypes = [vp, vp, vp, i64, C.c_int, C.c_int, C.c_float,
                                            C.c_float, C.c_float, vp, vp]
        lib.bv_plan_features.argtypes = [vp, vp, vp, vp]
        lib.bv_plan_set_band_model.argtypes = [vp, vp, C.c_int, C.c_int]
        lib.bv_plan_bands.argtypes = [vp, vp]
        lib.bv_plan_set_bands.argtypes = [vp, vp]
        lib.bv_band_labels.argtypes = [C.c_void_p, C.c_double, C.c_int, vp]
        lib.bv_band_model_get_state.argtypes = [vp, vp, vp, vp]
        lib.bv_band_model_set_state.argtypes = [vp, vp, vp, i64]
        lib.bv_band_model_read.argtypes = [C.c_char_p, vp]
        lib.bv_band_model_load.argtypes = [C.c_char_p, C.POINTER(vp)]
        lib.bv_band_model_write.argtypes = [vp, C.c_char_p]
        lib.bv_band_model_default_path.argtypes = [C.c_char_p, i64]
        _bound = True
    return lib


def _check(rc):
    pairhmm._check(rc)


class BandModel:
    """Device-resident MLP (bv_band_model_*)."""

    def __init__(self, seed: int = 0, params: np.ndarray | None = None):
        lib = _lib()
        h = C.c_void_p()
        _check(lib.bv_band_model_create(C.byref(h), C.c_uint64(seed)))
        self._h = h
        self._lib = lib
        self.meta: dict = {}
        if params is not None:
            self.set_params(params)

    @property
    def handle(self):
        return self._h

    def params(self) -> np.ndarray:
        out = np.zeros(N_PARAMS, np.float32)
        _check(self._lib.bv_band_model_get_params(self._h, out.ctypes.data))
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.shape == (N_PARAMS,)
        _check(self._lib.bv_band_model_set_params(self._h, p.ctypes.data))

    def predict(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros((len(X), 2), np.float32)
        _check(self._lib.bv_band_model_predict(self._h, X.ctypes.data, len(X), Y.ctypes.data))
        return Y

    def loss_grad(self, X, T, w=LOSS_W):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        loss = np.zeros(1, np.float32)
        grad = np.zeros(N_PARAMS, np.float32)
        _check(self._lib.bv_band_model_loss_grad(self._h, X.ctypes.data, T.ctypes.data, len(X),
                                                 w[0], w[1], loss.ctypes.data, grad.ctypes.data))
        return float(loss[0]), grad

    def fit(self, X, T, epochs=20, batch=1024, lr=2e-3, w=LOSS_W, seed=0):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        n = len(X)
        rng = np.random.default_rng(seed)
        order = np.concatenate([rng.permutation(n) for _ in range(epochs)]).astype(np.int64)
        losses = np.zeros(epochs, np.float32)
        _check(self._lib.bv_band_model_train(self._h, X.ctypes.data, T.ctypes.data, n, epochs,
                                             batch, lr, w[0], w[1], order.ctypes.data,
                                             losses.ctypes.data))
        return losses

    def state(self):
        """Adam's moments and step count (training resumes from them)."""
        m1 = np.zeros(N_PARAMS, np.float32)
        m2 = np.zeros(N_PARAMS, np.float32)
        step = C.c_int64()
        _check(self._lib.bv_band_model_get_state(self._h, m1.ctypes.data, m2.ctypes.data, C.byref(step)))
        return m1, m2, int(step.value)

    def set_state(self, m1, m2, step: int) -> None:
        m1 = np.ascontiguousarray(m1, np.float32)
        m2 = np.ascontiguousarray(m2, np.float32)
        _check(self._lib.bv_band_model_set_state(self._h, m1.ctypes.data, m2.ctypes.data, int(step)))

    def save(self, path, **meta) -> str:
        """An .npz with the parameters, their SHA-256, Adam's state and the
        training metadata; the parameters also go to a .bin beside it in the
        library's format (the file the C++ drop-in loads)."""
        p = self.params()
        m1, m2, step = self.state()
        digest = hashlib.sha256(p.tobytes()).hexdigest()
        m = dict(self.meta)
        m.update(meta)
        np.savez(path, params=p, version=np.int64(FORMAT_VERSION), sha256=np.array(digest),
                 meta=np.array(repr(sorted(m.items()))), adam_m1=m1, adam_m2=m2, adam_step=np.int64(step))
        write_model_file(Path(path).with_suffix(".bin"), p)
        return digest

    @classmethod
    def load(cls, path) -> "BandModel":
        d = np.load(path, allow_pickle=False)
        if int(d["version"]) != FORMAT_VERSION:
            raise ValueError(f"band model format {int(d['version'])}, expected {FORMAT_VERSION}")
        p = d["params"].astype(np.float32)
        if hashlib.sha256(p.tobytes()).hexdigest() != str(d["sha256"]):
            raise ValueError("band model parameters do not match their SHA-256")
        m = cls(params=p)
        if "adam_step" in d.files:
            m.set_state(d["adam_m1"], d["adam_m2"], int(d["adam_step"]))
        m.meta = {"source": str(path), "meta": str(d["meta"])}
        return m

    @classmethod
    def load_bin(cls, path=None) -> "BandModel":
        """A model from a file in the library's format (default: the one the
        library ships, next to it), read and checked by the library itself."""
        lib = _lib()
        h = C.c_void_p()
        _check(lib.bv_band_model_load(None if path is None else str(path).encode(), C.byref(h)))
        m = cls.__new__(cls)
        m._h, m._lib, m.meta = h, lib, {"source": str(path)}
        return m

    def close(self):
        if getattr(self, "_h", None):
            self._lib.bv_band_model_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_model_file(path, params) -> None:
    """The library's band model format (bv_band_model_read): magic, uint32
    version 1, uint32 parameter count, little-endian float32 parameters, and
    the SHA-256 of the parameter bytes."""
    p = np.ascontiguousarray(params, "<f4")
    assert p.shape == (N_PARAMS,)
    body = p.tobytes()
    Path(path).write_bytes(MODEL_MAGIC + np.array([1, N_PARAMS], "<u4").tobytes() + body
                           + hashlib.sha256(body).digest())


def read_model_file(path) -> np.ndarray:
    """Parameters of a file in the library's format, checked by the library
    (bv_band_model_read, host code only)."""
    out = np.zeros(N_PARAMS, np.float32)
    _check(_lib().bv_band_model_read(str(path).encode(), out.ctypes.data))
    return out


def plan_features(plan: pairhmm.Plan, with_ends: bool = False):
    """(X [n_pairs, 36], anchor [n_pairs]) computed on the device, and with
    with_ends the seed-anchored read span [n_pairs, 2] (first position, last
    end of seeds hitting within 8 of the anchor; -1, -1 without)."""
    n = plan.batch.n_pairs
    X = np.zeros((n, N_FEAT), np.float32)
    a = np.zeros(n, np.int32)
    e = np.zeros((n, 2), np.int32)
    _check(_lib().bv_plan_features(plan._h, X.ctypes.data, a.ctypes.data, e.ctypes.data))
    return (X, a, e) if with_ends else (X, a)


def set_plan_bands(plan: pairhmm.Plan, bands: np.ndarray | None) -> None:
    if bands is None:
        _check(_lib().bv_plan_set_bands(plan._h, None))
        return
    b = np.ascontiguousarray(bands, np.int32)
    assert b.shape == (plan.batch.n_pairs, 2)
    _check(_lib().bv_plan_set_bands(plan._h, b.ctypes.data))


MARGIN = 10      # diagonals added per side of the predicted span (narrow band iff the
                 # predicted span is at most 4 wide)
MAX_WIDTH = 68   # widest predicted span (with margins) given a band


def set_plan_model(plan: pairhmm.Plan, model: BandModel | None, margin=MARGIN, max_width=MAX_WIDTH) -> None:
    """Bands from the network, computed on the device at every run: 32 or
    64 diagonals centred on the predicted posterior span, chosen by the
    predicted span plus `margin` diagonals per side (32 when that fits between
    the narrow band's edge zones, 64 up to `max_width`); wider predictions run
    on the full matrix."""
    _check(_lib().bv_plan_set_band_model(plan._h, model.handle if model else None, margin, max_width))


def plan_bands(plan: pairhmm.Plan) -> np.ndarray:
    b = np.zeros((plan.batch.n_pairs, 2), np.int32)
    _check(_lib().bv_plan_bands(plan._h, b.ctypes.data))
    return b


def labels(batch: pairhmm.PackedBatch, mass=LABEL_MASS, n_threads=0) -> np.ndarray:
    """[n_pairs, 2] (o_lo, o_hi) posterior spans (host FP64 forward-backward)."""
    out = np.zeros((batch.n_pairs, 2), np.int32)
    cb = batch.cstruct()
    _check(_lib().bv_band_labels(C.byref(cb), mass, n_threads, out.ctypes.data))
    return out


def targets(lab: np.ndarray, anchor: np.ndarray) -> np.ndarray:
    """(centre - anchor) / 16 clipped to +-4 (bands further off are not
    supported by the seeds and stay on the full matrix), log2 width clipped
    to [0, 8]."""
    centre = (lab[:, 0] + lab[:, 1]) / 2.0
    width = lab[:, 1] - lab[:, 0] + 1
    t0 = np.clip((centre - anchor) / 16.0, -4.0, 4.0)
    t1 = np.clip(np.log2(np.maximum(width, 1)), 0.0, 8.0)
    return np.stack([t0, t1], 1).astype(np.float32)


BAND_LANES = 64   # diagonals the banded kernel evaluates per pair
HAP_MAX = 1024    # longest haplotype the banded kernel stages


def heuristic_bands(X: np.ndarray, anchor: np.ndarray, width: int = BAND_LANES) -> np.ndarray:
    """The band without the network: `width` (32 or 64) diagonals centred on
    the seed anchor, for every pair with seed hits (the same pairs the learned
    band considers)."""
    H = np.rint(X[:, 1] * 512).astype(np.int64)
    ok = (H <= HAP_MAX) & (X[:, 4:36].sum(1) > 0)
    b = np.stack([anchor - width // 2, anchor + width // 2 - 1], 1).astype(np.int32)
    b[~ok] = (1, 0)
    return b


NARROW = 32      # diagonals of the narrow band (two pairs per wavefront)
EDGE = 4         # edge-zone diagonals per side
END_SLACK = 8    # read bases the end-anchoring seeds may leave uncovered (narrow)
END_SLACK_WIDE = 32  # the same for the wide band


def seeds_ok(X: np.ndarray, ends: np.ndarray, width: int, slack: int) -> np.ndarray:
    """The seed conditions of a band of `width` diagonals centred on the
    anchor (band_kernels.hip, band_from_prediction_kernel): every seed hit
    within 64 of the anchor inside the band's edge zones, and seeds hitting
    next to the anchor reaching within `slack` bases of both read ends."""
    R = np.rint(X[:, 0] * 256).astype(np.int64)
    k_in = width // 2 - EDGE
    blo = -64 + 4 * np.arange(32)
    outside = (blo < -k_in) | (blo + 3 > k_in - 1)
    hits_in = ~((X[:, 4:36] > 0) & outside[None, :]).any(1)
    ok = hits_in & (ends[:, 0] >= 0) & (ends[:, 0] <= slack) & (ends[:, 1] >= R - slack)
    if width == BAND_LANES:
        ok &= 2 * (ends[:, 1] - ends[:, 0]) >= R
    return ok


def rule_bands(X: np.ndarray, anchor: np.ndarray, ends: np.ndarray, narrow: bool = False) -> np.ndarray:
    """The learned band's rules without the network (centre = anchor): wide
    where the wide band's seed conditions hold (with `narrow`, narrow where
    its own hold), else the full matrix."""
    b = heuristic_bands(X, anchor, BAND_LANES)
    ok = b[:, 0] <= b[:, 1]
    wide = ok & seeds_ok(X, ends, BAND_LANES, END_SLACK_WIDE)
    b[~wide] = (1, 0)
    if narrow:
        nar = ok & seeds_ok(X, ends, NARROW, END_SLACK)
        b[nar] = np.stack([anchor - NARROW // 2, anchor + NARROW // 2 - 1], 1)[nar]
    return b


def sample_pairs(workload: int, seed: int, n_pairs: int, pairs_per_region: int = 64):
    """A PackedBatch of about n_pairs (read, haplotype) pairs drawn from fresh
    regions of a workload generator, each pair as its own region."""
    rng = np.random.default_rng(seed)
    regions = []
    first = 0
    while sum(len(r[0]) for r in regions) < n_pairs:
        count = 64
        b = synth.generate(workload, seed=seed, first=first, count=count)
        first += count
        for g in range(b.n_regions):
            reads, haps = b.region(g)
            k = min(pairs_per_region, len(reads) * len(haps))
            for idx in rng.choice(len(reads) * len(haps), k, replace=False):
                r, h = divmod(int(idx), len(haps))
                regions.append(([reads[r]], [haps[h]]))
                if len(regions) >= n_pairs:
                    break
            if len(regions) >= n_pairs:
                break
    return pairhmm.PackedBatch.from_regions(regions[:n_pairs])


@dataclass
class TrainReport:
    n_train: int
    epochs: int
    losses: list
    seconds: dict = field(default_factory=dict)
    sha256: str = ""


def build_training_set(n_pairs=200_000, workload=1, seed=synth.SEED_BASE + 301, n_threads=0):
    """Features (device) and labels (host FP64 forward-backward) for n_pairs
    sampled from the workload generator."""
    t0 = time.perf_counter()
    batch = sample_pairs(workload, seed, n_pairs)
    t1 = time.perf_counter()
    with pairhmm.Plan(batch) as plan:
        X, anchor = plan_features(plan)
    t2 = time.perf_counter()
    lab = labels(batch, n_threads=n_threads)
    t3 = time.perf_counter()
    # pairs without seed hits have no anchor and always run on the full matrix
    keep = X[:, 4:36].sum(1) > 0
    X, anchor, lab = X[keep], anchor[keep], lab[keep]
    T = targets(lab, anchor)
    return X, T, anchor, lab, dict(generate=t1 - t0, features=t2 - t1, labels=t3 - t2)


def train(n_pairs=200_000, epochs=30, batch=1024, lr=2e-3, seed=7, out: str | Path | None = None,
          resume: str | Path | None = None, data=None) -> tuple[BandModel, TrainReport]:
    """Trains (or resumes training of) a band model on the GPU."""
    if data is None:
        X, T, anchor, lab, secs = build_training_set(n_pairs)
    else:
        X, T, anchor, lab, secs = data
    model = BandModel.load(resume) if resume else BandModel(seed=seed)
    done = 0
    if resume:
        meta = model.meta.get("meta", "")
        for part in meta.strip("[]").split("), ("):
            if "'epochs'" in part:
                done = int(part.split(",")[-1].strip(" )"))
    t0 = time.perf_counter()
    losses = model.fit(X, T, epochs=epochs, batch=batch, lr=lr, seed=seed + done)
    secs = dict(secs)
    secs["train"] = time.perf_counter() - t0
    rep = TrainReport(n_train=len(X), epochs=done + epochs, losses=[float(x) for x in losses],
                      seconds=secs)
    if out is not None:
        rep.sha256 = model.save(out, epochs=done + epochs, n_train=len(X), workload=1,
                                label_mass=LABEL_MASS)
    return model, rep


def _main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(description="train the band MLP on the GPU")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--pairs", type=int, default=200_000)
    t.add_argument("--epochs", type=int, default=30)
    t.add_argument("--batch", type=int, default=1024)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--seed", type=int, default=7)
    t.add_argument("--out", required=True)
    t.add_argument("--resume", default=None)
    args = ap.parse_args(argv)
    print(f"[band] building training set of {args.pairs} pairs", flush=True)
    data = build_training_set(args.pairs)
    print(f"[band] features/labels done: {data[4]}", flush=True)
    model, rep = train(n_pairs=args.pairs, epochs=args.epochs, batch=args.batch, lr=args.lr,
                       seed=args.seed, out=args.out, resume=args.resume, data=data)
    print(json.dumps(dict(n_train=rep.n_train, epochs=rep.epochs, first_loss=rep.losses[0],
                          last_loss=rep.losses[-1], seconds=rep.seconds, sha256=rep.sha256)),
          flush=True)


if __name__ == "__main__":
    _main()
