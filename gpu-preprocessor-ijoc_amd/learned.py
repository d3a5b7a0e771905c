"""Learned pass scheduler: training samples from presolve traces, training on
the GPU through the engine's MLP kernels (gp_mlp_forward / gp_mlp_loss_grad /
gp_mlp_adam_step), and the weight blob that gp_preprocess reads when
params.learned_schedule = 1.

One sample per scheduler round: the features gp_schedule_features computes at
round start, and for each of the 7 rule groups a label "the group fired in
this round and the live nnz fell in this round or the next one", masked to the
groups the round's phase runs.  The network is an MLP 32-64-64-7 in FP32; a
group runs when its logit reaches the group's threshold, chosen per group on
the training samples so that a target fraction of the productive group-rounds
is kept.
"""
from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

import gpupre as G

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS_PATH = os.path.join(HERE, "learned", "scheduler_mlp.f32")
META_PATH = os.path.join(HERE, "learned", "scheduler_mlp.json")

GROUP_OF_RULE = [0, 0, 0, 1, 2, 3, 3, 4, 5, 6, 6]
GROUP_NAMES = ["fixed+empty", "singleton_rows", "singleton_cols", "bounds+redundant", "dual_fix", "doubleton",
               "parallel"]
PHASE_GROUPS = {0: (0, 1, 2), 1: (3, 4), 2: (5,), 3: (6,)}


def trace_array(trace):
    n = max(1, len(trace))
    arr = (G.GpRoundTrace * n)()
    for i, t in enumerate(trace):
        arr[i] = t
    return arr


def features(trace, done, phase, nnz_now, nnz0, m0, n0, bound_only, limit) -> np.ndarray:
    """The engine's feature vector (host function of libgpupre.so)."""
    f = np.empty(G.GP_MLP_IN, dtype=np.float32)
    G.load().gp_schedule_features(trace_array(trace), int(done), int(phase), int(nnz_now), int(nnz0), int(m0),
                                  int(n0), int(bound_only), int(limit), f.ctypes.data_as(C.c_void_p))
    return f


def round_samples(trace, m0: int, n0: int):
    """(X[T][32], Y[T][7], M[T][7]) for the rounds of one plain presolve trace."""
    rounds = [t for t in trace if not t.final_sweep]
    T = len(rounds)
    X = np.zeros((T, G.GP_MLP_IN), dtype=np.float32)
    Y = np.zeros((T, G.GP_MLP_OUT), dtype=np.float32)
    M = np.zeros((T, G.GP_MLP_OUT), dtype=np.float32)
    if T == 0:
        return X, Y, M
    arr = trace_array(rounds)
    nnz0 = rounds[0].nnz_before
    lib = G.load()
    for t, tr in enumerate(rounds):
        lib.gp_schedule_features(arr, t, tr.phase, tr.nnz_before, nnz0, m0, n0, tr.bound_only_rounds,
                                 tr.bound_only_limit, X[t].ctypes.data_as(C.c_void_p))
        progress = tr.nnz_after < tr.nnz_before or (t + 1 < T and rounds[t + 1].nnz_after < rounds[t + 1].nnz_before)
        for g in PHASE_GROUPS.get(tr.phase, ()):
            M[t, g] = 1.0
            fired = any((tr.fired_mask >> r) & 1 for r in range(len(GROUP_OF_RULE)) if GROUP_OF_RULE[r] == g)
            Y[t, g] = 1.0 if (fired and progress) else 0.0
    return X, Y, M


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform weights, zero biases, in the blob's layout."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((G.GP_MLP_IN, G.GP_MLP_H1), (G.GP_MLP_H1, G.GP_MLP_H2), (G.GP_MLP_H2, G.GP_MLP_OUT)):
        lim = np.sqrt(6.0 / fan_in)
        parts.append(rng.uniform(-lim, lim, size=fan_out * fan_in))
        parts.append(np.zeros(fan_out))
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == G.GP_MLP_PARAMS
    return p


class GpuTrainer:
    """Adam on the GPU with the engine's forward/backward kernels."""

    def __init__(self, params: np.ndarray):
        assert params.size == G.GP_MLP_PARAMS
        self.P = G.DeviceArray.from_host(np.ascontiguousarray(params, dtype=np.float32))
        zeros = np.zeros(G.GP_MLP_PARAMS, dtype=np.float32)
        self.m = G.DeviceArray.from_host(zeros)
        self.v = G.DeviceArray.from_host(zeros)
        self.grad = G.DeviceArray.from_host(zeros)
        self.loss = G.DeviceArray.from_host(np.zeros(1, dtype=np.float32))
        self.t = 0
        self.data = None

    def set_data(self, X, Y, M):
        for a in self.data or ():
            a.free()
        self.n = X.shape[0]
        self.data = [G.DeviceArray.from_host(np.ascontiguousarray(a, dtype=np.float32)) for a in (X, Y, M)]

    def step(self, start: int, count: int, lr: float) -> float:
        X, Y, M = self.data
        lib = G.load()
        def at(a, width):
            return C.c_void_p(a.ptr.value + 4 * width * start)

        G.check(lib.gp_mlp_loss_grad(self.P.ptr, at(X, G.GP_MLP_IN), at(Y, G.GP_MLP_OUT), at(M, G.GP_MLP_OUT), count,
                                     self.loss.ptr, self.grad.ptr))
        self.t += 1
        G.check(lib.gp_mlp_adam_step(self.P.ptr, self.grad.ptr, self.m.ptr, self.v.ptr, self.t, lr, 0.9, 0.999,
                                     1e-8))
        return float(G.to_host(self.loss.ptr, 1, np.float32)[0])

    def logits(self, X) -> np.ndarray:
        d = G.DeviceArray.from_host(np.ascontiguousarray(X, dtype=np.float32))
        out = G.DeviceArray(4 * X.shape[0] * G.GP_MLP_OUT)
        G.check(G.load().gp_mlp_forward(self.P.ptr, d.ptr, X.shape[0], out.ptr))
        o = G.to_host(out.ptr, X.shape[0] * G.GP_MLP_OUT, np.float32).reshape(-1, G.GP_MLP_OUT)
        d.free()
        out.free()
        return o

    def params(self) -> np.ndarray:
        return G.to_host(self.P.ptr, G.GP_MLP_PARAMS, np.float32)

    def free(self):
        for a in [self.P, self.m, self.v, self.grad, self.loss] + list(self.data or ()):
            a.free()


def choose_thresholds(logits: np.ndarray, Y: np.ndarray, M: np.ndarray, keep: float) -> np.ndarray:
    """Per group, the largest logit threshold that keeps at least `keep` of the
    productive (label 1) group-rounds; groups with no positives always run."""
    thr = np.full(G.GP_MLP_OUT, -np.inf, dtype=np.float32)
    for g in range(G.GP_MLP_OUT):
        pos = logits[(M[:, g] > 0) & (Y[:, g] > 0), g]
        if pos.size == 0:
            continue
        thr[g] = np.float32(np.quantile(pos.astype(np.float64), 1.0 - keep, method="lower"))
    return thr


def make_blob(params: np.ndarray, thresholds: np.ndarray) -> np.ndarray:
    blob = np.concatenate([np.asarray(params, dtype=np.float32), np.asarray(thresholds, dtype=np.float32)])
    assert blob.size == G.GP_MLP_BLOB
    return blob


def save(blob: np.ndarray, meta: dict, weights_path: str = WEIGHTS_PATH, meta_path: str = META_PATH):
    os.makedirs(os.path.dirname(weights_path), exist_ok=True)
    np.asarray(blob, dtype="<f4").tofile(weights_path)
    with open(meta_path, "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)


def load_blob(path: str = WEIGHTS_PATH) -> np.ndarray:
    blob = np.fromfile(path, dtype="<f4")
    if blob.size != G.GP_MLP_BLOB:
        raise ValueError(f"{path}: {blob.size} floats, expected {G.GP_MLP_BLOB}")
    return blob.astype(np.float32)


def learned_params(blob: np.ndarray, **overrides) -> G.GpParams:
    """Parity parameters with the learned scheduler on; the blob must stay
    alive while the parameters are used (it is referenced, not copied)."""
    p = G.parity_params(learned_schedule=1, **overrides)
    p.scheduler_weights = blob.ctypes.data_as(C.c_void_p)
    return p


def rule_launches(trace) -> int:
    """Rule invocations of a run (enabled rules minus the skipped ones), from its trace."""
    per_phase = {0: (0, 1, 2, 3, 4), 1: (5, 6, 7), 2: (8,), 3: (9, 10)}
    n = 0
    for t in trace:
        for r in per_phase.get(t.phase, ()):
            n += 0 if (t.skipped_mask >> r) & 1 else 1
    return n
