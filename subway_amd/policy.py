"""The learned transfer/async scheduler: its network (trained and evaluated
by the library's HIP kernels), the features it reads and the parameter file.

Features (12, computed per global iteration from state the pipeline has):
  0 log2(1 + active vertices) / log2(V)     6 fraction of active vertices with degree 32..255
  1 active vertices / V                      7 fraction of active vertices with degree >= 256
  2 active edges / E                         8 device buffer edges / E
  3 log2(1 + active edges) / log2(E)         9 global iteration index / 16
  4 fraction of active vertices, degree < 4 10 last iteration's updates / E
  5 fraction of active vertices, degree 4..31 11 fraction of id ranges with active density >= 0.5
Heads (3): log2(next iteration's active-edge fraction) / 32; whole-range
density threshold; log2(inner-round cap) / 16.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from ._lib import lib, check

FEATURES = 12
HEADS = 3
PARAMS = 5211
NORM = 2 * FEATURES
_f32p = ctypes.POINTER(ctypes.c_float)


def _p(a):
    return a.ctypes.data_as(_f32p)


def init_params(seed: int = 0, mean=None, std=None) -> np.ndarray:
    """He-initialised weights; normalisation from the training features."""
    rng = np.random.default_rng(seed)
    p = np.zeros(PARAMS, np.float32)
    p[:FEATURES] = 0.0 if mean is None else mean
    p[FEATURES:NORM] = 1.0 if std is None else 1.0 / np.maximum(std, 1e-6)
    o = NORM
    for fan_in, fan_out in ((FEATURES, 64), (64, 64), (64, HEADS)):
        p[o:o + fan_in * fan_out] = rng.normal(0, np.sqrt(2.0 / fan_in), fan_in * fan_out)
        o += fan_in * fan_out + fan_out
    return p


def forward(params, x) -> np.ndarray:
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32).reshape(-1, FEATURES)
    out = np.empty((len(x), HEADS), np.float32)
    check(lib.subway_mlp_forward(_p(params), _p(x), len(x), _p(out)))
    return out


def grad(params, x, y, mask):
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32).reshape(-1, FEATURES)
    y = np.ascontiguousarray(y, np.float32).reshape(-1, HEADS)
    mask = np.ascontiguousarray(mask, np.float32).reshape(-1, HEADS)
    g = np.empty(PARAMS, np.float32)
    loss = ctypes.c_float()
    check(lib.subway_mlp_grad(_p(params), _p(x), _p(y), _p(mask), len(x), _p(g), ctypes.byref(loss)))
    return g, float(loss.value)


def train(params, x, y, mask, steps: int = 2000, lr: float = 3e-3):
    """Full-batch Adam on the GPU; returns (params, loss per step)."""
    params = np.array(params, np.float32, copy=True)
    x = np.ascontiguousarray(x, np.float32).reshape(-1, FEATURES)
    y = np.ascontiguousarray(y, np.float32).reshape(-1, HEADS)
    mask = np.ascontiguousarray(mask, np.float32).reshape(-1, HEADS)
    losses = np.empty(max(steps, 1), np.float32)
    check(lib.subway_mlp_train(_p(params), _p(x), _p(y), _p(mask), len(x), steps, lr, _p(losses)))
    return params, losses[:steps]


DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "scheduler_mlp.f32")


def default_params():
    """The scheduler network trained by subway_amd.train_scheduler and shipped
    with the package, or None when the package has none."""
    return load(DEFAULT_PATH) if os.path.exists(DEFAULT_PATH) else None


def save(path: str, params) -> None:
    np.asarray(params, np.float32).tofile(path)


def load(path: str) -> np.ndarray:
    p = np.fromfile(path, np.float32)
    if p.size != PARAMS:
        raise ValueError(f"{path}: expected {PARAMS} floats, found {p.size}")
    return p
