"""Test helpers: fixtures, the FP32 error metric and small seeded workloads."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def fixture(name):
    with open(os.path.join(HERE, "golden", name + ".json")) as f:
        return json.load(f)


def rel_err(y, y_star, den):
    """max_i |y_i - y*_i| / sum_j |M_ij x_j|; a row with a zero denominator must
    match exactly (it contributes inf otherwise)."""
    y = np.asarray(y, np.float64)
    y_star = np.asarray(y_star, np.float64)
    if y.shape != y_star.shape:
        raise ValueError(f"shapes differ: {y.shape} vs {y_star.shape}")
    den = np.asarray(den, np.float64)
    d = np.abs(y - y_star)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(den > 0, d / np.where(den > 0, den, 1.0), np.where(d == 0, 0.0, np.inf))
    return float(r.max()) if r.size else 0.0


def grammar_dict(f):
    """Fixture grammar -> arrays dict (alpha, kind, left, right, rep, C)."""
    return dict(alpha=f["alpha"], kind=np.array(f["kind"], np.int32), left=np.array(f["left"], np.int32),
                right=np.array(f["right"], np.int32), rep=np.array(f["rep"], np.int32),
                C=np.array(f["C"], np.int32))


def x_for(semiring, cols, seed, B=None):
    """Seeded right-hand sides in the forms the reference driver uses
    (gpu-engine/gpu_engine_test.cu:401-413): U[-1,1] for (+,x), 0/1 for Boolean,
    0..99 with 10% +infinity (1e9) for Tropical."""
    rng = np.random.default_rng(seed)
    shape = (cols,) if B is None else (cols, B)
    if semiring == 0:
        return rng.uniform(-1, 1, shape).astype(np.float32)
    if semiring == 1:
        return (rng.random(shape) < 0.5).astype(np.float32)
    x = rng.integers(0, 100, shape).astype(np.float32)
    x[rng.random(shape) < 0.1] = 1e9
    return x
