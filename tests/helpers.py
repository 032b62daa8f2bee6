"""Comparison helpers shared by the parity tests."""
import numpy as np


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b.ravel())
    d = np.linalg.norm((a - b).ravel())
    if nb == 0.0:
        return d
    return d / nb


def interior(a):
    return a[1:-1, 1:-1, 1:-1]


def bitwise_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


def describe(name, a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    diff = np.abs(a - b)
    idx = np.unravel_index(np.argmax(diff), diff.shape) if diff.size else ()
    return (f"{name}: rel_l2={rel_l2(a, b):.3e} max_abs={diff.max():.3e} "
            f"at {idx} gpu={a[idx] if diff.size else None} "
            f"oracle={b[idx] if diff.size else None}")
