"""Comparison helpers for parity tests."""
import numpy as np


def bitwise_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def max_rel(a, b, floor=1e-300):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.abs(a - b)
    return float(np.max(d / np.maximum(np.abs(b), floor))) if d.size else 0.0


def zone_rel(a, b):
    """max over zones and rows of |a-b| / |b_zone|, where |b_zone| is the norm
    of the zone's vector (the column of a (NQ, Nzones) array)."""
    a = np.asarray(a, dtype=np.float64).reshape(-1, np.shape(a)[-1])
    b = np.asarray(b, dtype=np.float64).reshape(-1, np.shape(b)[-1])
    norm = np.sqrt(np.sum(b * b, axis=0))
    return float(np.max(np.abs(a - b) / np.maximum(norm, 1e-300)))


def describe(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    n = int(np.sum(a.view(np.uint64) != b.view(np.uint64)))
    return f"{n} of {a.size} differ, max rel {max_rel(a, b):.3e}, max abs {float(np.max(np.abs(a-b))):.3e}"
