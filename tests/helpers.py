"""Shared assertions for the parity tests."""
import numpy as np


def same_f32(a, b):
    """Bitwise float32 equality, any NaN equal to any NaN."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    if a.shape != b.shape:
        return np.zeros(1, bool)
    return (np.isnan(a) & np.isnan(b)) | (a.view(np.uint32) == b.view(np.uint32))


def assert_same_f32(a, b, what=""):
    eq = same_f32(a, b)
    if not eq.all():
        bad = np.argwhere(~eq)
        i = tuple(bad[0])
        raise AssertionError(f"{what}: {len(bad)} mismatches, first at {i}: "
                             f"{np.asarray(a)[i]!r} vs {np.asarray(b)[i]!r}")


def assert_same_i32(a, b, what=""):
    a, b = np.asarray(a), np.asarray(b)
    if a.shape != b.shape:
        raise AssertionError(f"{what}: shape {a.shape} vs {b.shape}")
    if not np.array_equal(a, b):
        bad = np.argwhere(a != b)
        i = tuple(bad[0])
        raise AssertionError(f"{what}: {len(bad)} mismatches, first at {i}: {a[i]} vs {b[i]}")
