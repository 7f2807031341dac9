"""Shared helpers for the tests (graphs, tolerances)."""
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
LONG_MAX = np.iinfo(np.int64).max


def bc_close(got, want, rel=1e-9):
    """BC / delta agreement: |got - want| <= rel * max(|want|, floor), where
    floor = 1e-6 * max|want| covers entries whose value is a near-cancellation
    of much larger terms (summation order is the only difference allowed)."""
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    scale = np.max(np.abs(want)) if want.size else 0.0
    floor = max(scale * 1e-6, 1e-12)
    err = np.abs(got - want)
    lim = rel * np.maximum(np.abs(want), floor)
    bad = err > lim
    return (not bad.any()), (float((err / np.maximum(np.abs(want), floor)).max()) if want.size else 0.0)


def assert_bc_close(got, want, rel=1e-9, what="bc"):
    ok, worst = bc_close(got, want, rel)
    assert ok, f"{what}: worst relative error {worst:.3e} exceeds {rel}"


def karate_path():
    return os.path.join(GOLDEN, "graphs", "karate.graph")
