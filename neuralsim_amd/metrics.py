"""Spike-train comparison used to score the HH surrogate against a reference run.

A spike of neuron i at step s is *displaced* when the other run has no spike of neuron i within
`tol_ms` of it.  The displaced fraction pools both directions -- reference spikes the run under
test misses and spikes it adds -- over all spikes of both runs, so a spike moved by more than
the tolerance counts twice (once missing, once extra) out of two.
"""
import numpy as np


def _keys(steps, ids, span):
    return np.sort(np.asarray(ids, np.int64) * span + np.asarray(steps, np.int64))


def _unmatched(a, b, tol):
    """Elements of sorted a with no element of sorted b within tol."""
    if len(a) == 0:
        return 0
    if len(b) == 0:
        return len(a)
    j = np.searchsorted(b, a)
    left = np.abs(a - b[np.clip(j - 1, 0, len(b) - 1)])
    right = np.abs(b[np.clip(j, 0, len(b) - 1)] - a)
    return int(np.count_nonzero(np.minimum(left, right) > tol))


def spike_displacement(ref_steps, ref_ids, test_steps, test_ids, dt_ms, tol_ms=0.5):
    """Dict with the displaced fraction and its parts (see module docstring)."""
    tol = int(np.floor(tol_ms / dt_ms + 1e-9))
    hi = max(int(np.max(ref_steps, initial=0)), int(np.max(test_steps, initial=0)))
    span = hi + 4 * tol + 8  # keys of different neurons are further apart than tol
    a = _keys(ref_steps, ref_ids, span)
    b = _keys(test_steps, test_ids, span)
    miss = _unmatched(a, b, tol)
    extra = _unmatched(b, a, tol)
    total = len(a) + len(b)
    return {"ref_spikes": len(a), "test_spikes": len(b), "missing": miss, "extra": extra,
            "displaced_fraction": (miss + extra) / total if total else 0.0,
            "tol_steps": tol}
