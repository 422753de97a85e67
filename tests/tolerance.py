"""Parity criterion shared by the tests.

FP64 path: |x - x_ref| <= rtol * max(|x_ref|, M) per entry, where M is the
sum of absolute values of the terms that form the entry (the oracle returns
it beside H and rho).  Symmetry-zero entries (~1e-18) make a pure relative
test meaningless; M scales the bound to the entry's own summands.
"""
import numpy as np

RTOL_FP64 = 1e-12


def within(x, ref, mag, rtol=RTOL_FP64):
    x, ref, mag = np.asarray(x), np.asarray(ref), np.asarray(mag)
    bound = rtol * np.maximum(np.abs(ref), mag)
    bad = np.abs(x - ref) > bound
    if bad.any():
        i = int(np.argmax(np.abs(x - ref) - bound))
        return False, "entry %d: got %r want %r (bound %r), %d bad of %d" % (i, x.flat[i], ref.flat[i], bound.flat[i], int(bad.sum()), bad.size)
    return True, ""


def worst_ratio(x, ref, mag):
    x, ref, mag = np.asarray(x), np.asarray(ref), np.asarray(mag)
    den = np.maximum(np.maximum(np.abs(ref), mag), 1e-300)
    return float(np.max(np.abs(x - ref) / den)) if x.size else 0.0
