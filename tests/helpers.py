"""Shared helpers of the test suite (no GPU needed to import)."""
import sys

import numpy as np

import pyoracle as O


def orc_params(p):
    """sph_amd.Params -> oracle params (same fields, the oracle's names)."""
    return O.from_sph(p)


def announce(*what):
    """Print (and flush) the case a GPU step is about to run."""
    print("[case]", *what, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def assert_bitexact(got, want, what):
    g, w = bits(got), bits(want)
    # +0 and -0 compare equal as values; the bit test treats them as different
    diff = np.nonzero((g != w) & ~((got == 0) & (want == 0)))
    n_bad = diff[0].size
    assert n_bad == 0, (f"{what}: {n_bad} of {g.size} values differ; first at {diff[0][:5]}: "
                        f"got {np.asarray(got)[diff][:5]}, want {np.asarray(want)[diff][:5]}")
