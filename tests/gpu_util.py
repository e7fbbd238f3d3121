"""Shared helpers for the GPU parity tests (oracle side runs on the host)."""
import sys

import numpy as np

from oracle import oracle


def banner(msg):
    sys.stdout.write(f"[gpu-case] {msg}\n")
    sys.stdout.flush()


def oracle_tree(particles, bucket=16, max_nodes=None):
    """sortAndBuild on the oracle, returning sorted particles and the tree."""
    n = particles.shape[0]
    return oracle.sort_and_build(particles.reshape(-1), n, bucket=bucket, max_nodes=max_nodes)


def rel_err(a, b):
    """Per-particle |a - b| / |b| over acceleration vectors."""
    num = np.linalg.norm(a.astype(np.float64) - b.astype(np.float64), axis=1)
    den = np.linalg.norm(b.astype(np.float64), axis=1)
    return num / np.maximum(den, 1e-300)


def check_force_bar(acc, racc, tol=1e-5, floor=1e-2):
    """The default path's force bar against the transliterated walk: relative
    <= tol for every particle whose reference force is at least `floor` of
    the RMS force; where the force cancels below that (saddle points between
    galaxies, symmetric lattice points), a relative bound measures the float32
    rounding of the cancelling terms rather than the force, and the bound is
    tol of the RMS force instead.  Returns (max relative error over all
    particles, max relative error where the relative bound applies, max
    absolute error / RMS, number of particles in the absolute regime)."""
    a = acc[:, :3].astype(np.float64)
    b = racc[:, :3].astype(np.float64)
    mag = np.linalg.norm(b, axis=1)
    rms = float(np.sqrt(np.mean(mag ** 2)))
    diff = np.linalg.norm(a - b, axis=1)
    big = mag >= floor * rms
    rel = diff / np.maximum(mag, 1e-300)
    rel_big = float(rel[big].max(initial=0.0))
    abs_small = float(diff[~big].max(initial=0.0)) / rms
    assert rel_big <= tol, f"relative force error {rel_big:.3e} > {tol}"
    assert abs_small <= tol, f"force error {abs_small:.3e} of the RMS force > {tol}"
    return float(rel.max()), rel_big, float(diff.max()) / rms, int((~big).sum())
