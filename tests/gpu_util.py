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
