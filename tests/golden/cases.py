"""Loader for the golden vectors written by gen_golden.mjs from the reference's
src/physics/mortonTree.js, and the integer-exact input recipe used by the
large (digest-only) cases."""
import hashlib
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def manifest():
    with open(os.path.join(HERE, "manifest.json")) as f:
        return json.load(f)


def cases(kind="all"):
    out = []
    for c in manifest()["cases"]:
        if kind == "stored" and not c.get("stored"):
            continue
        if kind == "digest" and c.get("stored"):
            continue
        out.append(c)
    return out


def load(name, suffix, dtype):
    return np.fromfile(os.path.join(HERE, f"{name}.{suffix}"), dtype=dtype)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def input_of(case):
    """Stride-8 (or stride-4) f32 particle array of a case."""
    if case.get("recipe"):
        r = case["recipe"]
        assert r["kind"] == "exactCluster"
        return exact_cluster(case["N"], r["seed"], r["stride"])
    return load(case.get("input", case["name"]), "in.f32", np.float32)


class _XorShift32:
    """xorshift32, identical to rng() in gen_golden.mjs."""

    def __init__(self, seed):
        self.s = (seed & 0xFFFFFFFF) or 1

    def __call__(self):
        s = self.s
        s ^= (s << 13) & 0xFFFFFFFF
        s ^= s >> 17
        s ^= (s << 5) & 0xFFFFFFFF
        self.s = s
        return s


def exact_cluster(n, seed, stride):
    """exactCluster() of gen_golden.mjs.  Every value is a dyadic rational
    computed exactly in double and rounded once to f32, so numpy reproduces
    the JS Float32Array bit-for-bit."""
    r = _XorShift32(seed)
    a = np.zeros(n * stride, np.float32)
    centres = []
    for _ in range(6):
        centres.append([((r() >> 24) - 128) / 16, ((r() >> 24) - 128) / 16,
                        ((r() >> 26) - 32) / 16])
    for i in range(n):
        c = centres[r() % 6]
        sc = ((r() >> 29) + 1) * ((r() >> 29) + 1)
        for k in range(3):
            off = (((r() >> 16) - 32768) * sc) / 1048576
            a[i * stride + k] = np.float32(c[k] + off)
        a[i * stride + 3] = np.float32(((r() >> 20) + 1) / 4096)
        if stride == 8:
            for k in range(4, 7):
                a[i * stride + k] = np.float32(((r() >> 16) - 32768) / 65536)
            a[i * stride + 7] = -1.0 if r() % 10 < 3 else 0.0
    return a
