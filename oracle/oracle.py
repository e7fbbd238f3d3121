"""ctypes view of the CPU oracle (liboracle.so) — TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
and there only as the checker or the timed CPU baseline.  The HIP library
never links, loads or calls anything under oracle/.

What each function restates (reference file:line):
  compute_morton_codes  src/physics/mortonTree.js:43-62
  radix_sort_indices    src/physics/mortonTree.js:68-84
  tree_build            src/physics/mortonTree.js:107-182
  sort_and_build        src/physics/mortonTree.js:192-224
  walk                  src/physics/shaders/nbody_bh.wgsl:64-181
  refresh_f32           src/physics/shaders/nbody_bh.wgsl:288-313
  refresh_f64           src/physics/mortonTree.js:129-137,166-181 (FP64 order)
  pack_params           src/physics/NBodyGPU.js:253-275
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

u32p = ctypes.POINTER(ctypes.c_uint32)
f32p = ctypes.POINTER(ctypes.c_float)
f64p = ctypes.POINTER(ctypes.c_double)
i64p = ctypes.POINTER(ctypes.c_int64)


class Params(ctypes.Structure):
    """or_params: the 96-byte WGSL Params uniform plus exp(-gasDrag*dt)."""

    _fields_ = [
        ("dt", ctypes.c_float), ("G", ctypes.c_float), ("eps2", ctypes.c_float),
        ("n", ctypes.c_uint32), ("theta", ctypes.c_float),
        ("halo_mass", ctypes.c_float), ("halo_a", ctypes.c_float),
        ("gas_drag", ctypes.c_float), ("time", ctypes.c_float),
        ("aging_rate", ctypes.c_float), ("sf_density", ctypes.c_float),
        ("sf_eff", ctypes.c_float), ("sf_dt", ctypes.c_float),
        ("halo_count", ctypes.c_uint32), ("pad0", ctypes.c_float),
        ("pad1", ctypes.c_float), ("halo_c1", ctypes.c_float * 4),
        ("halo_c2", ctypes.c_float * 4), ("drag_damp", ctypes.c_float),
    ]


def default_threads():
    """Worker threads for the multi-threaded walks: every core this process may
    run on, capped by OMP_NUM_THREADS when that is set (a shared host's share)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = os.environ.get("OMP_NUM_THREADS", "")
    if cap.isdigit() and int(cap) > 0:
        n = min(n, int(cap))
    return max(n, 1)


def build():
    """Compile liboracle.so with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        L.or_part1by2.restype = ctypes.c_uint32
        L.or_pcg.restype = ctypes.c_uint32
        L.or_pcg.argtypes = [ctypes.c_uint32]
        L.or_rand01.restype = ctypes.c_float
        L.or_rand01.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        L.or_imf_sample.restype = ctypes.c_float
        L.or_imf_sample.argtypes = [ctypes.c_float, ctypes.c_float]
        L.or_part1by2.argtypes = [ctypes.c_uint32]
        L.or_js_toint32.restype = ctypes.c_int32
        L.or_js_toint32.argtypes = [ctypes.c_double]
        L.or_compute_morton_codes.restype = ctypes.c_double
        L.or_compute_morton_codes.argtypes = [f32p, ctypes.c_uint32, ctypes.c_uint32, u32p]
        L.or_radix_sort_indices.restype = None
        L.or_radix_sort_indices.argtypes = [u32p, u32p, ctypes.c_uint32]
        L.or_tree_build.restype = ctypes.c_int64
        L.or_tree_build.argtypes = [u32p, f32p, f32p, ctypes.c_uint32, ctypes.c_double,
                                    ctypes.c_uint32, ctypes.c_uint64, u32p, ctypes.c_uint64,
                                    u32p, u32p]
        L.or_sort_and_build.restype = ctypes.c_int64
        L.or_sort_and_build.argtypes = [f32p, f32p, ctypes.c_uint32, ctypes.c_int64,
                                        ctypes.c_uint32, ctypes.c_uint64, f32p, f32p, u32p,
                                        u32p, u32p, ctypes.c_uint64, u32p, u32p, f64p,
                                        i64p, i64p]
        L.or_pack_params.restype = None
        L.or_pack_params.argtypes = [ctypes.POINTER(Params)] + [ctypes.c_double] * 7 + [
            ctypes.c_uint32, ctypes.c_uint32, f64p, f64p]
        L.or_walk.restype = None
        L.or_walk.argtypes = [f32p, ctypes.c_uint32, u32p, ctypes.POINTER(Params), f32p,
                              f32p, f32p, f32p, u32p, ctypes.c_int]
        L.or_walk_subset.restype = None
        L.or_walk_subset.argtypes = [f32p, ctypes.c_uint32, u32p, ctypes.POINTER(Params),
                                     u32p, ctypes.c_uint32, f32p, u32p, ctypes.c_int]
        L.or_refresh_f32.restype = None
        L.or_refresh_f32.argtypes = [f32p, u32p, u32p, ctypes.c_uint32]
        L.or_refresh_f64.restype = None
        L.or_refresh_f64.argtypes = [f32p, u32p, ctypes.c_uint64]
        L.or_apply_permutation.restype = None
        L.or_apply_permutation.argtypes = [f32p, u32p, ctypes.c_uint32, ctypes.c_uint32, f32p]
        L.or_direct_sum_f64.restype = None
        L.or_direct_sum_f64.argtypes = [f32p, ctypes.c_uint32, ctypes.c_double,
                                        ctypes.c_double, u32p, ctypes.c_uint32, f64p,
                                        ctypes.c_int]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def part1by2(v):
    return lib().or_part1by2(v)


def js_toint32(d):
    return lib().or_js_toint32(d)


def compute_morton_codes(pos, n, stride):
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    codes = np.zeros(max(n, 1), np.uint32)
    size = lib().or_compute_morton_codes(_p(pos, f32p), n, stride, _p(codes, u32p))
    return codes[:n], size


def radix_sort_indices(codes):
    codes = np.ascontiguousarray(codes, dtype=np.uint32)
    n = codes.size
    perm = np.zeros(max(n, 1), np.uint32)
    lib().or_radix_sort_indices(_p(codes, u32p), _p(perm, u32p), n)
    return perm[:n]


def node_capacity(n):
    """Pool size of MortonTree (mortonTree.js:88)."""
    return 2 * n + 4096


def tree_build(sorted_codes, pos3, masses, root_size, bucket=16, max_nodes=None):
    n = sorted_codes.size
    cap = node_capacity(n)
    if max_nodes is None:
        max_nodes = cap
    cap = max(cap, max_nodes)
    nodes = np.zeros(cap * 16, np.uint32)
    lr = np.zeros(32, np.uint32)
    nl = np.zeros(1, np.uint32)
    nn = lib().or_tree_build(_p(np.ascontiguousarray(sorted_codes, np.uint32), u32p),
                             _p(np.ascontiguousarray(pos3, np.float32), f32p),
                             _p(np.ascontiguousarray(masses, np.float32), f32p), n,
                             root_size, bucket, max_nodes, _p(nodes, u32p), cap,
                             _p(lr, u32p), _p(nl, u32p))
    if nn < 0:
        raise RuntimeError("node pool exhausted")
    k = int(nl[0])
    return nodes[: nn * 16].copy(), [[int(lr[2 * i]), int(lr[2 * i + 1])] for i in range(min(k, 16))]


def sort_and_build(particles, n, stellar=None, track_idx2=-1, bucket=16, max_nodes=None):
    """sortAndBuild on stride-8 particles.  Returns a dict of numpy outputs."""
    particles = np.ascontiguousarray(particles, dtype=np.float32).reshape(-1)
    cap = node_capacity(n)
    if max_nodes is None:
        max_nodes = cap
    cap = max(cap, max_nodes)
    sorted_ = np.zeros(max(n, 1) * 8, np.float32)
    sst = np.zeros(max(n, 1) * 4, np.float32)
    perm = np.zeros(max(n, 1), np.uint32)
    scodes = np.zeros(max(n, 1), np.uint32)
    nodes = np.zeros(cap * 16, np.uint32)
    lr = np.zeros(32, np.uint32)
    nl = np.zeros(1, np.uint32)
    rs = np.zeros(1, np.float64)
    c1 = np.zeros(1, np.int64)
    c2 = np.zeros(1, np.int64)
    st = None if stellar is None else np.ascontiguousarray(stellar, np.float32).reshape(-1)
    nn = lib().or_sort_and_build(
        _p(particles, f32p), None if st is None else _p(st, f32p), n, track_idx2, bucket,
        max_nodes, _p(sorted_, f32p), _p(sst, f32p), _p(perm, u32p), _p(scodes, u32p),
        _p(nodes, u32p), cap, _p(lr, u32p), _p(nl, u32p), _p(rs, f64p), _p(c1, i64p),
        _p(c2, i64p))
    if nn < 0:
        raise RuntimeError("node pool exhausted")
    k = int(nl[0])
    return {
        "sorted": sorted_[: n * 8], "sorted_stellar": sst[: n * 4], "perm": perm[:n],
        "codes": scodes[:n], "nodes": nodes[: nn * 16].copy(), "num_nodes": int(nn),
        "level_ranges": [[int(lr[2 * i]), int(lr[2 * i + 1])] for i in range(min(k, 16))],
        "root_size": float(rs[0]), "central_idx": int(c1[0]), "central_idx2": int(c2[0]),
    }


def pack_params(n, dt=0.005, G=1.0, epsilon=0.02, theta=0.7, halo_mass=0.0, halo_a=1.0,
                gas_drag=0.0, halo_centers=(), time=0.0, aging_rate=25.0, sf_density=1e9,
                sf_eff=0.0, sf_dt=0.0):
    p = Params()
    c1 = (ctypes.c_double * 3)(*(halo_centers[0][:3] if len(halo_centers) > 0 else (0, 0, 0)))
    c2 = (ctypes.c_double * 3)(*(halo_centers[1][:3] if len(halo_centers) > 1 else (0, 0, 0)))
    lib().or_pack_params(ctypes.byref(p), dt, G, epsilon, theta, halo_mass, halo_a, gas_drag,
                         n, min(len(halo_centers), 2), c1, c2)
    # the remaining updateParams fields, each a Float32Array store (NBodyGPU.js:253-275)
    p.time, p.aging_rate, p.sf_density = float(time), float(aging_rate), float(sf_density)
    p.sf_eff, p.sf_dt = float(sf_eff), float(sf_dt)
    return p


def pcg(v):
    """pcg (nbody_bh.wgsl:206-210) over an array of u32."""
    v = np.ascontiguousarray(v, dtype=np.uint32)
    return np.array([lib().or_pcg(int(x)) for x in v], np.uint32)


def rand01(i, salt):
    i = np.ascontiguousarray(i, dtype=np.uint32)
    return np.array([lib().or_rand01(int(x), int(salt)) for x in i], np.float32)


def imf_sample(u1, u2):
    return np.float32(lib().or_imf_sample(float(u1), float(u2)))


def star_formation(particles, stellar, params):
    """star_formation (nbody_bh.wgsl:229-268) on copies: (particles, stellar)."""
    pr = np.ascontiguousarray(particles, dtype=np.float32).reshape(-1).copy()
    st = np.ascontiguousarray(stellar, dtype=np.float32).reshape(-1).copy()
    n = pr.size // 8
    lib().or_star_formation(_p(pr, f32p), _p(st, f32p), n, ctypes.byref(params))
    return pr.reshape(n, 8), st.reshape(n, 4)


def walk(particles, nodes, params, nthreads=None):
    """Full `main` kernel: returns (particles_out, energy, acc4, density, interactions)."""
    particles = np.ascontiguousarray(particles, dtype=np.float32).reshape(-1)
    n = particles.size // 8
    nodes = np.ascontiguousarray(nodes, dtype=np.uint32)
    pout = np.zeros(n * 8, np.float32)
    energy = np.zeros(n, np.float32)
    acc4 = np.zeros(n * 4, np.float32)
    dens = np.zeros(n, np.float32)
    inter = np.zeros(n, np.uint32)
    lib().or_walk(_p(particles, f32p), n, _p(nodes, u32p), ctypes.byref(params),
                  _p(pout, f32p), _p(energy, f32p), _p(acc4, f32p), _p(dens, f32p),
                  _p(inter, u32p), nthreads or default_threads())
    return pout, energy, acc4.reshape(n, 4), dens, inter


def walk_subset(particles, nodes, params, idx, nthreads=None):
    particles = np.ascontiguousarray(particles, dtype=np.float32).reshape(-1)
    n = particles.size // 8
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    acc4 = np.zeros(idx.size * 4, np.float32)
    inter = np.zeros(idx.size, np.uint32)
    lib().or_walk_subset(_p(particles, f32p), n, _p(np.ascontiguousarray(nodes, np.uint32), u32p),
                         ctypes.byref(params), _p(idx, u32p), idx.size, _p(acc4, f32p),
                         _p(inter, u32p), nthreads or default_threads())
    return acc4.reshape(-1, 4), inter


def refresh_f32(particles, nodes, level_ranges):
    nodes = np.array(nodes, dtype=np.uint32, copy=True)
    lr = np.zeros(32, np.uint32)
    for i, (a, b) in enumerate(level_ranges[:16]):
        lr[2 * i], lr[2 * i + 1] = a, b
    lib().or_refresh_f32(_p(np.ascontiguousarray(particles, np.float32).reshape(-1), f32p),
                         _p(nodes, u32p), _p(lr, u32p), min(len(level_ranges), 16))
    return nodes


def refresh_f64(particles, nodes):
    nodes = np.array(nodes, dtype=np.uint32, copy=True)
    lib().or_refresh_f64(_p(np.ascontiguousarray(particles, np.float32).reshape(-1), f32p),
                         _p(nodes, u32p), nodes.size // 16)
    return nodes


def apply_permutation(arr, perm, stride):
    arr = np.ascontiguousarray(arr, dtype=np.float32).reshape(-1)
    perm = np.ascontiguousarray(perm, dtype=np.uint32)
    out = np.zeros_like(arr)
    lib().or_apply_permutation(_p(arr, f32p), _p(perm, u32p), perm.size, stride, _p(out, f32p))
    return out


def direct_sum_f64(particles, G, eps2, idx, nthreads=None):
    particles = np.ascontiguousarray(particles, dtype=np.float32).reshape(-1)
    n = particles.size // 8
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    acc = np.zeros(idx.size * 3, np.float64)
    lib().or_direct_sum_f64(_p(particles, f32p), n, G, eps2, _p(idx, u32p), idx.size,
                            _p(acc, f64p), nthreads or default_threads())
    return acc.reshape(-1, 3)
