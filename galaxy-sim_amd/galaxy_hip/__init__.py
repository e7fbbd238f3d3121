"""Python host mirror of the reference's gravity-path call surface over the
C ABI of libgalaxy_hip.so (include/galaxy_hip.h).

  NBodyGPU            src/physics/NBodyGPU.js:14-369 (one galaxy on one GPU)
  GalaxyPool          many galaxies in fixed-capacity slots, every stage one
                      launch over all slots (no reference counterpart)
  GalaxyArena         a GalaxyPool whose slots take their ranges from a
                      byte-budget arena when filled (no reference counterpart)
  compute_morton_codes, radix_sort_indices, MortonTree, sort_and_build
                      src/physics/mortonTree.js:43-224 (executed on the GPU)

The HIP library is mandatory: importing this module on a machine without the
built library raises, and every call that needs a GPU raises GalaxyError when
none is usable.  There is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("GALAXY_HIP_LIB") or os.path.join(_PKG, "lib", "libgalaxy_hip.so")

GS_OK = 0
GS_E_INVALID = -1
GS_E_DEVICE = -2
GS_E_NOMEM = -3
GS_E_NODE_POOL = -4
GS_E_STATE = -5
GS_E_STACK = -6
GS_E_FULL = -7

# device bytes an arena charges per particle of a slot's (1024-aligned) range
# and per node of its 2n + 4096 node range
ARENA_PARTICLE_BYTES = 112
ARENA_NODE_BYTES = 36

FLAG_EXACT_FORCE = 0x1
FLAG_TIMING = 0x2
FLAG_NO_GRAPH = 0x4

RED_NAMES = ("ekin", "epot", "ehalo", "px", "py", "pz", "espec", "mass")
STAGE_NAMES = ("encode", "sort", "build", "refresh", "walk", "step")


class GalaxyError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"galaxy_hip error {code}: {msg}")
        self.code = code


class Config(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint32), ("bucket", ctypes.c_uint32),
                ("max_nodes", ctypes.c_uint64), ("flags", ctypes.c_uint32),
                ("device", ctypes.c_int32), ("walk_wg", ctypes.c_uint32),
                ("radix_bits", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 4)]


class PoolConfig(ctypes.Structure):
    _fields_ = [("nslots", ctypes.c_uint32), ("capacity", ctypes.c_void_p),
                ("bucket", ctypes.c_void_p), ("flags", ctypes.c_uint32),
                ("device", ctypes.c_int32), ("walk_wg", ctypes.c_uint32),
                ("radix_bits", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 4)]


class ArenaConfig(ctypes.Structure):
    _fields_ = [("budget_bytes", ctypes.c_uint64), ("nslots", ctypes.c_uint32),
                ("flags", ctypes.c_uint32), ("device", ctypes.c_int32),
                ("walk_wg", ctypes.c_uint32), ("radix_bits", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32 * 5)]


class Params(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("G", ctypes.c_double), ("epsilon", ctypes.c_double),
                ("theta", ctypes.c_double), ("halo_mass", ctypes.c_double),
                ("halo_a", ctypes.c_double), ("gas_drag", ctypes.c_double),
                ("time", ctypes.c_double), ("aging_rate", ctypes.c_double),
                ("sf_density", ctypes.c_double), ("sf_eff", ctypes.c_double),
                ("sf_dt", ctypes.c_double), ("halo_count", ctypes.c_uint32),
                ("pad_", ctypes.c_uint32), ("halo_centers", (ctypes.c_double * 3) * 2)]


_lib = None
_vp = ctypes.c_void_p
_u64 = ctypes.c_uint64
_u32 = ctypes.c_uint32


def lib():
    """Load libgalaxy_hip.so (built by galaxy_hip/build.py); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    sigs = {
        "gs_version": (ctypes.c_char_p, []),
        "gs_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
        "gs_params_default": (None, [ctypes.POINTER(Params)]),
        "gs_create": (ctypes.c_int, [ctypes.POINTER(Config), ctypes.POINTER(_vp)]),
        "gs_destroy": (None, [_vp]),
        "gs_last_error": (ctypes.c_char_p, [_vp]),
        "gs_upload": (ctypes.c_int, [_vp, _vp, _u64]),
        "gs_set_params": (ctypes.c_int, [_vp, ctypes.POINTER(Params)]),
        "gs_update_central_mass": (ctypes.c_int, [_vp, ctypes.c_float, _u32]),
        "gs_encode": (ctypes.c_int, [_vp]),
        "gs_sort": (ctypes.c_int, [_vp]),
        "gs_refresh": (ctypes.c_int, [_vp]),
        "gs_force": (ctypes.c_int, [_vp]),
        "gs_step": (ctypes.c_int, [_vp, _u32, _u32]),
        "gs_walk_step": (ctypes.c_int, [_vp]),
        "gs_radix_sort_indices_noalloc": (ctypes.c_int, [_vp, _u32, _vp, _vp]),
        "gs_sync": (ctypes.c_int, [_vp]),
        "gs_apply_permutation": (ctypes.c_int, [_vp, _vp, _u64]),
        "gs_upload_tree": (ctypes.c_int, [_vp, _vp, _u64, _vp, _u32]),
        "gs_download_particles": (ctypes.c_int, [_vp, _vp, _u64]),
        "gs_read_positions": (ctypes.c_int, [_vp, _vp, _u64]),
        "gs_download_codes": (ctypes.c_int, [_vp, _vp, _u64, ctypes.c_int]),
        "gs_download_perm": (ctypes.c_int, [_vp, _vp, _u64]),
        "gs_download_nodes": (ctypes.c_int, [_vp, _vp, _u64, ctypes.POINTER(_u64)]),
        "gs_tree_info": (ctypes.c_int, [_vp, ctypes.POINTER(_u64), _vp,
                                        ctypes.POINTER(_u32), ctypes.POINTER(ctypes.c_double)]),
        "gs_download_energy": (ctypes.c_int, [_vp, _vp, _u64]),
        "gs_download_accel": (ctypes.c_int, [_vp, _vp, _vp, _u64]),
        "gs_reduce": (ctypes.c_int, [_vp, _vp]),
        "gs_stage_times": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int]),
        "gs_count_interactions": (ctypes.c_int, [_vp, ctypes.POINTER(_u64), _vp]),
        "gs_create_pool": (ctypes.c_int, [ctypes.POINTER(PoolConfig), ctypes.POINTER(_vp)]),
        "gs_pool_info": (ctypes.c_int, [_vp, ctypes.POINTER(_u32), ctypes.POINTER(_u64)]),
        "gs_create_arena": (ctypes.c_int, [ctypes.POINTER(ArenaConfig), ctypes.POINTER(_vp)]),
        "gs_arena_info": (ctypes.c_int, [_vp, ctypes.POINTER(_u64), ctypes.POINTER(_u64),
                                         ctypes.POINTER(_u64)]),
        "gs_slot_upload": (ctypes.c_int, [_vp, _u32, _vp, _u64]),
        "gs_slot_clear": (ctypes.c_int, [_vp, _u32]),
        "gs_slot_count": (ctypes.c_int, [_vp, _u32, ctypes.POINTER(_u32)]),
        "gs_slot_set_params": (ctypes.c_int, [_vp, _u32, ctypes.POINTER(Params)]),
        "gs_slot_set_tree": (ctypes.c_int, [_vp, _u32, _u32, _u64]),
        "gs_slot_download_particles": (ctypes.c_int, [_vp, _u32, _vp, _u64]),
        "gs_slot_download_perm": (ctypes.c_int, [_vp, _u32, _vp, _u64]),
        "gs_slot_tree_info": (ctypes.c_int, [_vp, _u32, ctypes.POINTER(_u64), _vp,
                                             ctypes.POINTER(_u32),
                                             ctypes.POINTER(ctypes.c_double)]),
        "gs_slot_download_nodes": (ctypes.c_int, [_vp, _u32, _vp, _u64, ctypes.POINTER(_u64)]),
        "gs_slot_download_energy": (ctypes.c_int, [_vp, _u32, _vp, _u64]),
        "gs_slot_reduce": (ctypes.c_int, [_vp, _u32, _vp]),
        "gs_upload_stellar": (ctypes.c_int, [_vp, _vp, _u64]),
        "gs_download_stellar": (ctypes.c_int, [_vp, _vp, _u64]),
        "gs_slot_upload_stellar": (ctypes.c_int, [_vp, _u32, _vp, _u64]),
        "gs_slot_download_stellar": (ctypes.c_int, [_vp, _u32, _vp, _u64]),
        "gs_star_formation": (ctypes.c_int, [_vp]),
        "gs_locate": (ctypes.c_int, [_vp, _u32, ctypes.POINTER(_u32), _vp]),
        "gs_slot_locate": (ctypes.c_int, [_vp, _u32, _u32, ctypes.POINTER(_u32), _vp]),
        "gs_compute_morton_codes": (ctypes.c_int, [_vp, _vp, _u32, _u32, _vp,
                                                   ctypes.POINTER(ctypes.c_double)]),
        "gs_radix_sort_indices": (ctypes.c_int, [_vp, _vp, _u32, _vp]),
        "gs_tree_build": (ctypes.c_int, [_vp, _vp, _vp, _vp, _u32, ctypes.c_double, _vp, _u64,
                                         ctypes.POINTER(_u64), _vp, ctypes.POINTER(_u32)]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def exported_symbols():
    return [n for n in dir(lib()) if n.startswith("gs_")]


def device_count():
    n = ctypes.c_int(0)
    lib().gs_device_count(ctypes.byref(n))
    return n.value


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def make_params(dt, G, epsilon, theta=0.5, haloMass=0.0, haloA=1.0, gasDrag=0.0, time=0.0,
                agingRate=25.0, sfDensity=1e9, sfEff=0.0, sfDt=0.0, haloCenters=()):
    """The updateParams dictionary (NBodyGPU.js:253-275) as a gs_params struct."""
    p = Params(dt=dt, G=G, epsilon=epsilon, theta=theta, halo_mass=haloMass, halo_a=haloA,
               gas_drag=gasDrag, time=time, aging_rate=agingRate, sf_density=sfDensity,
               sf_eff=sfEff, sf_dt=sfDt, halo_count=min(len(haloCenters), 2))
    for h, c in enumerate(list(haloCenters)[:2]):
        for k in range(3):
            p.halo_centers[h][k] = float(c[k])
    return p


class NBodyGPU:
    """One galaxy resident on one GPU (NBodyGPU.js:14-369).

    particles: float32 array of N x 8 records [x,y,z,m, vx,vy,vz,pad].
    """

    def __init__(self, particles, bucket=16, max_nodes=0, exact_force=False, timing=False,
                 device=-1, walk_wg=0, radix_bits=0, graph=True):
        particles = np.ascontiguousarray(particles, dtype=np.float32).reshape(-1, 8)
        self.N = particles.shape[0]
        cfg = Config(n=self.N, bucket=bucket, max_nodes=max_nodes,
                     flags=(FLAG_EXACT_FORCE if exact_force else 0) |
                     (FLAG_TIMING if timing else 0) | (0 if graph else FLAG_NO_GRAPH),
                     device=device, walk_wg=walk_wg, radix_bits=radix_bits)
        h = ctypes.c_void_p()
        rc = lib().gs_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != GS_OK:
            raise GalaxyError(rc, "gs_create failed (no usable gfx950 device, bad config or "
                              "out of device memory)")
        self._h = h
        self.ready = True
        self._call("gs_upload", _ptr(particles), self.N)
        self.params = None

    # -- plumbing --
    def _call(self, name, *args):
        rc = getattr(lib(), name)(self._h, *args)
        if rc != GS_OK:
            raise GalaxyError(rc, lib().gs_last_error(self._h).decode())
        return rc

    def destroy(self):
        if getattr(self, "_h", None):
            lib().gs_destroy(self._h)
            self._h = None
            self.ready = False

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()

    # -- NBodyGPU surface --
    def upload(self, particles):
        particles = np.ascontiguousarray(particles, dtype=np.float32).reshape(-1, 8)
        self._call("gs_upload", _ptr(particles), particles.shape[0])

    def update_params(self, dt, G, epsilon, theta=0.5, haloMass=0.0, haloA=1.0, gasDrag=0.0,
                      time=0.0, agingRate=25.0, sfDensity=1e9, sfEff=0.0, sfDt=0.0,
                      haloCenters=()):
        """NBodyGPU.updateParams (NBodyGPU.js:253-275)."""
        p = make_params(dt, G, epsilon, theta, haloMass, haloA, gasDrag, time, agingRate,
                        sfDensity, sfEff, sfDt, haloCenters)
        self._call("gs_set_params", ctypes.byref(p))
        self.params = p

    def apply_permutation(self, perm):
        """applyPermutation (NBodyGPU.js:225-238): gather the particle records
        through perm[new] = old into the other ping-pong buffer."""
        perm = np.ascontiguousarray(perm, dtype=np.uint32)
        self._call("gs_apply_permutation", _ptr(perm), perm.size)

    def upload_tree(self, words, num_nodes, level_ranges):
        """uploadBHTree (NBodyGPU.js:288-301): a host-built tree in the 64-byte
        node layout for the current particle order."""
        words = np.ascontiguousarray(words, dtype=np.uint32).reshape(-1)
        lr = np.ascontiguousarray(np.asarray(level_ranges, np.uint32).reshape(-1))
        if int(num_nodes) < 1 or words.size < 16 * int(num_nodes):
            raise ValueError("tree words must hold 16 x num_nodes uint32 values")
        if lr.size % 2 or lr.size > 32:
            raise ValueError("level_ranges must be at most 16 [start, count] pairs")
        self._call("gs_upload_tree", _ptr(words), int(num_nodes), _ptr(lr) if lr.size else None,
                   lr.size // 2)

    def upload_stellar(self, stellar):
        """The stellar records (NBodyGPU.js:83-90): N x 4 f32 [mass, birth Myr,
        seed, dust] in the current particle order."""
        stellar = np.ascontiguousarray(stellar, dtype=np.float32).reshape(-1, 4)
        self._call("gs_upload_stellar", _ptr(stellar), stellar.shape[0])

    def stellar(self):
        """stellarBuffer (NBodyGPU.js:241-243), in the current particle order."""
        out = np.empty((self.N, 4), np.float32)
        self._call("gs_download_stellar", _ptr(out), self.N)
        return out

    def star_formation(self):
        """encodeStarFormation (NBodyGPU.js:277-285)."""
        self._call("gs_star_formation")

    def locate(self, old_index):
        """remapIndex (main.js:436-440): current index and (x, y, z, m) of the
        particle that had old_index before the last sort."""
        ni = _u32(0)
        pos = np.zeros(4, np.float32)
        self._call("gs_locate", int(old_index), ctypes.byref(ni), _ptr(pos))
        return ni.value, pos

    def update_central_mass(self, mass, idx=0):
        self._call("gs_update_central_mass", float(mass), int(idx))

    def encode(self):
        self._call("gs_encode")

    def sort(self):
        """Topology rebuild: codes, stable sort, gather, BFS build."""
        self._call("gs_sort")

    def refresh(self):
        """encodeRefresh (NBodyGPU.js:308-319)."""
        self._call("gs_refresh")

    def force(self):
        self._call("gs_force")

    def step(self, nsteps=1, sort_every=1):
        """nsteps x (refresh + step) with a rebuild every sort_every steps."""
        self._call("gs_step", nsteps, sort_every)

    def walk_step(self):
        """step (NBodyGPU.js:321-329): walk + integrate on the tree's current
        COMs, no refresh (encodeRefresh is refresh())."""
        self._call("gs_walk_step")

    def sync(self):
        self._call("gs_sync")

    # -- downloads --
    def particles(self):
        out = np.empty((self.N, 8), np.float32)
        self._call("gs_download_particles", _ptr(out), self.N)
        return out

    def read_positions(self):
        out = np.empty((self.N, 4), np.float32)
        self._call("gs_read_positions", _ptr(out), self.N)
        return out

    def codes(self, sorted=False):
        out = np.empty(self.N, np.uint32)
        self._call("gs_download_codes", _ptr(out), self.N, 1 if sorted else 0)
        return out

    def perm(self):
        out = np.empty(self.N, np.uint32)
        self._call("gs_download_perm", _ptr(out), self.N)
        return out

    def tree_info(self):
        nn = _u64(0)
        lr = np.zeros(32, np.uint32)
        nl = _u32(0)
        rs = ctypes.c_double(0)
        self._call("gs_tree_info", ctypes.byref(nn), _ptr(lr), ctypes.byref(nl), ctypes.byref(rs))
        return {"num_nodes": nn.value, "num_levels": nl.value, "root_size": rs.value,
                "level_ranges": [[int(lr[2 * k]), int(lr[2 * k + 1])]
                                 for k in range(min(nl.value, 16))]}

    def nodes(self):
        info = self.tree_info()
        out = np.empty(max(info["num_nodes"], 1) * 16, np.uint32)
        nn = _u64(0)
        self._call("gs_download_nodes", _ptr(out), info["num_nodes"], ctypes.byref(nn))
        return out[: nn.value * 16]

    def energy(self):
        out = np.empty(self.N, np.float32)
        self._call("gs_download_energy", _ptr(out), self.N)
        return out

    def accel(self):
        acc = np.empty((self.N, 4), np.float32)
        dens = np.empty(self.N, np.float32)
        self._call("gs_download_accel", _ptr(acc), _ptr(dens), self.N)
        return acc, dens

    def reduce(self):
        out = np.zeros(8, np.float64)
        self._call("gs_reduce", _ptr(out))
        return dict(zip(RED_NAMES, out.tolist()))

    def stage_times(self, reset=False):
        ms = np.zeros(6, np.float64)
        cnt = np.zeros(6, np.uint64)
        self._call("gs_stage_times", _ptr(ms), _ptr(cnt), 1 if reset else 0)
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(STAGE_NAMES)}

    def count_interactions(self, per_particle=False):
        tot = _u64(0)
        pp = np.empty(self.N, np.uint32) if per_particle else None
        self._call("gs_count_interactions", ctypes.byref(tot),
                   _ptr(pp) if pp is not None else None)
        return (tot.value, pp) if per_particle else tot.value


def track_halo_centers(centers, targets, k=0.1):
    """Halo-centre EMA of the frame loop (main.js:449-457): each centre moves
    the fraction k toward its target (the tracked central black hole), in
    double precision, in place; returns the centres."""
    for h in range(min(len(centers), len(targets))):
        for c in range(3):
            centers[h][c] += (float(targets[h][c]) - centers[h][c]) * k
    return centers


class GalaxyPool:
    """Many galaxies on one GPU in fixed-capacity slots.

    Every stage of a step is one launch over all slots and a step replays one
    captured graph, so a pool of small galaxies runs at the launch count of one
    galaxy.  Each slot holds 0 .. capacity particles with its own parameters
    and leaf capacity; its results equal those of the same galaxy run alone on
    an NBodyGPU.
    """

    def __init__(self, capacities, buckets=None, exact_force=False, timing=False, device=-1,
                 walk_wg=0, graph=True, radix_bits=0):
        caps = np.ascontiguousarray(capacities, dtype=np.uint32)
        self.nslots = int(caps.size)
        self.capacity = caps.copy()
        bk = None if buckets is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(buckets, np.uint32), caps.shape), dtype=np.uint32)
        cfg = PoolConfig(nslots=self.nslots, capacity=caps.ctypes.data,
                         bucket=bk.ctypes.data if bk is not None else None,
                         flags=(FLAG_EXACT_FORCE if exact_force else 0) |
                         (FLAG_TIMING if timing else 0) | (0 if graph else FLAG_NO_GRAPH),
                         device=device, walk_wg=walk_wg, radix_bits=radix_bits)
        h = ctypes.c_void_p()
        rc = lib().gs_create_pool(ctypes.byref(cfg), ctypes.byref(h))
        if rc != GS_OK:
            raise GalaxyError(rc, "gs_create_pool failed (no usable gfx950 device, bad config "
                              "or out of device memory)")
        self._h = h
        self.counts = np.zeros(self.nslots, np.int64)

    _call = NBodyGPU._call
    destroy = NBodyGPU.destroy
    __del__ = NBodyGPU.__del__
    __enter__ = NBodyGPU.__enter__
    __exit__ = NBodyGPU.__exit__

    def info(self):
        ns = _u32(0)
        nb = _u64(0)
        self._call("gs_pool_info", ctypes.byref(ns), ctypes.byref(nb))
        return {"nslots": ns.value, "device_bytes": nb.value}

    def upload(self, slot, particles, **params):
        """Place a galaxy (n x 8 records) in a slot; params as updateParams."""
        particles = np.ascontiguousarray(particles, dtype=np.float32).reshape(-1, 8)
        self._call("gs_slot_upload", slot, _ptr(particles), particles.shape[0])
        self.counts[slot] = particles.shape[0]
        if params:
            self.set_params(slot, **params)

    def clear(self, slot):
        self._call("gs_slot_clear", slot)
        self.counts[slot] = 0

    def upload_stellar(self, slot, stellar):
        stellar = np.ascontiguousarray(stellar, dtype=np.float32).reshape(-1, 4)
        self._call("gs_slot_upload_stellar", slot, _ptr(stellar), stellar.shape[0])

    def stellar(self, slot):
        n = int(self.counts[slot])
        out = np.empty((n, 4), np.float32)
        if n:
            self._call("gs_slot_download_stellar", slot, _ptr(out), n)
        return out

    def star_formation(self):
        self._call("gs_star_formation")

    def count(self, slot):
        n = _u32(0)
        self._call("gs_slot_count", slot, ctypes.byref(n))
        return n.value

    def set_params(self, slot, **kw):
        p = make_params(**kw)
        self._call("gs_slot_set_params", slot, ctypes.byref(p))

    def set_tree(self, slot, bucket, max_nodes=0):
        self._call("gs_slot_set_tree", slot, bucket, max_nodes)

    def sort(self):
        self._call("gs_sort")

    def refresh(self):
        self._call("gs_refresh")

    def force(self):
        self._call("gs_force")

    def step(self, nsteps=1, sort_every=1):
        self._call("gs_step", nsteps, sort_every)

    def sync(self):
        self._call("gs_sync")

    def particles(self, slot):
        n = int(self.counts[slot])
        out = np.empty((n, 8), np.float32)
        if n:
            self._call("gs_slot_download_particles", slot, _ptr(out), n)
        return out

    def perm(self, slot):
        n = int(self.counts[slot])
        out = np.empty(n, np.uint32)
        self._call("gs_slot_download_perm", slot, _ptr(out), n)
        return out

    def tree_info(self, slot):
        nn = _u64(0)
        lr = np.zeros(32, np.uint32)
        nl = _u32(0)
        rs = ctypes.c_double(0)
        self._call("gs_slot_tree_info", slot, ctypes.byref(nn), _ptr(lr), ctypes.byref(nl),
                   ctypes.byref(rs))
        return {"num_nodes": nn.value, "num_levels": nl.value, "root_size": rs.value,
                "level_ranges": [[int(lr[2 * k]), int(lr[2 * k + 1])]
                                 for k in range(min(nl.value, 16))]}

    def nodes(self, slot):
        info = self.tree_info(slot)
        out = np.empty(max(info["num_nodes"], 1) * 16, np.uint32)
        nn = _u64(0)
        self._call("gs_slot_download_nodes", slot, _ptr(out), info["num_nodes"],
                   ctypes.byref(nn))
        return out[: nn.value * 16]

    def energy(self, slot):
        n = int(self.counts[slot])
        out = np.empty(n, np.float32)
        self._call("gs_slot_download_energy", slot, _ptr(out), n)
        return out

    def reduce(self, slot):
        out = np.zeros(8, np.float64)
        self._call("gs_slot_reduce", slot, _ptr(out))
        return dict(zip(RED_NAMES, out.tolist()))

    def stage_times(self, reset=False):
        return NBodyGPU.stage_times(self, reset)

    def count_interactions(self):
        tot = _u64(0)
        self._call("gs_count_interactions", ctypes.byref(tot), None)
        return tot.value


def arena_charge(n):
    """Bytes an arena takes for a galaxy of n particles: its particle range
    (n rounded up to 1024) and its node range (2n + 4096, mortonTree.js:88)."""
    n = int(n)
    if n <= 0:
        return 0
    return -(-n // 1024) * 1024 * ARENA_PARTICLE_BYTES + (2 * n + 4096) * ARENA_NODE_BYTES


class GalaxyArena(GalaxyPool):
    """A slot pool whose device state is sized by a byte budget: `nslots`
    slots start empty, upload() places a galaxy in a free range of the arena
    (raising ArenaFull, with the slot left empty, when no range holds it) and
    clear() gives the range back.  Steps launch over the occupied slots only.
    Every GalaxyPool call works on its slots except the stellar records."""

    def __init__(self, budget_bytes, nslots, exact_force=False, timing=False, device=-1,
                 walk_wg=0, graph=True, radix_bits=0):
        cfg = ArenaConfig(budget_bytes=int(budget_bytes), nslots=int(nslots),
                          flags=(FLAG_EXACT_FORCE if exact_force else 0) |
                          (FLAG_TIMING if timing else 0) | (0 if graph else FLAG_NO_GRAPH),
                          device=device, walk_wg=walk_wg, radix_bits=radix_bits)
        h = ctypes.c_void_p()
        rc = lib().gs_create_arena(ctypes.byref(cfg), ctypes.byref(h))
        if rc != GS_OK:
            raise GalaxyError(rc, "gs_create_arena failed (no usable gfx950 device, a budget "
                              "below one range, bad config or out of device memory)")
        self._h = h
        self.nslots = int(nslots)
        self.budget = int(budget_bytes)
        self.capacity = None
        self.counts = np.zeros(self.nslots, np.int64)

    def upload(self, slot, particles, **params):
        try:
            GalaxyPool.upload(self, slot, particles, **params)
        except GalaxyError as e:
            self.counts[slot] = 0
            if e.code == GS_E_FULL:
                raise ArenaFull(e.code, str(e)) from None
            raise

    def arena_info(self):
        dev, used, fit = _u64(0), _u64(0), _u64(0)
        self._call("gs_arena_info", ctypes.byref(dev), ctypes.byref(used), ctypes.byref(fit))
        return {"device_bytes": dev.value, "used_bytes": used.value, "largest_fit": fit.value,
                "budget_bytes": self.budget}


class ArenaFull(GalaxyError):
    """No free range of the arena holds the galaxy (its budget is spent)."""


# ---- mortonTree.js surface (device-executed on host arrays) ----------------
_scratch = {}


def _scratch_sim(n, radix_bits=0):
    """A resident handle sized for n particles, reused across calls."""
    sim = _scratch.get((n, radix_bits))
    if sim is None:
        sim = NBodyGPU(np.zeros((n, 8), np.float32), radix_bits=radix_bits)
        _scratch.clear()
        _scratch[(n, radix_bits)] = sim
    return sim


def compute_morton_codes(pos, N, stride, codes=None):
    """computeMortonCodes (mortonTree.js:43-62) -> (codes, rootSize)."""
    pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1)
    out = np.empty(N, np.uint32) if codes is None else codes
    rs = ctypes.c_double(0)
    sim = _scratch_sim(N)
    sim._call("gs_compute_morton_codes", _ptr(pos), N, stride, _ptr(out), ctypes.byref(rs))
    return out, rs.value


def radix_sort_indices(codes, N=None, radix_bits=0):
    """radixSortIndices (mortonTree.js:68-84): the stable sorted index array
    (codes < 2^30; radix_bits 8 or 10 selects the device digit width)."""
    codes = np.ascontiguousarray(codes, dtype=np.uint32)
    N = codes.size if N is None else N
    out = np.empty(N, np.uint32)
    _scratch_sim(N, radix_bits)._call("gs_radix_sort_indices", _ptr(codes), N, _ptr(out))
    return out


class MortonTree:
    """MortonTree (mortonTree.js:86-185) built on the GPU."""

    def __init__(self, max_particles, bucket=16, max_nodes=None):
        self.max_particles = max_particles
        self.maxNodes = 2 * max_particles + 4096 if max_nodes is None else max_nodes
        self.bucket = bucket
        self.numNodes = 0
        self.levelRanges = []
        self.words = np.zeros(0, np.uint32)

    def build(self, codes, pos, masses, N, rootSize):
        codes = np.ascontiguousarray(codes, np.uint32)
        pos = np.ascontiguousarray(pos, np.float32)
        masses = np.ascontiguousarray(masses, np.float32)
        with NBodyGPU(np.zeros((N, 8), np.float32), bucket=self.bucket,
                      max_nodes=self.maxNodes) as sim:
            cap = 2 * N + 4096 if self.maxNodes <= 2 * N + 4096 else self.maxNodes
            nodes = np.zeros(cap * 16, np.uint32)
            nn = _u64(0)
            lr = np.zeros(32, np.uint32)
            nl = _u32(0)
            sim._call("gs_tree_build", _ptr(codes), _ptr(pos), _ptr(masses), N, float(rootSize),
                      _ptr(nodes), cap, ctypes.byref(nn), _ptr(lr), ctypes.byref(nl))
        self.numNodes = nn.value
        self.levelRanges = [[int(lr[2 * k]), int(lr[2 * k + 1])] for k in range(nl.value)]
        self.words = nodes[: nn.value * 16]
        return self

    @property
    def f32(self):
        return self.words.view(np.float32)

    @property
    def i32(self):
        return self.words.view(np.int32)

    @property
    def byteLength(self):
        return self.numNodes * 64


def sort_and_build(particle_data, N, track_idx2=-1, bucket=16):
    """sortAndBuild (mortonTree.js:192-224) on the GPU (particle records only;
    the 16-byte stellar render record is not part of this library)."""
    particle_data = np.ascontiguousarray(particle_data, np.float32).reshape(-1, 8)
    with NBodyGPU(particle_data, bucket=bucket) as sim:
        sim.sort()
        perm = sim.perm()
        sorted_ = sim.particles()
        codes = sim.codes(sorted=True)
        info = sim.tree_info()
        sim.refresh()
        nodes = sim.nodes()
    out = {"sorted": sorted_, "perm": perm, "codes": codes, "nodes": nodes,
           "num_nodes": info["num_nodes"], "level_ranges": info["level_ranges"],
           "root_size": info["root_size"], "masses": sorted_[:, 3].copy()}
    where0 = np.nonzero(perm == 0)[0]
    out["central_idx"] = int(where0[0]) if where0.size else 0
    if track_idx2 >= 0:
        w2 = np.nonzero(perm == track_idx2)[0]
        out["central_idx2"] = int(w2[0]) if w2.size else -1
    else:
        out["central_idx2"] = -1
    return out
