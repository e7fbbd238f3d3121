"""Seeded synthetic initial conditions for the benchmark and test workloads.

The reference's generators (src/physics/InitialConditions.js) draw from
Math.random and cannot be reproduced; these follow the same recipes with a
counter-based SplitMix64 stream so every run sees identical particles:

  disk(n, seed)      exponential disk p(r) ~ r e^{-r} (scale 1, truncated at 5;
                     cf. InitialConditions.js:129-135), sech^2 vertical profile
                     z = 0.1 atanh(2u-1) (cf. :233-235), 20 % of the mass in a
                     Hernquist bulge a = 0.2 (r = a sqrt(u)/(1-sqrt(u)),
                     cf. :116-121, truncated at 5), Kroupa-IMF-shaped masses
                     (cf. :18-37) normalised to a total of 1, circular
                     velocities from the sampled enclosed mass (+ halo), an
                     optional gas flag (pad = -1) on a fraction of the disk.
  plummer(n, seed)   Plummer sphere a = 1, r = a / sqrt(u^{-2/3} - 1),
                     isotropic velocities from the Plummer distribution
                     function (Aarseth-Henon-Wielen rejection).
  collision(n, seed) two disks of n/2, centres x = -4 and +4, disk B inclined
                     30 degrees about x, approaching at relative speed 0.5
                     (cf. InitialConditions.js:301-363).
Records are float32 [x, y, z, m, vx, vy, vz, pad].
"""
import numpy as np

SEED_BASE = 0x6A1A5EED00000000
_GOLD = np.uint64(0x9E3779B97F4A7C15)


def _mix(z):
    z = z + _GOLD
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


class Stream:
    """Counter-based SplitMix64: value j of stream `seed` is mix(seed*phi + j)."""

    def __init__(self, seed):
        with np.errstate(over="ignore"):
            self.base = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) * _GOLD
        self.ctr = 0

    def uniform(self, size):
        j = np.arange(self.ctr, self.ctr + size, dtype=np.uint64)
        self.ctr += size
        with np.errstate(over="ignore"):
            z = _mix(self.base + j)
        return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def open_uniform(self, size):
        u = self.uniform(size)
        return np.where(u <= 0.0, 1.0 / 9007199254740992.0, u)


def kroupa(st, size, m_max=100.0):
    """Kroupa (2001) two-segment IMF, inverse CDF (InitialConditions.js:18-37)."""
    low = (0.08 ** -0.3 - 0.5 ** -0.3) / 0.3
    hi_cut = max(min(m_max, 100.0), 0.5001)
    k2 = 0.5 ** (-1.3 + 2.3)
    hi = k2 * (0.5 ** -1.3 - hi_cut ** -1.3) / 1.3
    p_low = low / (low + hi)
    u = st.uniform(size)
    v = st.uniform(size)
    a, b = 0.08 ** -0.3, 0.5 ** -0.3
    m_low = (a - v * (a - b)) ** (-1.0 / 0.3)
    a1, b1 = 0.5 ** -1.3, hi_cut ** -1.3
    m_hi = (a1 - v * (a1 - b1)) ** (-1.0 / 1.3)
    return np.where(u < p_low, m_low, m_hi)


def _truncated(st, size, draw, rmax):
    out = draw(size)
    bad = out > rmax
    while bad.any():
        out[bad] = draw(int(bad.sum()))
        bad = out > rmax
    return out


def _iso_dirs(st, size):
    ct = 2.0 * st.uniform(size) - 1.0
    ph = 2.0 * np.pi * st.uniform(size)
    s = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
    return np.stack([s * np.cos(ph), s * np.sin(ph), ct], axis=1)


def _enclosed_vc(r_all, m_all, r_query, G=1.0, halo_mass=0.0, halo_a=1.0):
    order = np.argsort(r_all, kind="stable")
    rs = r_all[order]
    cm = np.cumsum(m_all[order])
    idx = np.searchsorted(rs, r_query, side="right")
    menc = np.where(idx > 0, cm[np.maximum(idx - 1, 0)], 0.0)
    rq = np.maximum(r_query, 1e-3)
    v2 = G * menc / rq
    if halo_mass > 0:
        v2 = v2 + G * halo_mass * rq / (rq + halo_a) ** 2
    return np.sqrt(np.maximum(v2, 0.0))


def disk(n, seed=SEED_BASE + 1, gas_fraction=0.0, halo_mass=0.0, halo_a=3.0, G=1.0,
         scale=1.0, rmax=5.0, z0=0.1, bulge_frac=0.2, bulge_a=0.2):
    st = Stream(seed)
    nb = int(round(n * bulge_frac))
    nd = n - nb
    # masses: Kroupa-shaped, normalised to total 1 (disk + bulge)
    m = kroupa(st, n)
    m = m / m.sum()
    pos = np.zeros((n, 3))
    # bulge: Hernquist, isotropic
    def hern(k):
        s = np.sqrt(st.uniform(k))
        return bulge_a * s / np.maximum(1.0 - s, 1e-9)
    rb = _truncated(st, nb, hern, rmax)
    pos[:nb] = _iso_dirs(st, nb) * rb[:, None]
    # disk: Gamma(2, scale) radii, sech^2 heights
    def gam(k):
        return -scale * np.log(st.open_uniform(k) * st.open_uniform(k))
    rd = _truncated(st, nd, gam, rmax)
    ph = 2.0 * np.pi * st.uniform(nd)
    zz = z0 * np.arctanh(np.clip(2.0 * st.uniform(nd) - 1.0, -1 + 1e-12, 1 - 1e-12))
    pos[nb:, 0] = rd * np.cos(ph)
    pos[nb:, 1] = rd * np.sin(ph)
    pos[nb:, 2] = zz
    r3 = np.linalg.norm(pos, axis=1)
    vel = np.zeros((n, 3))
    # disk: circular rotation about z with a 10 % dispersion
    vc = _enclosed_vc(r3, m, rd, G, halo_mass, halo_a)
    sig = 0.1 * vc
    g = st.open_uniform(nd * 3).reshape(nd, 3)
    g2 = st.uniform(nd * 3).reshape(nd, 3)
    gauss = np.sqrt(-2.0 * np.log(g)) * np.cos(2.0 * np.pi * g2)
    vel[nb:, 0] = -vc * np.sin(ph) + sig * gauss[:, 0]
    vel[nb:, 1] = vc * np.cos(ph) + sig * gauss[:, 1]
    vel[nb:, 2] = 0.5 * sig * gauss[:, 2]
    # bulge: isotropic dispersion vc / sqrt(3)
    vcb = _enclosed_vc(r3, m, rb, G, halo_mass, halo_a) / np.sqrt(3.0)
    gb = st.open_uniform(nb * 3).reshape(nb, 3)
    gb2 = st.uniform(nb * 3).reshape(nb, 3)
    gaussb = np.sqrt(-2.0 * np.log(gb)) * np.cos(2.0 * np.pi * gb2)
    vel[:nb] = vcb[:, None] * gaussb
    pad = np.zeros(n)
    if gas_fraction > 0:
        gas = st.uniform(nd) < gas_fraction
        pad[nb:][gas] = -1.0
    out = np.zeros((n, 8), np.float32)
    out[:, 0:3] = pos
    out[:, 3] = m
    out[:, 4:7] = vel
    out[:, 7] = pad
    return out


def plummer(n, seed=SEED_BASE + 4, a=1.0, G=1.0, rmax=10.0):
    st = Stream(seed)
    m = np.full(n, 1.0 / n)

    def rad(k):
        u = st.open_uniform(k)
        return a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
    r = _truncated(st, n, rad, rmax * a)
    pos = _iso_dirs(st, n) * r[:, None]
    # speed: q from g(q) = q^2 (1 - q^2)^3.5 by rejection, v = q v_esc
    q = np.zeros(n)
    todo = np.ones(n, bool)
    while todo.any():
        k = int(todo.sum())
        x = st.uniform(k)
        y = 0.1 * st.uniform(k)
        ok = y < x * x * (1.0 - x * x) ** 3.5
        idx = np.nonzero(todo)[0]
        q[idx[ok]] = x[ok]
        todo[idx[ok]] = False
    vesc = np.sqrt(2.0 * G / a) * (1.0 + (r / a) ** 2) ** -0.25
    vel = _iso_dirs(st, n) * (q * vesc)[:, None]
    out = np.zeros((n, 8), np.float32)
    out[:, 0:3] = pos
    out[:, 3] = m
    out[:, 4:7] = vel
    return out


def _rot_x(v, deg):
    c, s = np.cos(np.radians(deg)), np.sin(np.radians(deg))
    y = v[:, 1] * c - v[:, 2] * s
    z = v[:, 1] * s + v[:, 2] * c
    return np.stack([v[:, 0], y, z], axis=1)


def collision(n, seed=SEED_BASE + 3, offset=8.0, rel_speed=0.5, incline=30.0, **disk_kw):
    na = n // 2
    nb = n - na
    A = disk(na, seed=seed, **disk_kw).astype(np.float64)
    B = disk(nb, seed=seed ^ 0x5A5A5A5A, **disk_kw).astype(np.float64)
    B[:, 0:3] = _rot_x(B[:, 0:3], incline)
    B[:, 4:7] = _rot_x(B[:, 4:7], incline)
    A[:, 0] -= offset / 2
    B[:, 0] += offset / 2
    A[:, 4] += rel_speed / 2
    B[:, 4] -= rel_speed / 2
    return np.concatenate([A, B]).astype(np.float32)


def batch_sizes(count=64, lo=4096, hi=262144, seed=SEED_BASE + 4):
    """N_i = round(exp(U[ln lo, ln hi])) for the batched workload."""
    st = Stream(seed ^ 0xB47C4)
    u = st.uniform(count)
    return np.rint(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))).astype(np.int64)


def batch64(count=64, lo=4096, hi=262144, seed=SEED_BASE + 4):
    """The batched workload: `count` galaxies, N log-uniform in [lo, hi],
    each a disk or a Plummer sphere (a = 1) with probability 1/2.
    Returns a list of (preset, records)."""
    sizes = batch_sizes(count, lo, hi, seed)
    kinds = Stream(seed ^ 0x9E5E7).uniform(count) < 0.5
    out = []
    for k, (n, is_disk) in enumerate(zip(sizes.tolist(), kinds.tolist())):
        s = seed + 0x1000 * (k + 1)
        out.append(("disk", disk(n, seed=s)) if is_disk else ("plummer", plummer(n, seed=s)))
    return out


def serve_jobs(count, rate=20.0, lo=8192, hi=1048576, steps=200, seed=SEED_BASE + 5):
    """The served workload's job stream: Poisson arrivals (mean `rate` jobs/s,
    exponential gaps), N log-uniform in [lo, hi], preset uniform over disk /
    Plummer sphere / two-disk collision.  Returns dicts with the arrival time
    (s), n, preset, steps and a per-job seed; records come from job_records."""
    st = Stream(seed ^ 0x5E77E)
    gaps = -np.log(st.open_uniform(count)) / rate
    t = np.cumsum(gaps) - gaps[0]
    u = st.uniform(count)
    sizes = np.rint(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))).astype(np.int64)
    kinds = np.minimum((st.uniform(count) * 3).astype(np.int64), 2)
    names = ("disk", "plummer", "collision")
    return [dict(id=k, arrival=float(t[k]), n=int(sizes[k]), preset=names[kinds[k]],
                 steps=steps, seed=seed + 0x1000 * (k + 1)) for k in range(count)]


def job_records(job):
    n, s = job["n"], job["seed"]
    if job["preset"] == "disk":
        return disk(n, seed=s)
    if job["preset"] == "plummer":
        return plummer(n, seed=s)
    return collision(n, seed=s)


CONFIGS = {
    # name: (builder, params) — the workloads of the benchmark table
    "cpu_disk_16k": dict(n=16384, theta=0.7, epsilon=0.02, bucket=16, dt=0.005),
    "disk_500k": dict(n=500000, theta=0.7, epsilon=0.01, bucket=16, dt=0.005, steps=100,
                      gas_fraction=0.3, halo_mass=5.0, halo_a=3.0, gas_drag=0.25),
    "collision_4m": dict(n=4194304, theta=0.5, epsilon=0.005, bucket=32, dt=0.005),
    "batch64": dict(count=64, lo=4096, hi=262144, theta=0.6, epsilon=0.01, dt=0.005),
    "serve": dict(rate=20.0, lo=8192, hi=1048576, steps=200, theta=0.5, epsilon=0.01, dt=0.005,
                  hbm_budget=200e9),
}


def make(name, n=None):
    c = CONFIGS[name]
    n = n or c["n"]
    if name == "cpu_disk_16k":
        return disk(n, seed=SEED_BASE + 1)
    if name == "disk_500k":
        return disk(n, seed=SEED_BASE + 2, gas_fraction=c["gas_fraction"],
                    halo_mass=c["halo_mass"], halo_a=c["halo_a"])
    if name == "collision_4m":
        return collision(n, seed=SEED_BASE + 3)
    raise KeyError(name)
