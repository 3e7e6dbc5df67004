"""Seeded synthetic oriented point clouds for the benchmark workloads.

Every workload is generated from a counter-based SplitMix64 stream, so the
same (workload, seed) pair gives bit-identical fp32 points on any machine and
in any chunking.  Geometry is computed in float64 and rounded to float32 at
the end.

Workloads (sizes and shapes as used by the benchmark and the parity tests):

* ``sphere``          20,000 points area-uniform on a sphere of radius 0.4
                      centred at (0.5,0.5,0.5), exact outward normals,
                      positions jittered by N(0, 0.001^2) per axis (depth 6).
* ``torus``           200,000 points area-uniform on a torus (R=0.3, r=0.1,
                      z axis, centre 0.5); normals rotated by U[0,2] degrees
                      about a uniform axis perpendicular to the normal
                      (depth 7).
* ``displaced``       2,000,000 points on r(t,p)=0.35(1+0.12 sin5t cos3p),
                      area-proportional, analytic normals, N(0,0.002^2)
                      position noise (depth 9).
* ``batch64``         64 clouds, sizes round(exp U[ln 5k, ln 150k]), each a
                      random ellipsoid or torus with a uniform rotation
                      (depth 7).

The reference reads its clouds from files (PCD/PLY/TXT, mesh/PoissonReconstruction.cpp:63-186)
and estimates normals with PCL; this module replaces that front-end with
analytic oriented samples.
"""
import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_MASK64 = (1 << 64) - 1

SEED_BASE = 20261202


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


class SplitMix64:
    """Counter-based SplitMix64: draw k is mix(seed + k*gamma)."""

    def __init__(self, seed, stream=0):
        s = (int(seed) * 0x100000001B3 + int(stream) * 0xD6E8FEB86659FD93) & _MASK64
        self.state = int(_mix(np.array([s], dtype=np.uint64))[0])
        self.counter = 0

    def next_u64(self, n):
        with np.errstate(over="ignore"):
            k = np.arange(self.counter + 1, self.counter + 1 + n, dtype=np.uint64)
            z = np.uint64(self.state) + k * _GAMMA
            out = _mix(z)
        self.counter += n
        return out

    def uniform(self, n):
        """53-bit uniforms in [0, 1)."""
        return (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))

    def normal(self, n):
        """Box-Muller standard normals (two uniforms per pair)."""
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log1p(-u1))
        t = 2.0 * np.pi * u2
        z = np.concatenate([r * np.cos(t), r * np.sin(t)])
        return z[:n]

    def unit_vectors(self, n):
        g = self.normal(3 * n).reshape(3, n).T
        return g / np.linalg.norm(g, axis=1, keepdims=True)


def _rejection(rng, n, propose, accept_prob, batch_factor=2.0):
    """Draw proposals until n are accepted; deterministic in counter order."""
    chunks = []
    got = 0
    while got < n:
        m = max(1024, int((n - got) * batch_factor))
        cand = propose(m)
        u = rng.uniform(m)
        ok = u < accept_prob(cand)
        sel = {k: v[ok] for k, v in cand.items()}
        chunks.append(sel)
        got += int(ok.sum())
    out = {k: np.concatenate([c[k] for c in chunks])[:n] for k in chunks[0]}
    return out


def sphere(n=20000, seed=SEED_BASE + 1, radius=0.4, noise=0.001):
    rng = SplitMix64(seed, 1)
    d = rng.unit_vectors(n)
    nrng = SplitMix64(seed, 2)
    p = 0.5 + radius * d + noise * nrng.normal(3 * n).reshape(3, n).T
    return _f32(p), _f32(d)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _torus_points(rng, n, R, r):
    def propose(m):
        return {"theta": 2 * np.pi * rng.uniform(m), "phi": 2 * np.pi * rng.uniform(m)}

    s = _rejection(rng, n, propose, lambda c: (R + r * np.cos(c["theta"])) / (R + r))
    th, ph = s["theta"], s["phi"]
    rho = R + r * np.cos(th)
    p = np.stack([rho * np.cos(ph), rho * np.sin(ph), r * np.sin(th)], axis=1)
    nrm = np.stack([np.cos(th) * np.cos(ph), np.cos(th) * np.sin(ph), np.sin(th)], axis=1)
    return p, nrm


def _perturb_normals(rng, nrm, max_deg):
    n = nrm.shape[0]
    a = rng.unit_vectors(n)
    a = a - np.sum(a * nrm, axis=1, keepdims=True) * nrm
    a /= np.linalg.norm(a, axis=1, keepdims=True)
    ang = np.deg2rad(max_deg) * rng.uniform(n)
    axn = np.cross(a, nrm)
    out = nrm * np.cos(ang)[:, None] + axn * np.sin(ang)[:, None]
    return out / np.linalg.norm(out, axis=1, keepdims=True)


def torus(n=200000, seed=SEED_BASE + 2, R=0.3, r=0.1):
    rng = SplitMix64(seed, 1)
    p, nrm = _torus_points(rng, n, R, r)
    nrm = _perturb_normals(SplitMix64(seed, 2), nrm, 2.0)
    return _f32(p + 0.5), _f32(nrm)


def displaced_sphere(n=2000000, seed=SEED_BASE + 3, noise=0.002):
    rng = SplitMix64(seed, 1)
    a0, eps = 0.35, 0.12

    def geom(ct, ph):
        th = np.arccos(np.clip(ct, -1.0, 1.0))
        st = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
        rr = a0 * (1 + eps * np.sin(5 * th) * np.cos(3 * ph))
        r_t = a0 * eps * 5 * np.cos(5 * th) * np.cos(3 * ph)
        r_p = -a0 * eps * 3 * np.sin(5 * th) * np.sin(3 * ph)
        return th, st, rr, r_t, r_p

    bound = 0.40  # >= max of r*sqrt(r^2 + r_t^2 + (r_p/sin t)^2) (about 0.31)

    def propose(m):
        return {"ct": 2 * rng.uniform(m) - 1, "phi": 2 * np.pi * rng.uniform(m)}

    def acc(c):
        th, st, rr, r_t, r_p = geom(c["ct"], c["phi"])
        rp_s = np.where(st > 1e-12, r_p / np.maximum(st, 1e-12), 0.0)
        return rr * np.sqrt(rr * rr + r_t * r_t + rp_s * rp_s) / bound

    s = _rejection(rng, n, propose, acc, batch_factor=2.5)
    th, st, rr, r_t, r_p = geom(s["ct"], s["phi"])
    ph = s["phi"]
    rhat = np.stack([st * np.cos(ph), st * np.sin(ph), np.cos(th)], axis=1)
    that = np.stack([np.cos(th) * np.cos(ph), np.cos(th) * np.sin(ph), -st], axis=1)
    phat = np.stack([-np.sin(ph), np.cos(ph), np.zeros_like(ph)], axis=1)
    rp_s = np.where(st > 1e-12, r_p / np.maximum(st, 1e-12), 0.0)
    nrm = rr[:, None] * rhat - r_t[:, None] * that - rp_s[:, None] * phat
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    p = 0.5 + rr[:, None] * rhat
    p = p + noise * SplitMix64(seed, 2).normal(3 * n).reshape(3, n).T
    return _f32(p), _f32(nrm)


def _random_rotation(rng):
    q = rng.normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def ellipsoid(n, axes, rot, rng):
    a, b, c = axes
    inv2 = np.array([1 / a ** 2, 1 / b ** 2, 1 / c ** 2])
    bound = a * b * c * np.sqrt(inv2.max())

    def propose(m):
        return {"u": rng.unit_vectors(m)}

    def acc(cd):
        u = cd["u"]
        return a * b * c * np.sqrt((u * u) @ inv2) / bound

    s = _rejection(rng, n, propose, acc)
    u = s["u"]
    p = u * np.array([a, b, c])
    nrm = u / np.array([a, b, c])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    return 0.5 + p @ rot.T, nrm @ rot.T


def batch64(count=64, seed=SEED_BASE + 4, lo=5000, hi=150000):
    """64 independent clouds; returns a list of (xyz, normals) fp32 pairs."""
    meta = SplitMix64(seed, 0)
    out = []
    for i in range(count):
        u = meta.uniform(8)
        n = int(np.rint(np.exp(np.log(lo) + u[0] * (np.log(hi) - np.log(lo)))))
        rng = SplitMix64(seed, 100 + i)
        rot = _random_rotation(rng)
        if u[1] < 0.5:
            axes = 0.15 + 0.30 * u[2:5]
            p, nrm = ellipsoid(n, axes, rot, rng)
        else:
            R = 0.2 + 0.1 * u[2]
            r = 0.05 + 0.05 * u[3]
            p, nrm = _torus_points(rng, n, R, r)
            p, nrm = 0.5 + p @ rot.T, nrm @ rot.T
        out.append((_f32(p), _f32(nrm)))
    return out


def depth_for_size(n):
    """Octree depth of a request by its point count (depth 6 to 9)."""
    return 6 if n < 40000 else 7 if n < 150000 else 8 if n < 450000 else 9


def request_stream(count=120, rate_hz=20.0, seed=SEED_BASE + 5, lo=10000, hi=1000000):
    """A seeded stream of reconstruction requests: Poisson arrivals at rate_hz
    (exponential gaps), sizes log-uniform in [lo, hi], ellipsoids and tori as
    in batch64, depth by size.  Returns a list of (arrival_s, xyz, normals,
    depth), arrivals increasing from 0."""
    meta = SplitMix64(seed, 0)
    t = 0.0
    out = []
    for i in range(count):
        u = meta.uniform(8)
        if i:
            t += -np.log1p(-u[7]) / rate_hz
        n = int(np.rint(np.exp(np.log(lo) + u[0] * (np.log(hi) - np.log(lo)))))
        rng = SplitMix64(seed, 100 + i)
        rot = _random_rotation(rng)
        if u[1] < 0.5:
            p, nrm = ellipsoid(n, 0.15 + 0.30 * u[2:5], rot, rng)
        else:
            p, nrm = _torus_points(rng, n, 0.2 + 0.1 * u[2], 0.05 + 0.05 * u[3])
            p, nrm = 0.5 + p @ rot.T, nrm @ rot.T
        out.append((t, _f32(p), _f32(nrm), depth_for_size(n)))
    return out


WORKLOADS = {
    "sphere": (sphere, 6),
    "torus": (torus, 7),
    "displaced": (displaced_sphere, 9),
}


def workload(name, n=None):
    """Return (xyz, normals, depth) for a named single-cloud workload."""
    fn, depth = WORKLOADS[name]
    if n is None:
        p, nrm = fn()
    else:
        p, nrm = fn(n=n)
    return p, nrm, depth
