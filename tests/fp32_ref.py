"""Binary32 restatement of rayTriangle / raySphere (intersect.ts:20-85) and of
the exhaustive closest hit (intersectBruteForce, intersect.ts:423-468), used
to check the FP32 precision kernels: every step is one float32 numpy
operation (IEEE round to nearest) in the kernels' order, with no fused
multiply-add -- the same arithmetic the kernels are compiled to."""
import numpy as np

F = np.float32
INF = F(np.inf)


def tri_t(o, d, a, b, c, tmin, tmax):
    """(R, T) float32 hit distances of rays (o, d: (R,3)) against triangles
    (a, b, c: (T,3)); +inf where the test misses."""
    o = o[:, None, :]
    d = d[:, None, :]
    a, b, c = a[None], b[None], c[None]
    e1 = b - a
    e2 = c - a
    px = d[..., 1] * e2[..., 2] - d[..., 2] * e2[..., 1]
    py = d[..., 2] * e2[..., 0] - d[..., 0] * e2[..., 2]
    pz = d[..., 0] * e2[..., 1] - d[..., 1] * e2[..., 0]
    det = e1[..., 0] * px + e1[..., 1] * py + e1[..., 2] * pz
    miss = (det > F(-1e-12)) & (det < F(1e-12))
    t_ = o - a
    un = t_[..., 0] * px + t_[..., 1] * py + t_[..., 2] * pz
    qx = t_[..., 1] * e1[..., 2] - t_[..., 2] * e1[..., 1]
    qy = t_[..., 2] * e1[..., 0] - t_[..., 0] * e1[..., 2]
    qz = t_[..., 0] * e1[..., 1] - t_[..., 1] * e1[..., 0]
    vn = d[..., 0] * qx + d[..., 1] * qy + d[..., 2] * qz
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        inv = F(1) / det
        u = un * inv
        v = vn * inv
        t = (e2[..., 0] * qx + e2[..., 1] * qy + e2[..., 2] * qz) * inv
    miss |= (u < 0) | (u > 1) | (v < 0) | (u + v > 1) | (t <= tmin) | (t >= tmax)
    return np.where(miss, INF, t).astype(F)


def sphere_t(o, d, cen, r, tmin, tmax):
    """(R, S) float32 hit distances against spheres (centre (S,3), radius (S,))."""
    o = o[:, None, :]
    d = d[:, None, :]
    l_ = o - cen[None]
    a = d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2]
    hb = l_[..., 0] * d[..., 0] + l_[..., 1] * d[..., 1] + l_[..., 2] * d[..., 2]
    cc = l_[..., 0] * l_[..., 0] + l_[..., 1] * l_[..., 1] + l_[..., 2] * l_[..., 2] - r[None] * r[None]
    disc = hb * hb - a * cc
    with np.errstate(invalid="ignore", divide="ignore"):
        s = np.sqrt(disc)
        t0 = (-hb - s) / a
        t1 = (-hb + s) / a
    ok0 = (t0 > tmin) & (t0 < tmax)
    ok1 = (t1 > tmin) & (t1 < tmax)
    t = np.where(ok0, t0, np.where(ok1, t1, INF))
    return np.where(disc < 0, INF, t).astype(F)


def brute_closest(prims, rays, tmin=1e-4, chunk=256):
    """Closest primitive in packed order (first of equal t) and its t, for
    float32 prim records (n, 16) and rays (R, 6) float32."""
    prims = prims.reshape(-1, 16)
    sph = prims[:, 3] == 0
    R = rays.shape[0]
    best_p = np.full(R, -1, np.int64)
    best_t = np.full(R, np.inf, F)
    tmin = F(tmin)
    for s in range(0, R, chunk):
        o = rays[s:s + chunk, 0:3].astype(F)
        d = rays[s:s + chunk, 3:6].astype(F)
        tt = np.full((o.shape[0], prims.shape[0]), INF, F)
        ti = np.flatnonzero(~sph)
        if ti.size:
            tt[:, ti] = tri_t(o, d, prims[ti, 0:3], prims[ti, 4:7], prims[ti, 8:11], tmin, INF)
        si = np.flatnonzero(sph)
        if si.size:
            tt[:, si] = sphere_t(o, d, prims[si, 0:3], prims[si, 4], tmin, INF)
        p = np.argmin(tt, axis=1)
        t = tt[np.arange(o.shape[0]), p]
        best_p[s:s + chunk] = np.where(np.isfinite(t), p, -1)
        best_t[s:s + chunk] = t
    return best_p, best_t
