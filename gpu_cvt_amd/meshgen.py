"""Seeded synthetic meshes and area-uniform site sampling.

The reference ships no meshes (its runs load OBJ files that are not in the
repository), so every workload here is generated: unit icospheres, tori with
normal jitter, and radially displaced spheres with spherical-harmonic bands and
sharp plane-clamped creases.  All randomness comes from numpy's PCG64 with an
explicit seed, so the same call gives the same arrays on every machine that has
this numpy.

Meshes are returned as ``(V float64 (n,3), F int32 (m,3))`` with outward
(counter-clockwise seen from outside) triangles.
"""
from __future__ import annotations

import math

import zlib

import numpy as np


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(int(seed)))


# ------------------------------------------------------------------ icosphere
def icosphere(level: int):
    """Unit icosphere: 10*4^level + 2 vertices, 20*4^level triangles."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    V = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    F = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    for _ in range(level):
        nv = len(V)
        e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
        ek = np.sort(e, axis=1)
        key = ek[:, 0] * nv + ek[:, 1]
        uk, inv = np.unique(key, return_inverse=True)
        a, b = uk // nv, uk % nv
        mid = V[a] + V[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        m = inv.reshape(3, -1).T + nv  # midpoints of edges (01, 12, 20)
        V = np.concatenate([V, mid])
        f0, f1, f2 = F[:, 0], F[:, 1], F[:, 2]
        m01, m12, m20 = m[:, 0], m[:, 1], m[:, 2]
        F = np.concatenate([np.stack([f0, m01, m20], 1), np.stack([f1, m12, m01], 1),
                            np.stack([f2, m20, m12], 1), np.stack([m01, m12, m20], 1)])
    return V, F.astype(np.int32)


# ---------------------------------------------------------------------- torus
def torus(R: float = 1.0, r: float = 0.3, nu: int = 500, nv: int = 200, jitter: float = 0.0, seed: int = 1):
    """Torus on an nu x nv grid (nu*nv vertices, 2*nu*nv triangles).

    ``jitter``: std-dev of a Gaussian displacement along the vertex normal.
    """
    u = np.arange(nu) * (2 * np.pi / nu)
    v = np.arange(nv) * (2 * np.pi / nv)
    U, Vv = np.meshgrid(u, v, indexing="ij")
    cu, su, cv, sv = np.cos(U), np.sin(U), np.cos(Vv), np.sin(Vv)
    P = np.stack([(R + r * cv) * cu, (R + r * cv) * su, r * sv], -1).reshape(-1, 3)
    Nrm = np.stack([cv * cu, cv * su, sv], -1).reshape(-1, 3)
    if jitter > 0:
        P = P + Nrm * _rng(seed).normal(0.0, jitter, size=(len(P), 1))
    i = np.arange(nu)[:, None]
    j = np.arange(nv)[None, :]
    a = i * nv + j
    b = ((i + 1) % nu) * nv + j
    c = ((i + 1) % nu) * nv + (j + 1) % nv
    d = i * nv + (j + 1) % nv
    F = np.concatenate([np.stack([a, b, c], -1).reshape(-1, 3), np.stack([a, c, d], -1).reshape(-1, 3)])
    return P, F.astype(np.int32)


# ---------------------------------------------------------------- cube sphere
def cube_sphere(n: int):
    """Unit sphere from a cube with n x n quads per face: 12 n^2 triangles,
    6 n^2 + 2 vertices (shared edges merged)."""
    g = np.arange(n + 1)
    A, B = np.meshgrid(g, g, indexing="ij")
    A = A.reshape(-1)
    B = B.reshape(-1)
    Z = np.zeros_like(A)
    N_ = np.full_like(A, n)
    # integer lattice coordinates on the cube surface [0, n]^3, each face as (a,b) grid
    faces = [
        (np.stack([N_, A, B], 1)),          # +x : (a,b) -> (y,z)
        (np.stack([Z, B, A], 1)),           # -x
        (np.stack([B, N_, A], 1)),          # +y : (a,b) -> (z,x)
        (np.stack([A, Z, B], 1)),           # -y
        (np.stack([A, B, N_], 1)),          # +z : (a,b) -> (x,y)
        (np.stack([B, A, Z], 1)),           # -z
    ]
    allp = np.concatenate(faces)
    key = (allp[:, 0] * (n + 1) + allp[:, 1]) * (n + 1) + allp[:, 2]
    uk, inv = np.unique(key, return_inverse=True)
    lat = np.stack([uk // ((n + 1) ** 2), (uk // (n + 1)) % (n + 1), uk % (n + 1)], 1).astype(np.float64)
    P = lat * (2.0 / n) - 1.0
    # equi-angular warp keeps triangles closer to uniform size
    P = np.tan(P * (np.pi / 4))
    P /= np.linalg.norm(P, axis=1, keepdims=True)
    per = (n + 1) * (n + 1)
    tris = []
    ia = np.arange(n)[:, None]
    ib = np.arange(n)[None, :]
    for f in range(6):
        idx = inv[f * per:(f + 1) * per].reshape(n + 1, n + 1)
        q00 = idx[ia, ib]
        q10 = idx[ia + 1, ib]
        q11 = idx[ia + 1, ib + 1]
        q01 = idx[ia, ib + 1]
        t = np.concatenate([np.stack([q00, q10, q11], -1).reshape(-1, 3),
                            np.stack([q00, q11, q01], -1).reshape(-1, 3)])
        tris.append(t)
    F = np.concatenate(tris)
    # orient outward
    a, b, c = P[F[:, 0]], P[F[:, 1]], P[F[:, 2]]
    nrm = np.cross(b - a, c - a)
    flip = (nrm * (a + b + c)).sum(1) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    return P, F.astype(np.int32)


def cube(n: int):
    """Closed cube surface [-1, 1]^3, n x n quads per face (two triangles
    each), outward-oriented, edges welded: sharp edges and corners for the
    curvature tiers."""
    g = np.linspace(-1.0, 1.0, n + 1)
    pts, idx, faces = [], {}, []

    def vid(p):
        key = tuple(np.round(p, 12))
        if key not in idx:
            idx[key] = len(pts)
            pts.append(p)
        return idx[key]

    for axis in range(3):
        for sgn in (-1.0, 1.0):
            u, v = [a for a in range(3) if a != axis]
            for i in range(n):
                for j in range(n):
                    q = []
                    for di, dj in ((0, 0), (1, 0), (1, 1), (0, 1)):
                        p = np.zeros(3)
                        p[axis], p[u], p[v] = sgn, g[i + di], g[j + dj]
                        q.append(vid(p))
                    if (sgn > 0) == (axis != 1):
                        faces += [[q[0], q[1], q[2]], [q[0], q[2], q[3]]]
                    else:
                        faces += [[q[0], q[2], q[1]], [q[0], q[3], q[2]]]
    return np.array(pts), np.array(faces, np.int32)


def _legendre_table(ls, ct: np.ndarray) -> dict:
    """Orthonormal associated Legendre values N_lm P_l^m(ct) for every l in
    ``ls`` and 0 <= m <= l, by the standard three-term recurrence in m-columns
    (stable for the l <= 16 used here)."""
    ls = sorted(set(int(l) for l in ls))
    lmax = max(ls)
    st = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
    out = {}
    pmm = np.full_like(ct, math.sqrt(1.0 / (4 * math.pi)))
    for m in range(0, lmax + 1):
        if m > 0:
            pmm = -math.sqrt((2 * m + 1) / (2.0 * m)) * st * pmm
        p_lm2, p_lm1 = None, pmm
        for l in range(m, lmax + 1):
            if l == m:
                cur = pmm
            elif l == m + 1:
                cur = math.sqrt(2 * m + 3) * ct * pmm
            else:
                a = math.sqrt((4 * l * l - 1) / (l * l - m * m))
                b = math.sqrt(((l - 1) ** 2 - m * m) / (4 * (l - 1) ** 2 - 1))
                cur = a * (ct * p_lm1 - b * p_lm2)
            if l in ls:
                out[(l, m)] = cur
            p_lm2, p_lm1 = p_lm1, cur
    return out


def displaced_sphere(n: int, *, amplitude: float = 0.05, bands: int = 8, lmax: int = 16, creases: int = 12,
                     seed: int = 4, crease_seed: int = 5):
    """Cube-sphere (12 n^2 triangles) displaced radially by a seeded sum of
    ``bands`` random real spherical-harmonic bands (l <= lmax) scaled to peak
    ``amplitude``, then flattened against ``creases`` random planes, which
    leaves sharp creases where the surface meets each plane."""
    P, F = cube_sphere(n)
    rng = _rng(seed)
    ct = np.clip(P[:, 2], -1.0, 1.0)
    phi = np.arctan2(P[:, 1], P[:, 0])
    spec = []
    for _ in range(bands):
        l = int(rng.integers(1, lmax + 1))
        coef = rng.normal(size=2 * l + 1)
        spec.append((l, coef / np.linalg.norm(coef)))
    table = _legendre_table([l for l, _ in spec], ct)
    disp = np.zeros(len(P))
    for l, coef in spec:
        for k, m in enumerate(range(-l, l + 1)):
            am = abs(m)
            y = table[(l, am)]
            if m > 0:
                y = math.sqrt(2.0) * y * np.cos(am * phi)
            elif m < 0:
                y = math.sqrt(2.0) * y * np.sin(am * phi)
            disp += coef[k] * y
    disp *= amplitude / max(1e-12, np.abs(disp).max())
    X = P * (1.0 + disp)[:, None]
    if creases > 0:
        cr = _rng(crease_seed)
        for _ in range(creases):
            nrm = cr.normal(size=3)
            nrm /= np.linalg.norm(nrm)
            off = float(cr.uniform(0.80, 0.97))
            h = X @ nrm - off
            over = h > 0
            X[over] -= h[over, None] * nrm[None, :]
    return X, F


# ------------------------------------------------------------------- sampling
def triangle_areas(V, F) -> np.ndarray:
    a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    return 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1)


def sample_surface(V, F, n: int, seed: int) -> np.ndarray:
    """n area-uniform points: triangle by the area CDF, then barycentrics
    (1-sqrt(r1), sqrt(r1)(1-r2), sqrt(r1) r2)."""
    rng = _rng(seed)
    A = triangle_areas(V, F)
    cdf = np.cumsum(A)
    cdf /= cdf[-1]
    t = np.minimum(np.searchsorted(cdf, rng.random(n), side="right"), len(F) - 1)
    r1 = np.sqrt(rng.random(n))
    r2 = rng.random(n)
    w0, w1, w2 = 1.0 - r1, r1 * (1.0 - r2), r1 * r2
    a, b, c = V[F[t, 0]], V[F[t, 1]], V[F[t, 2]]
    return a * w0[:, None] + b * w1[:, None] + c * w2[:, None]


# --------------------------------------------------------------- named workloads
def workload(name: str):
    """Named synthetic workloads: returns (V, F, S, K, iters, freeze_on)."""
    if name == "icosphere_l5_2k":
        V, F = icosphere(5)
        S = sample_surface(V, F, 2000, seed=0)
        return V, F, S, 20, 50, False
    if name == "torus_200k_100k":
        V, F = torus(1.0, 0.3, 500, 200, jitter=1e-3, seed=1)
        S = sample_surface(V, F, 100_000, seed=2)
        return V, F, S, 24, 100, True
    if name == "dsphere_1m_500k":
        V, F = displaced_sphere(289, seed=4, crease_seed=5)
        S = sample_surface(V, F, 500_000, seed=6)
        return V, F, S, 32, 100, True
    if name == "dsphere_2m_1m":
        V, F = displaced_sphere(408, seed=9, crease_seed=9)
        S = sample_surface(V, F, 1_000_000, seed=10)
        return V, F, S, 24, 100, False
    raise KeyError(name)


def batch_workload(njobs: int = 256, seed: int = 7, nf_range=(5_000, 400_000), ns_range=(1_000, 100_000)):
    """Mixed jobs: mesh type uniform over {icosphere, torus, displaced sphere},
    triangle count log-uniform in nf_range, site count log-uniform in ns_range
    capped at triangles/2.  Meshes are shared between jobs whose (type, size)
    round to the same tessellation; sites are sampled per job.
    Returns a list of (V, F, S, tag)."""
    rng = _rng(seed)
    cache = {}
    jobs = []
    for k in range(njobs):
        typ = int(rng.integers(0, 3))
        nf = float(np.exp(rng.uniform(np.log(nf_range[0]), np.log(nf_range[1]))))
        ns = float(np.exp(rng.uniform(np.log(ns_range[0]), np.log(ns_range[1]))))
        if typ == 0:
            lvl = int(np.clip(round(np.log(nf / 20.0) / np.log(4.0)), 4, 7))
            key = ("ico", lvl)
            if key not in cache:
                cache[key] = icosphere(lvl)
        elif typ == 1:
            nv = max(8, int(round(np.sqrt(nf / 5.0) / 4.0)) * 4)
            key = ("torus", nv)
            if key not in cache:
                cache[key] = torus(1.0, 0.3, int(2.5 * nv), nv, jitter=1e-3, seed=nv)
        else:
            n = max(8, int(round(np.sqrt(nf / 12.0) / 8.0)) * 8)
            key = ("dsphere", n)
            if key not in cache:
                cache[key] = displaced_sphere(n, seed=n, crease_seed=n + 1)
        V, F = cache[key]
        nsites = int(min(ns, len(F) / 2))
        S = sample_surface(V, F, nsites, seed=seed * 100003 + k)
        jobs.append((V, F, S, f"{key[0]}{key[1]}"))
    return jobs


def serve_trace(n_arrivals: int = 2000, rate: float = 200.0, seed: int = 8, large_frac: float = 0.05,
                nf_range=(5_000, 400_000), ns_range=(1_000, 100_000)):
    """Request trace for the serving loop: Poisson arrivals at `rate` jobs/s;
    each request is drawn from the batched-form size distribution (mesh type
    uniform over icosphere / torus / displaced sphere, triangles log-uniform in
    nf_range, sites log-uniform in ns_range capped at triangles/2), or with
    probability large_frac is a large job: 1M sites on a 1,997,568-triangle
    displaced sphere, cycling through four variants (two displacement seeds x
    two site seeds).  Meshes are shared by requests whose size rounds to the
    same tessellation and site sets by requests whose site count rounds to the
    same of 16 log-spaced buckets per decade, so generation stays bounded.
    Returns a list of (t_arrive_s, V, F, S, tag), ordered by arrival."""
    rng = _rng(seed)
    t = 0.0
    mcache, scache = {}, {}
    large, lmesh = {}, {}
    n_large = 0
    out = []
    for k in range(n_arrivals):
        t += float(rng.exponential(1.0 / rate))
        if rng.uniform() < large_frac:
            v = n_large % 4
            n_large += 1
            if v not in large:
                m = v % 2
                if m not in lmesh:
                    lmesh[m] = displaced_sphere(408, seed=9 + 2 * m, crease_seed=9 + 3 * m)
                V, F = lmesh[m]
                large[v] = (V, F, sample_surface(V, F, 1_000_000, seed=10 + v))
            out.append((t, *large[v], f"large{v}"))
            continue
        typ = int(rng.integers(0, 3))
        nf = float(np.exp(rng.uniform(np.log(nf_range[0]), np.log(nf_range[1]))))
        ns = float(np.exp(rng.uniform(np.log(ns_range[0]), np.log(ns_range[1]))))
        if typ == 0:
            key = ("ico", int(np.clip(round(np.log(nf / 20.0) / np.log(4.0)), 4, 7)))
        elif typ == 1:
            key = ("torus", max(8, int(round(np.sqrt(nf / 5.0) / 4.0)) * 4))
        else:
            key = ("dsphere", max(8, int(round(np.sqrt(nf / 12.0) / 8.0)) * 8))
        if key not in mcache:
            if key[0] == "ico":
                mcache[key] = icosphere(key[1])
            elif key[0] == "torus":
                mcache[key] = torus(1.0, 0.3, int(2.5 * key[1]), key[1], jitter=1e-3, seed=key[1])
            else:
                mcache[key] = displaced_sphere(key[1], seed=key[1], crease_seed=key[1] + 1)
        V, F = mcache[key]
        nsites = int(min(ns, len(F) / 2))
        b = int(round(np.log10(max(nsites, 1)) * 16))
        skey = (key, b)
        if skey not in scache:
            n_b = int(min(round(10 ** (b / 16.0)), len(F) // 2))
            seed_s = (seed * 1_000_003 + zlib.crc32(repr(skey).encode())) & 0x7FFFFFFF
            scache[skey] = sample_surface(V, F, max(n_b, 64), seed=seed_s)
        out.append((t, V, F, scache[skey], f"{key[0]}{key[1]}"))
    return out
