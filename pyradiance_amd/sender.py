"""Sender-surface ray samples for rfluxmtx's sampling mode (reference:
util/rfluxmtx.c:587-1010 and :1115-1195).

For every bin of the sender's hemisphere basis (uniform, Shirley-Chiu,
Reinhart/Tregenza, Klems full/half/quarter) `sampcnt` rays are made, with
directions inside the bin and origins on the sender surfaces weighted by
their projected area.  Like the reference, each bin's samples are
stratified along a Hilbert curve: sample n of N draws x = (n + u)/N and
spreads it over 2 or 3 dimensions (SDmultiSamp, common/bsdf.c:531-557).

Random numbers come from numpy's generator, so matrices agree with the
reference in distribution, not sample for sample.  All of it is vectorised
over the samples; the rays then go to the GPU as one rcontrib call with
-c sampcnt.
"""

from __future__ import annotations

import math

import numpy as np

FTINY = 1e-6

# ANGLE_BASIS tables of common/bsdf_m.c:31-63: (theta_min degrees, azimuths)
KLEMS_BASES = {
    "1": ((0., 1), (5., 8), (15., 16), (25., 20), (35., 24), (45., 24), (55., 24), (65., 16), (75., 12), (90., 0)),
    "2": ((0., 1), (6.5, 8), (19.5, 12), (32.5, 16), (45.5, 20), (58.5, 12), (71.5, 8), (90., 0)),
    "4": ((0., 1), (9., 8), (27., 12), (45., 12), (63., 8), (90., 0)),
}
TNAZ = (30, 30, 24, 24, 18, 12, 6)      # Tregenza azimuth counts per row


def hilbert_i2c(n: int, nbits: int, idx: np.ndarray) -> np.ndarray:
    """Hilbert-curve index -> n-dimensional cell coordinates (nbits each).

    J. Skilling, "Programming the Hilbert curve", AIP Conf. Proc. 707 (2004):
    the index bits are dealt out to the n axes ("transpose" form), Gray
    decoded, and the per-level reflections/exchanges undone.  Returns
    (len(idx), n) uint64."""
    idx = np.asarray(idx, dtype=np.uint64)
    one = np.uint64(1)
    X = [np.zeros(idx.shape, np.uint64) for _ in range(n)]
    for j in range(nbits):
        for i in range(n):
            X[i] |= ((idx >> np.uint64(j * n + (n - 1 - i))) & one) << np.uint64(j)
    t = X[n - 1] >> one                      # Gray decode
    for i in range(n - 1, 0, -1):
        X[i] = X[i] ^ X[i - 1]
    X[0] = X[0] ^ t
    Q = 2
    while Q != (1 << nbits):                 # undo excess work
        P, Qv = np.uint64(Q - 1), np.uint64(Q)
        for i in range(n - 1, -1, -1):
            hit = (X[i] & Qv) != 0
            x0 = X[0]
            t = (x0 ^ X[i]) & P
            X[0] = np.where(hit, x0 ^ P, x0 ^ t)
            if i:
                X[i] = np.where(hit, X[i], X[i] ^ t)
        Q <<= 1
    return np.stack(X, axis=1)


def multi_samp(rng: np.random.Generator, n: int, x: np.ndarray) -> np.ndarray:
    """x in [0,1) -> points in [0,1)^n along a Hilbert curve, jittered within
    the final cell (what SDmultiSamp does, common/bsdf.c:531-557)."""
    x = np.clip(np.asarray(x, dtype=np.float64), 0.0, 0.999999999999999)
    if n == 1:
        return x[:, None]
    nbits = 63 // n
    idx = (x * float(1 << (nbits * n))).astype(np.uint64)
    c = hilbert_i2c(n, nbits, idx).astype(np.float64)
    return (c + rng.random(c.shape)) / float(1 << nbits)


def square2disk(sx: np.ndarray, sy: np.ndarray):
    """Shirley-Chiu concentric map of the unit square onto the unit disk
    (P. Shirley, K. Chiu, J. Graphics Tools 2(3), 1997)."""
    a, b = 2.0 * sx - 1.0, 2.0 * sy - 1.0
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(a > -b, np.where(a > b, a, b), np.where(a < b, -a, -b))
        phi = np.where(a > -b,
                       np.where(a > b, (math.pi / 4) * (b / a), (math.pi / 4) * (2.0 - a / b)),
                       np.where(a < b, (math.pi / 4) * (4.0 + b / a),
                                np.where(b != 0, (math.pi / 4) * (6.0 - a / b), 0.0)))
    phi = np.nan_to_num(phi)
    return r * np.cos(phi), r * np.sin(phi)


def _perpendicular(nrm) -> tuple:
    """Two unit tangents (u, v = n x u) of a unit normal."""
    n = np.asarray(nrm, dtype=np.float64)
    a = np.zeros(3)
    a[int(np.argmin(np.abs(n)))] = 1.0
    u = np.cross(a, n)
    u /= np.linalg.norm(u)
    return u, np.cross(n, u)


def _triangulate(p2: np.ndarray) -> list:
    """Ear clipping of a simple polygon given in 2D; returns index triples
    with the polygon's own winding."""
    n = len(p2)
    area2 = sum(p2[i, 0] * p2[(i + 1) % n, 1] - p2[(i + 1) % n, 0] * p2[i, 1] for i in range(n))
    sgn = 1.0 if area2 >= 0 else -1.0
    idx = list(range(n))
    tris = []

    def cross(o, a, b):
        return (a[0] - o[0]) * (b[1] - o[1]) - (a[1] - o[1]) * (b[0] - o[0])

    guard = 0
    while len(idx) > 3 and guard < 10 * n * n:
        guard += 1
        m = len(idx)
        for k in range(m):
            i0, i1, i2 = idx[(k - 1) % m], idx[k], idx[(k + 1) % m]
            a, b, c = p2[i0], p2[i1], p2[i2]
            if sgn * cross(a, b, c) <= 0:
                continue                        # reflex or degenerate corner
            inside = False
            for j in idx:
                if j in (i0, i1, i2):
                    continue
                q = p2[j]
                if (sgn * cross(a, b, q) >= 0 and sgn * cross(b, c, q) >= 0 and sgn * cross(c, a, q) >= 0):
                    inside = True
                    break
            if not inside:
                tris.append((i0, i1, i2))
                idx.pop(k)
                break
        else:
            raise ValueError("cannot triangulate polygon")
    tris.append(tuple(idx))
    return tris


class _SurfSampler:
    """Origins on one sender surface from a uniform variable (ssamp_poly,
    ssamp_ring of util/rfluxmtx.c:607-720)."""

    def __init__(self, s):
        self.s = s
        f = np.asarray(s.farg, dtype=np.float64)
        if s.styp == "ring":
            self.u, self.v = _perpendicular(s.nrm)
        elif s.styp == "polygon":
            verts = f.reshape(-1, 3)
            u, v = _perpendicular(s.nrm)
            p2 = np.stack([verts @ u, verts @ v], 1)
            tris = _triangulate(p2)
            a = np.array([(p2[j, 0] - p2[i, 0]) * (p2[k, 1] - p2[i, 1]) - (p2[k, 0] - p2[i, 0]) * (p2[j, 1] - p2[i, 1])
                          for i, j, k in tris]) / (2.0 * s.area)
            self.tris = np.array(tris)
            self.afrac = np.abs(a)
            self.verts = verts
        else:
            raise ValueError(f"illegal sender surface '{s.name}'")

    def origins(self, rng, x: np.ndarray) -> np.ndarray:
        s = self.s
        if s.styp == "ring":
            t = multi_samp(rng, 2, x)
            f = s.farg
            r = np.sqrt(t[:, 0] * s.area * (1.0 / math.pi) + f[6] * f[6])
            ph = 2.0 * math.pi * t[:, 1]
            c = np.asarray(f[:3])
            return c + (r * np.cos(ph))[:, None] * self.u + (r * np.sin(ph))[:, None] * self.v
        cum = np.cumsum(self.afrac)
        k = np.minimum(np.searchsorted(cum, x, side="left"), len(self.afrac) - 1)
        xl = (x - (cum[k] - self.afrac[k])) / self.afrac[k]
        t = multi_samp(rng, 2, xl)
        a = np.sqrt(t[:, 1])
        s0, s1 = t[:, 0] * a, 1.0 - a
        v0, v1, v2 = (self.verts[self.tris[k, j]] for j in range(3))
        return v0 + s0[:, None] * (v1 - v0) + s1[:, None] * (v2 - v0)


def _origins(samplers: list, rng, dirs: np.ndarray, x: np.ndarray) -> np.ndarray:
    """sample_origin() of util/rfluxmtx.c:722-767: a surface chosen by its
    projected area towards the ray direction, then a point on it."""
    if len(samplers) == 1:
        s = samplers[0].s
        if np.any(dirs @ np.asarray(s.nrm) >= FTINY):
            raise ValueError(f"internal - sample behind sender '{s.name}'")
        return samplers[0].origins(rng, x)
    proj = np.stack([-(dirs @ np.asarray(sp.s.nrm)) * sp.s.area for sp in samplers], 1)
    proj *= proj > 0
    tot = proj.sum(1)
    if np.any(tot < FTINY * FTINY):
        raise ValueError("internal - sample behind all sender elements!")
    want = tot * x
    cum = np.cumsum(proj, 1)
    k = np.minimum((want[:, None] > cum).sum(1), len(samplers) - 1)
    rows = np.arange(len(x))
    xl = (want - (cum[rows, k] - proj[rows, k])) / proj[rows, k]
    out = np.empty((len(x), 3))
    for i, sp in enumerate(samplers):
        m = k == i
        if m.any():
            out[m] = sp.origins(rng, xl[m])
    return out


def nbins(hemis: str, hsiz: int) -> int:
    """Number of sender bins (the `fp == NULL` queries of the samplers)."""
    h0 = hemis[0].lower()
    if h0 == "u" or hemis[0] == "1":
        return 1
    if h0 == "s" and hemis[1:2].lower() == "c":
        return hsiz * hsiz
    if h0 in "rt":
        rowmax = 7 * hsiz + 1
        return sum(1 if r >= rowmax - 1 else hsiz * TNAZ[r // hsiz] for r in range(rowmax))
    return sum(n for _, n in KLEMS_BASES[hemis[1]])


def sender_rays(p, sampcnt: int, rng: np.random.Generator) -> np.ndarray:
    """All sample rays, bin by bin: (nbins * sampcnt, 6).  `p` holds the
    prepared sender parameters: hemis (Klems normalised to k1/k2/k4), hsiz,
    nrm, udir, vdir, surfaces."""
    nb = nbins(p.hemis, p.hsiz)
    b = np.repeat(np.arange(nb), sampcnt)
    n = np.tile(np.arange(sampcnt), nb)
    x = (n + rng.random(len(n))) / sampcnt
    N, U, V = (np.asarray(v, dtype=np.float64) for v in (p.nrm, p.udir, p.vdir))
    h0 = p.hemis[0].lower()
    if h0 == "u" or p.hemis[0] == "1":            # sample_uniform, :769-795
        t = multi_samp(rng, 3, x)
        dx, dy = square2disk(t[:, 1], t[:, 2])
        dz = -np.sqrt(np.maximum(0.0, 1.0 - dx * dx - dy * dy))
        d = dx[:, None] * U + dy[:, None] * V + dz[:, None] * N
        xo = t[:, 0]
    elif h0 == "s" and p.hemis[1:2].lower() == "c":   # sample_shirchiu, :797-825
        t = multi_samp(rng, 3, x)
        hs = p.hsiz
        dx, dy = square2disk((b // hs + t[:, 1]) / hs, (b % hs + t[:, 2]) / hs)
        dz = np.sqrt(np.maximum(0.0, 1.0 - dx * dx - dy * dy))
        d = -(dx[:, None] * U + dy[:, None] * V + dz[:, None] * N)
        xo = t[:, 0]
    elif h0 in "rt":                               # sample_reinhart, :827-877
        hs = p.hsiz
        rowmax = 7 * hs + 1
        rah = (0.5 * math.pi) / (rowmax - 0.5)
        rn = np.array([1 if r >= rowmax - 1 else hs * TNAZ[r // hs] for r in range(rowmax)])
        start = np.concatenate([[0], np.cumsum(rn)])
        row = np.searchsorted(start, b, side="right") - 1
        col = b - start[row]
        t = multi_samp(rng, 3, x)
        t1 = np.where(row >= rowmax - 1, t[:, 1] * t[:, 1], t[:, 1])
        alt = (row + t1) * rah
        azi = (2.0 * math.pi) * (col + t[:, 2] - 0.5) / rn[row]
        c = np.cos(alt)
        d0, d1, d2 = np.sin(azi) * c, -np.cos(azi) * c, np.sqrt(np.maximum(0.0, 1.0 - c * c))
        d = -(d0[:, None] * U + d1[:, None] * V + d2[:, None] * N)
        xo = t[:, 0]
    else:                                          # sample_klems, :879-926 with fo_getvec
        lat = KLEMS_BASES[p.hemis[1]]
        t = multi_samp(rng, 2, x)
        nph = np.array([nphi for _, nphi in lat])
        start = np.concatenate([[0], np.cumsum(nph)])
        li = np.searchsorted(start, b, side="right") - 1
        ndx = b - start[li]
        rx = multi_samp(rng, 2, t[:, 1])
        tmin = np.radians(np.array([tm for tm, _ in lat]))
        cz = np.sqrt((1.0 - rx[:, 0]) * np.cos(tmin[li]) ** 2 + rx[:, 0] * np.cos(tmin[li + 1]) ** 2)
        azi = 2.0 * math.pi * (ndx + rx[:, 1] - 0.5) / nph[li]
        sz = np.sqrt(np.maximum(0.0, 1.0 - cz * cz))
        d0, d1, d2 = np.cos(azi) * sz, np.sin(azi) * sz, cz
        d = -(d0[:, None] * U + d1[:, None] * V + d2[:, None] * N)
        xo = t[:, 0]
    samplers = [_SurfSampler(s) for s in p.surfaces]
    o = _origins(samplers, rng, d, xo)
    return np.hstack([o, d])
