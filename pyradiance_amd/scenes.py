"""Seeded synthetic workloads: shoebox rooms, furniture, sensor grids, skies.

Every generator is deterministic for a given seed (numpy PCG64) and writes
plain Radiance scene text (``.rad``), which both the reference ``oconv`` and
this package's own loader read.  Geometry conventions follow Radiance:
polygon normals come from the right-hand rule over the vertex order
(common/face.c:17-24), +Y is north, +Z is up (util/reinhartb.cal:24-25).

The south wall carries its window as one polygon with a doubled seam edge,
which the crossing-number test of common/face.c:121-162 treats as a hole.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

__all__ = [
    "Room",
    "room_rad",
    "sun_rad",
    "sky_rad",
    "sun_direction",
    "sensor_grid",
    "box_quads_rad",
    "config1",
    "config2",
    "furniture_rad",
    "sky_matrix",
    "reinhart_nbins",
    "config3",
    "boxes_rad",
    "annual_wea",
    "config5",
]


@dataclass
class Room:
    width: float = 6.0      # x extent (m)
    depth: float = 4.0      # y extent (m)
    height: float = 3.0     # z extent (m)
    rho_floor: float = 0.2
    rho_wall: float = 0.5
    rho_ceiling: float = 0.7
    window: bool = True
    win_w: float = 2.0
    win_h: float = 1.2
    win_sill: float = 0.9
    glass_tau: float = 0.6
    extra: list = field(default_factory=list)


def _poly(mod: str, name: str, verts) -> str:
    flat = " ".join(f"{c:.9g}" for v in verts for c in v)
    return f"{mod} polygon {name}\n0\n0\n{3 * len(verts)} {flat}\n"


def room_rad(room: Room | None = None) -> str:
    """Six inward-facing plastic walls, optional glass window in the south wall."""
    r = room or Room()
    W, D, H = r.width, r.depth, r.height
    out = [
        f"void plastic floor_mat\n0\n0\n5 {r.rho_floor} {r.rho_floor} {r.rho_floor} 0 0\n",
        f"void plastic wall_mat\n0\n0\n5 {r.rho_wall} {r.rho_wall} {r.rho_wall} 0 0\n",
        f"void plastic ceiling_mat\n0\n0\n5 {r.rho_ceiling} {r.rho_ceiling} {r.rho_ceiling} 0 0\n",
        f"void glass window_glass\n0\n0\n3 {r.glass_tau} {r.glass_tau} {r.glass_tau}\n",
        _poly("floor_mat", "floor", [(0, 0, 0), (W, 0, 0), (W, D, 0), (0, D, 0)]),
        _poly("ceiling_mat", "ceiling", [(0, 0, H), (0, D, H), (W, D, H), (W, 0, H)]),
        _poly("wall_mat", "north_wall", [(0, D, 0), (W, D, 0), (W, D, H), (0, D, H)]),
        _poly("wall_mat", "west_wall", [(0, 0, 0), (0, D, 0), (0, D, H), (0, 0, H)]),
        _poly("wall_mat", "east_wall", [(W, 0, 0), (W, 0, H), (W, D, H), (W, D, 0)]),
    ]
    if r.window:
        x0 = 0.5 * (W - r.win_w)
        x1 = x0 + r.win_w
        z0 = r.win_sill
        z1 = z0 + r.win_h
        south = [(0, 0, 0), (0, 0, H), (W, 0, H), (W, 0, 0), (0, 0, 0),
                 (x0, 0, z0), (x1, 0, z0), (x1, 0, z1), (x0, 0, z1), (x0, 0, z0)]
        out.append(_poly("wall_mat", "south_wall", south))
        out.append(_poly("window_glass", "window",
                         [(x0, 0, z0), (x0, 0, z1), (x1, 0, z1), (x1, 0, z0)]))
    else:
        out.append(_poly("wall_mat", "south_wall",
                         [(0, 0, 0), (0, 0, H), (W, 0, H), (W, 0, 0)]))
    out.extend(r.extra)
    return "".join(out)


def sun_direction(alt_deg: float, azi_deg: float) -> np.ndarray:
    """Unit vector toward the sun; azimuth clockwise from north (+Y) toward east (+X)."""
    a = math.radians(alt_deg)
    z = math.radians(azi_deg)
    return np.array([math.sin(z) * math.cos(a), math.cos(z) * math.cos(a), math.sin(a)])


def sun_rad(direction, radiance=(1e6, 1e6, 1e6), angle=0.533) -> str:
    d = direction
    return (f"void light solar\n0\n0\n3 {radiance[0]:.9g} {radiance[1]:.9g} {radiance[2]:.9g}\n"
            f"solar source sun\n0\n0\n4 {d[0]:.12g} {d[1]:.12g} {d[2]:.12g} {angle}\n")


def sky_rad(sky=(50.0, 50.0, 50.0), ground=None) -> str:
    out = (f"void glow skyglow\n0\n0\n4 {sky[0]:.9g} {sky[1]:.9g} {sky[2]:.9g} 0\n"
           "skyglow source sky\n0\n0\n4 0 0 1 180\n")
    if ground is not None:
        out += (f"void glow groundglow\n0\n0\n4 {ground[0]:.9g} {ground[1]:.9g} {ground[2]:.9g} 0\n"
                "groundglow source ground\n0\n0\n4 0 0 -1 180\n")
    return out


def sensor_grid(nx: int, ny: int, x0: float, y0: float, pitch: float, z: float,
                normal=(0.0, 0.0, 1.0)) -> np.ndarray:
    """(nx*ny, 6) float64 array of origin+direction records, row-major in y then x."""
    xs = x0 + pitch * np.arange(nx)
    ys = y0 + pitch * np.arange(ny)
    gx, gy = np.meshgrid(xs, ys)
    n = gx.size
    rays = np.empty((n, 6))
    rays[:, 0] = gx.ravel()
    rays[:, 1] = gy.ravel()
    rays[:, 2] = z
    rays[:, 3:] = normal
    return rays


def box_quads_rad(mod: str, name: str, lo, hi, nu: int = 3, nv: int = 7) -> str:
    """Axis-aligned box with outward normals, each face split into nu*nv quads."""
    lo = np.asarray(lo, float)
    hi = np.asarray(hi, float)
    out = []
    # (fixed axis, value, u axis, v axis, outward sign)
    faces = [(0, lo[0], 1, 2, -1), (0, hi[0], 1, 2, +1),
             (1, lo[1], 0, 2, -1), (1, hi[1], 0, 2, +1),
             (2, lo[2], 0, 1, -1), (2, hi[2], 0, 1, +1)]
    k = 0
    for ax, val, ua, va, sgn in faces:
        us = np.linspace(lo[ua], hi[ua], nu + 1)
        vs = np.linspace(lo[va], hi[va], nv + 1)
        for i in range(nu):
            for j in range(nv):
                quad = []
                for (a, b) in ((0, 0), (1, 0), (1, 1), (0, 1)):
                    p = [0.0, 0.0, 0.0]
                    p[ax] = val
                    p[ua] = us[i + a]
                    p[va] = vs[j + b]
                    quad.append(tuple(p))
                # right-hand rule over (ua, va) gives +normal along the third axis
                # when (ua, va, ax) is a cyclic order; flip to keep normals outward.
                cyc = (ua, va, ax) in ((0, 1, 2), (1, 2, 0), (2, 0, 1))
                if (sgn > 0) != cyc:
                    quad = quad[::-1]
                out.append(_poly(mod, f"{name}.{k}", quad))
                k += 1
    return "".join(out)


def furniture_rad(rng: np.random.Generator, nboxes: int, room: Room,
                  nu: int = 3, nv: int = 7, prefix: str = "box") -> str:
    """Seeded axis-aligned boxes, sides U(0.3,1.5) m, clipped inside the room.

    Each box has its own gray plastic with rho ~ U(0.1, 0.8).  Box bottoms sit
    a few mm above the floor so no face is coplanar with a room surface.
    """
    out = []
    for b in range(nboxes):
        size = rng.uniform(0.3, 1.5, 3)
        size = np.minimum(size, [room.width - 0.02, room.depth - 0.02, room.height - 0.05])
        x = rng.uniform(0.01, room.width - size[0] - 0.01)
        y = rng.uniform(0.01, room.depth - size[1] - 0.01)
        z = rng.uniform(0.002, 0.01)
        rho = rng.uniform(0.1, 0.8)
        out.append(f"void plastic {prefix}_mat{b}\n0\n0\n5 {rho:.6g} {rho:.6g} {rho:.6g} 0 0\n")
        out.append(box_quads_rad(f"{prefix}_mat{b}", f"{prefix}{b}", (x, y, z),
                                 (x + size[0], y + size[1], z + size[2]), nu, nv))
    return "".join(out)


def config1(seed: int = 0):
    """rtrace -I -ab 0 -dj 0 -ds 0 workload: room + window + sun + glow sky, 11x7 sensors.

    Returns (rad_text, sensors (77,6), sun_dir).
    """
    rng = np.random.default_rng(seed)
    alt = rng.uniform(20.0, 60.0)
    azi = rng.uniform(120.0, 240.0)
    d = sun_direction(alt, azi)
    rad = room_rad() + sun_rad(d) + sky_rad()
    sensors = sensor_grid(11, 7, 0.5, 0.5, 0.5, 0.8)
    return rad, sensors, d


def config2(seed: int = 0, nboxes: int = 200, nsensors: int = 2500, nu: int = 3, nv: int = 7):
    """rcontrib -ab 4 -ad 4096 workload: room + window + tessellated boxes + binned sky.

    Sensors: the 60x40 interior lattice at 0.1 m pitch (x=0.05..5.95,
    y=0.05..3.95), then points of the half-offset lattice (x=0.1.., y=0.1..)
    until ``nsensors`` are reached, all at z=0.8 facing +Z.
    """
    rng = np.random.default_rng(seed)
    room = Room()
    rad = room_rad(room) + furniture_rad(rng, nboxes, room, nu, nv) + sky_rad((1, 1, 1), ground=(1, 1, 1))
    main = sensor_grid(60, 40, 0.05, 0.05, 0.1, 0.8)
    extra = sensor_grid(59, 39, 0.1, 0.1, 0.1, 0.8)
    sensors = np.concatenate([main, extra])[:nsensors]
    return rad, sensors


def reinhart_nbins(mf: int) -> int:
    """Number of sky bins of Reinhart MF:n (util/reinhartb.cal:37-47), ground excluded."""
    tnaz = [30, 30, 24, 24, 18, 12, 6]
    total = 0
    for r in range(7 * mf + 1):
        if r - (7 * mf - 0.5) > 0:
            total += 1
        else:
            total += mf * tnaz[int(math.floor((r + 0.5) / mf))]
    return total


def sky_matrix(seed: int, nbins: int, nhours: int = 8760, latitude: float = 37.8) -> np.ndarray:
    """Synthetic (nbins, nhours, 3) float32 sky matrix.

    Lognormal(0,1) patch radiances; every hour whose sun altitude at the given
    latitude is <= 0 is zeroed (simple solar-position model, local solar time).
    """
    rng = np.random.default_rng(seed)
    sky = rng.lognormal(0.0, 1.0, (nbins, nhours, 3)).astype(np.float32)
    h = np.arange(nhours)
    day = h // 24
    hour = (h % 24) + 0.5
    decl = np.radians(23.45) * np.sin(2 * np.pi * (284 + day + 1) / 365.0)
    lat = np.radians(latitude)
    ha = np.radians(15.0 * (hour - 12.0))
    sin_alt = np.sin(lat) * np.sin(decl) + np.cos(lat) * np.cos(decl) * np.cos(ha)
    sky[:, sin_alt <= 0.0, :] = 0.0
    return sky


def _box_quads(lo: np.ndarray, hi: np.ndarray, nu: int, nv: int) -> np.ndarray:
    """Vertices (nbox, 6*nu*nv, 4, 3) of boxes split as box_quads_rad() does
    (same faces, order and outward vertex order), for many boxes at once."""
    n = lo.shape[0]
    ti, tj = np.arange(nu + 1) / nu, np.arange(nv + 1) / nv
    corners = [(0, 0), (1, 0), (1, 1), (0, 1)]
    out = []
    for ax, side, ua, va, sgn in ((0, 0, 1, 2, -1), (0, 1, 1, 2, 1), (1, 0, 0, 2, -1),
                                  (1, 1, 0, 2, 1), (2, 0, 0, 1, -1), (2, 1, 0, 1, 1)):
        cyc = (ua, va, ax) in ((0, 1, 2), (1, 2, 0), (2, 0, 1))
        order = corners if (sgn > 0) == cyc else corners[::-1]
        val = hi[:, ax] if side else lo[:, ax]
        us = lo[:, ua, None] + (hi[:, ua] - lo[:, ua])[:, None] * ti
        vs = lo[:, va, None] + (hi[:, va] - lo[:, va])[:, None] * tj
        q = np.empty((n, nu, nv, 4, 3))
        for c, (a, b) in enumerate(order):
            q[:, :, :, c, ax] = val[:, None, None]
            q[:, :, :, c, ua] = us[:, a:a + nu][:, :, None]
            q[:, :, :, c, va] = vs[:, b:b + nv][:, None, :]
        out.append(q.reshape(n, nu * nv, 4, 3))
    return np.concatenate(out, axis=1)


def _quads_text(mods, quads: np.ndarray) -> bytes:
    """`mod polygon q` entries for quads (n, 4, 3) with modifier names mods[i]."""
    from . import _lib
    lines = _lib.format_ascii(quads.reshape(-1, 12), 12, "%.7g", " ", "").split(b"\n")
    heads = {m: f"{m} polygon q\n0\n0\n12 ".encode() for m in set(mods)}
    return b"".join([heads[m] + ln + b"\n" for m, ln in zip(mods, lines)])


def config5(seed: int = 0, storeys: int = 10, boxes_per_storey: int = 1980, nu: int = 3, nv: int = 7,
            sensors_per_storey: tuple = (50, 40)):
    """rcontrib -ab 5 -ad 8192 -lw 5e-5 workload: a 10-storey office of about
    5 M triangles (quads counted as two, as for config 2) under a binned sky.

    Floor plate 60 x 30 m, storey height 3.5 m (0.3 m slabs).  Per storey: a
    floor and a ceiling, facade bands (wall 0-0.9 m, glass 0.9-2.3 m = 40 % of
    the facade, wall above), a 4 m corridor along x with two-sided partition
    walls every 6 m, and `boxes_per_storey` seeded furniture boxes (sides
    0.4-1.6 m, heights 0.4-2.0 m) each split into nu x nv quads per face, with
    16 seeded gray plastics.  Sensors: a 50 x 40 grid per storey at 0.8 m
    above the floor, facing +Z (20,000 in all).  Sky and ground glow sources
    for Reinhart binning.  Returns (scene bytes, sensors (n, 6)).
    """
    rng = np.random.default_rng(seed)
    W, D, H, slab = 60.0, 30.0, 3.5, 0.3
    head = ["void plastic c5_floor\n0\n0\n5 .2 .2 .2 0 0\n", "void plastic c5_ceiling\n0\n0\n5 .7 .7 .7 0 0\n",
            "void plastic c5_wall\n0\n0\n5 .5 .5 .5 0 0\n", "void glass c5_glass\n0\n0\n3 .6 .6 .6\n"]
    for m in range(16):
        r = rng.uniform(0.1, 0.8)
        head.append(f"void plastic c5_box{m}\n0\n0\n5 {r:.6g} {r:.6g} {r:.6g} 0 0\n")
    quads, mods = [], []

    def quad(m, verts):
        quads.append(np.asarray(verts, float))
        mods.append(m)

    def wall2(m, p0, p1, z0, z1):      # vertical two-sided wall from p0 to p1 (x, y)
        a = [(p0[0], p0[1], z0), (p1[0], p1[1], z0), (p1[0], p1[1], z1), (p0[0], p0[1], z1)]
        quad(m, a)
        quad(m, a[::-1])

    for k in range(storeys):
        z0 = H * k
        zc = z0 + H - slab
        quad("c5_floor", [(0, 0, z0), (W, 0, z0), (W, D, z0), (0, D, z0)])
        quad("c5_ceiling", [(0, 0, zc), (0, D, zc), (W, D, zc), (W, 0, zc)])
        # facade, inward normals: south (y=0), east (x=W), north (y=D), west (x=0)
        ring = [((0, 0), (W, 0)), ((W, 0), (W, D)), ((W, D), (0, D)), ((0, D), (0, 0))]
        for (x0, y0), (x1, y1) in ring:
            for m, za, zb in (("c5_wall", z0, z0 + 0.9), ("c5_glass", z0 + 0.9, z0 + 2.3), ("c5_wall", z0 + 2.3, zc)):
                quad(m, [(x0, y0, za), (x0, y0, zb), (x1, y1, zb), (x1, y1, za)])
        # corridor walls (door gaps of 1 m every 6 m) and office partitions
        for y in (13.0, 17.0):
            for x in np.arange(0.0, W, 6.0):
                wall2("c5_wall", (x, y), (x + 5.0, y), z0, zc)
        for x in np.arange(6.0, W, 6.0):
            wall2("c5_wall", (x, 0.0), (x, 13.0), z0, zc)
            wall2("c5_wall", (x, 17.0), (x, D), z0, zc)
    env = np.stack(quads)
    # furniture: boxes resting 2-10 mm above each floor, inside the plate
    nb = storeys * boxes_per_storey
    size = np.stack([rng.uniform(0.4, 1.6, nb), rng.uniform(0.4, 1.6, nb), rng.uniform(0.4, 2.0, nb)], axis=1)
    lo = np.stack([rng.uniform(0.05, W - 0.05 - size[:, 0]), rng.uniform(0.05, D - 0.05 - size[:, 1]),
                   rng.uniform(0.002, 0.01, nb)], axis=1)
    lo[:, 2] += H * np.repeat(np.arange(storeys), boxes_per_storey)
    bq = _box_quads(lo, lo + size, nu, nv).reshape(-1, 4, 3)
    bmods = np.repeat(rng.integers(0, 16, nb), 6 * nu * nv)
    text = ("".join(head) + sky_rad((1, 1, 1), ground=(1, 1, 1))).encode()
    text += _quads_text(mods, env)
    text += _quads_text([f"c5_box{m}" for m in bmods], bq)
    nx, ny = sensors_per_storey
    xs = (np.arange(nx) + 0.5) * (W / nx)
    ys = (np.arange(ny) + 0.5) * (D / ny)
    gx, gy = np.meshgrid(xs, ys, indexing="ij")
    sens = []
    for k in range(storeys):
        s = np.zeros((nx * ny, 6))
        s[:, 0], s[:, 1], s[:, 2], s[:, 5] = gx.ravel(), gy.ravel(), H * k + 0.8, 1.0
        sens.append(s)
    return text, np.concatenate(sens)


def annual_wea(seed: int = 0, latitude: float = 37.62, longitude: float = 122.38, time_zone: float = 120.0) -> str:
    """Synthetic hourly weather tape for a whole year (8,760 records) in the
    WEA format gendaymtx reads (gen/gendaymtx.c:524-560): direct normal and
    diffuse horizontal irradiance (weather_data_file_units 1, W/m^2).

    Longitude and time zone are degrees west, as in WEA headers.  A cloud
    fraction drawn per hour scales direct normal irradiance; the diffuse part
    grows with cloud and sun altitude; night hours are zero.
    """
    rng = np.random.default_rng(seed)
    mdays = [31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31]
    lines = ["place synthetic", f"latitude {latitude}", f"longitude {longitude}", f"time_zone {time_zone}",
             "site_elevation 5.0", "weather_data_file_units 1"]
    lat = np.radians(latitude)
    jd = 0
    for mo, nd in enumerate(mdays, 1):
        for da in range(1, nd + 1):
            jd += 1
            dec = 0.4093 * np.sin(2 * np.pi / 365 * (jd - 81))
            for h in range(24):
                hr = h + 0.5
                st = hr - (longitude - time_zone) / 15.0
                alt = np.arcsin(np.sin(lat) * np.sin(dec) - np.cos(lat) * np.cos(dec) * np.cos(st * np.pi / 12))
                sa = max(np.sin(alt), 0.0)
                cloud = rng.random()
                dn = 900 * sa ** 0.5 * cloud if sa > 0.02 else 0.0
                dh = (60 * sa + 90 * sa * (1 - cloud) + 5 * rng.random()) if sa > 0.0 else 0.0
                lines.append(f"{mo} {da} {hr:.3f} {dn:.1f} {dh:.1f}")
    return "\n".join(lines) + "\n"


def boxes_rad(rng: np.random.Generator, nboxes: int, room: Room, mod_prefix: str = "obj",
              nmats: int = 8, alpha: float = 2.0, smin: float = 0.05, smax: float = 2.0) -> str:
    """`nboxes` axis-aligned boxes (6 quads each) with Pareto(alpha) sizes in
    [smin, smax] m, placed uniformly inside the room; `nmats` gray plastics.
    Vectorised text generation for large counts."""
    out = [f"void plastic {mod_prefix}_m{k}\n0\n0\n5 {r:.4g} {r:.4g} {r:.4g} 0 0\n"
           for k, r in enumerate(rng.uniform(0.1, 0.8, nmats))]
    if nboxes <= 0:
        return "".join(out)
    u = rng.random((nboxes, 3))
    size = smin * (1 - u) ** (-1.0 / alpha)
    size = np.minimum(size, [min(smax, room.width - 0.02), min(smax, room.depth - 0.02), min(smax, room.height - 0.05)])
    lo = np.empty((nboxes, 3))
    lo[:, 0] = rng.uniform(0.01, 1.0, nboxes) * (room.width - size[:, 0] - 0.02) + 0.01
    lo[:, 1] = rng.uniform(0.01, 1.0, nboxes) * (room.depth - size[:, 1] - 0.02) + 0.01
    lo[:, 2] = rng.uniform(0.0, 1.0, nboxes) * (room.height - size[:, 2] - 0.05) + 0.003
    hi = lo + size
    mats = rng.integers(0, nmats, nboxes)
    # unit-cube quads with outward normals (counter-clockwise seen from outside)
    faces = np.array([
        [[0, 0, 0], [0, 0, 1], [0, 1, 1], [0, 1, 0]],   # -x
        [[1, 0, 0], [1, 1, 0], [1, 1, 1], [1, 0, 1]],   # +x
        [[0, 0, 0], [1, 0, 0], [1, 0, 1], [0, 0, 1]],   # -y
        [[0, 1, 0], [0, 1, 1], [1, 1, 1], [1, 1, 0]],   # +y
        [[0, 0, 0], [0, 1, 0], [1, 1, 0], [1, 0, 0]],   # -z
        [[0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]],   # +z
    ], float)
    verts = lo[:, None, None, :] + faces[None] * size[:, None, None, :]     # (n, 6, 4, 3)
    coords = verts.reshape(nboxes * 6, 12)
    out.append(_polygon_records(coords, mats.repeat(6), f"{mod_prefix}_m", mod_prefix).decode())
    return "".join(out)


def _fixed6(v: np.ndarray) -> np.ndarray:
    """Non-negative values below 1000 as fixed-width 'ddd.dddddd' ASCII
    (uint8 array (..., 10)); the decimal value is v rounded to 1e-6, as '%.6f'
    gives it."""
    if v.size and (v.min() < 0 or v.max() >= 999.9999995):
        raise ValueError("coordinate outside [0, 1000)")
    q = np.rint(v * 1e6).astype(np.int64)
    digits = np.empty(v.shape + (10,), np.uint8)
    for k, pw in enumerate((10 ** 8, 10 ** 7, 10 ** 6)):
        digits[..., k] = 48 + (q // pw) % 10
    digits[..., 3] = ord(".")
    for k, pw in enumerate((10 ** 5, 10 ** 4, 10 ** 3, 10 ** 2, 10, 1)):
        digits[..., 4 + k] = 48 + (q // pw) % 10
    return digits


def _polygon_records(coords: np.ndarray, mats: np.ndarray, mat_prefix: str, name_prefix: str) -> bytes:
    """Quadrilateral polygons as scene text, built as fixed-width byte records
    with numpy (the float formatting of millions of vertices in Python is the
    slow part of the large workloads): 'MAT polygon NAME\n0\n0\n12 x y z ...'."""
    n = len(coords)
    if n == 0:
        return b""
    mw = len(str(int(mats.max())))
    nw = len(str(n - 1))
    head = np.frombuffer(f"{mat_prefix}{'0' * mw} polygon {name_prefix}{'0' * nw}\n0\n0\n12".encode(), np.uint8)
    rec = len(head) + 12 * 11 + 1
    buf = np.empty((n, rec), np.uint8)
    buf[:, :len(head)] = head
    def put_int(col, width, vals):
        for k in range(width):
            buf[:, col + k] = 48 + (vals // 10 ** (width - 1 - k)) % 10
    put_int(len(mat_prefix), mw, mats.astype(np.int64))
    put_int(len(f"{mat_prefix}{'0' * mw} polygon {name_prefix}"), nw, np.arange(n, dtype=np.int64))
    digits = _fixed6(coords)                          # (n, 12, 10)
    body = np.empty((n, 12, 11), np.uint8)
    body[:, :, 0] = ord(" ")
    body[:, :, 1:] = digits
    buf[:, len(head):len(head) + 132] = body.reshape(n, 132)
    buf[:, -1] = ord("\n")
    return buf.tobytes()


def config3(seed: int = 0, nscenes: int = 64, prims=(1_000, 500_000), sensors=(100, 20_000)):
    """Batched workload: `nscenes` shoebox rooms with seeded dimensions
    U(3,20) m, log-uniform primitive counts (6-quad boxes, Pareto sizes) and
    log-uniform sensor counts, each with its own sun.
    Returns a list of (rad_text, sensors (n,6))."""
    rng = np.random.default_rng(seed)
    res = []
    for k in range(nscenes):
        dims = rng.uniform(3.0, 20.0, 3)
        dims[2] = min(dims[2], 6.0)
        room = Room(width=dims[0], depth=dims[1], height=max(dims[2], 3.0),
                    win_w=min(2.0, dims[0] - 0.5), win_h=1.2, win_sill=0.9)
        nprim = int(np.exp(rng.uniform(np.log(prims[0]), np.log(prims[1]))))
        nsens = int(np.exp(rng.uniform(np.log(sensors[0]), np.log(sensors[1]))))
        alt, azi = rng.uniform(20, 60), rng.uniform(120, 240)
        rad = (room_rad(room) + boxes_rad(rng, max(nprim // 6, 1), room, f"s{k}o")
               + sun_rad(sun_direction(alt, azi)) + sky_rad((1, 1, 1), ground=(1, 1, 1)))
        # sensors: square lattice on the floor area at 0.8 m
        area = (room.width - 0.2) * (room.depth - 0.2)
        pitch = np.sqrt(area / nsens)
        nx = max(int((room.width - 0.2) / pitch), 1)
        ny = max(int(np.ceil(nsens / nx)), 1)
        g = sensor_grid(nx, ny, 0.1 + pitch / 2, 0.1 + pitch / 2, pitch, 0.8)[:nsens]
        g[:, 1] = np.minimum(g[:, 1], room.depth - 0.05)
        res.append((rad, g))
    return res


def _quad(p, u, v):
    p, u, v = (np.asarray(x, float) for x in (p, u, v))
    return [tuple(p), tuple(p + u), tuple(p + u + v), tuple(p + v)]


def config4_groups(room: Room | None = None):
    """The config-1 room surfaces split into 24 patches (name, vertices): floor,
    ceiling and the north, west and east walls 2 x 2 each, the south wall as
    the four rectangles around the window.  Every quad faces into the room."""
    r = room or Room()
    W, D, H = r.width, r.depth, r.height
    walls = [((0, 0, 0), (W, 0, 0), (0, D, 0)),      # floor, +z
             ((0, 0, H), (0, D, 0), (W, 0, 0)),      # ceiling, -z
             ((0, D, 0), (W, 0, 0), (0, 0, H)),      # north, -y
             ((0, 0, 0), (0, D, 0), (0, 0, H)),      # west, +x
             ((W, 0, 0), (0, 0, H), (0, D, 0))]      # east, -x
    quads = []
    for p, u, v in walls:
        p, u, v = (np.asarray(x, float) for x in (p, u, v))
        for i in range(2):
            for j in range(2):
                quads.append(_quad(p + u * (i / 2) + v * (j / 2), u / 2, v / 2))
    x0 = 0.5 * (W - r.win_w)
    x1 = x0 + r.win_w
    z0 = r.win_sill
    z1 = z0 + r.win_h
    for xa, xb, za, zb in [(0, x0, 0, H), (x1, W, 0, H), (x0, x1, 0, z0), (x0, x1, z1, H)]:
        quads.append(_quad((xa, 0, za), (0, 0, zb - za), (xb - xa, 0, 0)))
    return [(f"g{k:02d}", q) for k, q in enumerate(quads)]


def config4_theta(seed: int = 0, ngroups: int = 24):
    """(ngroups, 5) plastic arguments: RGB reflectance U(0.05, 0.9),
    specularity U(0, 0.05), roughness U(0, 0.2)."""
    rng = np.random.default_rng(seed)
    th = np.empty((ngroups, 5))
    th[:, :3] = rng.uniform(0.05, 0.9, (ngroups, 3))
    th[:, 3] = rng.uniform(0.0, 0.05, ngroups)
    th[:, 4] = rng.uniform(0.0, 0.2, ngroups)
    return th


def config4(theta: np.ndarray | None = None, seed: int = 0, nsensors: int = 1000):
    """Differentiable workload: the config-1 room with 24 plastic material
    groups, glass window, binned glow sky and ground (no sun), 1000 sensors
    on a 40 x 25 lattice at 0.8 m facing +Z.
    Returns (rad_text, sensors, group_names)."""
    room = Room()
    if theta is None:
        theta = config4_theta(seed)
    groups = config4_groups(room)
    out = []
    for (name, q), th in zip(groups, theta):
        out.append(f"void plastic {name}\n0\n0\n5 {th[0]:.9g} {th[1]:.9g} {th[2]:.9g} {th[3]:.9g} {th[4]:.9g}\n")
        out.append(_poly(name, f"{name}_surf", q))
    x0 = 0.5 * (room.width - room.win_w)
    z0 = room.win_sill
    out.append(f"void glass window_glass\n0\n0\n3 {room.glass_tau} {room.glass_tau} {room.glass_tau}\n")
    out.append(_poly("window_glass", "window", [(x0, 0, z0), (x0, 0, z0 + room.win_h),
                                                (x0 + room.win_w, 0, z0 + room.win_h), (x0 + room.win_w, 0, z0)]))
    out.append(sky_rad((1, 1, 1), ground=(1, 1, 1)))
    sens = sensor_grid(40, 25, 0.075, 0.08, 0.15, 0.8)
    sens[:, 1] = 0.08 + 0.16 * np.repeat(np.arange(25), 40)
    return "".join(out), sens[:nsensors], [g for g, _ in groups]
