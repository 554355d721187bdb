"""Closed forms of the Gaussian specular lobe that gaussamp() samples
(rt/normal.c:364-495), at normal incidence.

A sample perturbs the surface normal (reflection) or the ray direction
(transmission) by d = alpha * sqrt(-ln u) along a uniformly random azimuth,
u uniform in (0, 1].  Reflecting the incident ray about the perturbed
normal turns it by twice the normal's tilt, so the reflected polar angle is
theta = 2 atan(d); the transmitted one is theta = atan(d).  Hence
P(theta > x) = exp(-tan^2(x/2) / alpha^2) (reflection) and
exp(-tan^2(x) / alpha^2) (transmission): fractions per 12-degree altitude
band of a Reinhart MF:1 sky (rows of util/reinhartb.cal) follow.
"""

from __future__ import annotations

import numpy as np

# first bin of each Reinhart MF:1 row and the end (util/reinhartb.cal:37-49)
ROW_START = np.array([0, 30, 60, 84, 108, 126, 138, 144, 145])


def row_fractions(binned: np.ndarray) -> np.ndarray:
    """Per-row sums of per-bin values (..., 145) -> (..., 8)."""
    return np.stack([binned[..., a:b].sum(-1) for a, b in zip(ROW_START[:-1], ROW_START[1:])], -1)


def expected_rows(alpha: float, transmitted: bool = False) -> np.ndarray:
    """Probability of each altitude row (row r: altitude in [12 r, 12 r + 12)
    degrees, row 7 the cap above 84) for a lobe around the zenith."""
    def surv(theta_deg):
        t = np.radians(theta_deg)
        x = np.tan(t) if transmitted else np.tan(t / 2)
        return np.exp(-x * x / (alpha * alpha))
    # polar angle bounds of row r: (78 - 12 r, 90 - 12 r]
    p = np.array([surv(78.0 - 12 * r) - surv(90.0 - 12 * r) for r in range(7)] + [1.0 - surv(6.0)])
    return p / (1.0 - surv(90.0))         # samples below the surface are drawn again


def lobe_scene(alpha: float) -> str:
    """A rough mirror-white metal floor at z=0 and, below z=0, a rough clear
    trans pane at z=-2 (diffuse parts zero), under a glow sky (upper
    hemisphere) and a glow 'ground' (lower hemisphere), both radiance 1."""
    return (f"void metal lobem 0 0 5 1 1 1 1 {alpha}\n"
            "lobem polygon floor 0 0 12 -4 -4 0 4 -4 0 4 4 0 -4 4 0\n"
            f"void trans lobet 0 0 7 1 1 1 0 {alpha} 1 1\n"
            "lobet polygon pane 0 0 12 14 -4 -2 24 -4 -2 24 4 -2 14 4 -2\n"
            "void glow skyglow 0 0 4 1 1 1 0\nskyglow source sky 0 0 4 0 0 1 180\n"
            "void glow groundglow 0 0 4 1 1 1 0\ngroundglow source ground 0 0 4 0 0 -1 180\n")


# rays straight down onto the metal (reflected lobe around +z) and through
# the pane (transmitted lobe around -z)
REFLECT_RAY = np.array([0.0, 0.0, 1.0, 0.0, 0.0, -1.0])
TRANSMIT_RAY = np.array([19.0, 0.0, -1.0, 0.0, 0.0, -1.0])


def check(rows: np.ndarray, n: int, alpha: float, transmitted: bool, k: float = 5.0) -> None:
    """rows: observed fraction per row from n samples; binomial k-sigma."""
    p = expected_rows(alpha, transmitted)
    sd = np.sqrt(p * (1 - p) / n) + 1.0 / n
    assert np.all(np.abs(rows - p) <= k * sd), (rows, p, sd)
