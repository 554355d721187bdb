"""Three-sigma comparison of two sets of independent Monte Carlo estimates.

Each set gives, per entry (sensor, bin, channel), the mean and the sample
variance of R independent estimates.  The difference of the means of entry
k has standard error se_k = sqrt(var_a/R_a + var_b/R_b); z_k = |mean_a -
mean_b| / se_k.  Two correct implementations of the same estimator give
z-scores close to |N(0,1)|: about 0.3-0.5 % above 3 and a mean z^2 near 1
(two independent reference rcontrib runs of the config-2 room, 96 sensors x
16 estimates: 0.44 % above 3, mean z^2 1.10, aggregate difference 0.04 %).
Entries where both sets have zero variance must have equal means.
"""

from __future__ import annotations

import numpy as np


def summary(mean_a, var_a, ra, mean_b, var_b, rb) -> dict:
    mean_a, mean_b = np.asarray(mean_a, float), np.asarray(mean_b, float)
    se = np.sqrt(np.asarray(var_a, float) / ra + np.asarray(var_b, float) / rb)
    ok = se > 0
    z = np.abs(mean_a - mean_b)[ok] / se[ok]
    tot_a, tot_b = float(mean_a.sum()), float(mean_b.sum())
    return {
        "n": int(ok.sum()),
        "const_mismatch": int(((mean_a != mean_b) & ~ok).sum()),
        "frac_over_3": float((z > 3).mean()) if z.size else 0.0,
        "mean_z2": float((z ** 2).mean()) if z.size else 1.0,
        "max_z": float(z.max()) if z.size else 0.0,
        "aggregate": abs(tot_a - tot_b) / max(abs(tot_a), 1e-300),
    }


def check(s: dict, frac: float = 0.01, z2=(0.6, 1.6), aggregate: float = 0.01) -> None:
    assert s["const_mismatch"] == 0, s
    assert s["frac_over_3"] <= frac, s
    assert z2[0] <= s["mean_z2"] <= z2[1], s
    assert s["aggregate"] < aggregate, s
