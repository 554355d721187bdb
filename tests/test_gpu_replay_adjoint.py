"""Path-replay adjoint of the differentiable irradiance workload (the
config-1 room with 24 plastic groups and a binned sky): gradients against
central finite differences with common random numbers, at the workload's
size too; the roughness derivative against the closed form of the Gaussian
lobe; metal, trans and shadow-ray (direct light) deposits; the sky gradient
against the coefficient matrix, and the torch.autograd wrapper."""

import numpy as np
import pytest

from pyradiance_amd import scenes

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def L():
    from pyradiance_amd import _lib
    _lib.lib()
    return _lib


def _sky_mods(L):
    m = L.RgModifier()
    m.name, m.bin_kind, m.nbins, m.mf, m.rhs = b"skyglow", 1, 145, 1.0, 1.0
    m.nrm[:] = (0, 0, -1)
    m.up[:] = (0, 1, 0)
    g = L.RgModifier()
    g.name, g.bin_kind, g.nbins, g.cbin = b"groundglow", 0, 1, 0
    return [m, g]


def _params(L, ab, ad, seed, roulette, minweight=1e-7):
    p = L.default_params(True)
    p.ambounce, p.ambdiv, p.seed = ab, ad, seed
    # Without roulette (-lr > 0) a ray is dropped below -lw; a specular
    # child's weight is proportional to the specularity, so a finite step
    # moves paths across that cut-off and the difference quotient picks up
    # their jumps (at -lw 1e-4 and h = 1e-3 the specular-path part reads 1.40
    # against 0.99 at h = 1e-4 and 1e-5: the cut-off, not the derivative).
    # The cut-off is therefore set below every path weight here.  With
    # roulette on, a weight limit that the second-bounce rays fall below.
    p.minweight = 5e-3 if roulette else minweight
    p.maxdepth = -10 if roulette else 8
    return p


def _scene(L, nsens, ab, ad, seed=3, roulette=False, minweight=1e-7):
    from pyradiance_amd.adjoint import DiffScene
    theta = scenes.config4_theta(0)
    rad, sens, groups = scenes.config4(theta)
    sc = DiffScene(rad, sens[:nsens], groups, _sky_mods(L), _params(L, ab, ad, seed, roulette, minweight))
    return sc, theta


def _loss_parts(sc, theta, w, sky):
    sc.set_theta(theta)
    dc = sc.coefficients()
    e = np.einsum("sbc,bc->sc", dc, sky)
    return float((w * e).sum()), dc


# h = 1e-3 as specified for the workload.  With roulette on, survival
# decisions (rand < weight / -lw) flip under any practical step for the
# parameters that set a ray's weight (the largest reflectance channel, the
# specularity): each flip moves the loss by a whole path, so those
# differences are noisy (on the CPU oracle one specularity derivative reads
# -0.0003, 0.003 and 0.026 at h = 1e-4, 1e-5, 1e-6).  The roulette case
# therefore allows an absolute error of 2 % of the largest gradient; the
# parameters that do not set a weight agree to all digits there too.
@pytest.mark.parametrize("roulette,h,floor", [(False, 1e-3, 0.005), (True, 1e-4, 0.02)])
def test_material_gradient_matches_central_differences(L, roulette, h, floor):
    _say(f"adjoint vs central differences, roulette={roulette}, h={h}")
    sc, theta = _scene(L, 120, 2, 256, roulette=roulette)
    rng = np.random.default_rng(5)
    sky = rng.lognormal(0, 0.5, (146, 3))
    w = rng.uniform(0.5, 1.5, (120, 3))            # linear loss: dLoss/dE = w
    sc.set_theta(theta)
    g = sc.adjoint(w, sky)
    assert g.shape == (24, 5)
    gmax = np.abs(g).max()
    assert gmax > 0
    checked = 0
    for grp in (1, 5, 9, 13, 18, 21):     # specularities away from -st 0.02
        for q in (0, 1, 2, 3):
            tp, tm = theta.copy(), theta.copy()
            tp[grp, q] += h
            tm[grp, q] -= h
            fd = (_loss_parts(sc, tp, w, sky)[0] - _loss_parts(sc, tm, w, sky)[0]) / (2 * h)
            ad = g[grp, q]
            tol = 0.03 * max(abs(fd), abs(ad)) + floor * gmax
            assert abs(fd - ad) <= tol, (grp, q, fd, ad)
            checked += 1
    assert checked == 24
    sc.close()


def test_sky_gradient_is_the_coefficient_matrix(L):
    import torch
    from pyradiance_amd.adjoint import irradiance
    _say("sky gradient = coefficient matrix")
    sc, theta = _scene(L, 60, 2, 128)
    sky = torch.tensor(np.random.default_rng(1).lognormal(0, 0.5, (146, 3)), dtype=torch.float64,
                       device="cuda", requires_grad=True)
    th = torch.tensor(theta, dtype=torch.float64, device="cuda", requires_grad=True)
    e = irradiance(th, sky, sc)
    wt = torch.rand(e.shape, dtype=e.dtype, device=e.device)
    (wt * e).sum().backward()
    sc.set_theta(theta)
    dc = sc.coefficients()
    ref = np.einsum("sbc,sc->bc", dc, wt.cpu().numpy())
    assert np.allclose(sky.grad.cpu().numpy(), ref, rtol=1e-5, atol=1e-9)
    # material gradient through autograd equals the direct adjoint call
    g = sc.adjoint(wt.cpu().numpy(), sky.detach().cpu().numpy())
    assert np.allclose(th.grad.cpu().numpy(), g, rtol=1e-9, atol=1e-12)
    # forward agrees with DC x sky
    e_ref = np.einsum("sbc,bc->sc", dc, sky.detach().cpu().numpy())
    assert np.allclose(e.detach().cpu().numpy(), e_ref, rtol=1e-5)
    sc.close()


def test_adjoint_replays_the_forward_paths(L):
    """With dLoss/dE = 1 on one channel and L = 1 on one column, a group's
    reflectance gradient of that channel times rho equals the part of the
    coefficients that passes through that group (log-derivative identity,
    no roulette)."""
    _say("adjoint path identity")
    sc, theta = _scene(L, 40, 2, 64)
    sky = np.ones((146, 3))
    w = np.zeros((40, 3))
    w[:, 0] = 1.0
    sc.set_theta(theta)
    g = sc.adjoint(w, sky)
    # at -ab 2 a path meets at most one diffuse group vertex, so E_red is
    # linear in each group's rho_red: rho * dE/drho sums to the part of E_red
    # that was reflected diffusely once
    dc = sc.coefficients()
    e_red = dc[:, :, 0].sum()
    assert np.all(g[:, 0] >= 0)
    total = float((g[:, 0] * theta[:, 0]).sum())
    assert 0 < total < e_red
    sc.close()


def _lobe_mods(L):
    sky = _sky_mods(L)[0]
    g = L.RgModifier()
    g.name, g.bin_kind, g.nbins, g.mf, g.rhs = b"groundglow", 1, 145, 1.0, 1.0
    g.nrm[:] = (0, 0, 1)
    g.up[:] = (0, 1, 0)
    return [sky, g]


@pytest.mark.parametrize("alpha", [0.08, 0.2])
def test_roughness_gradient_closed_form(L, alpha):
    """The roughness derivative of a Gaussian lobe's binned output against
    the derivative of its closed form (tests/lobes.py): a rough white metal
    mirror (reflection) and a rough clear trans pane (transmission) seen at
    normal incidence, -ab 0 -st 0, loss = sum over altitude rows of
    v_row x (row's share).  The score-function estimate averaged over 8
    seeds of 2^17 samples must sit within 4 standard errors of
    d/dalpha sum_row v_row P_row(alpha).  The other columns are exact:
    own = rho x rspec (metal) and rho x trans x tspec (trans) at all-ones
    arguments give d/drho_red = d/dspec = E (metal) and d/drho_red = E,
    d/dspec = -E, d/dtrans = d/dtspec = E (trans)."""
    import lobes
    _say(f"roughness gradient vs closed-form lobe, alpha {alpha}")
    n, seeds = 1 << 17, 8
    c = L.Context(0)
    c.load_text(lobes.lobe_scene(alpha)).commit()
    mods = _lobe_mods(L)
    v = np.array([0.3, 1.7, -0.8, 2.2, 0.5, -1.1, 1.4, 0.9])       # weight per altitude row
    vb = np.zeros(145)
    for r, (a, b) in enumerate(zip(lobes.ROW_START[:-1], lobes.ROW_START[1:])):
        vb[a:b] = v[r]
    dlde = np.array([[1.0, 0.0, 0.0]])
    for ray, trans, name, half in ((lobes.REFLECT_RAY, False, "lobem", 0), (lobes.TRANSMIT_RAY, True, "lobet", 1)):
        mat = c.material_index(name)
        lcol = np.zeros((290, 3))
        lcol[145 * half:145 * half + 145, 0] = vb
        ga = []
        for s in range(seeds):
            p = L.default_params(True)
            p.ambounce, p.specthresh, p.specjitter, p.minweight, p.seed = 0, 0.0, 1.0, 1e-6, 101 + s
            c.set_params(p)
            g = c.contrib_adjoint(np.tile(ray, (n, 1)), mods, dlde, lcol.ravel(), [mat], n, 0)[0]
            e = c.contrib(np.tile(ray, (n, 1)), mods, n, 0).reshape(290, 3)[:, 0] @ lcol[:, 0]
            if trans:
                assert np.allclose(g[[0, 3, 5, 6]], [e, -e, e, e], rtol=1e-6, atol=1e-12), (g, e)
            else:
                assert np.allclose(g[[0, 3]], [e, e], rtol=1e-6, atol=1e-12), (g, e)
                assert g[5] == 0.0 and g[6] == 0.0
            ga.append(g[4])
        h = 1e-5
        want = (v @ lobes.expected_rows(alpha + h, trans) - v @ lobes.expected_rows(alpha - h, trans)) / (2 * h)
        ga = np.array(ga)
        mean, se = ga.mean(), ga.std(ddof=1) / np.sqrt(seeds)
        _say(f"  {'T' if trans else 'R'}: adjoint {mean:.4f} +- {se:.4f}, closed form {want:.4f}")
        assert abs(mean - want) <= 4 * se + 1e-3 * abs(want), (mean, se, want)
        assert se < 0.1 * abs(want)                   # the check has power
    c.close()


MIXED = {2: "metal", 9: "metal", 5: "trans", 21: "trans"}     # group -> material type


def _mixed_scene(L, nsens, ab, ad, specthresh, seed=3):
    """Config 4's room with two groups turned to metal and two to trans
    (transmissivity 0.3, transmitted specularity 0.5) and a sun tracked
    beside the sky and ground, so shadow-ray deposits enter the loss."""
    import re
    from pyradiance_amd.adjoint import DiffScene
    theta = np.zeros((24, 7))
    theta[:, :5] = scenes.config4_theta(0)
    rad, sens, groups = scenes.config4(theta[:, :5])
    for gi, kind in MIXED.items():
        name = groups[gi]
        args = theta[gi, :5]
        if kind == "trans":
            theta[gi, 5:] = (0.3, 0.5)
            args = theta[gi]
        txt = f"void {kind} {name}\n0\n0\n{len(args)} " + " ".join(f"{x:.9g}" for x in args) + "\n"
        rad, k = re.subn(rf"void plastic {name}\n0\n0\n5 [^\n]*\n", txt, rad)
        assert k == 1
    rad += scenes.sun_rad(scenes.sun_direction(40.0, 180.0), radiance=(1e4, 1e4, 1e4))
    sun = L.RgModifier()
    sun.name, sun.bin_kind, sun.nbins, sun.cbin = b"solar", 0, 1, 0
    p = _params(L, ab, ad, seed, False)
    p.specthresh = specthresh
    sc = DiffScene(rad, sens[:nsens], groups, _sky_mods(L) + [sun], p)
    assert sc.width == 7
    return sc, theta


# -st 0.02 (lobe samples drawn): every column but the roughness, whose
# derivative is a likelihood-ratio score that finite differences with common
# random numbers do not see.  -st 1 (every specular part folded into the
# ambient value, no lobe samples): the roughness then only enters dirnorm()'s
# lobe on shadow rays, a smooth function, so it is checked too.
@pytest.mark.parametrize("st,cols", [(0.02, (0, 1, 2, 3, 5, 6)), (1.0, (0, 1, 2, 3, 4, 5, 6))])
def test_metal_trans_and_direct_gradients_match_central_differences(L, st, cols):
    _say(f"metal/trans/direct adjoint vs central differences, -st {st}")
    h = 1e-3
    sc, theta = _mixed_scene(L, 120, 2, 256, st)
    rng = np.random.default_rng(7)
    sky = np.concatenate([rng.lognormal(0, 0.5, (146, 3)), [[2.0, 1.5, 1.0]]])
    w = rng.uniform(0.5, 1.5, (120, 3))
    sc.set_theta(theta)
    dc = sc.coefficients()
    assert dc[:, 146].sum() > 0                      # the sun is reached by shadow rays
    g = sc.adjoint(w, sky)
    assert g.shape == (24, 7)
    gmax = np.abs(g).max()
    checked = 0
    for grp in (1, 2, 5, 9, 13, 21):
        for q in cols:
            if q >= 5 and MIXED.get(grp) != "trans":
                assert g[grp, q] == 0.0
                continue
            if q == 4 and theta[grp, 4] ** 2 <= 1e-6:
                continue
            tp, tm = theta.copy(), theta.copy()
            tp[grp, q] += h
            tm[grp, q] -= h
            fd = (_loss_parts(sc, tp, w, sky)[0] - _loss_parts(sc, tm, w, sky)[0]) / (2 * h)
            ad = g[grp, q]
            tol = 0.03 * max(abs(fd), abs(ad)) + 0.005 * gmax
            assert abs(fd - ad) <= tol, (grp, q, fd, ad)
            checked += 1
    assert checked >= 24
    sc.close()


def test_config4_workload_gradient(L):
    """Config 4 at its workload: 1000 sensors, -ab 3 -ad 2048.  Reflectance
    and specularity gradients of every group against central differences
    with common random numbers (h = 1e-3, no roulette); the roughness
    gradient, a likelihood-ratio estimate, against central differences
    (h = 1e-2) over 8 seeds, per group within 4 combined standard errors."""
    _say("config 4 workload: adjoint vs central differences")
    # -ad 2048 starts the first bounce at weight 1/2025, and chains of
    # specular children (weight x specularity each) reach a 1e-7 cut-off,
    # where a finite step moves paths across it: the cut-off goes to 1e-12.
    sc, theta = _scene(L, 1000, 3, 2048, minweight=1e-12)
    rng = np.random.default_rng(11)
    sky = rng.lognormal(0, 0.5, (146, 3))
    w = rng.uniform(0.5, 1.5, (1000, 3))
    sc.set_theta(theta)
    g = sc.adjoint(w, sky)
    gmax = np.abs(g[:, :4]).max()
    h = 1e-3
    worst = 0.0
    for grp in range(24):
        if abs(theta[grp, 3] - 0.02) < 2 * h:        # specularity at the -st cut-off
            continue
        for q in range(4):
            if theta[grp, q] < 2 * h:                    # the step would clip at zero
                continue
            tp, tm = theta.copy(), theta.copy()
            tp[grp, q] += h
            tm[grp, q] -= h
            fd = (_loss_parts(sc, tp, w, sky)[0] - _loss_parts(sc, tm, w, sky)[0]) / (2 * h)
            err = abs(fd - g[grp, q])
            worst = max(worst, err / (0.03 * max(abs(fd), abs(g[grp, q])) + 0.005 * gmax))
            assert err <= 0.03 * max(abs(fd), abs(g[grp, q])) + 0.005 * gmax, (grp, q, fd, g[grp, q])
    _say(f"  rho/spec: worst error / tolerance {worst:.3f}")
    # roughness: 8 seeds of adjoint and of central differences
    hr, seeds = 1e-2, 8
    rough = [k for k in range(24) if theta[k, 4] > 2 * hr]
    ga, fa = [], []
    for s in range(seeds):
        p = _params(L, 3, 2048, 50 + s, False, 1e-12)
        sc.ctx.set_params(p)
        sc.set_theta(theta)
        ga.append(sc.adjoint(w, sky)[rough, 4])
        f = []
        for grp in rough:
            tp, tm = theta.copy(), theta.copy()
            tp[grp, 4] += hr
            tm[grp, 4] -= hr
            f.append((_loss_parts(sc, tp, w, sky)[0] - _loss_parts(sc, tm, w, sky)[0]) / (2 * hr))
        fa.append(f)
    ga, fa = np.array(ga), np.array(fa)
    se = np.sqrt(ga.var(0, ddof=1) / seeds + fa.var(0, ddof=1) / seeds)
    z = (ga.mean(0) - fa.mean(0)) / np.maximum(se, 1e-30)
    _say(f"  roughness, {len(rough)} groups: |z| max {np.abs(z).max():.2f}, mean z^2 {np.mean(z * z):.2f}; "
         f"groups with |adjoint| > 2 se: {int((np.abs(ga.mean(0)) > 2 * ga.std(0, ddof=1) / np.sqrt(seeds)).sum())}")
    assert np.all(np.abs(z) <= 4.0), z
    assert np.mean(z * z) < 3.0
    sc.close()
