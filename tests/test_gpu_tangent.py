"""Derivatives of the TTM energy and gradient with respect to the model's
vdW, polarisability, Thole and DMS-correction parameters (pimd_ttm_param_grad).

Fitting sets of gas-phase 5-water clusters (in a 40 A box) and periodic
32-water boxes, with targets from the FP64 forward model at parameters
perturbed by 2 % (Gaussian, seeded).  Checked against central finite
differences (h = 1e-6 relative) of the GPU forward model and of the oracle
with its parameters overridden; the bar is 1e-6 relative.  A central
difference carries a rounding error of about eps |E| / h, which exceeds that
bar for a parameter as small as dms_param3 (0.012, so h = 1.2e-8 against
energies of hundreds of kcal/mol); the h = 1e-6 check allows that bound on
top, and a five-point difference with h = 1e-4 relative (truncation error
O(h^4), rounding 100 times smaller) checks the parameters at 1e-7.  The
five-point check leaves out the parameters in the argument of the Thole
charge term's Gamma(2/3, x) (polfacO/H/M, aCCaCD): the reference's gammq2_3
(math.f90:188-285) switches from its series to its continued fraction at
x = 5/3 and stops each at a data-dependent number of terms, so the function
jumps by up to ~3e-7 relative along those parameters, and points 4e-4 apart
straddle such jumps.  The value parts of the tangent evaluation must
reproduce the forward energies (1e-10).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FD_REL = 1e-6
FD5_REL = 1e-7
EPS = np.finfo(np.float64).eps


POLAR = (5, 6, 7)   # polarisabilities: must not go negative
GAMMA_ARG = (8, 9, 10, 12)   # polfacO/H/M, aCCaCD: in the argument of gammq2_3
# the size of each parameter in the model that uses it, for the steps of
# parameters that are zero in the other model (vdwA, vdwB of TTM3-F, the DMS
# corrections of TTM2.1-F)
SCALE = np.array([1.33e6, 3.63e5, 2.15e3, 1e13, 13.2, 0.837, 0.496, 1.444, 0.837, 0.496, 0.837,
                  0.3, 0.2, 0.5, 0.9578, 0.012])


def _h(theta, j, rel):
    return rel * (abs(theta[j]) if theta[j] != 0.0 else SCALE[j])


def _pm(theta, j):
    """(theta+, theta-, their distance): central differences with h = 1e-6
    relative.  A zero polarisability (those of O and H in TTM3-F, which do not
    enter at all) takes a forward difference instead."""
    tp, tm = theta.copy(), theta.copy()
    h = _h(theta, j, 1e-6)
    if theta[j] == 0.0 and j in POLAR:
        tp[j] += h
        return tp, tm, h
    tp[j] += h
    tm[j] -= h
    return tp, tm, 2 * h


def _five_point(theta, j):
    """Points theta + m h e_j (m = 2, 1, -1, -2) with h = 1e-4 relative
    (forward points for a zero polarisability) for the five-point first
    derivative."""
    h = _h(theta, j, 1e-4)
    ms = (4, 3, 1, 0) if theta[j] == 0.0 and j in POLAR else (2, 1, -1, -2)
    pts = []
    for m in ms:
        t = theta.copy()
        t[j] += m * h
        pts.append(t)
    return pts, h


def _groups(model, ncl=6, nbox=4, seed=5000, polar_eps=1e-12):
    from pimd_amd.fit import FitGroup
    from pimd_amd.params import TTMParams
    from pimd_amd.systems import distorted_box, water_cluster
    from pimd_amd.ttm import TTMForceField
    cl = np.array([water_cluster(5, seed=seed + i) for i in range(ncl)])
    bx = []
    for i in range(nbox):
        r, L = distorted_box(32, seed=seed + 500 + i)
        bx.append(r)
    bx = np.array(bx)
    common = dict(pot_model=model, eps_ewald=1e-8, polar_eps=polar_eps, polar_maxiter=100)
    pcl = TTMParams(nwaters=5, nbeads=ncl, box=(40.0, 40.0, 40.0), **common)
    pbx = TTMParams(nwaters=32, nbeads=nbox, box=(L, L, L), rc=4.9, **common)
    ffc, ffb = TTMForceField(pcl), TTMForceField(pbx)
    theta0 = ffc.ttm_params()
    rng = np.random.default_rng(seed)
    theta_t = theta0 * (1.0 + 0.02 * rng.standard_normal(theta0.size))
    pes0 = ffc.pes_params()
    groups = []
    for ff, R in ((ffc, cl), (ffb, bx)):
        ff.set_ttm_params(theta_t)
        r = ff.full_bead_forces(R, R[0])
        groups.append(FitGroup(ff, R, r["Upott"], r["dRRt"]))
        ff.set_ttm_params(theta0)
    return groups, theta0, pes0


@pytest.fixture(scope="module", params=[3, 2], ids=["ttm3f", "ttm21f"])
def fitset(request):
    from pimd_amd.fit import loss_and_grad
    groups, theta0, pes0 = _groups(request.param)
    full = np.concatenate([pes0, theta0])
    L0, g = loss_and_grad(groups, full, method="tangent")
    return request.param, groups, theta0, pes0, L0, g[pes0.size:]


def _forward(groups, theta):
    out = []
    for grp in groups:
        grp.ff.set_ttm_params(theta)
        r = grp.ff.full_bead_forces(grp.RRt, grp.RRc)
        out.append((r["Upott"], r["dRRt"]))
    for grp in groups:
        grp.ff.set_ttm_params(theta)
    return out


def _central(plus, minus, groups, h2):
    """[L(theta+) - L(theta-)] / h2 summed as (a+ - a-)(a+ + a- - 2 target),
    and a bound on its rounding error: each a carries about eps |a| (a few
    eps for sums of many terms)."""
    tot, noise = 0.0, 0.0
    for (Ep, Gp), (Em, Gm), grp in zip(plus, minus, groups):
        wE, wG = Ep + Em - 2 * grp.E_target, Gp + Gm - 2 * grp.G_target
        tot += np.sum((Ep - Em) * wE)
        tot += np.sum((Gp - Gm) * wG)
        noise += np.sum(np.abs(Ep) * np.abs(wE)) + np.sum(np.abs(Gp) * np.abs(wG))
    return tot / h2, 8 * EPS * noise / h2


def test_tangent_values_match_forward(fitset):
    model, groups, theta0, pes0, L0, g = fitset
    for grp in groups:
        grp.ff.set_ttm_params(theta0)
        r = grp.ff.full_bead_forces(grp.RRt, grp.RRc)
        P, na = grp.RRt.shape[:2]
        _, E, dE = grp.ff.ttm_param_grad(grp.RRt, np.zeros(P), np.zeros((P, na, 3)))
        rel = np.max(np.abs(E - r["Upott"]) / np.abs(r["Upott"]))
        print(f"model {model} N {na // 3}: tangent energy vs forward rel {rel:.2e}", flush=True)
        assert rel < 1e-10


def test_energy_derivatives_vs_fd(fitset):
    """dE_b/dtheta_j of every configuration against central differences of the
    forward energies."""
    model, groups, theta0, pes0, L0, g = fitset
    for grp in groups:
        P, na = grp.RRt.shape[:2]
        grp.ff.set_ttm_params(theta0)
        _, _, dE = grp.ff.ttm_param_grad(grp.RRt, np.zeros(P), np.zeros((P, na, 3)))
        for j in range(theta0.size):
            tp, tm, h2 = _pm(theta0, j)
            grp.ff.set_ttm_params(tp)
            Ep = grp.ff.full_bead_forces(grp.RRt, grp.RRc)["Upott"]
            grp.ff.set_ttm_params(tm)
            Em = grp.ff.full_bead_forces(grp.RRt, grp.RRc)["Upott"]
            fd = (Ep - Em) / h2
            E5 = []
            pts, h = _five_point(theta0, j)
            for t in pts:
                grp.ff.set_ttm_params(t)
                E5.append(grp.ff.full_bead_forces(grp.RRt, grp.RRc)["Upott"])
            fd5 = (8 * (E5[1] - E5[2]) - (E5[0] - E5[3])) / (12 * h)
            scale = np.max(np.abs(fd))
            err = np.max(np.abs(fd - dE[:, j]))
            err5 = np.max(np.abs(fd5 - dE[:, j]))
            noise = 4 * EPS * np.max(np.abs(Ep)) / h2
            noise5 = 6 * EPS * np.max(np.abs(Ep)) / h
            print(f"model {model} N {na // 3} param {j:2d} |dE| {scale:.3e} err {err:.2e} "
                  f"(rounding bound {noise:.1e}) five-point err {err5:.2e}", flush=True)
            if scale == 0.0:
                assert np.all(dE[:, j] == 0.0) and np.all(fd5 == 0.0), j
            else:
                assert err < FD_REL * scale + noise, j
                if j not in GAMMA_ARG:
                    assert err5 < FD5_REL * scale + noise5, j
        grp.ff.set_ttm_params(theta0)


def test_loss_gradient_vs_fd_of_gpu_forward(fitset):
    model, groups, theta0, pes0, L0, g = fitset
    assert L0 > 0 and np.all(np.isfinite(g))
    for j in range(theta0.size):
        tp, tm, h2 = _pm(theta0, j)
        fd, noise = _central(_forward(groups, tp), _forward(groups, tm), groups, h2)
        _forward(groups, theta0)
        err = abs(fd - g[j])
        print(f"model {model} param {j:2d} tangent {g[j]: .10e} fd {fd: .10e} "
              f"rel {err / max(abs(g[j]), 1e-300):.2e} (rounding bound {noise:.1e})", flush=True)
        if fd == 0.0 and g[j] == 0.0:
            continue
        assert err < FD_REL * abs(g[j]) + noise, j


def test_loss_gradient_vs_fd_of_oracle(fitset):
    """The same loss gradient against central differences of the oracle (the
    C restatement of the reference) with its parameters overridden, on the
    cluster group."""
    import oracle
    model, groups, theta0, pes0, L0, g = fitset
    grp = groups[0]
    sys_ = oracle.System(grp.ff.params)
    assert np.array_equal(sys_.ttm_params(), theta0)
    P, na = grp.RRt.shape[:2]
    grp.ff.set_ttm_params(theta0)
    gc = grp.ff.ttm_param_grad(grp.RRt, *_weights(grp))[0]

    def fwd(theta):
        sys_.set_ttm_params(theta)
        o = sys_.forces(grp.RRt, grp.RRc, t=1)
        return [(o["En"], o["dRR"])]

    for j in range(theta0.size):
        tp, tm, h2 = _pm(theta0, j)
        fd, noise = _central(fwd(tp), fwd(tm), [grp], h2)
        err = abs(fd - gc[j])
        print(f"oracle model {model} param {j:2d} tangent {gc[j]: .10e} fd {fd: .10e} "
              f"(rounding bound {noise:.1e})", flush=True)
        if fd == 0.0 and gc[j] == 0.0:
            continue
        assert err < FD_REL * abs(gc[j]) + noise, j


def _weights(grp):
    r = grp.ff.full_bead_forces(grp.RRt, grp.RRc)
    return 2.0 * (r["Upott"] - grp.E_target), 2.0 * (r["dRRt"] - grp.G_target)


def test_fit_step_reduces_loss(fitset):
    """One gradient step of the torch driver over PES and TTM parameters
    together lowers the loss."""
    import torch
    from pimd_amd.fit import loss_and_grad, pes_loss
    model, groups, theta0, pes0, L0, g = fitset
    full = np.concatenate([pes0, theta0])
    th = torch.tensor(full, dtype=torch.float64, requires_grad=True)
    loss = pes_loss(th, groups, method="tangent")
    loss.backward()
    gall = th.grad.numpy()
    assert np.allclose(gall[pes0.size:], g, rtol=1e-12, atol=0)
    # step in the TTM parameters only, scaled per parameter
    d = np.zeros_like(full)
    d[pes0.size:] = -g * theta0 ** 2
    slope = float(np.dot(gall, d))
    assert slope < 0
    L1, _ = loss_and_grad(groups, full + 1e-3 * L0 / abs(slope) * d, grad=False)
    loss_and_grad(groups, full, grad=False)
    print("loss", L0, "->", L1, flush=True)
    assert L1 < L0


def _dms_h(c, j, rel):
    return rel * max(abs(c[j]), 1e-3)


def _dms_points(ff, grp, c0, j, h):
    """Energies and gradients at c0 + m h e_j, m = 2, 1, -1, -2."""
    out = []
    for m in (2, 1, -1, -2):
        c = c0.copy()
        c[j] += m * h
        ff.set_dms_coefs(c)
        r = ff.full_bead_forces(grp.RRt, grp.RRc)
        out.append((r["Upott"], r["dRRt"]))
    ff.set_dms_coefs(c0)
    return out


def test_dms_coef_energy_derivatives_vs_fd(fitset):
    """dE_b/dcoefD_j (pimd_dms_coef_grad) of every configuration against the
    five-point difference (h = 1e-3 relative; the charges are linear in the
    coefficients and the energy smooth in the charges, so the truncation
    error, O(h^4), is far below the bar) of the forward energies, for all 84
    coefficients of the dipole moment surface.  The bar is normwise over the
    84 derivatives: the largest error below 1e-7 of the largest derivative
    (some derivatives are 1e-5 of the largest, and their differences sit at
    the noise of the forward energies, ~1e-9 kcal/mol).  The same tangent
    sweep's value parts reproduce the forward energies.  (A central
    difference at h = 1e-6 relative drowns in that noise for coefficients of
    order 1e-3.)"""
    model, groups, theta0, pes0, L0, g = fitset
    for grp in groups:
        P, na = grp.RRt.shape[:2]
        grp.ff.set_ttm_params(theta0)
        c0 = grp.ff.dms_coefs()
        r0 = grp.ff.full_bead_forces(grp.RRt, grp.RRc)
        _, E, dE = grp.ff.dms_coef_grad(grp.RRt, np.zeros(P), np.zeros((P, na, 3)))
        assert np.max(np.abs(E - r0["Upott"]) / np.abs(r0["Upott"])) < 1e-10
        fd = np.zeros_like(dE)
        for j in range(c0.size):
            h = _dms_h(c0, j, 1e-3)
            E2, E1, Em1, Em2 = [e for e, _ in _dms_points(grp.ff, grp, c0, j, h)]
            fd[:, j] = (8 * (E1 - Em1) - (E2 - Em2)) / (12 * h)
        err, scale = np.max(np.abs(fd - dE)), np.max(np.abs(fd))
        big = np.abs(fd) > 1e-3 * scale
        rel_big = np.max(np.abs(fd - dE)[big] / np.abs(fd)[big])
        print(f"model {model} N {na // 3}: 84 DMS coefficients, max |dE/dc| {scale:.3e}, "
              f"largest error {err:.2e} ({err / scale:.1e} of it); componentwise "
              f"{rel_big:.1e} where |dE/dc| > 1e-3 of the largest", flush=True)
        assert err < FD5_REL * scale
        again = grp.ff.full_bead_forces(grp.RRt, grp.RRc)
        assert np.array_equal(again["dRRt"], r0["dRRt"])


def test_dms_coef_loss_gradient_vs_fd(fitset):
    """The loss gradient with respect to the DMS coefficients against the
    five-point difference (h = 1e-3 relative) of the loss over both groups
    (energies and gradients), each difference L(c + m h) - L(c - m h) summed
    as (a+ - a-)(a+ + a- - 2 target); normwise over the 84 components, 1e-6
    of the largest."""
    model, groups, theta0, pes0, L0, g = fitset
    for grp in groups:
        grp.ff.set_ttm_params(theta0)
    c0 = groups[0].ff.dms_coefs()
    gc = np.zeros(c0.size)
    for grp in groups:
        gc += grp.ff.dms_coef_grad(grp.RRt, *_weights(grp))[0]
    fds = np.zeros(c0.size)
    for j in range(c0.size):
        h = _dms_h(c0, j, 1e-3)
        D = {1: 0.0, 2: 0.0}
        for grp in groups:
            (E2, G2), (E1, G1), (Em1, Gm1), (Em2, Gm2) = _dms_points(grp.ff, grp, c0, j, h)
            for m, (Ep, Gp, Em, Gm) in ((1, (E1, G1, Em1, Gm1)), (2, (E2, G2, Em2, Gm2))):
                wE, wG = Ep + Em - 2 * grp.E_target, Gp + Gm - 2 * grp.G_target
                D[m] += np.sum((Ep - Em) * wE) + np.sum((Gp - Gm) * wG)
        fds[j] = (8 * D[1] - D[2]) / (12 * h)
    err, scale = np.max(np.abs(fds - gc)), np.max(np.abs(fds))
    print(f"model {model}: loss gradient over 84 DMS coefficients, max {scale:.3e}, "
          f"largest error {err:.2e} ({err / scale:.1e} of it)", flush=True)
    assert err < FD_REL * scale


def _gm_forward(groups, gm):
    out = []
    for grp in groups:
        grp.ff.set_gamma_m(gm)
        r = grp.ff.full_bead_forces(grp.RRt, grp.RRc)
        out.append((r["Upott"], r["dRRt"]))
    return out


def test_gamma_m_derivatives_vs_fd(fitset):
    """dE_b/dgammaM of every configuration and the loss gradient along gammaM
    (pimd_gamma_m_grad: the M sites move with it, pot_ttm.f90:74-79, and the
    charges are placed with it, pot_ttm.f90:134-137) against central
    differences (h = 1e-6 relative) of the GPU forward model at 1e-6 relative
    plus the rounding bound.  No five-point check: the M sites' distances
    enter the argument of the Thole charge term's Gamma(2/3, x), which jumps
    where gammq2_3 changes its number of terms.  The value parts reproduce
    the forward energies."""
    model, groups, theta0, pes0, L0, g = fitset
    gm0 = groups[0].ff.gamma_m()
    assert gm0 == (0.46 if model == 3 else 0.426706882)
    h = 1e-6 * gm0
    gsum = 0.0
    for grp in groups:
        grp.ff.set_ttm_params(theta0)
        P, na = grp.RRt.shape[:2]
        r0 = grp.ff.full_bead_forces(grp.RRt, grp.RRc)
        _, E, dE = grp.ff.gamma_m_grad(grp.RRt, np.zeros(P), np.zeros((P, na, 3)))
        assert np.max(np.abs(E - r0["Upott"]) / np.abs(r0["Upott"])) < 1e-10
        (Ep, _), = _gm_forward([grp], gm0 + h)
        (Em, _), = _gm_forward([grp], gm0 - h)
        grp.ff.set_gamma_m(gm0)
        fd = (Ep - Em) / (2 * h)
        scale, err = np.max(np.abs(fd)), np.max(np.abs(fd - dE))
        noise = 4 * EPS * np.max(np.abs(Ep)) / (2 * h)
        print(f"model {model} N {na // 3}: |dE/dgammaM| {scale:.3e} err {err:.2e} "
              f"(rounding bound {noise:.1e})", flush=True)
        assert err < FD_REL * scale + noise
        gsum += grp.ff.gamma_m_grad(grp.RRt, *_weights(grp))[0]
    fd, noise = _central(_gm_forward(groups, gm0 + h), _gm_forward(groups, gm0 - h), groups, 2 * h)
    _gm_forward(groups, gm0)
    err = abs(fd - gsum)
    print(f"model {model} loss gradient along gammaM: tangent {gsum: .10e} fd {fd: .10e} "
          f"rel {err / abs(gsum):.2e} (rounding bound {noise:.1e})", flush=True)
    assert err < FD_REL * abs(gsum) + noise


def test_gamma_m_vs_oracle(fitset):
    """On the cluster group: the loss gradient along gammaM against central
    differences of the oracle with its gammaM overridden, and the forward
    model at a changed gammaM against the oracle at the same value (1e-10
    relative energies, 1e-8 gradient)."""
    import oracle
    from helpers import E_REL, F_ABS, max_abs, rel_err
    model, groups, theta0, pes0, L0, g = fitset
    grp = groups[0]
    grp.ff.set_ttm_params(theta0)
    sys_ = oracle.System(grp.ff.params)
    gm0 = grp.ff.gamma_m()
    assert sys_.gamma_m() == gm0
    gc = grp.ff.gamma_m_grad(grp.RRt, *_weights(grp))[0]

    def fwd(gm):
        sys_.set_gamma_m(gm)
        o = sys_.forces(grp.RRt, grp.RRc, t=1)
        return [(o["En"], o["dRR"])]

    h = 1e-6 * gm0
    fd, noise = _central(fwd(gm0 + h), fwd(gm0 - h), [grp], 2 * h)
    err = abs(fd - gc)
    print(f"oracle model {model} gammaM: tangent {gc: .10e} fd {fd: .10e} "
          f"(rounding bound {noise:.1e})", flush=True)
    assert err < FD_REL * abs(gc) + noise
    gm1 = 0.9 * gm0
    sys_.set_gamma_m(gm1)
    grp.ff.set_gamma_m(gm1)
    try:
        r = grp.ff.full_bead_forces(grp.RRt, grp.RRc)
        o = sys_.forces(grp.RRt, grp.RRc, t=1)
    finally:
        grp.ff.set_gamma_m(gm0)
    e, f = rel_err(r["Upott"], o["En"]), max_abs(r["dRRt"], o["dRR"])
    print(f"forward at gammaM {gm1}: energy rel {e:.2e} gradient {f:.2e}", flush=True)
    assert e < E_REL and f < F_ABS


def test_gamma_m_set_refuses_out_of_range(fitset):
    from pimd_amd.ttm import PimdError
    model, groups, theta0, pes0, L0, g = fitset
    ff = groups[0].ff
    gm0 = ff.gamma_m()
    for bad in (1.0, -0.1, float("nan")):
        with pytest.raises(PimdError):
            ff.set_gamma_m(bad)
    assert ff.gamma_m() == gm0


def test_fit_vector_with_dms_and_gamma_m(fitset):
    """The fitting driver over all 351 parameters (250 PES, 16 TTM, 84 DMS,
    gammaM): the same loss at the same parameters, and its TTM and gammaM
    components equal to the separate tangent sweeps'."""
    from pimd_amd.fit import NPARAM_GM, loss_and_grad
    model, groups, theta0, pes0, L0, g = fitset
    c0, gm0 = groups[0].ff.dms_coefs(), groups[0].ff.gamma_m()
    full = np.concatenate([pes0, theta0, c0, [gm0]])
    assert full.size == NPARAM_GM
    L, gall = loss_and_grad(groups, full, method="tangent")
    assert abs(L - L0) <= 1e-12 * L0
    assert np.allclose(gall[pes0.size:pes0.size + theta0.size], g, rtol=1e-12, atol=0)
    ggm = sum(grp.ff.gamma_m_grad(grp.RRt, *_weights(grp))[0] for grp in groups)
    assert abs(gall[-1] - ggm) <= 1e-12 * abs(ggm)
