"""The adjoint of the fitting loss with respect to the 16 TTM parameters, the
84 DMS coefficients and gammaM (pimd_ttm_adjoint: one sweep per
configuration) against the tangent sweeps of the same parameters
(pimd_ttm_param_grad, pimd_dms_coef_grad, pimd_gamma_m_grad), which are
themselves checked against central differences of the GPU forward model and
of the oracle (test_gpu_tangent.py).

The adjoint takes the derivative of the converged energy as its explicit
derivative at fixed dipoles (the induction energy is stationary in them) and
carries the gradient term along the atoms' direction gw as a dual number,
with the screening functions' position derivatives taken from the chain the
reference's gradient uses; the tangents differentiate the SOR iterations as
executed.  At the fitting sets' SOR tolerance (polar_eps 1e-12) the two agree
to 1e-8 relative componentwise wherever a component exceeds 1e-6 of the
largest of its block (1e-9 measured), and to 1e-9 of the block's largest
normwise.  gammaM comes from its tangent sweep in both.  Same fitting sets as
test_gpu_tangent.py: 5-water clusters and 32-water boxes, both models."""
import numpy as np
import pytest

from test_gpu_tangent import _groups, _weights

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[3, 2], ids=["ttm3f", "ttm21f"])
def adjset(request):
    groups, theta0, pes0 = _groups(request.param)
    return request.param, groups, theta0


def _tangents(grp, ew, gw):
    gt, E, dt = grp.ff.ttm_param_grad(grp.RRt, ew, gw)
    gd, _, dd = grp.ff.dms_coef_grad(grp.RRt, ew, gw)
    gg, _, dg = grp.ff.gamma_m_grad(grp.RRt, ew, gw)
    return (np.concatenate([gt, gd, [gg]]), E,
            np.concatenate([dt, dd, dg[:, None]], axis=1))


BLOCKS = {"TTM": slice(0, 16), "DMS": slice(16, 100), "gammaM": slice(100, 101)}


def _compare(what, a, t, model):
    for name, sl in BLOCKS.items():
        x, y = a[..., sl], t[..., sl]
        big = np.max(np.abs(y))
        err = np.max(np.abs(x - y))
        sig = np.abs(y) > 1e-6 * big
        rel = np.max(np.abs(x - y)[sig] / np.abs(y)[sig]) if np.any(sig) else 0.0
        print(f"model {model} {what} {name}: largest {big:.3e}, normwise err {err / big:.2e}, "
              f"componentwise {rel:.2e}", flush=True)
        assert err <= 1e-9 * big, (what, name)
        assert rel <= 1e-8, (what, name)


def test_adjoint_energy_derivatives_match_tangents(adjset):
    model, groups, theta0 = adjset
    for grp in groups:
        grp.ff.set_ttm_params(theta0)
        P, na = grp.RRt.shape[:2]
        z, zg = np.zeros(P), np.zeros((P, na, 3))
        ga, Ea, da = grp.ff.ttm_adjoint(grp.RRt, z, zg)
        _, Et, dt = _tangents(grp, z, zg)
        assert np.max(np.abs(Ea - Et) / np.abs(Et)) < 1e-12
        assert np.allclose(ga, da.sum(axis=0) * 0.0)   # zero weights: zero gradient
        _compare(f"N {na // 3} dE/dtheta", da, dt, model)


def test_adjoint_loss_gradient_matches_tangents(adjset):
    """The loss gradient over both groups (energy and gradient residuals as
    weights) from the adjoint against the tangents'."""
    model, groups, theta0 = adjset
    ga, gt = 0.0, 0.0
    for grp in groups:
        grp.ff.set_ttm_params(theta0)
        ew, gw = _weights(grp)
        ga = ga + grp.ff.ttm_adjoint(grp.RRt, ew, gw)[0]
        gt = gt + _tangents(grp, ew, gw)[0]
    _compare("loss gradient", ga, gt, model)


def test_adjoint_in_fit_driver(adjset):
    """loss_and_grad over 351 parameters with the adjoint equals the tangents'
    gradient at the same bar."""
    from pimd_amd.fit import loss_and_grad
    model, groups, theta0 = adjset
    ff = groups[0].ff
    full = np.concatenate([ff.pes_params(), theta0, ff.dms_coefs(), [ff.gamma_m()]])
    La, ga = loss_and_grad(groups, full, method="adjoint")
    Lt, gt = loss_and_grad(groups, full, method="tangent")
    assert La == Lt
    assert np.array_equal(ga[:250], gt[:250])
    _compare("fit driver", ga[250:], gt[250:], model)


def test_adjoint_loss_gradient_vs_fd(adjset):
    """The adjoint's loss gradient over the 16 TTM parameters against central
    differences (h = 1e-6 relative) of the loss of the GPU forward model over
    both groups, at 1e-6 relative plus the rounding bound (the bar of
    test_gpu_tangent.py)."""
    from test_gpu_tangent import FD_REL, _central, _forward, _pm
    model, groups, theta0 = adjset
    g = 0.0
    for grp in groups:
        grp.ff.set_ttm_params(theta0)
        g = g + grp.ff.ttm_adjoint(grp.RRt, *_weights(grp))[0]
    for j in range(theta0.size):
        tp, tm, h2 = _pm(theta0, j)
        fd, noise = _central(_forward(groups, tp), _forward(groups, tm), groups, h2)
        _forward(groups, theta0)
        err = abs(fd - g[j])
        print(f"model {model} param {j:2d} adjoint {g[j]: .10e} fd {fd: .10e} "
              f"rel {err / max(abs(g[j]), 1e-300):.2e} (rounding bound {noise:.1e})", flush=True)
        if fd == 0.0 and g[j] == 0.0:
            continue
        assert err < FD_REL * abs(g[j]) + noise, j



def test_adjoint_in_two_halves(adjset):
    """pimd_ttm_adjoint_start / _finish on both groups' contexts, both
    started before either is finished, give pimd_ttm_adjoint's results bit
    for bit (16 and 100 parameters); the fit driver's overlapped sweeps equal
    its one-group-at-a-time sweeps; a second start before the finish, a
    blocking sweep in between, and a finish with nothing started are
    refused."""
    from pimd_amd.fit import loss_and_grad
    from pimd_amd.ttm import PimdError
    model, groups, theta0 = adjset
    ws, ref = [], []
    for grp in groups:
        grp.ff.set_ttm_params(theta0)
        ew, gw = _weights(grp)
        ws.append((ew, gw))
        ref.append([grp.ff.ttm_adjoint(grp.RRt, ew, gw, n) for n in (16, 100)])
    for k, n in enumerate((16, 100)):
        for grp, (ew, gw) in zip(groups, ws):
            grp.ff.ttm_adjoint_start(grp.RRt, ew, gw, n)
        for i, grp in enumerate(groups):
            got = grp.ff.ttm_adjoint_finish()
            for a, b in zip(got, ref[i][k]):
                assert np.array_equal(a, b), (n, i)
    ff = groups[0].ff
    with pytest.raises(PimdError):
        ff.ttm_adjoint_finish()
    ff.ttm_adjoint_start(groups[0].RRt, *ws[0], 16)
    with pytest.raises(PimdError):
        ff.ttm_adjoint_start(groups[0].RRt, *ws[0], 16)
    with pytest.raises(PimdError):
        ff.ttm_adjoint(groups[0].RRt, *ws[0], 16)
    assert np.array_equal(ff.ttm_adjoint_finish()[0], ref[0][0][0])
    for full in (np.concatenate([ff.pes_params(), theta0]),
                 np.concatenate([ff.pes_params(), theta0, ff.dms_coefs()])):
        L1, g1 = loss_and_grad(groups, full)
        L0, g0 = 0.0, 0.0
        for grp in groups:
            l, g = grp._evaluate(full, True, "adjoint")
            L0 += l
            g0 = g0 + g
        assert L1 == L0 and np.array_equal(g1, g0)
