"""Stream-track loss gradient (progenitor flow -> Fardal release -> particle
flow -> nearest-track-point loss) against central finite differences of the
same pipeline run on the oracle (h = 1e-6 relative)."""

import numpy as np
import pytest

import oracle
import galax_amd.dynamics as gd
from galax_amd.adjoint import MilkyWayParams
from galax_amd.dynamics import FardalStreamDF


def _tidal_radius_oracle(pot, x, v, mass):
    """King (1962) from the oracle's Hessian (forward-mode jets of Phi)."""
    from galax_amd.units import G

    H = oracle.evaluate(pot, "hessian", x)
    rh = x / np.linalg.norm(x, axis=-1, keepdims=True)
    d2 = np.einsum("ni,nij,nj->n", rh, H, rh)
    om = np.linalg.norm(np.cross(x, v), axis=-1) / np.sum(x * x, axis=-1)
    return np.cbrt(G * mass / (om**2 - d2))


def _loss_oracle(theta, prob):
    """The whole stream loss on the oracle, the tidal radius recomputed at
    theta from the oracle's progenitor orbit."""
    pot = MilkyWayParams.potential(theta)
    solver = gd.OrbitSolver(gd.Dopri8(), stepsize_controller=gd.ConstantStepSize(), max_steps=None)
    ts_desc = np.ascontiguousarray(prob.ts[::-1])
    prog, st, *_ = oracle.integrate(pot, solver, prob.w_prog[None], 0.0, float(prob.ts[0]), ts=ts_desc,
                                    dt0=prob.dt)
    assert np.all(st == 0)
    prog = prog[0][::-1]
    r_t = _tidal_radius_oracle(pot, prog[:, :3], prog[:, 3:], prob.mass)
    xl, vl, xt, vt = FardalStreamDF.release(prog[:, :3], prog[:, 3:], r_t, prob.k)
    w0 = np.concatenate([np.concatenate([xl, vl], 1), np.concatenate([xt, vt], 1)], 0)
    t0 = np.concatenate([prob.ts, prob.ts])
    out, st, *_ = oracle.integrate(pot, solver, w0, t0, 0.0, dt0=prob.dt)
    assert np.all(st == 0)
    pos = out[:, 0, :3]
    d2 = ((pos[:, None, :] - prob.target[None]) ** 2).sum(-1)
    return float(d2.min(1).mean())


def test_release_torch_matches_numpy():
    torch = pytest.importorskip("torch")
    from galax_amd.streamfit import release_torch

    rng = np.random.default_rng(3)
    x = rng.normal(size=(20, 3)) * 10
    v = rng.normal(size=(20, 3)) * 0.2
    r_t = rng.uniform(0.01, 0.1, 20)
    k = FardalStreamDF().draw_k(4, 20)
    got = release_torch(torch.as_tensor(x), torch.as_tensor(v), torch.as_tensor(r_t), torch.as_tensor(k)).numpy()
    xl, vl, xt, vt = FardalStreamDF.release(x, v, r_t, k)
    want = np.concatenate([np.concatenate([xl, vl], 1), np.concatenate([xt, vt], 1)], 0)
    np.testing.assert_allclose(got, want, rtol=1e-14, atol=1e-15)


def test_radial_second_derivative_matches_oracle_hessian():
    """The closed-form d2Phi/dr2 of the Milky Way model (torch, used for the
    tidal radius inside the loss) against r^T H r of the oracle's Hessian, at
    the default parameters and at perturbed ones, and its theta-gradient
    against central differences of the oracle."""
    torch = pytest.importorskip("torch")
    from galax_amd.streamfit import mw_d2phi_dr2

    rng = np.random.default_rng(11)
    x = rng.normal(size=(200, 3)) * np.array([12.0, 12.0, 2.0])
    for scale in (np.ones(6), np.array([1.3, 0.8, 1.1, 0.7, 1.5, 0.9])):
        th = MilkyWayParams.default * scale
        H = oracle.evaluate(MilkyWayParams.potential(th), "hessian", x)
        rh = x / np.linalg.norm(x, axis=-1, keepdims=True)
        want = np.einsum("ni,nij,nj->n", rh, H, rh)
        got = mw_d2phi_dr2(torch.as_tensor(x), torch.as_tensor(th)).numpy()
        np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-18)
    th0 = MilkyWayParams.default.copy()
    tt = torch.tensor(th0, requires_grad=True)
    xs = torch.as_tensor(x[:20])
    mw_d2phi_dr2(xs, tt).sum().backward()
    for j in range(6):
        h = 1e-6 * th0[j]
        tp, tm = th0.copy(), th0.copy()
        tp[j] += h
        tm[j] -= h
        rh = x[:20] / np.linalg.norm(x[:20], axis=-1, keepdims=True)
        f = lambda t: np.einsum("ni,nij,nj->", rh, oracle.evaluate(MilkyWayParams.potential(t), "hessian", x[:20]), rh)
        fd = (f(tp) - f(tm)) / (2 * h)
        assert float(tt.grad[j]) == pytest.approx(fd, rel=1e-6), j


@pytest.mark.gpu
def test_stream_loss_gradient_matches_oracle_fd():
    import torch

    from galax_amd.streamfit import make_problem, stream_loss

    print("[case] streamfit small", flush=True)
    prob = make_problem(n_release=48, span=300.0, seed=6, n_target=30, K=16)
    theta0 = MilkyWayParams.default.copy()
    theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    L = stream_loss(theta, prob)
    L.backward()
    g = theta.grad.numpy()
    L_ref = _loss_oracle(theta0, prob)
    assert float(L) == pytest.approx(L_ref, rel=1e-10)
    for j, name in enumerate(MilkyWayParams.names):
        h = 1e-6 * theta0[j]
        tp, tm = theta0.copy(), theta0.copy()
        tp[j] += h
        tm[j] -= h
        fd = (_loss_oracle(tp, prob) - _loss_oracle(tm, prob)) / (2 * h)
        print(name, g[j], fd, flush=True)
        assert g[j] == pytest.approx(fd, rel=1e-5, abs=1e-9 * abs(L_ref) / theta0[j]), name
