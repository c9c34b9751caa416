"""Stream-track loss gradient (progenitor flow -> Fardal release -> particle
flow -> nearest-track-point loss) against central finite differences of the
same pipeline run on the oracle (h = 1e-6 relative)."""

import numpy as np
import pytest

import oracle
import galax_amd.dynamics as gd
from galax_amd.adjoint import MilkyWayParams
from galax_amd.dynamics import FardalStreamDF


def _loss_oracle(theta, prob):
    pot = MilkyWayParams.potential(theta)
    solver = gd.OrbitSolver(gd.Dopri8(), stepsize_controller=gd.ConstantStepSize(), max_steps=None)
    ts_desc = np.ascontiguousarray(prob.ts[::-1])
    prog, st, *_ = oracle.integrate(pot, solver, prob.w_prog[None], 0.0, float(prob.ts[0]), ts=ts_desc,
                                    dt0=prob.dt)
    assert np.all(st == 0)
    prog = prog[0][::-1]
    xl, vl, xt, vt = FardalStreamDF.release(prog[:, :3], prog[:, 3:], prob.r_t, prob.k)
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
