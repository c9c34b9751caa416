"""Discrete adjoint of fixed-step integration against central finite
differences of the oracle's forward integration (h = 1e-6 relative)."""

import numpy as np
import pytest

import oracle
import galax_amd.dynamics as gd
import galax_amd.potential as gp
from galax_amd.adjoint import MilkyWayParams
from galax_amd.units import G
from galax_amd import workloads


def test_milky_way_parameter_mapping():
    pot = MilkyWayParams.potential(MilkyWayParams.default)
    ref = gp.MilkyWayPotential()
    for (a, _), (b, _) in zip(pot._rows(), ref._rows()):
        assert a.kind == b.kind and list(a.p) == list(b.p)


def _loss_oracle(pot, solver, w0, t0, t1, ts, dt0, C):
    out, st, *_ = oracle.integrate(pot, solver, w0, t0, t1, ts=ts, dt0=dt0)
    assert np.all(st == 0)
    return float(np.sum(out * C))


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["dopri8", "dopri5"])
def test_adjoint_initial_state_and_parameters(method):
    import torch

    from galax_amd.adjoint import adjoint_orbits

    meth = gd.Dopri8() if method == "dopri8" else gd.Dopri5()
    solver = gd.OrbitSolver(meth, stepsize_controller=gd.ConstantStepSize(), max_steps=None)
    theta = MilkyWayParams.default.copy()
    pot = MilkyWayParams.potential(theta)
    n = 6
    w0 = workloads.disk_ics(n, 17)
    t0 = np.array([0.0, -3.3, 5.0, -100.0, 0.0, 12.0])
    t1 = np.array([157.7, 120.0, -60.5, 33.0, 200.0, 12.0 + 1e-3])
    ts = np.stack([np.linspace(a, b, 5) for a, b in zip(t0, t1)])  # interior and end saves
    dt0 = 1.0 if method == "dopri8" else 0.25
    C = np.random.default_rng(4).normal(size=(n, 5, 6))
    print(f"[case] adjoint {method} n={n}", flush=True)
    w0_bar, raw, st = adjoint_orbits(pot, w0, t0, t1, torch.as_tensor(C), dt0=dt0, ts=ts,
                                     method=meth, ntheta=MilkyWayParams.ntheta)
    assert int(st.abs().sum()) == 0
    w0_bar = w0_bar.cpu().numpy()
    g_theta = MilkyWayParams.grad_from_table(raw.sum(0)).cpu().numpy()
    # initial-state gradient, per orbit, by central differences
    for i in range(n):
        for k in range(6):
            hk = 1e-6 * max(abs(w0[i, k]), 1e-2)
            wp, wm = w0.copy(), w0.copy()
            wp[i, k] += hk
            wm[i, k] -= hk
            fd = (_loss_oracle(pot, solver, wp[i:i+1], t0[i], t1[i], ts[i:i+1], dt0, C[i:i+1])
                  - _loss_oracle(pot, solver, wm[i:i+1], t0[i], t1[i], ts[i:i+1], dt0, C[i:i+1])) / (2 * hk)
            assert w0_bar[i, k] == pytest.approx(fd, rel=2e-6, abs=1e-7 * np.abs(w0_bar[i]).max())
    # parameter gradient
    for j, name in enumerate(MilkyWayParams.names):
        hj = 1e-6 * theta[j]
        tp, tm = theta.copy(), theta.copy()
        tp[j] += hj
        tm[j] -= hj
        fd = (_loss_oracle(MilkyWayParams.potential(tp), solver, w0, t0, t1, ts, dt0, C)
              - _loss_oracle(MilkyWayParams.potential(tm), solver, w0, t0, t1, ts, dt0, C)) / (2 * hj)
        print(name, g_theta[j], fd, flush=True)
        assert g_theta[j] == pytest.approx(fd, rel=1e-5), name


@pytest.mark.gpu
def test_adjoint_single_component_kinds():
    import torch

    from galax_amd.adjoint import adjoint_orbits

    solver = gd.OrbitSolver(gd.Dopri8(), stepsize_controller=gd.ConstantStepSize(), max_steps=None)
    w0 = workloads.disk_ics(3, 23)
    C = np.random.default_rng(5).normal(size=(3, 1, 6))
    makers = {
        "kepler": (lambda p: gp.KeplerPotential(m_tot=p[0]), [1e11]),
        "hernquist": (lambda p: gp.HernquistPotential(m_tot=p[0], r_s=p[1]), [1e12, 5.0]),
        "plummer": (lambda p: gp.PlummerPotential(m_tot=p[0], r_s=p[1]), [1e11, 2.0]),
        "nfw": (lambda p: gp.NFWPotential(m=p[0], r_s=p[1]), [5e11, 12.0]),
        "mn": (lambda p: gp.MiyamotoNagaiPotential(m_tot=p[0], a=p[1], b=p[2]), [7e10, 3.0, 0.3]),
    }
    for name, (mk, params) in makers.items():
        print(f"[case] adjoint kind {name}", flush=True)
        pot = mk(params)
        _, raw, st = adjoint_orbits(pot, w0, 0.0, 80.0, torch.as_tensor(C), dt0=0.5, ntheta=3)
        raw = raw.sum(0).cpu().numpy()
        for j in range(len(params)):
            h = 1e-6 * params[j]
            pp, pm = list(params), list(params)
            pp[j] += h
            pm[j] -= h
            fd = (_loss_oracle(mk(pp), solver, w0, 0.0, 80.0, None, 0.5, C)
                  - _loss_oracle(mk(pm), solver, w0, 0.0, 80.0, None, 0.5, C)) / (2 * h)
            got = raw[j] * (G if j == 0 else 1.0)  # table slot 0 is G m
            assert got == pytest.approx(fd, rel=1e-5), (name, j)
