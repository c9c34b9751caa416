"""N-body field (galax NBodyField) and its integration as one system: the
oracle against the reference docstring, the exact two-body solution and
momentum conservation; the GPU against the oracle."""

import math

import numpy as np
import pytest

import oracle
import galax_amd.dynamics as gd
import galax_amd.potential as gp
from galax_amd.units import G, G_SOLARSYSTEM, KMS_SOLARSYSTEM

from test_oracle import _kepler_exact


def _doctest_setup(golden):
    e = golden["nbody_doctest"]
    q0 = np.array(e["q0_au"])
    p0 = np.array(e["p0_kms"]) * KMS_SOLARSYSTEM
    gm = G_SOLARSYSTEM * np.array(e["masses_msun"])
    return e, np.concatenate([q0, p0], 1), gm


def test_solarsystem_constants():
    # G M_sun in AU^3/yr^2 is 4 pi^2 to the accuracy of the Julian year vs
    # the sidereal year (about 2e-5)
    assert G_SOLARSYSTEM == pytest.approx(4 * math.pi**2, rel=5e-5)
    assert KMS_SOLARSYSTEM == pytest.approx(0.21094952656969870, rel=1e-15)


def test_oracle_nbody_docstring(golden):
    e, w0, gm = _doctest_setup(golden)
    solver = gd.OrbitSolver()
    out, st, ns, na = oracle.nbody_integrate(gp.NullPotential(), gm, e["eps_au"], solver, w0, e["t0_yr"], e["t1_yr"])
    assert st == 0
    np.testing.assert_allclose(out[:, 0, :3], e["q1_au_round4"], atol=5e-5)


def test_oracle_two_body_matches_kepler():
    """Unsoftened two bodies: the separation follows the Kepler orbit of
    G (m1 + m2) exactly; the centre of mass moves uniformly."""
    m = np.array([3e10, 1e10])
    gm = G * m
    x1, x2 = np.array([-2.0, 0.3, 0.1]), np.array([6.0, -0.2, 0.0])
    v1, v2 = np.array([0.0, -0.02, 0.001]), np.array([0.01, 0.06, 0.0])
    w0 = np.array([np.concatenate([x1, v1]), np.concatenate([x2, v2])])
    solver = gd.OrbitSolver(gd.Dopri8(), stepsize_controller=gd.ConstantStepSize(), max_steps=None)
    T = 300.0
    out, st, *_ = oracle.nbody_integrate(gp.NullPotential(), gm, 0.0, solver, w0, 0.0, T, dt0=0.25)
    assert st == 0
    rel0 = w0[1] - w0[0]
    exact = _kepler_exact(G * m.sum(), rel0, T)
    got = out[1, 0] - out[0, 0]
    np.testing.assert_allclose(got, exact, rtol=0, atol=1e-9 * np.abs(exact).max())
    com = (m[:, None] * out[:, 0, :3]).sum(0) / m.sum()
    com_exact = (m[:, None] * w0[:, :3]).sum(0) / m.sum() + T * (m[:, None] * w0[:, 3:]).sum(0) / m.sum()
    np.testing.assert_allclose(com, com_exact, atol=1e-11)


def test_oracle_momentum_conserved():
    rng = np.random.default_rng(3)
    n = 6
    m = rng.uniform(1e9, 5e9, n)
    w0 = np.concatenate([rng.normal(0, 2.0, (n, 3)), rng.normal(0, 0.05, (n, 3))], 1)
    solver = gd.OrbitSolver(gd.Dopri8(), stepsize_controller=gd.PIDController(rtol=1e-9, atol=1e-9), max_steps=None)
    out, st, ns, _ = oracle.nbody_integrate(gp.NullPotential(), G * m, 0.05, solver, w0, 0.0, 200.0)
    assert st == 0 and ns > 10
    P0 = (m[:, None] * w0[:, 3:]).sum(0)
    P1 = (m[:, None] * out[:, 0, 3:]).sum(0)
    assert np.max(np.abs(P1 - P0)) < 1e-12 * (m[:, None] * np.abs(w0[:, 3:])).sum()


# ------------------------------------------------------------------- GPU
def _cluster(n, seed, mmax=1e6):
    rng = np.random.default_rng(seed)
    centre = np.array([8.0, 0.0, 0.2])
    vc = np.array([0.0, 0.23, 0.0])
    x = centre + rng.normal(0, 0.05, (n, 3))
    v = vc + rng.normal(0, 0.002, (n, 3))
    m = rng.uniform(0.01 * mmax, mmax, n)
    return np.concatenate([x, v], 1), m


# Trajectory comparisons need a well-conditioned system.  With softening
# 1-2 pc and masses up to 1e6 Msun this cluster has unresolved close
# encounters: a 1e-13 perturbation of the oracle's own initial state grows
# to 4e-7 in 10 Myr and 1e-2 in 20 Myr, and the oracle's forward-backward
# round trip over 30 Myr misses by kpc.  Softening at the cluster's size
# (50 pc) and masses up to 1e5 Msun keep such a perturbation at 4e-13.
SMOOTH = dict(eps=0.05, mmax=1e5)


@pytest.mark.gpu
def test_gpu_nbody_field_matches_direct_sum():
    import torch

    w0, m = _cluster(700, 1)
    mw = gp.MilkyWayPotential()
    field = gd.NBodyField(masses=m, eps=1e-3, external_potential=mw)
    print("[case] nbody field n=700", flush=True)
    _, acc = field(12.0, (w0[:, :3], w0[:, 3:]))
    acc = acc.cpu().numpy()
    x = w0[:, :3]
    d = x[None, :, :] - x[:, None, :]  # d[j, i] = x_i - x_j
    r2 = (d * d).sum(-1) + 1e-6
    w = G * m[None, :] / (r2 * np.sqrt(r2))
    np.fill_diagonal(w, 0.0)
    ref = (w[..., None] * d).sum(1) - oracle.evaluate(mw, "gradient", x, np.full(len(x), 12.0))
    err = np.linalg.norm(acc - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert err.max() < 1e-12, err.max()


@pytest.mark.gpu
def test_gpu_nbody_docstring(golden):
    e, w0, gm = _doctest_setup(golden)
    field = gd.fields.NBodyField(masses=np.array(e["masses_msun"]), eps=e["eps_au"],
                                 external_potential=gp.NullPotential(units="solarsystem"))
    print("[case] nbody docstring", flush=True)
    soln = gd.OrbitSolver().solve(field, (w0[:, :3], w0[:, 3:]), e["t0_yr"], e["t1_yr"])
    assert soln.status == 0
    np.testing.assert_allclose(soln.ys[0][-1], e["q1_au_round4"], atol=5e-5)
    ref, st, ns, na = oracle.nbody_integrate(gp.NullPotential(), gm, e["eps_au"], gd.OrbitSolver(), w0, 0.0, 2.0)
    assert soln.nsteps == ns
    np.testing.assert_allclose(soln.ys[0][-1], ref[:, 0, :3], rtol=0, atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("controller", ["constant", "pid"])
def test_gpu_nbody_matches_oracle(controller):
    w0, m = _cluster(37, 2, SMOOTH["mmax"])
    eps = SMOOTH["eps"]
    mw = gp.MilkyWayPotential()
    if controller == "constant":
        solver = gd.OrbitSolver(gd.Dopri8(), stepsize_controller=gd.ConstantStepSize(), max_steps=None)
        kw = {"dt0": 0.05}
    else:
        solver = gd.OrbitSolver(gd.Dopri8(), stepsize_controller=gd.PIDController(rtol=1e-9, atol=1e-9),
                                max_steps=None)
        kw = {}
    ts = np.linspace(0.0, 60.0, 7)
    print(f"[case] nbody vs oracle {controller}", flush=True)
    field = gd.NBodyField(masses=m, eps=eps, external_potential=mw)
    soln = gd.solve_nbody(field, (w0[:, :3], w0[:, 3:]), 0.0, 60.0, saveat=ts, solver=solver, **kw)
    ref, st, ns, na = oracle.nbody_integrate(mw, G * m, eps, solver, w0, 0.0, 60.0, ts=ts, **kw)
    assert soln.status == 0 and st == 0
    got = np.concatenate([np.transpose(soln.ys[0], (1, 0, 2)), np.transpose(soln.ys[1], (1, 0, 2))], -1)
    rel = np.abs(got - ref) / np.abs(ref).max(axis=(1, 2), keepdims=True)
    print("steps", soln.nsteps, ns, "max rel", rel.max(), flush=True)
    if controller == "constant":
        assert soln.nsteps == ns == 1200
        assert rel.max() < 1e-12
    elif soln.nsteps == ns:
        assert rel.max() < 1e-9
    else:
        # an accept/reject decision within round-off of the threshold
        # flipped: the two runs then differ at the tolerance level
        assert rel.max() < 1e-6


@pytest.mark.gpu
def test_gpu_nbody_backward_and_bad_sizes():
    from galax_amd import _lib

    w0, m = _cluster(20, 4, SMOOTH["mmax"])
    field = gd.NBodyField(masses=m, eps=SMOOTH["eps"], external_potential=gp.MilkyWayPotential())
    solver = gd.OrbitSolver(gd.Dopri8(), stepsize_controller=gd.ConstantStepSize(), max_steps=None)
    print("[case] nbody backward", flush=True)
    fwd = gd.solve_nbody(field, (w0[:, :3], w0[:, 3:]), 0.0, 30.0, solver=solver, dt0=0.1)
    wf = np.concatenate([fwd.ys[0][-1], fwd.ys[1][-1]], 1)
    back = gd.solve_nbody(field, (wf[:, :3], wf[:, 3:]), 30.0, 0.0, solver=solver, dt0=0.1)
    np.testing.assert_allclose(back.ys[0][-1], w0[:, :3], atol=1e-9)
    big = gd.NBodyField(masses=np.ones(5000), eps=1e-3)
    with pytest.raises(_lib.GalaxHipError):
        gd.solve_nbody(big, (np.zeros((5000, 3)), np.zeros((5000, 3))), 0.0, 1.0, solver=solver, dt0=0.5)
