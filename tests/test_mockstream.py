"""Mock streams: the Fardal release geometry, King's tidal radius against the
Jacobi closed form, and the GPU release integration against the oracle."""

import numpy as np
import pytest

import oracle
import galax_amd.dynamics as gd
import galax_amd.potential as gp
from galax_amd.units import G, KMS


def test_fardal_release_geometry():
    """fardal15.py:86-92: trailing particle at x + r_t (k_r r_hat + k_z z_hat),
    velocity v + Omega r_t (k_vphi phi_hat + k_vz z_hat); leading mirrored in
    the radial and azimuthal parts."""
    rng = np.random.default_rng(0)
    x = rng.normal(0, 10, (64, 3))
    v = rng.normal(0, 0.2, (64, 3))
    r_t = rng.uniform(0.01, 0.1, 64)
    k = gd.FardalStreamDF().draw_k(3, 64)
    xl, vl, xt, vt = gd.FardalStreamDF.release(x, v, r_t, k)
    r_hat = x / np.linalg.norm(x, axis=1, keepdims=True)
    z_hat = np.cross(x, v)
    z_hat /= np.linalg.norm(z_hat, axis=1, keepdims=True)
    om = np.linalg.norm(np.cross(x, v), axis=1) / np.sum(x * x, axis=1)
    dt_, dl = xt - x, xl - x
    np.testing.assert_allclose(np.sum(dt_ * r_hat, 1), r_t * k[:, 0], rtol=1e-12)
    np.testing.assert_allclose(np.sum(dl * r_hat, 1), -r_t * k[:, 0], rtol=1e-12)
    np.testing.assert_allclose(np.sum(dt_ * z_hat, 1), r_t * k[:, 2], rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(np.sum(dl * z_hat, 1), r_t * k[:, 2], rtol=1e-10, atol=1e-14)
    dvt = vt - v
    np.testing.assert_allclose(np.sum(dvt * z_hat, 1), om * r_t * k[:, 3], rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(np.linalg.norm(dvt, axis=1), om * r_t * np.hypot(k[:, 1], k[:, 3]), rtol=1e-12)
    # k statistics follow the reference's distributions
    kk = gd.FardalStreamDF().draw_k(5, 200_000)
    assert abs(kk[:, 0].mean() - 2.0) < 0.01 and abs(kk[:, 0].std() - 0.5) < 0.01
    assert abs(kk[:, 2].std() - 0.5) < 0.01 and abs(np.mean(kk[:, 1] / kk[:, 0]) - 0.3) < 0.01


def test_king_radius_is_jacobi_radius_for_a_point_mass():
    """For a circular orbit in a Kepler potential, Omega^2 = GM/r^3 and
    d2Phi/dr2 = -2GM/r^3, so r_t = r (m / 3M)^(1/3)."""
    M, m = 1e11, 1e5
    pot = gp.KeplerPotential(m_tot=M)
    r = np.array([5.0, 10.0, 30.0])
    x = np.stack([r, np.zeros(3), np.zeros(3)], -1)
    v = np.stack([np.zeros(3), np.sqrt(G * M / r), np.zeros(3)], -1)
    H = oracle.evaluate(pot, "hessian", x)
    rhat = x / r[:, None]
    d2 = np.einsum("ni,nij,nj->n", rhat, H, rhat)
    rt = np.cbrt(G * m / (gd.omega(x, v) ** 2 - d2))
    np.testing.assert_allclose(rt, r * (m / (3 * M)) ** (1 / 3), rtol=1e-12)


@pytest.mark.gpu
def test_tidal_radius_on_gpu_matches_closed_form():
    M, m = 1e11, 1e5
    pot = gp.KeplerPotential(m_tot=M)
    r = np.array([5.0, 10.0, 30.0])
    x = np.stack([r, np.zeros(3), np.zeros(3)], -1)
    v = np.stack([np.zeros(3), np.sqrt(G * M / r), np.zeros(3)], -1)
    np.testing.assert_allclose(gd.tidal_radius(pot, x, v, mass=m), r * (m / (3 * M)) ** (1 / 3), rtol=1e-12)


@pytest.mark.gpu
def test_stream_release_integration_parity():
    """A 2x300-particle stream over 1 Gyr with the progenitor's own Plummer
    potential riding the progenitor's dense solution: the single-launch GPU
    release integration (MockStreamGenerator.run, track staged in LDS) agrees
    with the oracle integrating the same ICs in the same field."""
    mw = gp.MilkyWayPotential()
    prog = gp.PlummerPotential(m_tot=1e5, r_s=0.01)
    gen = gd.MockStreamGenerator(gd.FardalStreamDF(), mw, progenitor_potential=prog)
    ts = np.linspace(-1000.0, 0.0, 300)
    w0 = gd.PhaseSpaceCoordinate(q=np.array([10.0, 0.0, 2.0]), p=np.array([0.0, 0.2, 0.05]), t=-1000.0)
    print("[case] stream 600 particles", flush=True)
    stream, prog_f = gen.run(7, ts, w0, 1e5)
    assert np.all(stream.status == 0)
    assert stream["lead"].q.shape == (300, 3)
    # rebuild the same ICs and field, integrate with the oracle
    prog_o = gd.evaluate_orbit(mw, w0, ts, integrator=gen.progenitor_integrator, dense=True)
    prog_o.q = prog_o.q.reshape(-1, 3)
    prog_o.p = prog_o.p.reshape(-1, 3)
    ics = gd.FardalStreamDF().sample(7, mw, prog_o, 1e5)
    w_all = np.concatenate([ics["lead"].w(), ics["trail"].w()])
    field = gen.stream_potential(prog_o)
    assert type(field["progenitor"].translation.track).__name__ == "DenseTrack"
    ref, st, _ns, _ = oracle.integrate(field, gen.stream_integrator.dynamics_solver, w_all,
                                   np.concatenate([ts, ts]), ts[-1] + 1e-3, max_steps=None)
    got = np.concatenate([stream["lead"].w(), stream["trail"].w()])
    from test_gpu_parity import check_adaptive, noise_floor

    floor = noise_floor(field, gen.stream_integrator.dynamics_solver, w_all, np.concatenate([ts, ts]),
                        ts[-1] + 1e-3, ref, _ns, max_steps=None)
    check_adaptive(got[:, None, :], ref, stream.nsteps, _ns, floor)


@pytest.mark.gpu
def test_stream_knot_track_integration_parity():
    """The same stream field with the progenitor's track given as knots (its
    saved positions and velocities, cubic Hermite), what stream_potential
    builds from an orbit without a dense solution: GPU against the oracle."""
    mw = gp.MilkyWayPotential()
    prog = gp.PlummerPotential(m_tot=1e5, r_s=0.01)
    gen = gd.MockStreamGenerator(gd.FardalStreamDF(), mw, progenitor_potential=prog)
    ts = np.linspace(-1000.0, 0.0, 300)
    w0 = gd.PhaseSpaceCoordinate(q=np.array([10.0, 0.0, 2.0]), p=np.array([0.0, 0.2, 0.05]), t=-1000.0)
    prog_o = gd.evaluate_orbit(mw, w0, ts, integrator=gen.progenitor_integrator)
    prog_o.q = prog_o.q.reshape(-1, 3)
    prog_o.p = prog_o.p.reshape(-1, 3)
    field = gen.stream_potential(prog_o)
    assert type(field["progenitor"].translation.track).__name__ == "Track"
    ics = gd.FardalStreamDF().sample(7, mw, prog_o, 1e5)
    w_all = np.concatenate([ics["lead"].w(), ics["trail"].w()])
    t0 = np.concatenate([ts, ts])
    solver = gen.stream_integrator.dynamics_solver
    print("[case] stream 600 particles, knot track", flush=True)
    res = gd.integrate_orbits(field, w_all, t0, np.full(t0.shape, ts[-1] + 1e-3), solver=solver)
    got = res.out.cpu().numpy()
    ref, st, _ns, _ = oracle.integrate(field, solver, w_all, t0, ts[-1] + 1e-3, max_steps=None)
    from test_gpu_parity import check_adaptive, noise_floor

    floor = noise_floor(field, solver, w_all, t0, ts[-1] + 1e-3, ref, _ns, max_steps=None)
    check_adaptive(got, ref, res.nsteps.cpu().numpy(), _ns, floor)


@pytest.mark.gpu
def test_device_release_matches_fardal_sample():
    """MockStreamGenerator.run samples the release on the device (King
    radius from the device Hessian, Fardal's formula in torch, the host k
    draws): the same ICs as FardalStreamDF.sample to round-off."""
    import torch

    from galax_amd.dynamics._mockstream import release_torch, tidal_radius_device

    mw = gp.MilkyWayPotential()
    ts = np.linspace(-500.0, 0.0, 2000)
    w0 = gd.PhaseSpaceCoordinate(q=np.array([12.0, 1.0, 3.0]), p=np.array([0.01, 0.21, 0.03]), t=-500.0)
    prog_o = gd.evaluate_orbit(mw, w0, ts)
    prog_o.q = np.asarray(prog_o.q).reshape(-1, 3)
    prog_o.p = np.asarray(prog_o.p).reshape(-1, 3)
    want = gd.FardalStreamDF().sample(11, mw, prog_o, 2e4)
    want = np.concatenate([want["lead"].w(), want["trail"].w()])
    dev = dict(device="cuda", dtype=torch.float64)
    x, v = torch.as_tensor(prog_o.q, **dev), torch.as_tensor(prog_o.p, **dev)
    r_t = tidal_radius_device(mw, x, v, torch.full((ts.size,), 2e4, **dev), torch.as_tensor(ts, **dev))
    k = torch.as_tensor(gd.FardalStreamDF().draw_k(11, ts.size), **dev)
    got = release_torch(x, v, r_t, k).cpu().numpy()
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)
