"""galax's experimental API (dynamics/_src/experimental/): time parsing as
the reference's doctests state it, integrate_orbit against its docstring and
the oracle, and StreamSimulator's two stages against the oracle."""

import numpy as np
import pytest

import oracle
import galax_amd.dynamics as gd
import galax_amd.potential as gp
from galax_amd.dynamics import experimental as gdx
from galax_amd.dynamics._mockstream import FardalStreamDF, tidal_radius


def say(*a):
    print("[case]", *a, flush=True)


# ------------------------------------------------------------------ CPU
def test_parse_t0_t1_saveat_follows_the_reference():
    # experimental/integrate.py:46-92
    t0, t1, ts = gdx.parse_t0_t1_saveat(0.0, 1.0, None, dense=False)
    assert (t0, t1) == (0.0, 1.0) and ts.tolist() == [1.0]
    with pytest.raises(ValueError, match="t0, t1 must be specified"):
        gdx.parse_t0_t1_saveat(None, 1.0, None, dense=False)
    with pytest.raises(ValueError, match="t0, t1 must be specified"):
        gdx.parse_t0_t1_saveat(0.0, None, None, dense=False)
    t0, t1, ts = gdx.parse_t0_t1_saveat(0.0, None, np.array(0.5), dense=False)
    assert (t0, t1) == (0.0, 0.5) and ts.tolist() == [0.0, 0.5]
    with pytest.raises(ValueError, match="t0 must be specified"):
        gdx.parse_t0_t1_saveat(None, None, np.array(0.5), dense=False)
    t0, t1, ts = gdx.parse_t0_t1_saveat(0.0, 1.0, np.array([0.1, 0.5, 0.75]), dense=False)
    assert (t0, t1) == (0.0, 1.0) and ts.shape == (3,)
    t0, t1, ts = gdx.parse_t0_t1_saveat(None, None, np.array([0.1, 0.5, 0.75]), dense=False)
    assert (float(t0), float(t1)) == (0.1, 0.75)
    with pytest.raises(ValueError, match=r"saveat must be \[t0, ..., t1\]"):
        gdx.parse_t0_t1_saveat(None, None, np.array([0.5]), dense=False)


def test_default_solvers_are_the_references():
    s = gdx.default_solver
    assert isinstance(s.solver, gd.Dopri8) and s.max_steps == 10_000
    c = s.stepsize_controller
    assert (c.rtol, c.atol, c.dtmin, c.dtmax, c.force_dtmin) == (1e-7, 1e-7, 0.3, None, True)
    s5 = gdx.stream.default_solver
    assert isinstance(s5.solver, gd.Dopri5) and s5.stepsize_controller == c and s5.max_steps == 10_000


def test_fardal2015_release_matches_the_formula():
    """Given the k-draws, the release states are the Fardal (2015) formula
    with King's tidal radius (experimental/df.py:136-166), for both arms."""
    x = np.array([[15.0, 0.0, 0.0], [3.0, -8.0, 1.0]])
    v = np.array([[0.0, 0.22, 0.01], [0.15, 0.05, -0.02]])
    r_t = np.array([0.05, 0.02])
    k = np.array([[2.1, 0.4, 0.3, -0.2], [1.7, 0.2, -0.1, 0.4]])
    xl, vl, xt, vt = FardalStreamDF.release(x, v, r_t, k)
    for i in range(2):
        r_hat = x[i] / np.linalg.norm(x[i])
        z_hat = np.cross(x[i], v[i]) / np.linalg.norm(np.cross(x[i], v[i]))
        phi = v[i] - np.dot(v[i], r_hat) * r_hat
        phi_hat = phi / np.linalg.norm(phi)
        om = np.linalg.norm(np.cross(x[i], v[i])) / np.dot(x[i], x[i])
        kr, kvphi, kz, kvz = k[i]
        np.testing.assert_allclose(xl[i], x[i] - r_t[i] * (kr * r_hat - kz * z_hat), rtol=1e-15, atol=1e-16)
        np.testing.assert_allclose(xt[i], x[i] + r_t[i] * (kr * r_hat + kz * z_hat), rtol=1e-15, atol=1e-16)
        np.testing.assert_allclose(vl[i], v[i] - om * r_t[i] * (kvphi * phi_hat - kvz * z_hat), rtol=1e-14, atol=1e-17)
        np.testing.assert_allclose(vt[i], v[i] + om * r_t[i] * (kvphi * phi_hat + kvz * z_hat), rtol=1e-14, atol=1e-17)


def test_fardal2015_draws_follow_the_distributions():
    k = gdx.Fardal2015DF().draw_k(np.random.default_rng(1), 400_000)
    kr, kvphi, kz, kvz = k.T
    assert abs(kr.mean() - 2.0) < 0.005 and abs(kr.std() - 0.5) < 0.005
    assert abs(kz.mean()) < 0.005 and abs(kvz.std() - 0.5) < 0.005
    # k_vphi = k_r N(0.3, 0.5): mean 0.6, variance E[k_r^2] 0.25 + 0.09 Var k_r
    assert abs(kvphi.mean() - 0.6) < 0.01
    assert abs(kvphi.var() - ((4.25) * 0.25 + 0.09 * 0.25)) < 0.03


# ------------------------------------------------------------------ GPU
@pytest.mark.gpu
def test_integrate_orbit_docstring(golden):
    e = golden["integrator"]["experimental_integrate_orbit_nfw"]
    pot = gp.NFWPotential(m=1e12, r_s=15.0)
    xv0 = (np.array(e["q0"]), np.array(e["p0"]))
    saveat = np.linspace(0.0, 10.0, 100)
    say("experimental integrate_orbit docstring")
    sol = gdx.integrate_orbit(pot, xv0, t0=0, t1=10, saveat=saveat)
    assert sol.ys[0].shape == (100, 3) and sol.result == 0
    for k, row in e["rows"].items():
        np.testing.assert_allclose(sol.ys[0][int(k)], row["q"], atol=5.5e-9)
        np.testing.assert_allclose(sol.ys[1][int(k)], row["p"], atol=5.5e-9)
    # a field, and a batch with per-orbit save times
    f = gdx.integrate_orbit(gd.HamiltonianField(pot), xv0, t0=0, t1=10, saveat=saveat)
    np.testing.assert_array_equal(f.ys[0], sol.ys[0])
    qb = np.stack([xv0[0], xv0[0] * 1.1])
    pb = np.stack([xv0[1], xv0[1] * 0.9])
    tsb = np.stack([saveat, saveat * 0.5])
    b = gdx.integrate_orbit(pot, (qb, pb), t0=0.0, t1=np.array([10.0, 5.0]), saveat=tsb)
    assert b.ys[0].shape == (2, 100, 3)
    np.testing.assert_array_equal(b.ys[0][0], sol.ys[0])
    ref, st, ns, _ = oracle.integrate(pot, gdx.default_solver, np.concatenate([qb, pb], 1), np.zeros(2),
                                      np.array([10.0, 5.0]), ts=tsb)
    np.testing.assert_allclose(b.ys[0], ref[..., :3], rtol=0, atol=1e-10)
    # dense: the interpolation evaluates between the knots
    d = gdx.integrate_orbit(pot, xv0, t0=0, t1=10, dense=True)
    assert d.ys is None
    q, p = d.evaluate(saveat[[25, 72]])
    # the same steps and the same interpolant as the saves: round-off apart
    np.testing.assert_allclose(q, sol.ys[0][[25, 72]], rtol=0, atol=1e-12)
    np.testing.assert_allclose(p, sol.ys[1][[25, 72]], rtol=0, atol=1e-12)
    q_end, _ = d.evaluate(10.0)
    np.testing.assert_allclose(q_end[0], sol.ys[0][-1], atol=1e-12)
    assert np.all(np.isnan(d.evaluate(10.5)[0]))  # beyond the integrated interval


@pytest.mark.gpu
def test_stream_simulator_matches_oracle():
    """init: the progenitor at the sorted release times (the oracle's solve of
    the same orbit) and the Fardal release there; run: every particle from
    its release time to t1 (the oracle's independent solves)."""
    pot = gp.HernquistPotential(m_tot=1e12, r_s=10.0)
    qp0 = (np.array([15.0, 0.0, 0.0]), np.array([0.0, 0.225, 0.0]))
    release = np.linspace(-1500.0, -150.0, 300)[::-1]  # unsorted on purpose
    sim = gdx.StreamSimulator()
    say("StreamSimulator init")
    ics = sim.init(pot, qp0, 0.0, release_times=release, Msat=1e5, key=3)
    assert np.all(np.diff(ics.release_times) > 0) and ics.qp_lead[0].shape == (300, 3)
    solver = gdx.stream.default_solver
    w0 = np.concatenate(qp0)[None]
    a, *_ = oracle.integrate(pot, solver, w0, 0.0, ics.release_times[0])
    b, st, *_ = oracle.integrate(pot, solver, a[:, -1], ics.release_times[0], ics.release_times[-1],
                                 ts=ics.release_times)
    xs, vs = b[0, :, :3], b[0, :, 3:]
    r_t = tidal_radius(pot, xs, vs, mass=np.full(300, 1e5), t=ics.release_times)
    k = gdx.Fardal2015DF().draw_k(3, 300)
    xl, vl, xt, vt = FardalStreamDF.release(xs, vs, r_t, k)
    scale = np.abs(xs).max()
    np.testing.assert_allclose(ics.qp_lead[0], xl, rtol=0, atol=1e-8 * scale)
    np.testing.assert_allclose(ics.qp_trail[1], vt, rtol=0, atol=1e-8 * np.abs(vs).max())
    say("StreamSimulator run")
    (xl1, vl1), (xt1, vt1) = sim.run(pot, ics, t1=0.0)
    assert xl1.shape == (300, 3)
    w = np.concatenate([np.concatenate(ics.qp_lead, 1), np.concatenate(ics.qp_trail, 1)])
    t0 = np.concatenate([ics.release_times, ics.release_times])
    ref, st, ns, _ = oracle.integrate(pot, solver, w, t0, np.zeros(600))
    got = np.concatenate([np.concatenate([xl1, vl1], 1), np.concatenate([xt1, vt1], 1)])
    e = np.linalg.norm(got - ref[:, 0], axis=1) / np.linalg.norm(ref[:, 0], axis=1)
    print("median", np.median(e), "max", e.max(), flush=True)
    assert np.median(e) < 1e-10 and e.max() < 1e-5
