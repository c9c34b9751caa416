"""Dense-solution tracks: a centre that rides an integrated orbit's own dense
output (TimeDependentTranslationParameter.from_dense; the progenitor's
track in MockStreamGenerator).

CPU: the oracle's dense track is pinned by closed forms -- records whose
stage accelerations follow a polynomial in time must give back the exact
polynomial trajectory (the interpolants of both methods reproduce cubic
motion), forward and backward in time, with the end positions held.
GPU: the device's power-basis pieces (field kernels, generic lookups) and
the LDS-staged pieces of the mock-stream integrator against the oracle."""

import types

import numpy as np
import pytest

import oracle
import galax_amd.dynamics as gd
import galax_amd.potential as gp
from galax_amd import _abi

# Dopri8 / Dopri5 nodes (Prince & Dormand RK8(7)13M; Dormand & Prince 5(4))
_C8 = [0.0, 1 / 18, 1 / 12, 1 / 8, 5 / 16, 3 / 8, 59 / 400, 93 / 200, 5490023248 / 9719169821, 13 / 20,
       1201146811 / 1299019798, 1.0, 1.0, 1.0]  # 13 stages, then the end stage
_C5 = [0.0, 1 / 5, 3 / 10, 4 / 5, 8 / 9, 1.0, 1.0]


def _motion(tau, q0, v0, a, j):
    return q0 + v0 * tau + a * tau ** 2 / 2 + j * tau ** 3 / 6


def _records(method, t0, hs, q0, v0, a, j):
    """Step records of cubic motion q(t) = q0 + v0 s + a s^2/2 + j s^3/6
    (s = t - t0) for steps hs from t0: each holds the step's start state
    and the acceleration a + j s at its stage times."""
    c = _C8 if method == _abi.GX_DOPRI8 else _C5
    rec = np.zeros((len(hs), _abi.GX_DENSE_RECORD))
    ts = t0
    for k, h in enumerate(hs):
        s = ts - t0
        rec[k, 0], rec[k, 1] = ts, h
        rec[k, 2:5] = _motion(s, q0, v0, a, j)
        rec[k, 5:8] = v0 + a * s + j * s ** 2 / 2
        for l, cl in enumerate(c):
            rec[k, 8 + 3 * l:11 + 3 * l] = a + j * (s + cl * h)
        ts = ts + h
    return rec


def _field(rec, method):
    sol = types.SimpleNamespace(rec=rec[None], nrec=np.array([rec.shape[0]]), method=method)
    track = gp.TimeDependentTranslationParameter.from_dense(sol)
    return gp.TranslatedPotential(gp.KeplerPotential(m_tot=1e9), track)


def _kepler_acc(x, centre):
    from galax_amd.units import G

    d = x - centre
    r = np.linalg.norm(d, axis=-1, keepdims=True)
    return -G * 1e9 * d / r ** 3


@pytest.mark.parametrize("method", [_abi.GX_DOPRI8, _abi.GX_DOPRI5])
@pytest.mark.parametrize("backward", [False, True])
def test_oracle_dense_track_reproduces_cubic_motion(method, backward):
    q0, v0 = np.array([1.0, -2.0, 0.5]), np.array([0.3, 0.1, -0.2])
    a, j = np.array([0.02, -0.01, 0.03]), np.array([-0.004, 0.002, 0.001])
    hs = np.array([0.7, 1.3, 0.4, 2.0, 0.9])
    if backward:
        hs = -hs
    rec = _records(method, 10.0, hs, q0, v0, a, j)
    pot = _field(rec, method)
    t_end = 10.0 + hs.sum()
    lo, hi = min(10.0, t_end), max(10.0, t_end)
    t = np.concatenate([np.linspace(lo, hi, 41), [lo - 3.0, hi + 2.0]])
    x = np.array([5.0, 4.0, -3.0])
    got = oracle.evaluate(pot, "acceleration", np.broadcast_to(x, (t.size, 3)), t)
    s = np.clip(t, lo, hi) - 10.0
    centre = _motion(s[:, None], q0, v0, a, j)
    np.testing.assert_allclose(got, _kepler_acc(x, centre), rtol=1e-12, atol=0)


def test_oracle_dense_track_detects_a_wrong_stage_acceleration():
    """The check above is sensitive to the records: one stage acceleration
    off by 1e-6 moves the track inside that step."""
    q0, v0 = np.array([1.0, -2.0, 0.5]), np.array([0.3, 0.1, -0.2])
    a, j = np.array([0.02, -0.01, 0.03]), np.array([-0.004, 0.002, 0.001])
    rec = _records(_abi.GX_DOPRI8, 0.0, [1.0, 1.0], q0, v0, a, j)
    rec[0, 8 + 3 * 7] += 1e-6
    pot = _field(rec, _abi.GX_DOPRI8)
    x = np.array([[5.0, 4.0, -3.0]])
    got = oracle.evaluate(pot, "acceleration", x, np.array([0.5]))
    want = _kepler_acc(x, _motion(0.5, q0, v0, a, j))
    assert np.max(np.abs(got - want) / np.abs(want)) > 1e-9


def _progenitor(ts, dt0=None, fixed=False):
    mw = gp.MilkyWayPotential()
    w0 = gd.PhaseSpaceCoordinate(q=np.array([12.0, 0.0, 1.0]), p=np.array([0.0, 0.21, 0.03]), t=float(ts[0]))
    if fixed:
        integ = gd.Integrator(gd.OrbitSolver(gd.Dopri8(), stepsize_controller=gd.ConstantStepSize(),
                                             max_steps=None), diffeq_kw={"dt0": dt0})
    else:
        integ = gd.Integrator(gd.OrbitSolver(gd.Dopri8(), stepsize_controller=gd.PIDController(rtol=1e-8, atol=1e-8),
                                             max_steps=None))
    return mw, gd.evaluate_orbit(mw, w0, ts, integrator=integ, dense=True)


def _stream_field(mw, prog_o, m=1e5, b=0.01):
    track = gp.TimeDependentTranslationParameter.from_dense(prog_o)
    return gp.CompositePotential(host=mw, progenitor=gp.TranslatedPotential(gp.PlummerPotential(m_tot=m, r_s=b),
                                                                            track))


def _satellite_orbits(prog_o, ts, every):
    """Orbits that start 0.5-1.5 kpc from the progenitor (here a 1e9 Msun
    Plummer sphere of scale 0.3 kpc) at every `every`-th time of the first
    half of ts: its pull matters to them, but they do not fall through its
    core, so that a fixed-step comparison is not swamped by chaos."""
    q = np.asarray(prog_o.q).reshape(-1, 3)
    p = np.asarray(prog_o.p).reshape(-1, 3)
    sel = np.arange(0, ts.size // 2, every)
    rng = np.random.default_rng(2)
    d = rng.normal(0.0, 1.0, (sel.size, 3))
    d *= rng.uniform(0.5, 1.5, (sel.size, 1)) / np.linalg.norm(d, axis=1, keepdims=True)
    w0 = np.concatenate([q[sel] + d, p[sel] + rng.normal(0.0, 0.01, (sel.size, 3))], -1)
    return w0, ts[sel]


@pytest.mark.gpu
@pytest.mark.parametrize("backward", [False, True])
def test_dense_track_field_matches_oracle(backward):
    """gx_acceleration on the Milky Way plus a Plummer sphere riding a
    progenitor's dense solution (integrated forward, or backward in time),
    at times across and beyond the solve, against the oracle's own dense
    output of the same records."""
    ts = np.linspace(-2000.0, 0.0, 400)
    mw, prog_o = _progenitor(ts[::-1].copy() if backward else ts)
    assert (float(prog_o.interpolant.dense.rec[0, 0, 1]) < 0) == backward
    field = _stream_field(mw, prog_o)
    rng = np.random.default_rng(5)
    t = np.concatenate([rng.uniform(-2000.0, 0.0, 3000), [-2500.0, -2000.0, 0.0, 1.0]])
    centre = np.array(prog_o(t).q, dtype=np.float64).reshape(-1, 3)
    centre[t < -2000.0] = centre[t == -2000.0][0]
    centre[t > 0.0] = centre[t == 0.0][0]
    x = centre + rng.normal(0.0, 0.05, centre.shape)
    print(f"[case] dense track field, 3004 points, backward={backward}", flush=True)
    got = np.asarray(field.acceleration(x, t))
    want = oracle.evaluate(field, "acceleration", x, t)
    err = np.linalg.norm(got - want, axis=-1) / np.linalg.norm(want, axis=-1)
    assert err.max() < 1e-11, err.max()


@pytest.mark.gpu
@pytest.mark.parametrize("backward", [False, True])
def test_lds_track_integration_matches_oracle_fixed_step(backward):
    """The mock-stream integrator with the track staged in LDS (a few dozen
    pieces) at a fixed step: every stage's track lookup against the oracle.
    The orbits circle a 1e9 Msun progenitor, whose pull changes them by more
    than 1e-3 over the run (checked), so a misplaced track would show far
    above the 1e-11 bound; the differences that remain are round-off grown
    by the encounters (1e-14 to 2e-12).  Backward: the progenitor and the
    orbits integrated from the present into the past (pieces of negative
    step, staged in reverse)."""
    ts = np.linspace(-300.0, 0.0, 300)
    if backward:
        ts = ts[::-1].copy()
    mw, prog_o = _progenitor(ts)
    field = _stream_field(mw, prog_o, m=1e9, b=0.3)
    nrec = int(prog_o.interpolant.dense.nrec[0])
    assert 8 <= nrec <= 512, nrec
    w0, t0 = _satellite_orbits(prog_o, ts, 2)
    t_end = float(ts[-1])
    solver = gd.OrbitSolver(gd.Dopri8(), stepsize_controller=gd.ConstantStepSize(), max_steps=None)
    print(f"[case] LDS track fixed step, {w0.shape[0]} orbits, {nrec} pieces, backward={backward}", flush=True)
    res = gd.integrate_orbits(field, w0, t0, np.full(t0.shape, t_end), solver=solver, dt0=0.5)
    got = res.out[:, 0, :].cpu().numpy()
    ref, st, _, _ = oracle.integrate(field, solver, w0, t0, t_end, dt0=0.5, max_steps=None)
    assert np.all(res.status.cpu().numpy() == 0) and np.all(st == 0)
    rel = np.linalg.norm(got - ref[:, -1], axis=-1) / np.linalg.norm(ref[:, -1], axis=-1)
    assert rel.max() < 1e-11, rel.max()
    free, _, _, _ = oracle.integrate(mw, solver, w0, t0, t_end, dt0=0.5, max_steps=None)
    pull = np.linalg.norm(free[:, -1] - ref[:, -1], axis=-1) / np.linalg.norm(ref[:, -1], axis=-1)
    assert np.median(pull) > 1e-3, np.median(pull)


@pytest.mark.gpu
def test_dense_track_too_long_for_lds_uses_device_memory():
    """A progenitor solved in 3,000 fixed steps gives more pieces than LDS
    holds: the field then looks its pieces up in device memory, with the
    same results as the oracle."""
    ts = np.linspace(-300.0, 0.0, 300)
    mw, prog_o = _progenitor(ts, dt0=0.1, fixed=True)
    assert int(prog_o.interpolant.dense.nrec[0]) == 3000
    field = _stream_field(mw, prog_o, m=1e9, b=0.3)
    w0, t0 = _satellite_orbits(prog_o, ts, 5)
    solver = gd.OrbitSolver(gd.Dopri8(), stepsize_controller=gd.ConstantStepSize(), max_steps=None)
    print(f"[case] dense track in device memory, {w0.shape[0]} orbits", flush=True)
    res = gd.integrate_orbits(field, w0, t0, np.full(t0.shape, 0.0), solver=solver, dt0=1.0)
    got = res.out[:, 0, :].cpu().numpy()
    ref, _, _, _ = oracle.integrate(field, solver, w0, t0, 0.0, dt0=1.0, max_steps=None)
    rel = np.linalg.norm(got - ref[:, -1], axis=-1) / np.linalg.norm(ref[:, -1], axis=-1)
    assert rel.max() < 1e-11, rel.max()


@pytest.mark.gpu
def test_dense_track_rejects_broken_records():
    """Records that are not one contiguous solve are refused (EINVAL)."""
    from galax_amd._lib import GalaxHipError

    ts = np.linspace(-500.0, 0.0, 300)
    mw, prog_o = _progenitor(ts)
    dense = prog_o.interpolant.dense
    nr = int(dense.nrec[0])
    rec = dense.rec[0, :nr].cpu().numpy().copy()
    rec[nr // 2, 0] += 1e-3  # a gap
    bad = types.SimpleNamespace(rec=rec[None], nrec=np.array([nr]), method=dense.method)
    f = gp.TranslatedPotential(gp.PlummerPotential(m_tot=1e5, r_s=0.01),
                               gp.TimeDependentTranslationParameter.from_dense(bad))
    with pytest.raises(GalaxHipError):
        f.acceleration(np.array([[1.0, 0.0, 0.0]]), 0.0)
