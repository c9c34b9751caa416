"""Position-only tracks (TimeDependentTranslationParameter(t, q)): the
not-a-knot cubic spline that the reference's interpax CubicSpline builds
(xfm/translate.py:185-199), pinned to SciPy's CubicSpline, whose default
boundary condition is the same not-a-knot rule; on the device through the
track records, against SciPy's spline and the oracle's untranslated field."""

import numpy as np
import pytest

import galax_amd.potential as gp
from galax_amd.potential._composite import not_a_knot_slopes

scipy_interp = pytest.importorskip("scipy.interpolate")


def _knots(rng, n):
    t = np.cumsum(rng.uniform(0.2, 2.0, n)) - 5.0  # uneven spacing
    y = np.stack([10 * np.cos(0.3 * t), 8 * np.sin(0.21 * t), 0.5 * t + rng.normal(size=n) * 0.05], -1)
    return t, y


def _hermite(t, q, v, tq):
    """The device's interval record: power-basis cubic in s = (t - t_k)/h."""
    k = np.clip(np.searchsorted(t, tq, side="right") - 1, 0, t.size - 2)
    h = (t[k + 1] - t[k])[:, None]
    qa, qb, va, vb = q[k], q[k + 1], h * v[k], h * v[k + 1]
    c0, c1 = qa, va
    c2 = 3.0 * (qb - qa) - 2.0 * va - vb
    c3 = 2.0 * (qa - qb) + va + vb
    s = ((tq - t[k]) * (1.0 / (t[k + 1] - t[k])))[:, None]
    return c0 + s * (c1 + s * (c2 + s * c3))


@pytest.mark.parametrize("n", [2, 3, 4, 5, 8, 40, 2000])
def test_not_a_knot_slopes_match_scipy(n):
    rng = np.random.default_rng(n)
    t, y = _knots(rng, n)
    got = not_a_knot_slopes(t, y)
    want = scipy_interp.CubicSpline(t, y).derivative()(t)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("n", [3, 4, 9, 300])
def test_spline_track_equals_scipy_cubic_spline_between_knots(n):
    rng = np.random.default_rng(100 + n)
    t, y = _knots(rng, n)
    tr = gp.TimeDependentTranslationParameter(t, y)
    tq = rng.uniform(t[0], t[-1], 500)
    got = _hermite(tr.track.t, tr.track.q, tr.track.v, tq)
    want = scipy_interp.CubicSpline(t, y)(tq)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())


def test_unsorted_knots_are_sorted_first():
    rng = np.random.default_rng(7)
    t, y = _knots(rng, 12)
    perm = rng.permutation(t.size)
    a = gp.TimeDependentTranslationParameter(t, y).track
    b = gp.TimeDependentTranslationParameter(t[perm], y[perm]).track
    np.testing.assert_array_equal(a.t, b.t)
    np.testing.assert_allclose(a.v, b.v, rtol=1e-14)


@pytest.mark.gpu
def test_spline_track_on_gpu_matches_scipy_and_oracle():
    """A Plummer sphere riding a position-only track: the device field at
    (x, t) equals the oracle's untranslated Plummer field at x - spline(t),
    with spline(t) from SciPy."""
    import oracle

    print("[case] spline track on the GPU", flush=True)
    rng = np.random.default_rng(3)
    t, y = _knots(rng, 60)
    base = gp.PlummerPotential(m_tot=1e10, r_s=0.5)
    pot = gp.TranslatedPotential(base, gp.TimeDependentTranslationParameter(t, y))
    tq = rng.uniform(t[0], t[-1], 400)
    x = rng.normal(size=(400, 3)) * 5.0
    got = np.asarray(pot.acceleration(x, tq))
    shift = scipy_interp.CubicSpline(t, y)(tq)
    want = oracle.evaluate(base, "acceleration", x - shift)
    np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-12 * np.abs(want).max())


def test_callable_translation_is_tabulated_on_given_knots():
    fn = lambda t: np.stack([np.cos(0.05 * t), np.sin(0.05 * t), 0.01 * t], -1) * 7.0
    t = np.linspace(-100.0, 100.0, 201)
    tr = gp.TimeDependentTranslationParameter.from_callable(fn, t)
    tq = np.random.default_rng(1).uniform(-100.0, 100.0, 300)
    np.testing.assert_allclose(_hermite(tr.track.t, tr.track.q, tr.track.v, tq),
                               scipy_interp.CubicSpline(t, fn(t))(tq), rtol=0, atol=1e-11)
    # a callable passed straight to TranslatedPotential is refused with the remedy
    pot = gp.TranslatedPotential(gp.PlummerPotential(m_tot=1e9, r_s=0.1), fn)
    with pytest.raises(TypeError, match="from_callable"):
        pot._rows()
    with pytest.raises(ValueError, match="positions"):
        gp.TimeDependentTranslationParameter.from_callable(lambda t: t, t)
