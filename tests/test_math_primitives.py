"""The library's short FP64 primitives (hardware estimate + cubic
refinement for 1/x and 1/sqrt(x); the reduced atanh-series log1p; the
square root, division by 6 and N-th roots built on them) against
extended-precision references on the host (x87 80-bit long double)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ulp_err(got, ref_ld):
    """|got - ref| in units of the float64 spacing at ref."""
    ref64 = ref_ld.astype(np.float64)
    spacing = np.spacing(np.abs(ref64))
    return np.abs((got.astype(np.longdouble) - ref_ld) / spacing.astype(np.longdouble)).astype(np.float64)


def test_primitives_within_an_ulp():
    import torch

    from galax_amd import _abi, _lib

    k = _abi.GX_MATH_CHECK_OUTPUTS
    rng = np.random.default_rng(0)
    pos = np.exp2(rng.uniform(-40, 40, 200_000))
    wide = np.exp2(rng.uniform(-1000, 1000, 50_000))
    small = np.concatenate([np.exp2(rng.uniform(-60, -1, 50_000)), -np.exp2(rng.uniform(-60, -0.001, 50_000))])
    x = np.concatenate([pos, wide, small, [1e-300, 1.0, 2.0, 0.5, 6.0, 1e300, 2.2250738585072014e-308]])
    xd = torch.as_tensor(x, device="cuda")
    out = torch.empty((x.size, k), dtype=torch.float64, device="cuda")
    print(f"[case] math primitives n={x.size}", flush=True)
    _lib.check(_lib.lib().gx_math_check(_lib.ptr(xd), x.size, _lib.ptr(out), _lib.stream_ptr()), "gx_math_check")
    o = out.cpu().numpy()
    xl = x.astype(np.longdouble)
    posm = x > 0
    e_rcp = _ulp_err(o[posm, 0], 1 / xl[posm])
    e_rsq = _ulp_err(o[posm, 1], 1 / np.sqrt(xl[posm]))
    e_div = _ulp_err(o[posm, 3], 1 / xl[posm])
    lm = x > -1
    e_log = _ulp_err(o[lm, 2], np.log1p(xl[lm]))
    e_sqrt = _ulp_err(o[posm, 4], np.sqrt(xl[posm]))
    e_div6 = _ulp_err(o[:, 5], xl / 6)
    e_r9 = _ulp_err(o[posm, 6], np.power(xl[posm], np.longdouble(1) / 9))
    e_r6 = _ulp_err(o[posm, 7], np.power(xl[posm], np.longdouble(1) / 6))
    e_r8 = _ulp_err(o[posm, 8], np.power(xl[posm], np.longdouble(1) / 8))
    print("max ulp: rcp", e_rcp.max(), "rsqrt", e_rsq.max(), "div", e_div.max(), "log1p", e_log.max(),
          "sqrt", e_sqrt.max(), "div6", e_div6.max(), "root9", e_r9.max(), "root6", e_r6.max(),
          "root8", e_r8.max(), flush=True)
    assert e_rcp.max() <= 1.0
    assert e_rsq.max() <= 1.0
    assert e_div.max() <= 1.0
    assert e_log.max() <= 2.0
    assert e_sqrt.max() <= 0.6
    assert e_div6.max() <= 1.0
    assert e_r9.max() <= 1.25 and e_r6.max() <= 1.25
    # three chained reciprocal square roots and a reciprocal (the step
    # controller's factor exponent 1/8)
    assert e_r8.max() <= 2.5


def test_primitive_special_values():
    """The values the step controller and the first-step rule rely on at
    zero, infinity and NaN."""
    import torch

    from galax_amd import _abi, _lib

    k = _abi.GX_MATH_CHECK_OUTPUTS
    x = np.array([0.0, np.inf, np.nan, 1e-310])
    xd = torch.as_tensor(x, device="cuda")
    out = torch.empty((x.size, k), dtype=torch.float64, device="cuda")
    _lib.check(_lib.lib().gx_math_check(_lib.ptr(xd), x.size, _lib.ptr(out), _lib.stream_ptr()), "gx_math_check")
    o = out.cpu().numpy()
    # sqrt: 0 -> 0, inf -> inf, NaN -> NaN, subnormal -> 0
    assert o[0, 4] == 0.0 and o[1, 4] == np.inf and np.isnan(o[2, 4]) and o[3, 4] == 0.0
    # x / 6 as the division gives it
    assert o[0, 5] == 0.0 and o[1, 5] == np.inf and np.isnan(o[2, 5])
    # roots as pow gives them
    for j in (6, 7):
        assert o[0, j] == 0.0 and o[1, j] == np.inf and np.isnan(o[2, j])
        assert o[3, j] == pytest.approx(1e-310 ** (1 / (9 if j == 6 else 6)), rel=1e-15)
