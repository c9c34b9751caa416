"""The library's short FP64 primitives (hardware estimate + cubic
refinement for 1/x and 1/sqrt(x); the reduced atanh-series log1p) against
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

    from galax_amd import _lib

    rng = np.random.default_rng(0)
    pos = np.exp2(rng.uniform(-40, 40, 200_000))
    small = np.concatenate([np.exp2(rng.uniform(-60, -1, 50_000)), -np.exp2(rng.uniform(-60, -0.001, 50_000))])
    x = np.concatenate([pos, small, [1e-300, 1.0, 2.0, 0.5, 1e300]])
    xd = torch.as_tensor(x, device="cuda")
    out = torch.empty((x.size, 4), dtype=torch.float64, device="cuda")
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
    print("max ulp: rcp", e_rcp.max(), "rsqrt", e_rsq.max(), "div", e_div.max(), "log1p", e_log.max(), flush=True)
    assert e_rcp.max() <= 1.0
    assert e_rsq.max() <= 1.0
    assert e_div.max() <= 1.0
    assert e_log.max() <= 2.0
