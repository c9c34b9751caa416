"""The allocation-free integrator entry point (gx_integrate_noalloc): its
workspace is a fixed per-GPU device array, so it must give exactly what
gx_integrate gives with a workspace of its own, call after call (the array
keeps the last call's contents), and refuse batches beyond its capacity."""

import ctypes

import numpy as np
import pytest

import galax_amd.dynamics as gd
import galax_amd.potential as gp
from galax_amd import _abi, _lib, workloads

pytestmark = pytest.mark.gpu


def _call(fn, pot, cfg, w0, t1, T, n, with_ws):
    import torch

    w0d = torch.as_tensor(w0[:n], device="cuda").contiguous()
    t0d = torch.zeros(n, dtype=torch.float64, device="cuda")
    t1d = torch.as_tensor(np.asarray(t1[:n], dtype=np.float64), device="cuda").contiguous()
    out = torch.full((n, T, 6), float("nan"), dtype=torch.float64, device="cuda")
    status = torch.empty(n, dtype=torch.int32, device="cuda")
    nsteps = torch.empty(n, dtype=torch.int64, device="cuda")
    naccept = torch.empty(n, dtype=torch.int64, device="cuda")
    args = [pot._handle(), ctypes.addressof(cfg), _lib.ptr(w0d), _lib.ptr(t0d), _lib.ptr(t1d), 1, None, T,
            _abi.GX_TS_LINSPACE, n, None, _lib.ptr(out), _lib.ptr(status), _lib.ptr(nsteps), _lib.ptr(naccept)]
    if with_ws:
        ws = torch.empty((int(_lib.lib().gx_integrate_workspace_bytes(n)) + 7) // 8, dtype=torch.float64,
                         device="cuda")
        rc = fn(*args, _lib.ptr(ws), _lib.stream_ptr())
    else:
        rc = fn(*args, _lib.stream_ptr())
    torch.cuda.synchronize()
    return rc, out.cpu().numpy(), status.cpu().numpy(), nsteps.cpu().numpy()


@pytest.mark.parametrize("adaptive", [True, False])
def test_noalloc_matches_the_workspace_entry_point(adaptive):
    lib = _lib.lib()
    pot = gp.MilkyWayPotential()
    w0, _, T1 = workloads.config_uneven(n=3000, seed=21)
    T1 = np.minimum(T1, 400.0)
    if adaptive:
        solver = gd.OrbitSolver(gd.Dopri8(), stepsize_controller=gd.PIDController(rtol=1e-8, atol=1e-10),
                                max_steps=None)
        cfg = solver.cfg()
    else:
        solver = gd.OrbitSolver(gd.Dopri8(), stepsize_controller=gd.ConstantStepSize(), max_steps=None)
        cfg = solver.cfg(dt0=0.9)
    print("[case] noalloc", "adaptive" if adaptive else "fixed", flush=True)
    # two calls in a row, the second smaller: the shared array keeps the
    # first call's contents and must not leak into the second
    for n in (3000, 1100):
        rc, out, st, ns = _call(lib.gx_integrate_noalloc, pot, cfg, w0, T1, 8, n, False)
        assert rc == 0
        rc2, out2, st2, ns2 = _call(lib.gx_integrate, pot, cfg, w0, T1, 8, n, True)
        assert rc2 == 0
        np.testing.assert_array_equal(out, out2)
        np.testing.assert_array_equal(st, st2)
        np.testing.assert_array_equal(ns, ns2)
        assert not np.isnan(out).any()


def test_noalloc_refuses_batches_beyond_its_capacity():
    lib = _lib.lib()
    cap = int(lib.gx_integrate_noalloc_max_orbits())
    assert cap == ((4 << 20) - 256) // 32
    pot = gp.MilkyWayPotential()
    cfg = gd.OrbitSolver(gd.Dopri8(), stepsize_controller=gd.ConstantStepSize(), max_steps=None).cfg(dt0=1.0)
    import torch

    w = torch.zeros((1, 6), dtype=torch.float64, device="cuda")
    t = torch.zeros(1, dtype=torch.float64, device="cuda")
    o = torch.empty((1, 1, 6), dtype=torch.float64, device="cuda")
    # argument checks come first for sizes the kernels would read past: the
    # entry point refuses n above its capacity before touching any pointer
    rc = lib.gx_integrate_noalloc(pot._handle(), ctypes.addressof(cfg), _lib.ptr(w), _lib.ptr(t), _lib.ptr(t), 0,
                                  None, 1, _abi.GX_TS_FINAL, cap + 1, None, _lib.ptr(o), None, None, None,
                                  _lib.stream_ptr())
    assert rc == _abi.GX_ETOOBIG
    # an empty batch is a no-op and a negative size an argument error, as
    # for gx_integrate
    args = (pot._handle(), ctypes.addressof(cfg), _lib.ptr(w), _lib.ptr(t), _lib.ptr(t), 0, None, 1,
            _abi.GX_TS_FINAL)
    tail = (None, _lib.ptr(o), None, None, None, _lib.stream_ptr())
    assert lib.gx_integrate_noalloc(*args, 0, *tail) == _abi.GX_OK
    assert lib.gx_integrate_noalloc(*args, -1, *tail) == _abi.GX_EINVAL
