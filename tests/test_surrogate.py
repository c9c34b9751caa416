"""The surrogate kernels: the BF16 MFMA GEMM against float64 products of the
same BF16 operands (every epilogue, strided and split-K shapes), training
gradients against torch autograd, Adam against its formula, and the
lock-step integrator against the persistent FP64 integrator."""

import math

import numpy as np
import pytest

from galax_amd import _abi, workloads

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    return torch


def _bf(x):
    torch = _torch()
    return x.to(torch.bfloat16)


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def _gemm_bound(a, b):
    """Round-off bound of an fp32-accumulated product: K * 2^-23 * |a||b|."""
    return a.shape[1] * 2.0 ** -23 * (np.abs(a) @ np.abs(b)) + 1e-30


@pytest.mark.parametrize("M,N,K", [(300, 200, 72), (64, 64, 32), (1, 3, 256), (129, 65, 40)])
def test_gemm_forward_epilogues(M, N, K):
    import galax_amd.surrogate as S

    torch = _torch()
    print(f"[case] gemm {M}x{N}x{K}", flush=True)
    g = torch.Generator().manual_seed(M * 7 + N)
    A = _bf(torch.randn(M, K, generator=g)).cuda()
    W = _bf(torch.randn(N, K, generator=g)).cuda()  # B(k, n) = W[n][k]
    bias = torch.randn(N, generator=g).float().cuda()
    a = A.double().cpu().numpy()
    b = W.double().cpu().numpy().T
    ref = a @ b
    bound = _gemm_bound(a, b)
    C = torch.empty(M, N, dtype=torch.float32, device="cuda")
    S.gemm(A, K, 1, W, 1, K, M, N, K, _abi.GX_EPI_BIAS, bias=bias, C=C, ldc=N)
    zref = ref + bias.double().cpu().numpy()
    err = np.abs(C.double().cpu().numpy() - zref)
    assert np.all(err <= bound + 2.0 ** -23 * np.abs(zref)), err.max()
    # bias + SiLU: Z in fp32 and silu(Z) in bf16
    Z = torch.empty(M, N, dtype=torch.float32, device="cuda")
    H = torch.empty(M, N, dtype=torch.bfloat16, device="cuda")
    S.gemm(A, K, 1, W, 1, K, M, N, K, _abi.GX_EPI_BIAS_SILU, bias=bias, C=Z, ldc=N, H=H, ldh=N)
    assert np.all(np.abs(Z.double().cpu().numpy() - zref) <= bound + 2.0 ** -23 * np.abs(zref))
    hs = _silu(Z.double().cpu().numpy())
    np.testing.assert_allclose(H.double().cpu().numpy(), hs, rtol=2.0 ** -8, atol=1e-6)
    # SiLU backward
    Zin = torch.randn(M, N, generator=g).float().cuda()
    D = torch.empty(M, N, dtype=torch.bfloat16, device="cuda")
    S.gemm(A, K, 1, W, 1, K, M, N, K, _abi.GX_EPI_DSILU, H=D, ldh=N, Z=Zin, ldz=N)
    dref = ref * _dsilu(Zin.double().cpu().numpy())
    dbound = bound * np.abs(_dsilu(Zin.double().cpu().numpy()))
    assert np.all(np.abs(D.double().cpu().numpy() - dref) <= dbound + 2.0 ** -8 * np.abs(dref) + 1e-6)


@pytest.mark.parametrize("ksplit", [1, 4, 37])
def test_gemm_transposed_operands_split_k(ksplit):
    """dW = dZ^T H shape: both operands strided along the reduction."""
    import galax_amd.surrogate as S

    torch = _torch()
    Bn, M, N = 5000, 96, 72
    print(f"[case] gemm transposed ksplit={ksplit}", flush=True)
    g = torch.Generator().manual_seed(ksplit)
    dZ = _bf(torch.randn(Bn, M, generator=g)).cuda()  # A(m, k) = dZ[k][m]
    H = _bf(torch.randn(Bn, N, generator=g)).cuda()  # B(k, n) = H[k][n]
    a = dZ.double().cpu().numpy().T
    b = H.double().cpu().numpy()
    ref = a @ b
    C0 = torch.randn(M, N, generator=g).float().cuda()
    C = C0.clone()
    S.gemm(dZ, 1, M, H, N, 1, M, N, Bn, _abi.GX_EPI_STORE, C=C, ldc=N, accumulate=1, ksplit=ksplit)
    want = ref + C0.double().cpu().numpy()
    bound = _gemm_bound(a, b) + ksplit * 2.0 ** -23 * (np.abs(ref) + np.abs(want))
    assert np.all(np.abs(C.double().cpu().numpy() - want) <= bound)


def test_gemm_scalar_strides_and_bad_arguments():
    import galax_amd.surrogate as S
    from galax_amd import _lib

    torch = _torch()
    g = torch.Generator().manual_seed(3)
    A = _bf(torch.randn(50, 2 * 33, generator=g)).cuda()  # every other column: sak = 2
    W = _bf(torch.randn(33, 20, generator=g)).cuda()  # B(k, n) = W[k][n]: sbk = 20, sbn = 1
    C = torch.empty(50, 20, dtype=torch.float32, device="cuda")
    S.gemm(A, 66, 2, W, 20, 1, 50, 20, 33, _abi.GX_EPI_STORE, C=C, ldc=20)
    a = A.double().cpu().numpy()[:, ::2]
    b = W.double().cpu().numpy()
    assert np.all(np.abs(C.double().cpu().numpy() - a @ b) <= _gemm_bound(a, b))
    with pytest.raises(_lib.GalaxHipError):
        S.gemm(A, 66, 2, W, 20, 1, 50, 20, 33, _abi.GX_EPI_BIAS, C=C, ldc=20)  # no bias
    with pytest.raises(_lib.GalaxHipError):
        S.gemm(A, 66, 2, W, 20, 1, 50, 20, 33, _abi.GX_EPI_BIAS_SILU, bias=C[0], ldc=20, ksplit=2)


@pytest.mark.parametrize("n", [1, 63, 64, 1000])
def test_fused_inference_matches_layer_by_layer(n):
    """gx_mlp_forward_fused (one kernel, activations in LDS) against the
    layer-by-layer GEMM path on the same BF16 weights: both accumulate in
    FP32 and round activations to BF16, so they differ only where an
    accumulation order moves a BF16 rounding."""
    import galax_amd.surrogate as S

    torch = _torch()
    mlp = S.MLP(hidden=256, depth=4, seed=3)
    gen = torch.Generator(device="cuda").manual_seed(1)
    for w in mlp.b:
        w.copy_(torch.randn(w.shape, generator=gen, device="cuda") * 0.1)
    X = torch.zeros((n, S.K_IN), dtype=torch.bfloat16, device="cuda")
    X[:, :5] = torch.randn((n, 5), generator=gen, device="cuda").to(torch.bfloat16)
    Y0 = torch.empty((n, 3), dtype=torch.float32, device="cuda")
    Y1 = torch.full((n, 3), float("nan"), dtype=torch.float32, device="cuda")
    mlp.forward(X, n, Y0)
    mlp.forward_fused(X, n, Y1)
    d = (Y1 - Y0).abs().max().item()
    scale = Y0.abs().max().item()
    print(f"n={n} max |fused - layered| {d:.3e} of {scale:.3e}", flush=True)
    assert torch.isfinite(Y1).all()
    assert d <= 2e-2 * scale


def test_adam_matches_formula():
    import galax_amd.surrogate as S
    from galax_amd import _lib

    torch = _torch()
    g = torch.Generator().manual_seed(9)
    n = 1000
    p = torch.randn(n, generator=g).float().cuda()
    grad = torch.randn(n, generator=g).float().cuda()
    m = torch.zeros(n, device="cuda")
    v = torch.zeros(n, device="cuda")
    pb = torch.empty(n, dtype=torch.bfloat16, device="cuda")
    P, Mm, V = p.double().cpu().numpy(), np.zeros(n), np.zeros(n)
    for step in (1, 2, 3):
        rc = _lib.lib().gx_mlp_adam(S._p(p), S._p(grad), S._p(m), S._p(v), S._p(pb), n, 1e-2, 0.9, 0.999, 1e-8,
                                    step, S._s())
        _lib.check(rc, "adam")
        G = grad.double().cpu().numpy()
        Mm = 0.9 * Mm + 0.1 * G
        V = 0.999 * V + 0.001 * G * G
        P = P - 1e-2 * (Mm / (1 - 0.9 ** step)) / (np.sqrt(V / (1 - 0.999 ** step)) + 1e-8)
    np.testing.assert_allclose(p.double().cpu().numpy(), P, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(pb.double().cpu().numpy(), P, rtol=2.0 ** -8, atol=1e-6)


@pytest.mark.parametrize("M", [1, 7, 1000, 65536 + 3])
def test_column_sums_match_torch(M):
    """gx_mlp_colsum (bias gradients): column sums of a BF16 matrix with a
    row stride, accumulated onto the output, against float64 sums; rows
    not a multiple of the kernel's chunk or of its four-row unroll."""
    torch = _torch()
    from galax_amd import _lib
    import galax_amd.surrogate as S

    g = torch.Generator().manual_seed(M)
    N, ld = 67, 72
    A = torch.randn(M, ld, generator=g).to(torch.bfloat16)
    out = torch.full((N,), 0.25, dtype=torch.float32, device="cuda")
    Ad = A.cuda()
    _lib.check(_lib.lib().gx_mlp_colsum(S._p(Ad), ld, M, N, S._p(out), S._s()), "gx_mlp_colsum")
    want = 0.25 + A[:, :N].double().sum(0)
    got = out.cpu().double()
    assert torch.allclose(got, want, rtol=1e-5, atol=1e-5 * (M ** 0.5)), (got - want).abs().max()


def test_mlp_gradients_match_autograd():
    """Gradients of the pseudo-Huber loss from the library's backward pass
    against torch autograd in float32 on the same BF16 weights; activations
    and cotangents are BF16 in the library, so the match is to BF16 level."""
    import galax_amd.surrogate as S

    torch = _torch()
    n = 2048
    mlp = S.MLP(hidden=64, depth=3, seed=1)
    g = torch.Generator().manual_seed(2)
    X = torch.zeros(n, S.K_IN)
    X[:, :5] = torch.randn(n, 5, generator=g)
    Xb = _bf(X).cuda()
    T = torch.randn(n, 3, generator=g).float().cuda() * 0.5
    Y = torch.empty(n, 3, device="cuda")
    dY = torch.empty(n, S.K_IN, dtype=torch.bfloat16, device="cuda")
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    mlp.forward(Xb, n, Y, train=True)
    from galax_amd import _lib

    _lib.check(_lib.lib().gx_mlp_huber(S._p(Y), 3, S._p(T), n, 1.0, 1.0 / n, S._p(dY), S.K_IN, S._p(loss), S._s()),
               "huber")
    mlp.backward(Xb, n, dY)
    # torch reference
    Ws = [w.float().clone().requires_grad_(True) for w in mlp.Wb]
    bs = [b.clone().requires_grad_(True) for b in mlp.b]
    h = Xb.float()
    for l in range(mlp.depth):
        h = torch.nn.functional.silu(h @ Ws[l].t() + bs[l])
    y = h @ Ws[-1].t() + bs[-1]
    np.testing.assert_allclose(Y.cpu().numpy(), y.detach().cpu().numpy(), rtol=3e-2, atol=3e-2 * float(y.abs().max()))
    r = y - T
    L = (torch.sqrt(1 + r * r) - 1).sum() / n
    np.testing.assert_allclose(float(loss.item()) / n, float(L), rtol=2e-2)
    L.backward()
    for got, p in zip(mlp.g, Ws + bs):
        ref = p.grad
        rel = float((got - ref).norm() / ref.norm().clamp_min(1e-30))
        assert rel < 3e-2, rel


def _lockstep_setup(n=512):
    import galax_amd.surrogate as S
    import galax_amd.potential as gp

    torch = _torch()
    mw = gp.MilkyWayPotential()
    mlp = S.MLP(hidden=64, depth=2, seed=3)
    # a network whose output is exactly zero: the surrogate field is the base
    mlp.W[-1].zero_()
    mlp.Wb[-1].zero_()
    mlp.b[-1].zero_()
    field = S.SurrogateField(full=mw, base=mw, mlp=mlp, omega=0.04, scale=1e-3)
    w0 = workloads.surrogate_orbits(n, 11)
    return S, field, torch.as_tensor(w0, device="cuda")


def test_lockstep_matches_persistent_integrator_and_graph_replays():
    import galax_amd.dynamics as gd

    S, field, w0 = _lockstep_setup()
    solver = gd.OrbitSolver(gd.Dopri8(), stepsize_controller=gd.ConstantStepSize(), max_steps=None)
    print("[case] lockstep zero network vs persistent", flush=True)
    ref = gd.integrate_orbits(field.base, w0, 0.0, 60.0, solver=solver, dt0=0.5).out[:, -1].cpu().numpy()
    w_graph, info = S.integrate_lockstep(field, w0, 0.0, 60.0, 0.5, check_every=7, use_graph=True)
    w_eager, _ = S.integrate_lockstep(field, w0, 0.0, 60.0, 0.5, check_every=7, use_graph=False)
    assert info["flagged"] == 0 and abs(info["t_end"] - 60.0) < 1e-12
    # graph replays (with eager checks between them) reproduce the eager run
    assert np.array_equal(w_graph.cpu().numpy(), w_eager.cpu().numpy())
    got = w_graph.cpu().numpy()
    rel = np.abs(got - ref) / np.abs(ref).max(axis=1, keepdims=True)
    assert rel.max() < 1e-12, rel.max()


def test_lockstep_exact_gate_matches_full_field():
    """With every orbit flagged, the lock-step path is the exact field of the
    full table (bar + rotating subhaloes) and matches the persistent
    integrator on it."""
    import galax_amd.dynamics as gd
    import galax_amd.surrogate as S

    torch = _torch()
    full, base, omega = workloads.barred_mw_with_subhaloes(seed=4, n_sub=8)
    mlp = S.MLP(hidden=64, depth=2, seed=5)
    field = S.SurrogateField(full=full, base=base, mlp=mlp, omega=omega, scale=1e-3)
    w0 = torch.as_tensor(workloads.surrogate_orbits(256, 12), device="cuda")
    solver = gd.OrbitSolver(gd.Dopri8(), stepsize_controller=gd.ConstantStepSize(), max_steps=None)
    print("[case] lockstep exact gate", flush=True)
    ref = gd.integrate_orbits(full, w0, 0.0, 40.0, solver=solver, dt0=0.5).out[:, -1].cpu().numpy()
    got, info = S.integrate_lockstep(field, w0, 0.0, 40.0, 0.5, exact=True)
    got = got.cpu().numpy()
    rel = np.abs(got - ref) / np.abs(ref).max(axis=1, keepdims=True)
    assert rel.max() < 1e-11, rel.max()


def test_surrogate_learns_the_perturbation():
    """A short training run on a small sample learns part of the perturbing
    acceleration: the median error of the total field drops below that of
    ignoring the perturbers (64 compact subhaloes are beyond what a small
    coordinate MLP resolves, so the gain is partial; the gate covers the
    rest)."""
    import galax_amd.surrogate as S

    torch = _torch()
    full, base, omega = workloads.barred_mw_with_subhaloes(seed=4)
    print("[case] surrogate training", flush=True)
    field, info = S.train_surrogate(full, base, omega, workloads.surrogate_orbits(4096, 4), epochs=6, batch=8192,
                                    hidden=128, depth=3, lr=3e-3)
    assert info["loss_history"][-1] < info["loss_history"][0]
    w = workloads.surrogate_orbits(4096, 5)
    xyz = torch.as_tensor(w[:, :3], device="cuda").contiguous()
    t = torch.as_tensor(np.random.default_rng(0).uniform(0, 1000, len(w)), device="cuda")
    exact = S._field(full, "acc", xyz, t)
    sur = field.acceleration(xyz, t)
    mwa = S._field(base, "acc", xyz, t)
    e_sur = ((sur - exact).norm(dim=1) / exact.norm(dim=1)).median().item()
    e_mw = ((mwa - exact).norm(dim=1) / exact.norm(dim=1)).median().item()
    print(f"median rel error: surrogate {e_sur:.3e}, ignoring perturbers {e_mw:.3e}", flush=True)
    assert e_sur < 0.8 * e_mw


def test_surrogate_configuration_against_the_adaptive_reference():
    """The learned-surrogate configuration at reduced size (16,384 training
    orbits, 2,048 new orbits, 1 Gyr, gate on): trained net, lock-step
    integration, the exact field at the same fixed step, and the reference
    solution --
    the exact field with adaptive Dopri8 at rtol 1e-8, whose Jacobi energy
    holds to round-off.  Reports median and max position error and Jacobi
    drift; the trained network must beat the Milky Way model alone (the
    method without it) in median position error after 1 Gyr."""
    import galax_amd.surrogate as S

    print("[case] surrogate configuration vs adaptive reference", flush=True)
    rep = S.barred_halo_experiment(n_train=16384, n_bench=2048, t1=1000.0, epochs=20)
    e = rep["pos_rel_err_vs_reference"]
    print({k: rep[k] for k in ("reference_ms", "exact_fixed_ms", "surrogate_ms", "surrogate_speedup_vs_exact_fixed_step",
                               "flagged_orbits")}, flush=True)
    print("position error", e, "\njacobi drift", rep["jacobi_drift_median_max"], flush=True)
    assert np.isfinite(e["surrogate"]["max"]) and np.isfinite(e["exact_fixed_step"]["max"])
    assert rep["jacobi_drift_median_max"]["reference"][0] < 1e-9
    assert e["surrogate"]["median"] < e["milky_way_only"]["median"]
