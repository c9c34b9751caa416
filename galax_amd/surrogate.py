"""Learned acceleration surrogate: a BF16 MLP on the matrix cores standing in
for the perturbing part of a field, with an exact FP64 fallback gate.

The reference evaluates every field exactly (potential/_src/base.py:170-179).
This module is an extension for fields whose exact evaluation is expensive:
here a rotating bar plus 64 co-rotating Plummer subhaloes added to the Milky
Way model.  Those perturbers rotate rigidly, so their field is static in the
rotating frame; the network maps rotating-frame position features to the
rotating-frame perturbing acceleration, and the Milky Way part stays exact:

    a(x, t) = a_MW(x) + s Rz(Omega t) MLP((x', y', z, R, r) / L),
    x' = Rz(-Omega t) x.

Everything runs in libgalaxhip.so (gx_mlp_*, gx_sur_*, gx_lockstep_*):
training (forward, backward, pseudo-Huber loss, Adam), inference, and a
lock-step fixed-step Dopri8 integrator whose 13 stages per step are captured
once in a HIP graph and replayed.  The gate integrates with the exact field
(i) outside the trained ball |x| > r_max and (ii) for orbits whose surrogate
acceleration, checked against the exact field every `check_every` steps,
is off by more than `gate_tol` (sticky per orbit).
"""

from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass

import numpy as np

from . import _abi, _lib

__all__ = ["MLP", "SurrogateField", "train_surrogate", "integrate_lockstep", "method_nodes"]

K_IN = 8  # padded input width: 5 features + 3 zeros (16-byte rows)


def _L():
    return _lib.lib()


def _s():
    return _lib.stream_ptr()


def _p(x):
    return _lib.ptr(x)


def method_nodes(method: int = _abi.GX_DOPRI8) -> np.ndarray:
    c = (ctypes.c_double * 16)()
    n = _L().gx_method_nodes(method, c, 16)
    if n < 0:
        raise _lib.GalaxHipError("gx_method_nodes: unknown method")
    return np.array(c[:n])


def gemm(A, sam, sak, B, sbk, sbn, M, N, K, epi, *, bias=None, C=None, ldc=0, H=None, ldh=0, Z=None, ldz=0,
         accumulate=0, ksplit=1):
    """C/H = epilogue(A(m,k) B(k,n)); see gx_mlp_gemm."""
    rc = _L().gx_mlp_gemm(_p(A), sam, sak, _p(B), sbk, sbn, M, N, K, epi, accumulate, ksplit, _p(bias), _p(C),
                          ldc, _p(H), ldh, _p(Z), ldz, _s())
    _lib.check(rc, "gx_mlp_gemm")


def _ksplit(M, N, K):
    tiles = ((M + 63) // 64) * ((N + 63) // 64)
    return int(max(1, min(1024 // tiles, K // 256, 65535)))


class MLP:
    """5 -> hidden^depth -> 3 with SiLU; FP32 master weights W[l] (out, in),
    BF16 copies for the GEMMs, Adam moments."""

    def __init__(self, hidden: int = 256, depth: int = 4, n_in: int = 5, seed: int = 0):
        torch = _lib.require_gpu()
        self.hidden, self.depth, self.n_in = hidden, depth, n_in
        dims = [K_IN] + [hidden] * depth + [3]
        self.dims = dims
        gen = torch.Generator().manual_seed(seed)
        self.W, self.b, self.Wb = [], [], []
        for l in range(len(dims) - 1):
            fan_in = n_in if l == 0 else dims[l]
            w = torch.randn(dims[l + 1], dims[l], generator=gen, dtype=torch.float32) * math.sqrt(2.0 / fan_in)
            if l == len(dims) - 2:
                w *= 0.1
            if l == 0:
                w[:, n_in:] = 0.0
            w = w.cuda().contiguous()
            self.W.append(w)
            self.b.append(torch.zeros(dims[l + 1], dtype=torch.float32, device="cuda"))
            self.Wb.append(w.to(torch.bfloat16).contiguous())
        self.params = self.W + self.b
        self.m = [torch.zeros_like(p) for p in self.params]
        self.v = [torch.zeros_like(p) for p in self.params]
        self.g = [torch.zeros_like(p) for p in self.params]
        self.step = 0
        self._bufs = {}

    # ------------------------------------------------------------ buffers
    def _buf(self, key, shape, dtype):
        torch = _lib.torch()
        b = self._bufs.get(key)
        if b is None or b.shape != shape or b.dtype != dtype:
            b = torch.empty(shape, dtype=dtype, device="cuda")
            self._bufs[key] = b
        return b

    def n_params(self) -> int:
        return sum(int(p.numel()) for p in self.params)

    def flops_per_eval(self) -> int:
        """Multiply-adds x 2 of one evaluation at the true widths."""
        d = [self.n_in] + [self.hidden] * self.depth + [3]
        return 2 * sum(d[i] * d[i + 1] for i in range(len(d) - 1))

    # ------------------------------------------------------------ forward
    def forward(self, X, n, Y, *, train=False):
        """Y[n][3] (fp32) = MLP(X[n][K_IN] bf16).  With train=True keeps the
        pre-activations Z[l] and activations H[l] for backward."""
        torch = _lib.torch()
        h = self.hidden
        prev, kin = X, K_IN
        for l in range(self.depth):
            if train:
                Z = self._buf(("Z", l), (n, h), torch.float32)
                H = self._buf(("H", l), (n, h), torch.bfloat16)
            else:
                Z = None
                H = self._buf(("I", l % 2), (n, h), torch.bfloat16)
            gemm(prev, kin, 1, self.Wb[l], 1, kin, n, h, kin, _abi.GX_EPI_BIAS_SILU, bias=self.b[l], C=Z, ldc=h,
                 H=H, ldh=h)
            prev, kin = H, h
        gemm(prev, h, 1, self.Wb[-1], 1, h, n, 3, h, _abi.GX_EPI_BIAS, bias=self.b[-1], C=Y, ldc=3)
        return Y

    def forward_fused(self, X, n, Y):
        """Y[n][3] = MLP(X[n][K_IN]) in one kernel (gx_mlp_forward_fused:
        activations stay on chip); hidden width 256 only."""
        if self.hidden != 256:
            raise ValueError("the fused inference kernel is built for hidden width 256")
        W = (ctypes.c_void_p * (self.depth + 1))(*[int(w.data_ptr()) for w in self.Wb])
        b = (ctypes.c_void_p * (self.depth + 1))(*[int(x.data_ptr()) for x in self.b])
        rc = _L().gx_mlp_forward_fused(_p(X), K_IN, K_IN, n, self.depth, W, b, 3, _p(Y), 3, _s())
        _lib.check(rc, "gx_mlp_forward_fused")
        return Y

    def infer(self, X, n, Y):
        """Inference: the fused kernel for hidden width 256, else layer by
        layer."""
        return self.forward_fused(X, n, Y) if self.hidden == 256 else self.forward(X, n, Y)

    # ------------------------------------------------------------ training
    def backward(self, X, n, dY):
        """Gradients g (W then b) of the loss whose output cotangent is
        dY[n][K_IN] (bf16, first 3 columns used) after forward(train=True)."""
        torch = _lib.torch()
        h, L = self.hidden, self.depth
        for g in self.g:
            g.zero_()
        gW, gb = self.g[: L + 1], self.g[L + 1 :]
        # output layer
        Hl = self._bufs[("H", L - 1)]
        gemm(dY, 1, K_IN, Hl, h, 1, 3, h, n, _abi.GX_EPI_STORE, C=gW[L], ldc=h, ksplit=_ksplit(3, h, n))
        _lib.check(_L().gx_mlp_colsum(_p(dY), K_IN, n, 3, _p(gb[L]), _s()), "gx_mlp_colsum")
        dZ = self._buf(("dZ", 0), (n, h), torch.bfloat16)
        gemm(dY, K_IN, 1, self.Wb[L], h, 1, n, h, 3, _abi.GX_EPI_DSILU, H=dZ, ldh=h, Z=self._bufs[("Z", L - 1)],
             ldz=h)
        for l in range(L - 1, -1, -1):
            Hin, kin = (self._bufs[("H", l - 1)], h) if l > 0 else (X, K_IN)
            gemm(dZ, 1, h, Hin, kin, 1, h, kin, n, _abi.GX_EPI_STORE, C=gW[l], ldc=kin, ksplit=_ksplit(h, kin, n))
            _lib.check(_L().gx_mlp_colsum(_p(dZ), h, n, h, _p(gb[l]), _s()), "gx_mlp_colsum")
            if l > 0:
                dZn = self._buf(("dZ", (L - l) % 2), (n, h), torch.bfloat16)
                gemm(dZ, h, 1, self.Wb[l], kin, 1, n, h, h, _abi.GX_EPI_DSILU, H=dZn, ldh=h,
                     Z=self._bufs[("Z", l - 1)], ldz=h)
                dZ = dZn

    def adam(self, lr, b1=0.9, b2=0.999, eps=1e-8):
        self.step += 1
        nW = len(self.W)
        for i, (p, g, m, v) in enumerate(zip(self.params, self.g, self.m, self.v)):
            pb = self.Wb[i] if i < nW else None
            rc = _L().gx_mlp_adam(_p(p), _p(g), _p(m), _p(v), _p(pb), p.numel(), lr, b1, b2, eps, self.step, _s())
            _lib.check(rc, "gx_mlp_adam")

    def train_step(self, X, T, n, lr, delta=1.0, loss=None):
        """One Adam step on the pseudo-Huber loss of MLP(X) against T[n][3];
        accumulates the summed loss into `loss` (a device double)."""
        torch = _lib.torch()
        Y = self._buf("Y", (n, 3), torch.float32)
        dY = self._buf("dY", (n, K_IN), torch.bfloat16)
        self.forward(X, n, Y, train=True)
        rc = _L().gx_mlp_huber(_p(Y), 3, _p(T), n, delta, 1.0 / n, _p(dY), K_IN, _p(loss), _s())
        _lib.check(rc, "gx_mlp_huber")
        self.backward(X, n, dY)
        self.adam(lr)


# ------------------------------------------------------------ the field
@dataclass
class SurrogateField:
    """The surrogate field: exact `base` + network perturbation, gated to the
    exact `full` field.  `scale` converts network outputs to kpc/Myr^2,
    `length` normalises positions, `omega` is the perturbers' pattern speed."""

    full: object
    base: object
    mlp: MLP
    omega: float
    scale: float
    length: float = 50.0
    r_max: float = 45.0

    def features(self, xyz, t, t_stride, dt, n, X):
        rc = _L().gx_sur_inputs(_p(xyz), _p(t), t_stride, dt, n, 1.0 / self.length, self.omega, _p(X), K_IN, _s())
        _lib.check(rc, "gx_sur_inputs")

    def acceleration_into(self, xyz, t, t_stride, dt, n, acc, *, X, Y, flag=None, r_max=None):
        self.features(xyz, t, t_stride, dt, n, X)
        self.mlp.infer(X, n, Y)
        rc = _L().gx_sur_combine(self.base._handle(), self.full._handle(), _p(xyz), _p(t), t_stride, dt, n, _p(Y),
                                 3, self.scale, self.omega, self.r_max if r_max is None else r_max, _p(flag),
                                 _p(acc), _s())
        _lib.check(rc, "gx_sur_combine")

    def acceleration(self, xyz, t):
        """Surrogate acceleration at points xyz (n,3) and times t (n,) or scalar."""
        torch = _lib.require_gpu()
        xyz = torch.as_tensor(xyz, dtype=torch.float64, device="cuda").reshape(-1, 3).contiguous()
        n = xyz.shape[0]
        t = torch.as_tensor(t, dtype=torch.float64, device="cuda").reshape(-1).contiguous()
        stride = 0 if t.numel() == 1 else 1
        X = torch.empty((n, K_IN), dtype=torch.bfloat16, device="cuda")
        Y = torch.empty((n, 3), dtype=torch.float32, device="cuda")
        acc = torch.empty((n, 3), dtype=torch.float64, device="cuda")
        self.acceleration_into(xyz, t, stride, 0.0, n, acc, X=X, Y=Y)
        return acc


def _field(pot, what, xyz, t):
    """gx_acceleration / gx_potential_eval at points with per-point times."""
    torch = _lib.torch()
    n = xyz.shape[0]
    out = torch.empty((n, 3) if what == "acc" else (n,), dtype=torch.float64, device="cuda")
    fn = _L().gx_acceleration if what == "acc" else _L().gx_potential_eval
    stride = 0 if t.numel() == 1 else 1
    _lib.check(fn(pot._handle(), _p(xyz), _p(t), stride, n, _p(out), _s()), "field")
    return out


def training_set(full, base, omega, orbits, *, t1=1000.0, nsave=64, dt=1.0, scale=None):
    """Samples (x, t) along exact orbits of `full` (fixed-step Dopri8, nsave
    saves each) with features X and targets T = Rz(-Omega t)(a_full - a_base)
    / scale.  Returns (X bf16 (N, K_IN), T fp32 (N, 3), scale, xyz, t)."""
    torch = _lib.require_gpu()
    from .dynamics import ConstantStepSize, Dopri8, OrbitSolver, integrate_orbits

    solver = OrbitSolver(Dopri8(), stepsize_controller=ConstantStepSize(), max_steps=None)
    ts = np.linspace(0.0, t1, nsave)
    res = integrate_orbits(full, orbits, 0.0, t1, ts=ts, solver=solver, dt0=dt)
    if int(res.status.abs().sum()) != 0:
        raise RuntimeError("training orbits failed")
    xyz = res.out[..., :3].reshape(-1, 3).contiguous()
    N = xyz.shape[0]
    tt = torch.as_tensor(ts, device="cuda").repeat(orbits.shape[0]).contiguous()
    a_full = _field(full, "acc", xyz, tt)
    a_base = _field(base, "acc", xyz, tt)
    T64 = torch.empty((N, 3), dtype=torch.float32, device="cuda")
    if scale is None:
        d = (a_full - a_base).abs().flatten()
        scale = float(torch.quantile(d[torch.randperm(d.numel(), device="cuda")[:1_000_000]], 0.99))
    rc = _L().gx_sur_targets(_p(tt), 1, N, _p(a_full), _p(a_base), omega, 1.0 / scale, _p(T64), _s())
    _lib.check(rc, "gx_sur_targets")
    return xyz, tt, T64, scale


def train_surrogate(full, base, omega, orbits, *, epochs=8, batch=65536, lr=2e-3, seed=0, hidden=256, depth=4,
                    length=50.0, r_max=45.0, log=None):
    """Builds the training set from exact orbits and trains the MLP with the
    library's kernels (Adam, cosine learning-rate decay).  Returns
    (SurrogateField, info dict)."""
    torch = _lib.require_gpu()
    t_start = time.perf_counter()
    xyz, tt, T, scale = training_set(full, base, omega, orbits)
    N = xyz.shape[0]
    mlp = MLP(hidden=hidden, depth=depth, seed=seed)
    field = SurrogateField(full=full, base=base, mlp=mlp, omega=omega, scale=scale, length=length, r_max=r_max)
    X = torch.empty((N, K_IN), dtype=torch.bfloat16, device="cuda")
    field.features(xyz, tt, 1, 0.0, N, X)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t_start
    gen = torch.Generator(device="cuda").manual_seed(seed)
    steps_per_epoch = N // batch
    total = epochs * steps_per_epoch
    loss = torch.zeros(1, dtype=torch.float64, device="cuda")
    history = []
    t0 = time.perf_counter()
    k = 0
    for ep in range(epochs):
        perm = torch.randperm(N, device="cuda", generator=gen)
        loss.zero_()
        for s in range(steps_per_epoch):
            idx = perm[s * batch : (s + 1) * batch]
            Xb = X.index_select(0, idx)
            Tb = T.index_select(0, idx)
            lr_k = lr * 0.5 * (1.0 + math.cos(math.pi * k / total))
            mlp.train_step(Xb, Tb, batch, lr_k, loss=loss)
            k += 1
        ep_loss = float(loss.item()) / (steps_per_epoch * batch)
        history.append(ep_loss)
        if log:
            log(f"[surrogate] epoch {ep + 1}/{epochs} loss {ep_loss:.4e}")
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    info = {
        "samples": int(N),
        "scale": scale,
        "epochs": epochs,
        "batch": batch,
        "steps": k,
        "loss_history": history,
        "data_seconds": t_data,
        "train_seconds": t_train,
        "train_flop": 3 * mlp.flops_per_eval() * N * epochs,
    }
    return field, info


# ------------------------------------------------------------ integration
def integrate_lockstep(field: SurrogateField, w0, t0: float, t1: float, dt: float, *, check_every: int = 50,
                       gate_tol: float = 1e-2, use_graph: bool = True, exact: bool = False):
    """Fixed-step Dopri8 of n orbits with the surrogate field, all orbits in
    lock-step (shared time).  One step = 13 stages x (stage points, features,
    MLP, combine) + update, captured in a HIP graph and replayed.  With
    exact=True every orbit is flagged from the start (the gate's exact path
    everywhere).  Returns (w (n,6) at t1, info)."""
    torch = _lib.require_gpu()
    w0 = torch.as_tensor(w0, dtype=torch.float64, device="cuda").reshape(-1, 6)
    n = int(w0.shape[0])
    nsteps = int(round((t1 - t0) / dt))
    if nsteps < 1 or abs(nsteps * dt - (t1 - t0)) > 1e-9 * max(1.0, abs(t1 - t0)):
        raise ValueError("t1 - t0 must be a positive whole number of steps")
    c = method_nodes(_abi.GX_DOPRI8)
    q = w0[:, :3].contiguous()
    p = w0[:, 3:].contiguous()
    t = torch.tensor([t0], dtype=torch.float64, device="cuda")
    G = torch.empty((13, n, 3), dtype=torch.float64, device="cuda")
    Qs = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    X = torch.empty((n, K_IN), dtype=torch.bfloat16, device="cuda")
    Y = torch.empty((n, 3), dtype=torch.float32, device="cuda")
    flag = torch.full((n,), 1 if exact else 0, dtype=torch.int32, device="cuda")
    a_chk = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    err = torch.empty((n,), dtype=torch.float32, device="cuda")
    lib = _L()

    def one_step():
        for s in range(13):
            _lib.check(lib.gx_lockstep_stage(s, _p(q), _p(p), _p(G), n, dt, _p(Qs), _s()), "gx_lockstep_stage")
            field.acceleration_into(Qs, t, 0, float(c[s] * dt), n, G[s], X=X, Y=Y, flag=flag)
        _lib.check(lib.gx_lockstep_update(_p(q), _p(p), _p(G), n, dt, _p(t), _s()), "gx_lockstep_update")

    def check():
        field.acceleration_into(q, t, 0, 0.0, n, a_chk, X=X, Y=Y, flag=None, r_max=float("inf"))
        rc = lib.gx_sur_check(field.full._handle(), _p(q), _p(t), 0, n, _p(a_chk), gate_tol, _p(flag), _p(err), _s())
        _lib.check(rc, "gx_sur_check")

    # warm the buffers (the MLP allocates its activations on first use)
    graph = None
    done = 0
    if not exact:
        check()
    if use_graph and nsteps > 1:
        one_step()
        done = 1
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            one_step()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    while done < nsteps:
        if graph is not None:
            graph.replay()
        else:
            one_step()
        done += 1
        if not exact and check_every and done % check_every == 0 and done < nsteps:
            check()
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1)
    steps_timed = nsteps - (1 if graph is not None else 0)
    info = {
        "steps": nsteps,
        "ms_timed": ms,
        "steps_timed": steps_timed,
        "flagged": int(flag.sum().item()) if not exact else n,
        "t_end": float(t.item()),
    }
    return torch.cat([q, p], dim=1), info


def jacobi_energy(pot, w, t: float, omega: float):
    """E_J = |v|^2 / 2 + Phi(x, t) - Omega L_z, conserved in a field that
    rotates rigidly at Omega about z."""
    torch = _lib.torch()
    xyz = w[:, :3].contiguous()
    phi = _field(pot, "phi", xyz, torch.tensor([t], dtype=torch.float64, device="cuda"))
    v = w[:, 3:]
    lz = w[:, 0] * v[:, 1] - w[:, 1] * v[:, 0]
    return 0.5 * (v * v).sum(1) + phi - omega * lz


def _position_error(w, ref):
    return ((w[:, :3] - ref[:, :3]).norm(dim=1) / ref[:, :3].norm(dim=1)).cpu().numpy()


def barred_halo_experiment(n_train: int = 262_144, n_bench: int = 262_144, t1: float = 1000.0, dt: float = 1.0,
                           epochs: int = 8, ref_rtol: float = 1e-8, use_graph: bool = True, log=None):
    """Trains the surrogate on exact orbits (seed 4) and integrates new
    orbits (seed 5) over t1 with (a) the surrogate (lock-step fixed step dt,
    gate on), (b) the exact FP64 field at the same fixed step, and (c) the
    reference solution: the exact field with adaptive Dopri8 at rtol
    ref_rtol, whose Jacobi-energy drift sits at round-off.  Position errors of (a) and (b)
    are measured against (c); so is the Milky Way model alone (bar and
    subhaloes ignored), which is what the network is meant to improve on.
    Times are HIP events around each whole integration.  use_graph=False
    launches the lock-step kernels one by one instead of replaying a
    captured step (for tools that do not follow graph replays)."""
    torch = _lib.require_gpu()
    from . import workloads
    from .dynamics import ConstantStepSize, Dopri8, OrbitSolver, PIDController, integrate_orbits

    full, base, omega = workloads.barred_mw_with_subhaloes(seed=4)
    field, tinfo = train_surrogate(full, base, omega, workloads.surrogate_orbits(n_train, 4), epochs=epochs,
                                   log=log)
    w0 = torch.as_tensor(workloads.surrogate_orbits(n_bench, 5), device="cuda")
    fixed = OrbitSolver(Dopri8(), stepsize_controller=ConstantStepSize(), max_steps=None)
    ref_solver = OrbitSolver(Dopri8(), stepsize_controller=PIDController(rtol=ref_rtol, atol=1e-10),
                             max_steps=None)

    def timed(fn):
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        r = fn()
        e1.record()
        torch.cuda.synchronize()
        return r, e0.elapsed_time(e1)

    integrate_orbits(full, w0[:1024], 0.0, 10.0, solver=fixed, dt0=dt)  # warm
    res_fixed, ms_exact = timed(lambda: integrate_orbits(full, w0, 0.0, t1, solver=fixed, dt0=dt))
    res_ref, ms_ref = timed(lambda: integrate_orbits(full, w0, 0.0, t1, solver=ref_solver))
    res_mw, _ = timed(lambda: integrate_orbits(base, w0, 0.0, t1, solver=ref_solver))
    (w_sur, sinfo), ms_sur_call = timed(lambda: integrate_lockstep(field, w0, 0.0, t1, dt, use_graph=use_graph))
    ms_sur = sinfo["ms_timed"] * sinfo["steps"] / max(sinfo["steps_timed"], 1)
    ref = res_ref.out[:, -1, :]
    w_exact = res_fixed.out[:, -1, :]
    e_sur = _position_error(w_sur, ref)
    e_fixed = _position_error(w_exact, ref)
    e_mw = _position_error(res_mw.out[:, -1, :], ref)
    ej0 = jacobi_energy(full, w0, 0.0, omega)

    def drift(w):
        d = ((jacobi_energy(full, w, t1, omega) - ej0) / ej0.abs()).abs()
        return float(d.median()), float(d.max())

    steps = n_bench * sinfo["steps"]

    def stats(e):
        return {"median": float(np.median(e)), "p99": float(np.quantile(e, 0.99)), "max": float(e.max())}

    return {
        "orbits": n_bench,
        "t1_myr": t1,
        "steps_per_orbit": sinfo["steps"],
        "training": {k: v for k, v in tinfo.items() if k != "loss_history"} | {"final_loss": tinfo["loss_history"][-1]},
        "reference": f"exact field, adaptive Dopri8 rtol {ref_rtol:g} atol 1e-10 (independent orbits)",
        "reference_ms": ms_ref,
        "reference_attempted_steps": int(res_ref.nsteps.sum()),
        "exact_fixed_ms": ms_exact,
        "surrogate_ms": ms_sur,
        "surrogate_ms_basis": "lock-step graph replays timed with HIP events; the first step (graph capture) "
                              "is scaled in at the mean replay time" if use_graph else
                              "lock-step launches timed with HIP events (no graph)",
        "surrogate_ms_whole_call": ms_sur_call,
        "exact_fixed_orbit_steps_per_s": steps / (ms_exact * 1e-3),
        "surrogate_orbit_steps_per_s": steps / (ms_sur * 1e-3),
        "surrogate_speedup_vs_exact_fixed_step": ms_exact / ms_sur,
        "surrogate_speedup_vs_reference": ms_ref / ms_sur,
        "mlp_tflops": 13 * steps * field.mlp.flops_per_eval() / (ms_sur * 1e-3) / 1e12,
        "flagged_orbits": sinfo["flagged"],
        "pos_rel_err_vs_reference": {"surrogate": stats(e_sur), "exact_fixed_step": stats(e_fixed),
                                 "milky_way_only": stats(e_mw)},
        "jacobi_drift_median_max": {"reference": drift(ref), "exact_fixed_step": drift(w_exact),
                                    "surrogate": drift(w_sur)},
    }
