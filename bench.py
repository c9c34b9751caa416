#!/usr/bin/env python
"""Benchmark: orbit RK-steps/s for 1,048,576 Milky Way orbits in FP64.

Workload (one "step" = one pass of the hot path over the whole batch):
1,048,576 orbits in MilkyWayPotential, R ~ U(0.5, 50) kpc, 25% velocity
dispersion, per-orbit integration time log-uniform in 0.1..5 Gyr, adaptive
Dopri8 (rtol 1e-8, atol 1e-10), 64 save times per orbit, one persistent
launch with a longest-first work queue per pass.  Passes are independent
batches; several are kept in flight on separate streams (each launch taking
a fraction of the device), so that one batch's bulk fills the tail of the
previous one, which waits on its single longest orbit.  Inputs are resident
in HBM before the timed region.  Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

# Hardware queues of this process (HIP reads the variable when it starts):
# the batches in flight each need their own queue to run side by side, so
# the process asks for sixteen when its environment offers fewer.  With
# eight, eight batches in flight shared queues and a pass took 30 ms; with
# sixteen, 8 in flight ran at 18.7-18.9 ms against 19.2-19.4 for 6 in
# flight on eight queues.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or "4") < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))


def host_cpu() -> dict:
    """The host's CPU model (as lscpu reports it, from /proc/cpuinfo), the
    cores this process may run on, and the cores a one-GPU job is allotted
    (OMP_NUM_THREADS where the job sets it)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.lower().startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    visible = len(os.sched_getaffinity(0))
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return {"model": model, "cores_visible": visible, "cores_allotted": min(visible, share) if share > 0 else visible}


def cpu_baseline(pot, solver, w0, T, nsave, out_gpu, nsteps_gpu, seconds: float, threads: int, seed: int = 12345):
    """The oracle on a bounded sample of the same workload (a seeded random
    subset of the orbits, integrated in chunks until `seconds` elapse); also
    compares those orbits with the GPU's output."""
    import numpy as np

    import oracle

    rng = np.random.default_rng(seed)
    idx = rng.permutation(w0.shape[0])
    steps = 0
    done = 0
    errs = []
    same = []
    t_start = time.perf_counter()
    chunk = 4 * threads
    while time.perf_counter() - t_start < seconds and done < w0.shape[0]:
        sel = idx[done : done + chunk]
        out, status, nsteps, _ = oracle.integrate(pot, solver, w0[sel], 0.0, T[sel], T=nsave, nthreads=threads)
        steps += int(nsteps.sum())
        done += len(sel)
        g = out_gpu[sel]
        rel = np.linalg.norm(g - out, axis=-1) / np.linalg.norm(out, axis=-1)
        errs.append(np.nanmax(rel, axis=-1))
        same.append(nsteps == nsteps_gpu[sel])
    elapsed = time.perf_counter() - t_start
    e = np.concatenate(errs)
    parity = {
        "orbits": int(e.size),
        "same_step_count": float(np.mean(np.concatenate(same))),
        "median_rel_err": float(np.median(e)),
        "frac_rel_err_below_1e-9": float(np.mean(e < 1e-9)),
        "frac_rel_err_below_1e-6": float(np.mean(e < 1e-6)),
        "p99_rel_err": float(np.quantile(e, 0.99)),
        "max_rel_err": float(e.max()),
        "note": "per-orbit max over saves of |w_gpu - w_oracle| / |w_oracle|; orbits whose accept/reject "
        "sequence flips on round-off separate at the tolerance level and chaotic ones then grow apart",
    }
    return steps / elapsed, done, elapsed, parity


def stream_report(steps: int) -> dict:
    """Mock-stream release integration: progenitor with r_peri = 10 and
    r_apo = 25 kpc in the Milky Way model (planar, seeded phase), 50,000
    release times over 4 Gyr, both arms, the progenitor's Plummer sphere
    riding its dense solution; each particle-integration pass (one launch)
    is timed with HIP events and the median reported."""
    import numpy as np
    import torch

    import galax_amd.dynamics as gd
    import galax_amd.potential as gp

    torch.cuda.set_device(0)
    mw = gp.MilkyWayPotential()
    phi = lambda r: float(np.asarray(mw.potential(np.array([[r, 0.0, 0.0]]), 0.0)).reshape(-1)[0])
    r_apo, r_peri = 25.0, 10.0
    vt = np.sqrt(2.0 * (phi(r_apo) - phi(r_peri)) / ((r_apo / r_peri) ** 2 - 1.0))
    ang = np.random.Generator(np.random.PCG64(3)).uniform(0, 2 * np.pi)
    c, s_ = np.cos(ang), np.sin(ang)
    ts = np.linspace(-4000.0, 0.0, 50_000)
    w0 = gd.PhaseSpaceCoordinate(q=np.array([r_apo * c, r_apo * s_, 0.0]), p=np.array([-vt * s_, vt * c, 0.0]),
                                 t=0.0)
    solver = gd.OrbitSolver(gd.Dopri8(), stepsize_controller=gd.PIDController(rtol=1e-8, atol=1e-8),
                            max_steps=None)
    integ = gd.Integrator(solver)
    gen = gd.MockStreamGenerator(gd.FardalStreamDF(), mw, progenitor_integrator=integ, stream_integrator=integ,
                                 progenitor_potential=gp.PlummerPotential(m_tot=1e5, r_s=0.01))
    # the present-day state above is integrated back to the first release time
    back = gd.evaluate_orbit(mw, w0, np.array([0.0, ts[0]]), integrator=integ)
    w_start = gd.PhaseSpaceCoordinate(q=np.asarray(back.q).reshape(-1, 3)[-1], p=np.asarray(back.p).reshape(-1, 3)[-1],
                                      t=ts[0])
    # MockStreamGenerator.run end to end (progenitor orbit, release sampling,
    # stream field, particle integration, read-back): the first call in the
    # process also loads the kernels' code objects; later calls are what a
    # long-running caller sees
    runs = []
    for _ in range(1 + max(steps, 1)):
        torch.cuda.synchronize()
        t_run = time.perf_counter()
        stream, _ = gen.run(3, ts, w_start, 1e5)
        torch.cuda.synchronize()
        runs.append((time.perf_counter() - t_run) * 1e3)
    # the release integration alone, on device-resident ICs, in the field
    # run() builds (the progenitor's dense solution as the track)
    prog_o = gd.evaluate_orbit(mw, w_start, ts, integrator=integ, dense=True)
    prog_o.q = np.asarray(prog_o.q).reshape(-1, 3)
    prog_o.p = np.asarray(prog_o.p).reshape(-1, 3)
    ics = gd.FardalStreamDF().sample(3, mw, prog_o, 1e5)
    w_all = torch.as_tensor(np.concatenate([ics["lead"].w(), ics["trail"].w()]), device="cuda")
    t0 = torch.as_tensor(np.concatenate([ts, ts]), device="cuda")
    t1 = torch.full_like(t0, ts[-1] + 1e-3)
    field = gen.stream_potential(prog_o)
    from galax_amd.dynamics._mockstream import release_order

    order = torch.as_tensor(release_order(t0.cpu().numpy(), t1.cpu().numpy()), device="cuda")
    gd.integrate_orbits(field, w_all[:256], t0[:256], t1[:256], solver=solver)
    torch.cuda.synchronize()
    # each pass timed on its own: one particle's serial steps bound a pass,
    # and passes of the same work on one device spread by a quarter
    passes = []
    for _ in range(max(steps, 1)):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        res = gd.integrate_orbits(field, w_all, t0, t1, solver=solver, order=order, validate_order=False)
        e1.record()
        torch.cuda.synchronize()
        passes.append(e0.elapsed_time(e1))
    ms = float(np.median(passes))
    nst = res.nsteps.cpu().numpy()
    return {
        "report": "mock-stream release integration (100,000 particles, 4 Gyr of releases)",
        "particles": int(w_all.shape[0]),
        "failed": int((res.status.cpu().numpy() != 0).sum()),
        "attempted_steps": int(nst.sum()),
        "max_steps_one_particle": int(nst.max()),
        "ms_release_integration": ms,
        "ms_release_integration_passes": [round(x, 2) for x in passes],
        "progenitor_track_pieces": int(prog_o.interpolant.dense.nrec[0]),
        "orbit_steps_per_s": float(nst.sum()) / (ms * 1e-3),
        "ms_generator_run_end_to_end": float(np.mean(runs[1:])),
        "ms_generator_run_end_to_end_runs": [round(r, 2) for r in runs[1:]],
        "ms_generator_run_first_call": runs[0],
        "stream_failed_in_run": int((stream.status != 0).sum()),
    }


def c1_report(steps: int) -> dict:
    """The reference's entry point on the correctness configuration:
    gd.evaluate_orbit over 1,024 Milky Way orbits (R ~ U(4, 20) kpc),
    0-1000 Myr, 101 save times, (a) Dopri8 at a fixed 0.5 Myr step (2,000
    steps) and (b) the default Integrator (Dopri8, PID rtol = atol = 1e-7),
    which galax solves as one ODE over the batch (shared steps).  Inputs are
    resident on the device; timed with HIP events.  Beside it, the oracle on
    the same inputs (all orbits, 16 threads) as the CPU baseline and the
    checker."""
    import numpy as np
    import torch

    import galax_amd.dynamics as gd
    import galax_amd.potential as gp
    import oracle
    from galax_amd import workloads

    torch.cuda.set_device(0)
    pot = gp.MilkyWayPotential()
    w0, _, _, ts = workloads.config_plain_small()
    wd = torch.as_tensor(w0, device="cuda")
    fixed = gd.Integrator(gd.OrbitSolver(gd.Dopri8(), stepsize_controller=gd.ConstantStepSize(), max_steps=None),
                          diffeq_kw={"dt0": 0.5})
    rep = {"report": "gd.evaluate_orbit, 1,024 Milky Way orbits, 0-1000 Myr, 101 save times",
           "orbits": int(w0.shape[0]), "saves": int(ts.size)}
    for name, integ, shared in (("fixed_dt_0.5", fixed, False), ("default_pid_1e-7_shared_steps", gd.Integrator(), True)):
        o = gd.evaluate_orbit(pot, wd, ts, integrator=integ)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(max(steps, 1)):
            o = gd.evaluate_orbit(pot, wd, ts, integrator=integ)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / max(steps, 1)
        got = torch.cat([o.q, o.p], dim=-1).cpu().numpy()
        nst = int(np.asarray(o.nsteps.cpu().numpy() if hasattr(o.nsteps, "cpu") else o.nsteps).max())
        t_cpu = time.perf_counter()
        ref, st, ns, _ = oracle.integrate(pot, integ.dynamics_solver, w0, ts[0], ts[-1], ts=ts,
                                          dt0=integ.diffeq_kw.get("dt0"), shared=shared, nthreads=16)
        cpu_ms = (time.perf_counter() - t_cpu) * 1e3
        rel = np.linalg.norm(got - ref, axis=-1) / np.linalg.norm(ref, axis=-1)
        rep[name] = {
            "ms": ms,
            "attempted_steps_per_orbit": nst,
            "orbit_steps_per_s": w0.shape[0] * nst / (ms * 1e-3),
            "cpu_oracle_ms": cpu_ms,
            "cpu_oracle_threads": 16,
            "oracle_same_steps": bool(int(ns.max()) == nst),
            "max_rel_err_vs_oracle": float(rel.max()),
        }
    return rep


def streamfit_report(steps: int) -> dict:
    """Forward + backward of the stream-track loss, timed over `steps`
    repetitions, with the gradient checked against central differences
    (h = 1e-6 relative) of the GPU forward pass."""
    import numpy as np
    import torch

    from galax_amd.adjoint import MilkyWayParams
    from galax_amd.streamfit import make_problem, stream_loss

    torch.cuda.set_device(0)
    prob = make_problem(n_release=10_000, span=3000.0, seed=6, n_target=200, dt=1.0, K=64)
    theta0 = MilkyWayParams.default.copy()

    def fwd_bwd():
        th = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
        L = stream_loss(th, prob)
        L.backward()
        return float(L), th.grad.numpy().copy()

    fwd_bwd()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(max(steps, 1)):
        L, g = fwd_bwd()
    torch.cuda.synchronize()
    ms_grad = (time.perf_counter() - t) / max(steps, 1) * 1e3
    t = time.perf_counter()
    with torch.no_grad():
        stream_loss(torch.tensor(theta0), prob)
    torch.cuda.synchronize()
    ms_fwd = (time.perf_counter() - t) * 1e3
    fd = []
    with torch.no_grad():
        for j in range(6):
            h = 1e-6 * theta0[j]
            tp, tm = theta0.copy(), theta0.copy()
            tp[j] += h
            tm[j] -= h
            fd.append((float(stream_loss(torch.tensor(tp), prob)) - float(stream_loss(torch.tensor(tm), prob)))
                      / (2 * h))
    fd = np.array(fd)
    # the two most sensitive parameters at other step sizes: a min-distance
    # loss over 20,000 particles has kinks where a particle's nearest track
    # point changes, which central differences straddle now and then
    fd_h = {}
    with torch.no_grad():
        for rel_h in (1e-5, 1e-7):
            for j in (0, 1):
                h = rel_h * theta0[j]
                tp, tm = theta0.copy(), theta0.copy()
                tp[j] += h
                tm[j] -= h
                fd_h[f"{MilkyWayParams.names[j]}@{rel_h:g}"] = (
                    float(stream_loss(torch.tensor(tp), prob)) - float(stream_loss(torch.tensor(tm), prob))
                ) / (2 * h)
    return {
        "report": "stream-track loss gradient by discrete adjoint (checkpoint every 64 steps)",
        "particles": prob.n_particles,
        "release_span_myr": 3000.0,
        "dt_myr": 1.0,
        "loss": L,
        "grad": dict(zip(MilkyWayParams.names, g.tolist())),
        "grad_fd": dict(zip(MilkyWayParams.names, fd.tolist())),
        "max_rel_diff_vs_fd": float(np.max(np.abs(g - fd) / np.abs(fd))),
        "rel_diff_vs_fd": dict(zip(MilkyWayParams.names, (np.abs(g - fd) / np.abs(fd)).tolist())),
        "grad_fd_other_steps": fd_h,
        "ms_forward": ms_fwd,
        "ms_forward_backward": ms_grad,
        "ms_central_fd_6_params": 12 * ms_fwd,
    }


# Independent batches kept in flight when the launch policy is not tuned
# (each on its own stream, each launch taking a quarter of the device):
# the tuner's starting point, measured best with sixteen hardware queues.
IN_FLIGHT = 8


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=48,
                    help="timed passes (48: the pipeline fill and drain are a small share of the timed region)")
    ap.add_argument("--warmup", type=int, default=6)
    ap.add_argument("--n", type=int, default=1_048_576)
    ap.add_argument("--saves", type=int, default=64)
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="oracle time on all allotted cores (a third of it again on one core)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the launch-policy search (galax_amd.tune) and use the fixed defaults")
    ap.add_argument("--single-launch", type=int, default=3,
                    help="after the timed region, passes timed as one launch on the whole device (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="surrogate workload: launch the lock-step kernels one by one instead of replaying a "
                         "captured step (for profilers that do not follow graph replays)")
    ap.add_argument(
        "--workload",
        choices=("uneven", "uniform", "longest", "copy", "surrogate", "streamfit", "stream", "c1"),
        default="uneven",
        help="uneven: the benchmark configuration; uniform: the same orbits with every "
        "integration time set to 200 Myr (balanced work: chip throughput); longest: the "
        "single most expensive orbit of the configuration alone (per-step latency); copy: a "
        "plain 1 GiB device-to-device copy (calibrates HBM byte counters, no JSON line); surrogate: "
        "train the BF16 MLP surrogate of a barred Milky Way with 64 co-rotating subhaloes on 262,144 "
        "exact orbits, then integrate 262,144 new orbits over 1 Gyr with it and with the exact FP64 field "
        "(prints its own report line); streamfit: gradient of a stream-track loss (20,000 particles "
        "released over 3 Gyr, fixed-step Dopri8 h = 1 Myr) w.r.t. six Milky Way parameters by the "
        "discrete adjoint, timed and checked against central differences (prints its own report line); "
        "stream: a 100,000-particle mock stream (50,000 release times over 4 Gyr, Fardal release, "
        "adaptive Dopri8 rtol 1e-8, the progenitor's Plummer riding its track; prints its own report line); "
        "c1: gd.evaluate_orbit on 1,024 Milky Way orbits over 1 Gyr with 101 saves, fixed step 0.5 Myr and the "
        "default adaptive Integrator (shared steps), beside the oracle (prints its own report line)",
    )
    ap.add_argument("--dump-steps", default=None, help="save per-orbit attempted steps and queue order (.npz)")
    ap.add_argument("--blocks-per-cu", type=int, default=None,
                    help="resident workgroups per compute unit that one integrator launch takes "
                         "(gx_set_launch_shape; default: the library's)")
    ap.add_argument("--launch-parts", type=int, default=None,
                    help="each integrator launch takes 1/parts of the device (gx_set_launch_fraction; "
                         "default: the number of passes in flight, at most 4, for the throughput workload)")
    ap.add_argument("--in-flight", type=int, default=None,
                    help="passes (independent batches) in flight on that many streams (1: back to back; "
                    "default 6 for the throughput workload, 1 for the diagnostic ones)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            f"bench.py measures one GPU (got --gpus {args.gpus}); independent orbits shard "
            "across GPUs without communication, so run it once per GPU",
            file=sys.stderr,
        )
        return 2

    import numpy as np
    import torch

    import galax_amd.dynamics as gd
    import galax_amd.potential as gp
    from galax_amd import _lib, flops, workloads

    torch.cuda.set_device(0)
    if args.workload == "copy":
        nbytes = 1 << 30
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(max(args.steps, 1)):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] copied {nbytes} bytes {max(args.steps, 1)} times (read {nbytes}, write {nbytes} each)",
              flush=True)
        return 0
    if args.workload == "surrogate":
        from galax_amd import surrogate

        rep = surrogate.barred_halo_experiment(use_graph=not args.no_graph,
                                               log=lambda m: print(m, file=sys.stderr, flush=True))
        print(json.dumps({"report": "learned-surrogate field vs exact FP64 (barred MW + 64 subhaloes)", **rep}),
              flush=True)
        return 0
    if args.workload == "stream":
        print(json.dumps(stream_report(args.steps)), flush=True)
        return 0
    if args.workload == "c1":
        print(json.dumps(c1_report(args.steps)), flush=True)
        return 0
    if args.workload == "streamfit":
        print(json.dumps(streamfit_report(args.steps)), flush=True)
        return 0
    if args.blocks_per_cu is not None:
        _lib.check(_lib.lib().gx_set_launch_shape(0, args.blocks_per_cu, None, None), "gx_set_launch_shape")
    pot = gp.MilkyWayPotential()
    n = args.n
    w0, t0, T = workloads.config_uneven(n=n, seed=2)
    order = workloads.cost_order(w0, T)
    solver = gd.OrbitSolver(gd.Dopri8(), stepsize_controller=gd.PIDController(rtol=1e-8, atol=1e-10), max_steps=None)
    if args.workload == "uniform":
        T = np.full(n, 200.0)
    elif args.workload == "longest":
        # the orbit that takes the most steps (measured: the cost model only
        # ranks it near the head of the queue)
        probe = gd.integrate_orbits(pot, w0, t0, T, solver=solver, order=order)
        sel = [int(np.argmax(probe.nsteps.cpu().numpy()))]
        w0, t0, T = w0[sel], t0[sel], T[sel]
        n = 1
        order = np.zeros(1, dtype=np.int64)

    w0d = torch.as_tensor(w0, device="cuda")
    t0d = torch.as_tensor(t0, device="cuda")
    Td = torch.as_tensor(T, device="cuda")
    od = torch.as_tensor(order, device="cuda")
    pot._handle()

    if not np.array_equal(np.sort(order), np.arange(n)):
        raise SystemExit("queue order is not a permutation")

    # Passes are independent batches.  With `in_flight` streams (default
    # priority), pass k+1 starts while pass k finishes its few longest
    # orbits: the persistent kernel's workgroups retire as its queue
    # empties, and the next launch's workgroups take their place.  Each
    # stream has its own output and workspace.
    ws_bytes = int(_lib.lib().gx_integrate_workspace_bytes(n))

    def buffers(k):
        return ([torch.empty((n, args.saves, 6), dtype=torch.float64, device="cuda") for _ in range(k)],
                [torch.empty((ws_bytes + 7) // 8, dtype=torch.float64, device="cuda") for _ in range(k)])

    tuning = None
    explicit = args.in_flight is not None or args.launch_parts is not None or args.blocks_per_cu is not None
    # The streams are created once, before any timing: HIP maps each new
    # stream onto one of the process's hardware queues in turn, and streams
    # created per tuning trial left the timed run's streams sharing queues
    # (28 ms per pass instead of 19).
    from galax_amd import tune

    all_streams = [torch.cuda.Stream() for _ in range(max(tune.MAX_IN_FLIGHT, args.in_flight or 1))]
    if args.workload == "uneven" and not explicit and not args.no_tune:
        # the library's launch policy for this batch, searched on this device
        # (cached per device, build and workload)
        outs, wss = buffers(tune.MAX_IN_FLIGHT)

        def launch(j):
            gd.integrate_orbits(pot, w0d, t0d, Td, T=args.saves, solver=solver, order=od, validate_order=False,
                                out=outs[j], workspace=wss[j])

        tuning = tune.tune_batches(torch, launch, f"uneven:n={n}:saves={args.saves}:rtol=1e-8",
                                   max_in_flight=tune.MAX_IN_FLIGHT, streams=all_streams,
                                   log=lambda m: print(m, file=sys.stderr, flush=True))
        nfl = int(tuning["settings"]["in_flight"])
        parts = int(tuning["settings"]["parts"])
        outs, wss = outs[:nfl], wss[:nfl]
        torch.cuda.synchronize()
    else:
        nfl = max(1, args.in_flight if args.in_flight is not None else (IN_FLIGHT if args.workload == "uneven" else 1))
        parts = args.launch_parts if args.launch_parts is not None else min(nfl, 4)
        _lib.check(_lib.lib().gx_set_launch_fraction(parts, None), "gx_set_launch_fraction")
        outs, wss = buffers(nfl)
    streams = all_streams[:nfl]

    def run(k, launch_events=None):
        # one pass on stream k % nfl: gx_integrate's queue reset, first-step
        # prepass and the persistent k_integrate launch
        j = k % nfl
        with torch.cuda.stream(streams[j]):
            if launch_events is not None:
                launch_events[0].record()
            r = gd.integrate_orbits(pot, w0d, t0d, Td, T=args.saves, solver=solver, order=od,
                                    validate_order=False, out=outs[j], workspace=wss[j])
            if launch_events is not None:
                launch_events[1].record()
        return r

    main_stream = torch.cuda.current_stream()
    for k in range(args.warmup):
        for s_ in streams:
            s_.wait_stream(main_stream)
        res = run(k)
    torch.cuda.synchronize()
    print(f"[bench] warmup done, timing {args.steps} passes over {n} orbits, {nfl} in flight", file=sys.stderr,
          flush=True)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    per_launch = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                  for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_wall = time.perf_counter()
    ev0.record()
    for s_ in streams:
        s_.wait_stream(main_stream)
    results = []
    for k in range(args.steps):
        results.append(run(k, per_launch[k]))
    for s_ in streams:
        main_stream.wait_stream(s_)
    ev1.record()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t_wall) / args.steps
    ms = ev0.elapsed_time(ev1) / args.steps
    launch_ms = [a.elapsed_time(b) for a, b in per_launch]
    res = results[-1]
    out = outs[(args.steps - 1) % nfl]
    status = res.status.cpu().numpy()
    nsteps = res.nsteps.cpu().numpy()
    naccept = res.naccept.cpu().numpy()
    total_steps = int(nsteps.sum())
    value = total_steps / (ms * 1e-3)
    n_dense = n * max(args.saves - 2, 0)  # interior saves: the dense output
    work = flops.f_step("dopri8") * total_steps + flops.f_dense("dopri8") * n_dense
    # the dominant kernel's own rate: its work per launch over its mean launch
    # duration (events on its stream; overlapping launches each count whole)
    launch_avg = float(np.mean(launch_ms))
    achieved = work / (launch_avg * 1e-3) / 1e12
    traffic = flops.MEASURED_TRAFFIC_BYTES.get(args.workload) if n == args.n and args.saves == 64 else None
    line = {
        "metric": "orbit RK-steps/s (attempted Dopri8 steps, 1M Milky Way orbits, FP64)",
        "value": value,
        "unit": "orbit-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded PCG64 initial conditions)",
        "config": {
            "workload": "batched uneven exact FP64: 1,048,576 MW orbits, R~U(0.5,50) kpc, 25% dispersion, "
            "T_i log-uniform 0.1-5 Gyr, adaptive Dopri8 rtol 1e-8 atol 1e-10, 64 saves per orbit",
            "orbits": n,
            "saves_per_orbit": args.saves,
            "attempted_steps": total_steps,
            "accepted_steps": int(naccept.sum()),
            "failed_orbits": int((status != 0).sum()),
            "wall_ms_per_step": wall * 1e3,
            "max_steps_one_orbit": int(nsteps.max()),
            "diagnostic_workload": args.workload,
            "in_flight": nfl,
            "launch_parts": parts,
            "blocks_per_cu": args.blocks_per_cu,
            "launch_ms_mean": launch_avg,
            "launch_ms": [round(x, 3) for x in launch_ms],
            "tuning": None if tuning is None else {
                "settings": tuning["settings"],
                "ms_per_pass_in_search": tuning["ms_per_pass"],
                "trials": len(tuning["trials"]),
                "seconds": tuning["seconds"],
                "cached": tuning["cached"],
                "search": tuning["search"],
                "fixed": tuning["fixed"],
            },
        },
        "roofline": {
            "bound": "fp64_valu",
            "pipe": "FP64 vector ALU (no matrix instructions; MI355X FP64 vector peak 78.6 TFLOP/s)",
            "achieved": achieved,
            "peak": flops.PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / flops.PEAK_FP64_TFLOPS,
            "traffic": traffic,
            "achieved_device": work / (ms * 1e-3) / 1e12,
            "frac_device": work / (ms * 1e-3) / 1e12 / flops.PEAK_FP64_TFLOPS,
            "device_basis": "FLOP per pass / wall time per pass with the batches in flight; each launch "
                            "holds 1/launch_parts of the device, so the per-launch rate above understates "
                            "the device's",
            "traffic_source": "profiles/k_integrate_mw_uneven_traffic.txt" if traffic else None,
            "achieved_basis": "FLOP per k_integrate launch / mean launch duration",
            "flop_per_step": flops.f_step("dopri8"),
            "flop_per_interior_save": flops.f_dense("dopri8"),
            "interior_saves": n_dense,
        },
    }
    if args.workload == "uneven" and args.single_launch > 0:
        # the configuration as stated: one persistent launch over the whole
        # batch on the whole device, passes back to back
        prev = __import__("ctypes").c_int32()
        _lib.check(_lib.lib().gx_set_launch_fraction(1, __import__("ctypes").byref(prev)), "gx_set_launch_fraction")
        run(0)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(streams[0]):
            e0.record()
            for _ in range(args.single_launch):
                r1 = gd.integrate_orbits(pot, w0d, t0d, Td, T=args.saves, solver=solver, order=od,
                                         validate_order=False, out=outs[0], workspace=wss[0])
            e1.record()
        torch.cuda.synchronize()
        ms1 = e0.elapsed_time(e1) / args.single_launch
        steps1 = int(r1.nsteps.sum())
        work1 = flops.f_step("dopri8") * steps1 + flops.f_dense("dopri8") * n_dense
        line["single_launch"] = {
            "in_flight": 1,
            "launch_parts": 1,
            "passes": args.single_launch,
            "ms_per_pass": ms1,
            "orbit_steps_per_s": steps1 / (ms1 * 1e-3),
            "achieved_tflops": work1 / (ms1 * 1e-3) / 1e12,
            "frac": work1 / (ms1 * 1e-3) / 1e12 / flops.PEAK_FP64_TFLOPS,
        }
        _lib.check(_lib.lib().gx_set_launch_fraction(parts, None), "gx_set_launch_fraction")
    if args.dump_steps:
        np.savez_compressed(args.dump_steps, nsteps=nsteps, order=order, T=T)
    if not args.no_cpu:
        cpu = host_cpu()
        threads = cpu["cores_allotted"]
        out_np = out.cpu().numpy()
        cps, done, el, parity = cpu_baseline(pot, solver, w0, T, args.saves, out_np, nsteps, args.cpu_seconds,
                                             threads)
        cps1, done1, el1, _ = cpu_baseline(pot, solver, w0, T, args.saves, out_np, nsteps,
                                           args.cpu_seconds / 3.0, 1, seed=54321)
        line["cpu_baseline"] = {
            "value": cps,
            "unit": "orbit-steps/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{done} orbits drawn at random from the same workload, {el:.1f} s on {threads} threads "
                      f"(every core allotted to this one-GPU job)",
            "cpu_model": cpu["model"],
            "cores_visible": cpu["cores_visible"],
            "one_core": {"value": cps1, "sample": f"{done1} other orbits of the workload, {el1:.1f} s on 1 thread"},
        }
        line["parity_sample"] = parity
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
