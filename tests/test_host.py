"""Host-side logic and the C ABI, without a GPU: structure layouts against
the C header, the library's exported symbols, potential flattening, units,
solver configuration and workload generation."""

import ctypes
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from galax_amd import _abi, _lib
from galax_amd.dynamics import ConstantStepSize, Dopri5, Dopri8, OrbitSolver, PIDController
import galax_amd.potential as gp
from galax_amd.units import G, KMS, Q
from galax_amd import workloads

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "galax_amd.h"


def _declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^(?:int|int64_t)\s+(gx_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    if not _lib.LIB_PATH.exists():
        pytest.skip("libgalaxhip.so not built")
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    names = _declared_functions()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_library_shares_torchs_hip_runtime():
    """Loading the library before anything imported torch must still leave
    one HIP runtime in the process: torch bundles its own libamdhip64 under
    the system runtime's soname, and two runtimes side by side do not share
    device memory or streams."""
    if not _lib.LIB_PATH.exists():
        pytest.skip("libgalaxhip.so not built")
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "from galax_amd import _lib; _lib.lib()\n"
        "print(sorted({l.split()[-1] for l in open('/proc/self/maps') if 'libamdhip64' in l}))\n" % str(ROOT)
    )
    res = subprocess.run(["python3", "-c", code], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr[-2000:]
    mapped = eval(res.stdout.strip().splitlines()[-1])
    assert len(mapped) == 1, mapped


def test_stale_library_is_refused_whole(tmp_path, monkeypatch):
    """A library that lacks an entry point of this package is refused before
    any signature is declared, and the loader keeps no handle: a later call
    must not reach a function with default (32-bit) argument conversion."""
    src = tmp_path / "stale.c"
    src.write_text("int gx_potential_create(void){return 0;}\n")
    so = tmp_path / "libstale.so"
    subprocess.run(["gcc", "-shared", "-fPIC", "-o", str(so), str(src)], check=True)
    monkeypatch.setattr(_lib, "LIB_PATH", so)
    monkeypatch.setattr(_lib, "_lib", None)
    for _ in range(2):
        with pytest.raises(_lib.GalaxHipError, match="older than this package"):
            _lib.lib()
        assert _lib._lib is None


def test_ctypes_signatures_match_header():
    """Every declared entry point has a ctypes signature with the same number
    of arguments (a mismatch would pass garbage through the ABI)."""
    text = HEADER.read_text()
    protos = dict(
        (m.group(1), m.group(2))
        for m in re.finditer(r"^(?:int|int64_t)\s+(gx_\w+)\s*\(([^)]*)\)\s*;", text, flags=re.M | re.S)
    )
    assert set(protos) == set(_lib.SIGNATURES), set(protos) ^ set(_lib.SIGNATURES)
    for name, args in protos.items():
        nargs = 0 if args.strip() in ("", "void") else args.count(",") + 1
        assert nargs == len(_lib.SIGNATURES[name][1]), name


def test_constants_match_header():
    """Every integer constant of the header (#define and enum members) that
    the Python mirror names has the header's value, and the mirror names the
    ones the host code passes through the ABI."""
    text = HEADER.read_text()
    consts = {m.group(1): int(m.group(2)) for m in re.finditer(r"^#define\s+(GX_\w+)\s+\(?(-?\d+)\)?", text, flags=re.M)}
    for body in re.findall(r"enum\s+\w+\s*\{([^}]*)\}", text, flags=re.S):
        for m in re.finditer(r"(GX_\w+)\s*=\s*(-?\d+)", body):
            consts[m.group(1)] = int(m.group(2))
    mirrored = {k: getattr(_abi, k) for k in consts if hasattr(_abi, k)}
    assert {k: consts[k] for k in mirrored} == mirrored
    for k in ("GX_MATH_CHECK_OUTPUTS", "GX_MAX_GL_ORDER", "GX_VOGELSBERGER_NFW", "GX_FRAME_TRACK",
              "GX_DOPRI5", "GX_PID", "GX_NONFINITE", "GX_TS_PER_ORBIT"):
        assert k in mirrored, k


def test_struct_layouts_match_header(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "galax_amd.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu\\n\", sizeof(gx_component),"
        " offsetof(gx_component,p), offsetof(gx_component,f), sizeof(gx_track), sizeof(gx_solver_cfg),"
        " offsetof(gx_solver_cfg,dt0), offsetof(gx_solver_cfg,force_dtmin)); return 0;}\n"
    )
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [
        ctypes.sizeof(_abi.GxComponent),
        _abi.GxComponent.p.offset,
        _abi.GxComponent.f.offset,
        ctypes.sizeof(_abi.GxTrack),
        ctypes.sizeof(_abi.GxSolverCfg),
        _abi.GxSolverCfg.dt0.offset,
        _abi.GxSolverCfg.force_dtmin.offset,
    ]
    assert got == want


def test_units_and_constants():
    # G from astropy 7.1 constants in kpc^3 / (Msun Myr^2)
    Gsi, msun, kpc, myr = 6.6743e-11, 1.988409870698051e30, 3.0856775814913673e19, 3.15576e13
    assert G == pytest.approx(Gsi * msun * myr**2 / kpc**3, rel=1e-15)
    assert KMS == pytest.approx(1e3 * myr / kpc, rel=1e-15)
    # 220 km/s as printed by the reference (orbit/solver.py:310)
    assert Q(220.0, "km/s").to_value() == pytest.approx(0.22499668, abs=5e-9)
    assert Q(1.0, "Gyr").to_value("Myr") == 1000.0
    assert Q(70, "pc").to_value() == pytest.approx(0.07)
    with pytest.raises(ValueError):
        Q(1.0, "furlong")


def test_milky_way_table():
    comps, tracks = gp.MilkyWayPotential()._table()
    kinds = [c.kind for c in comps]
    assert kinds == [_abi.GX_MIYAMOTO_NAGAI, _abi.GX_NFW, _abi.GX_HERNQUIST, _abi.GX_HERNQUIST]
    assert comps[0].p[0] == G * 6.8e10 and comps[0].p[1] == 3.0 and comps[0].p[2] == 0.28
    assert comps[1].p[0] == G * 5.4e11 and comps[1].p[1] == 15.62
    assert comps[3].p[1] == 0.07
    assert tracks == []
    mw2 = gp.MilkyWayPotential(halo={"m": 1e12})
    assert mw2._table()[0][1].p[0] == G * 1e12 and mw2._table()[0][1].p[1] == 15.62


def test_composite_and_translation_tables():
    bar = gp.MonariEtAl2016BarPotential(R0=8.0, v0=Q(220.0, "km/s"))
    assert bar.Omega == pytest.approx(52.2 * KMS)
    assert bar.phi_b == pytest.approx(np.deg2rad(25.0))
    tr = gp.TimeDependentTranslationParameter(np.linspace(0, 10, 11), np.zeros((11, 3)))
    pot = gp.CompositePotential(
        a=gp.KeplerPotential(m_tot=1.0),
        b=gp.TranslatedPotential(gp.PlummerPotential(m_tot=2.0, r_s=0.1), [1.0, 2.0, 3.0]),
        c=gp.TranslatedPotential(gp.PlummerPotential(m_tot=3.0, r_s=0.1), gp.RotatingTranslation([5.0, 0, 0], 0.04)),
        d=gp.TranslatedPotential(gp.PlummerPotential(m_tot=4.0, r_s=0.1), tr),
        e=bar,
    )
    comps, tracks = pot._table()
    assert [c.frame for c in comps] == [
        _abi.GX_FRAME_NONE, _abi.GX_FRAME_OFFSET, _abi.GX_FRAME_ROTATING, _abi.GX_FRAME_TRACK, _abi.GX_FRAME_NONE]
    assert list(comps[1].f[:3]) == [1.0, 2.0, 3.0]
    assert comps[2].f[3] == 0.04
    assert comps[3].track == 0 and len(tracks) == 1
    with pytest.raises(ValueError):
        gp.TranslatedPotential(pot["b"], [0.0, 0.0, 1.0])._rows()
    s = gp.KeplerPotential(m_tot=1.0) + gp.HernquistPotential(m_tot=1.0, r_s=1.0)
    assert len(s) == 2


def test_solver_configuration():
    c = OrbitSolver().cfg()
    assert (c.method, c.controller, c.max_steps) == (_abi.GX_DOPRI8, _abi.GX_PID, 2**16)
    assert (c.rtol, c.atol, c.icoeff, c.pcoeff, c.dcoeff) == (1e-8, 1e-8, 1.0, 0.0, 0.0)
    assert (c.safety, c.factormin, c.factormax, c.dt0) == (0.9, 0.2, 10.0, 0.0)
    c = OrbitSolver(Dopri5(), stepsize_controller=ConstantStepSize(), max_steps=None).cfg(dt0=-0.5)
    assert (c.method, c.controller, c.dt0, c.max_steps) == (_abi.GX_DOPRI5, _abi.GX_CONSTANT, 0.5, 0)
    with pytest.raises(ValueError):
        OrbitSolver(stepsize_controller=ConstantStepSize()).cfg()
    c = OrbitSolver(stepsize_controller=PIDController(rtol=1e-6, atol=1e-9, dtmin=0.3)).cfg(dt0=2.0)
    assert (c.dtmin, c.dt0, c.force_dtmin) == (0.3, 2.0, 1)


def test_workloads_are_seeded_and_shaped():
    w0, t0, t1, ts = workloads.config_plain_small()
    assert w0.shape == (1024, 6) and ts[0] == 0.0 and ts[-1] == 1000.0 and len(ts) == 100
    np.testing.assert_array_equal(w0, workloads.config_plain_small()[0])
    R = np.hypot(w0[:, 0], w0[:, 1])
    assert R.min() >= 4.0 and R.max() <= 20.0
    # speeds scatter around the circular speed (10% dispersion)
    assert 0.15 < np.median(np.linalg.norm(w0[:, 3:], axis=1)) < 0.3
    w, t0, T = workloads.config_uneven(n=4096)
    assert T.min() >= 100.0 and T.max() <= 5000.0
    order = workloads.cost_order(w, T)
    assert sorted(order.tolist()) == list(range(4096))


def test_numpy_gradient_matches_oracle():
    import oracle

    rng = np.random.default_rng(3)
    x = rng.normal(0, 10, (64, 3))
    np.testing.assert_allclose(
        workloads.mw_gradient_np(x), oracle.evaluate(gp.MilkyWayPotential(), "gradient", x), rtol=1e-12
    )


def _kernel_resources(tmp_path):
    """Kernel metadata (VGPRs, spills, scratch) of every kernel in the built
    library's gfx950 code objects: the .hip_fatbin section holds one offload
    bundle per source file."""
    tools = Path("/opt/rocm/llvm/bin")
    need = [tools / "llvm-objcopy", tools / "clang-offload-bundler", tools / "llvm-readelf"]
    if not _lib.LIB_PATH.exists() or not all(t.exists() for t in need):
        pytest.skip("library or ROCm LLVM tools missing")
    fat = tmp_path / "fat.bin"
    subprocess.run([str(need[0]), f"--dump-section=.hip_fatbin={fat}", str(_lib.LIB_PATH), str(tmp_path / "stripped.so")],
                   check=True, capture_output=True)
    data = fat.read_bytes()
    magic = b"__CLANG_OFFLOAD_BUNDLE__"
    starts = [m.start() for m in re.finditer(re.escape(magic), data)]
    out = {}
    for i, s in enumerate(starts):
        part = tmp_path / f"b{i}.bin"
        part.write_bytes(data[s:(starts[i + 1] if i + 1 < len(starts) else len(data))])
        dev = tmp_path / f"b{i}.o"
        subprocess.run([str(need[1]), "--unbundle", "--type=o", f"--input={part}",
                        "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", f"--output={dev}"], check=True, capture_output=True)
        notes = subprocess.run([str(need[2]), "--notes", str(dev)], check=True, capture_output=True, text=True).stdout
        for blk in notes.split("- .agpr_count")[1:]:
            name = re.search(r"\.name:\s+(\S+)", blk).group(1)
            get = lambda k: int(re.search(rf"\.{k}:\s+(\d+)", blk).group(1))
            out[name] = (get("vgpr_count"), get("vgpr_spill_count"), get("private_segment_fixed_size"))
    return out


def test_integrator_kernels_keep_their_occupancy(tmp_path):
    """The Milky Way and frame-less table integrators stay within 256 VGPRs
    (two waves per SIMD; above it a SIMD holds one wave of them and the
    throughput halves): a guard against additions to the field code that
    the compiler inlines into every kernel.  The tracked-Plummer integrator
    on a knot track runs one wave per SIMD by design (its batched track
    loads take the registers) and must stay within that wave's 512 without
    scratch memory (its few spills go to accumulation registers).  On an
    LDS-staged dense track it fits 256, so that two workgroups share a CU
    and a 100,000-particle release starts every particle at once."""
    res = _kernel_resources(tmp_path)
    # key: (VGPR limit, VGPR spill limit or None, scratch memory allowed)
    want = {"EvalMW": (256, 0, False), "EvalTableILb0E": (256, None, True),
            "EvalMilkyWayTracked": (512, None, False), "EvalMilkyWayPieces": (256, None, False)}
    for key, (max_vgpr, max_spill, scratch_ok) in want.items():
        ks = {n: v for n, v in res.items() if "k_integrate" in n and "Dopri8" in n and key in n}
        assert ks, key
        for n, (vgpr, spill, scratch) in ks.items():
            assert vgpr <= max_vgpr, (n, vgpr)
            if max_spill is not None:
                assert spill <= max_spill, (n, spill)
            if not scratch_ok:
                assert scratch == 0, (n, scratch)


def test_caller_buffers_are_checked():
    """integrate_orbits rejects an output or workspace buffer that the
    kernels would overrun or misread (checked before any launch)."""
    import torch

    from galax_amd.dynamics._solver import _check_out, _check_workspace

    with pytest.raises(ValueError, match="shape"):
        _check_out(torch, torch.empty((4, 3, 6), dtype=torch.float64), 5, 3)
    with pytest.raises(ValueError, match="float64 CUDA"):
        _check_out(torch, torch.empty((5, 3, 6), dtype=torch.float64), 5, 3)
    with pytest.raises(ValueError, match="float64 CUDA"):
        _check_out(torch, torch.empty((5, 3, 6), dtype=torch.float32), 5, 3)
    with pytest.raises(ValueError, match="bytes"):
        _check_workspace(torch, torch.empty(10, dtype=torch.float64), 81)
    with pytest.raises(ValueError, match="CUDA"):
        _check_workspace(torch, torch.empty(11, dtype=torch.float64), 81)
    with pytest.raises(ValueError):
        _check_out(torch, np.zeros((5, 3, 6)), 5, 3)


def test_autotuner_keeps_the_start_unless_the_winner_wins_again():
    """A setting that won the search on one noisy trial replaces the start
    only if it is faster again, by the margin, in alternating re-measurements."""
    from galax_amd.tune import START, confirm_winner

    lucky = dict(START, in_flight=10)
    calls = []

    def flat(c):  # the winner's first trial was luck: both settings equal
        calls.append(dict(c))
        return 20.0

    chosen, v, rec = confirm_winner(flat, START, lucky, 15.0, 0.02)
    assert chosen == START and v == 20.0 and rec["kept"] == START
    assert calls == [START, lucky, START, lucky]  # alternating
    faster = lambda c: 18.0 if c["in_flight"] == 10 else 20.0
    chosen, v, rec = confirm_winner(faster, START, lucky, 18.0, 0.02)
    assert chosen == lucky and v == 18.0
    assert confirm_winner(flat, START, dict(START), 19.0, 0.02) == (START, 19.0, None)


def test_autotuner_search_finds_the_minimum_and_measures_each_setting_once():
    from galax_amd.tune import SPACE, START, coordinate_descent

    calls = []

    def cost(c):
        calls.append(tuple(sorted(c.items())))
        # a separable bowl with its minimum off the start point
        return (1.0 + abs(c["in_flight"] - 4) + 0.5 * abs(c["parts"] - 3) + (c["blocks_per_cu"] != 2)
                + 0.1 * (c["head_lanes"] != 1024) + 0.2 * (c["chunk"] != 1))

    best, v, trials = coordinate_descent(cost, SPACE, START, margin=0.02)
    assert best == {"in_flight": 4, "parts": 3, "blocks_per_cu": 2, "head_lanes": 1024, "chunk": 1}
    assert v == pytest.approx(1.0)
    assert len(calls) == len(set(calls)) == len(trials)
    # a change within the margin is not taken
    best2, _, _ = coordinate_descent(lambda c: 1.0 - 0.001 * (c["chunk"] == 4), SPACE, START, margin=0.02)
    assert best2 == START
