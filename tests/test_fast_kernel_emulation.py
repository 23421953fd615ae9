"""The fast-math tiled grid kernel's tile logic (feriphys_amd/csrc/grid_fast_core.hpp) run on the
CPU against the oracle, before any GPU: every tile's three phases in barrier order with each LDS
index bounds-checked and the LDS filled with NaN (tests/native/fast_emul.cpp).  The same bars as
the GPU parity tests of the FAST form: one-step forces within 1e-5 of the largest force component,
one-step states within 1e-5 relative; grids that do not fill their last tiles, both tile shapes,
pins at the corners and along an edge, tiles inside the grid (the interior instantiation) and at its
border, non-zero starting forces, Euler and RK4.  (Obstacles are left
to the GPU tests: the collision functions are the exact kernels' own.)"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import pyoracle as orc
from feriphys_amd import workloads as wl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
F32 = np.float32


@pytest.fixture(scope="module")
def emul(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("emul") / "fast_emul.so")
    subprocess.run(["/opt/rocm/bin/hipcc", "-x", "hip", "--cuda-host-only", "--offload-arch=gfx950", "-std=c++17",
                    "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
                    os.path.join(ROOT, "tests", "native", "fast_emul.cpp"), "-o", so], check=True)
    lib = C.CDLL(so)
    fp = C.POINTER(C.c_float)
    lib.emul_fast.argtypes = [C.c_uint32, C.c_uint32, C.c_float, fp, fp, fp, fp, fp, C.c_int32,
                              C.POINTER(C.c_uint32), C.c_int32, fp, C.c_int32, C.c_int32]
    lib.emul_fast.restype = C.c_int
    return lib


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def run_emul(lib, scene, rec, mode, shape, force_init=False):
    """One call of the emulated kernel on cloth 0 of `scene` from records `rec`: returns the records
    after the step (mode 0) or with the forces (mode 1)."""
    c = scene.cloths[0]
    cf = scene.config
    n = c.rows * c.cols
    state = np.ascontiguousarray(rec[:, 1:7].T.astype(F32))
    out = np.zeros_like(state)
    forces = np.ascontiguousarray(rec[:, 7:10].T.astype(F32)) if force_init else np.zeros((3, n), F32)
    bits = np.zeros((n + 31) // 32, np.uint32)
    for p in c.pins:
        bits[p >> 5] |= np.uint32(1 << (p & 31))
    pos = np.array([c.position[0], c.position[2] if c.plane == 1 else c.position[1], 0.0], F32)
    kc6 = np.array([c.tk, c.tc, c.sk, c.sc, c.bk, c.bc], F32)
    dt = F32(cf.dt)
    cfg = np.array([dt, orc.duration_roundtrip(dt), *cf.gravity, *cf.wind, cf.drag_coefficient, cf.lift_coefficient,
                    cf.coefficient_of_restitution, cf.coefficient_of_friction, rec[0, 0]], F32)
    assert (rec[:, 0] == rec[0, 0]).all()
    lib.emul_fast(c.rows, c.cols, F32(c.spacing), _p(pos), _p(kc6), _p(state), _p(out), _p(forces),
                  1 if force_init else 0, bits.ctypes.data_as(C.POINTER(C.c_uint32)), cf.integration, _p(cfg), mode,
                  shape)
    res = np.array(rec, F32, copy=True)
    if mode == 0:
        res[:, 1:7] = out.T
        res[:, 7:10] = 0.0 if force_init else rec[:, 7:10]
    else:
        res[:, 7:10] = forces.T
    return res


def oracle_cfg(scene):
    c = scene.config
    return orc.Config(c.integration, c.dt, c.gravity, c.wind, c.lift_coefficient, c.drag_coefficient,
                      c.coefficient_of_restitution, c.coefficient_of_friction)


def moving(scene, steps):
    o = orc.build_scene(scene)
    o.step(steps)
    return o.mesh(0).points()


def force_rel_err(f, ref):
    return float(np.abs(f.astype(np.float64) - ref).max()) / max(float(np.abs(ref).max()), 1e-30)


CASES = [(32, 32, "corners", 30), (32, 32, "edge", 60), (5, 3, "edge", 10), (2, 2, "edge", 5),
         (19, 129, "edge", 20), (37, 70, "corners", 15), (3, 40, "edge", 10), (16, 16, "edge", 25),
         (31, 31, "corners", 25), (30, 62, "edge", 12), (64, 96, "edge", 40)]


@pytest.mark.parametrize("shape", [0, 1])
@pytest.mark.parametrize("rows,cols,pins,steps", CASES)
def test_emulated_forces_and_step(emul, rows, cols, pins, steps, shape):
    sc = wl.config0(rows=rows, cols=cols, pins=pins)
    rec = moving(sc, steps)
    # forces
    o = orc.build_scene(sc)
    o.mesh(0).set_points(rec)
    o.mesh(0).accumulate_forces(oracle_cfg(sc))
    fr = o.mesh(0).points()[:, 7:10]
    fe = run_emul(emul, sc, rec, 1, shape)[:, 7:10]
    assert np.isfinite(fe).all()
    assert force_rel_err(fe, fr) <= 1e-5
    assert (fe[sc.cloths[0].pins] == 0).all()
    # one step
    o = orc.build_scene(sc)
    o.mesh(0).set_points(rec)
    o.step(1)
    sr = o.mesh(0).points()
    se = run_emul(emul, sc, rec, 0, shape)
    assert np.isfinite(se).all()
    dx = float(np.abs(se[:, 1:4].astype(np.float64) - sr[:, 1:4]).max())
    dv = float(np.abs(se[:, 4:7].astype(np.float64) - sr[:, 4:7]).max())
    assert dx <= 1e-5 * max(1.0, float(np.abs(sr[:, 1:4]).max()))
    assert dv <= 1e-5 * max(float(np.abs(sr[:, 4:7]).max()), 1e-3)
    pins_ = sc.cloths[0].pins
    assert (se[pins_, 1:7] == rec[pins_, 1:7]).all()


@pytest.mark.parametrize("integration", [0, 1])
def test_emulated_step_from_nonzero_forces(emul, integration):
    """A step that starts from force records left by the split API (accumulate_forces) adds them and
    clears them, as Simulation::step after accumulate_forces would (springy/simulation.rs:27-45)."""
    sc = wl.config0(rows=20, cols=23, pins="edge")
    sc.config.integration = integration
    rec = moving(sc, 20)
    rng = np.random.default_rng(3)
    rec[:, 7:10] = rng.normal(scale=0.01, size=(rec.shape[0], 3)).astype(F32)
    o = orc.build_scene(sc)
    o.mesh(0).set_points(rec)
    o.step(1)
    sr = o.mesh(0).points()
    se = run_emul(emul, sc, rec, 0, 0, force_init=True)
    dx = float(np.abs(se[:, 1:4].astype(np.float64) - sr[:, 1:4]).max())
    dv = float(np.abs(se[:, 4:7].astype(np.float64) - sr[:, 4:7]).max())
    assert dx <= 1e-5 * max(1.0, float(np.abs(sr[:, 1:4]).max()))
    assert dv <= 1e-5 * max(float(np.abs(sr[:, 4:7]).max()), 1e-3)
    assert (se[:, 7:10] == 0).all()


def test_emulated_wind_dominated_faces(emul):
    """No gravity, a strong wind, random velocities and an undeformed cloth with undamped struts: the
    strut forces are exactly zero, so the face forces (drag, lift, corner weights) are the whole
    force and are held to the 1e-5 bar on their own."""
    sc = wl.config0(rows=24, cols=24, pins="edge")
    c = sc.cloths[0]
    sc.cloths[0] = wl.cloth_spec(c.rows, c.cols, tensile=(500.0, 0.0), shear=(300.0, 0.0), bend=(50.0, 0.0),
                                 pins=c.pins)
    sc.config.gravity = (0.0, 0.0, 0.0)
    sc.config.wind = (3.0, -1.0, 2.0)
    sc.jitter = 0.0
    o = orc.build_scene(sc)
    rec = o.mesh(0).points()
    rng = np.random.default_rng(5)
    rec[:, 4:7] = rng.normal(scale=0.5, size=(rec.shape[0], 3)).astype(F32)  # velocities only
    o = orc.build_scene(sc)
    o.mesh(0).set_points(rec)
    o.mesh(0).accumulate_forces(oracle_cfg(sc))
    fr = o.mesh(0).points()[:, 7:10]
    o2 = orc.build_scene(sc)
    o2.mesh(0).set_points(rec)
    o2.mesh(0).apply_strut_forces()
    assert (o2.mesh(0).points()[:, 7:10] == 0).all()  # the struts contribute nothing
    assert np.abs(fr).max() > 1e-3
    fe = run_emul(emul, sc, rec, 1, 0)[:, 7:10]
    assert force_rel_err(fe, fr) <= 1e-5


def test_emulated_staging_stays_in_bounds_under_address_sanitizer(tmp_path):
    """The emulation built with an address sanitizer (host code only) and driven over 100 grids of
    sides 2-34, both tile shapes and both modes (tests/native/fast_emul_asan.cpp): every state
    plane in an allocation of its own, so that a staging load past a plane's end (the clamped loads
    of lanes past the staged region or outside the grid) or a pin-word read past the mask would be
    reported; none is."""
    exe = str(tmp_path / "fast_emul_asan")
    native = os.path.join(ROOT, "tests", "native")
    subprocess.run(["/opt/rocm/bin/hipcc", "-x", "hip", "--cuda-host-only", "--offload-arch=gfx950", "-std=c++17", "-O1",
                    "-g", "-ffp-contract=off", "-fno-fast-math",
                    *"-Xarch_host -fsanitize=address -Xarch_host -fno-omit-frame-pointer".split(),
                    os.path.join(native, "fast_emul.cpp"), os.path.join(native, "fast_emul_asan.cpp"), "-o", exe],
                   check=True, stderr=subprocess.DEVNULL)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1"))
    assert r.returncode == 0 and r.stdout.strip() == "ok", (r.stdout[-2000:], r.stderr[-4000:])
