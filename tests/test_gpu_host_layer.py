"""Drop-in check of the host layer on the device: a scene JSON run through
Simulation::importScene -> GpuSolver (the sphg_run front end) ends in the
same state, bit for bit, as the same scene built through the Python binding
of the C ABI."""
import json
import os
import subprocess

import numpy as np
import pytest

from sph_amd import configs, sphg

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "sph_amd", "sphg_run")
SCENES = os.path.join(ROOT, "sph_amd", "scenes")
STEPS = 5


def _cli_state(tmp_path, name):
    out = tmp_path / "final.json"
    p = subprocess.run([RUN, "-f", os.path.join(SCENES, name), "--steps", str(STEPS), "--final", str(out)],
                       capture_output=True, text=True, timeout=110)
    assert p.returncode == 0, p.stderr
    summary = json.loads(p.stdout.strip().splitlines()[-1])
    assert summary["steps"] == STEPS
    with open(out) as f:
        st = json.load(f)
    return np.array(st["positions"], dtype=np.float64), np.array(st["velocities"], dtype=np.float64), summary


def _api_state(ctx, spec):
    s = sphg.Scene(ctx, spec.params, spec.fluid, spec.boundary)
    s.step(STEPS)
    x, v = s.get(sphg.FIELD_X), s.get(sphg.FIELD_V)
    s.close()
    return x, v


def _pcisph_drop():
    r, h = 0.025, 0.1
    fluid = sphg.fluid_spheres((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), r, [(0.5, 0.5, 0.6)], [0.25])
    boundary = np.concatenate([sphg.sample_box((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), h),
                               sphg.sample_sphere((0.5, 0.5, 0.15), 0.1, r)])
    p = sphg.SceneParams(method=sphg.PCISPH, particle_radius=r, time_step=5e-4, min_time_step=5e-4,
                         max_time_step=5e-4, fps=50.0, cfl_factor=1.0, eta=0.01, min_iterations=3,
                         max_iterations=100, viscosity_method=sphg.VISCOSITY_XSPH, viscosity=0.01,
                         surface_tension_method=sphg.SURFACE_TENSION_AKINCI, surface_tension=0.05,
                         adhesion_method=sphg.ADHESION_AKINCI, adhesion=0.1)
    return configs.SceneSpec("drop", p, fluid, np.ascontiguousarray(boundary))


@pytest.mark.parametrize("name", ["dam_break_dfsph_small.json", "sphere_drop_pcisph.json", "dam_break_wcsph_32k.json"])
def test_scene_json_matches_c_abi(ctx, tmp_path, name):
    if name.startswith("dam_break_dfsph"):
        spec = configs.small_scene(sphg.DFSPH, n_side=8, spacing=0.05, jitter=0.01, seed=7)
    elif name.startswith("sphere_drop"):
        spec = _pcisph_drop()
    else:
        spec = configs.wcsph_dam_32k()
    xc, vc, _ = _cli_state(tmp_path, name)
    xa, va = _api_state(ctx, spec)
    assert xc.shape == xa.shape
    assert np.array_equal(xc, xa.astype(np.float64))
    assert np.array_equal(vc, va.astype(np.float64))


def test_viscosity_methods_through_scene_json(ctx, tmp_path):
    """viscosityMethod 0 and 1 (standard, artificial) selected by the scene
    file end in the same state as through the C ABI."""
    with open(os.path.join(SCENES, "dam_break_dfsph_small.json")) as f:
        base = json.load(f)
    for method in (sphg.VISCOSITY_STANDARD, sphg.VISCOSITY_ARTIFICIAL):
        d = json.loads(json.dumps(base))
        d["Fluid"]["viscosityMethod"] = method
        d["Fluid"]["FluidModel"][0].update({"viscosity": 0.05, "boundaryViscosity": 0.02})
        path = tmp_path / f"visc{method}.json"
        path.write_text(json.dumps(d))
        out = tmp_path / f"final{method}.json"
        p = subprocess.run([RUN, "-f", str(path), "--steps", str(STEPS), "--final", str(out)],
                           capture_output=True, text=True, timeout=110)
        assert p.returncode == 0, p.stderr
        with open(out) as f:
            xc = np.array(json.load(f)["positions"], dtype=np.float64)
        spec = configs.small_scene(sphg.DFSPH, n_side=8, spacing=0.05, jitter=0.01, seed=7, viscosity_method=method,
                                   viscosity=0.05, boundary_viscosity=0.02)
        xa, _ = _api_state(ctx, spec)
        assert np.array_equal(xc, xa.astype(np.float64)), method


def test_cube_boundary_method_scene_json(ctx, tmp_path):
    """boundaryMethod 0 in a scene file: the reference's loader creates no
    boundary model, so the fluid has no boundary at all; the same through
    the C ABI (boundary_method cube, no boundary particles, no clamp)."""
    with open(os.path.join(SCENES, "dam_break_dfsph_small.json")) as f:
        d = json.load(f)
    d["Configuration"]["boundaryMethod"] = 0
    path = tmp_path / "cube.json"
    path.write_text(json.dumps(d))
    out = tmp_path / "final_cube.json"
    p = subprocess.run([RUN, "-f", str(path), "--steps", str(STEPS), "--final", str(out)],
                       capture_output=True, text=True, timeout=110)
    assert p.returncode == 0, p.stderr
    with open(out) as f:
        xc = np.array(json.load(f)["positions"], dtype=np.float64)
    spec = configs.small_scene(sphg.DFSPH, n_side=8, spacing=0.05, jitter=0.01, seed=7)
    from dataclasses import replace
    spec.params = replace(spec.params, boundary_method=sphg.BOUNDARY_CUBE)
    spec.boundary = np.zeros((0, 3), np.float32)
    xa, _ = _api_state(ctx, spec)
    assert np.array_equal(xc, xa.astype(np.float64))


def test_emitter_scene_json(ctx, tmp_path):
    """sph_amd/scenes/inflow_dfsph.json: a square emitter shooting along -x
    and a circle emitter shooting down (quaternion (0, 1, 0, 0): 180 degrees
    about x).  The scene file through the host layer and the same scene
    through the C ABI end in bit-identical states; the frame lists every
    slot (numParticles) with the ones not yet emitted at zero, as the
    reference's writer does (src/SceneLoader.cpp:224-262)."""
    steps = 40
    out = tmp_path / "final_inflow.json"
    p = subprocess.run([RUN, "-f", os.path.join(SCENES, "inflow_dfsph.json"), "--steps", str(steps), "--final",
                        str(out)], capture_output=True, text=True, timeout=110)
    assert p.returncode == 0, p.stderr
    with open(out) as f:
        st = json.load(f)
    spec = configs.small_scene(sphg.DFSPH, n_side=6, spacing=0.05, jitter=0.01, seed=7, box=(1.1, 0.525, 0.7))
    ems = [sphg.Emitter(type=0, num_particles=400, position=(0.9, 0.26, 0.45), velocity=2.0,
                        rotation=(0.7071067811865476, 0.0, -0.7071067811865476, 0.0), start_time=0.0, width=0.16,
                        height=0.16, spacing=1.0),
           sphg.Emitter(type=1, num_particles=200, position=(0.55, 0.26, 0.6), velocity=1.5,
                        rotation=(0.0, 1.0, 0.0, 0.0), start_time=0.01, width=0.1, height=0.0, spacing=1.0)]
    s = sphg.Scene(ctx, spec.params, spec.fluid, spec.boundary, emitters=ems)
    s.step(steps)
    n = s.n
    xa, va = s.get(sphg.FIELD_X), s.get(sphg.FIELD_V)
    cap = s.capacity
    s.close()
    assert st["numParticles"] == cap == spec.n + 600
    assert st["numActiveParticles"] == n and n > spec.n
    xc = np.array(st["positions"], dtype=np.float64)
    vc = np.array(st["velocities"], dtype=np.float64)
    assert np.array_equal(xc[:n], xa.astype(np.float64)) and np.array_equal(vc[:n], va.astype(np.float64))
    assert np.all(xc[n:] == 0.0) and np.all(vc[n:] == 0.0)
    # the circle emitter fired too: its velocity (0, 0, 1.5) turned to -z
    # (the square's particles fall from rest along z: above -0.6 m/s here)
    assert np.any(va[spec.n:, 2] < -1.2)


@pytest.mark.parametrize("name,steps,min_frames", [("dam_break_dfsph_small.json", 60, 3),
                                                   ("dam_break_wcsph_32k.json", 100, 2)])
def test_frames_cadence_and_content(ctx, tmp_path, name, steps, min_frames):
    """Frames of `sphg_run -o true` (SceneLoader::writeFluid at
    TimeManager::hasToSave, src/SceneLoader.cpp:204-268,
    src/TimeManager.h:101-111), written by the asynchronous path: each
    frame holds, bit for bit, the C-ABI state at the frame's time, and the
    frames follow the cadence rule replayed on the C-ABI time sequence.  The
    WCSPH scene's fixed step lets Simulation::run hand the steps between
    frames to the device in runs (38 of every 40 here), and the replay steps
    one at a time."""
    import shutil
    scene = tmp_path / "dam.json"
    shutil.copy(os.path.join(SCENES, name), scene)
    p = subprocess.run([RUN, "-f", str(scene), "-o", "true", "--steps", str(steps)], capture_output=True, text=True,
                       timeout=110)
    assert p.returncode == 0, p.stderr
    assert json.loads(p.stdout.strip().splitlines()[-1])["steps"] == steps
    frames = sorted(tmp_path.glob("dam_*.json"))
    # replay: time after each step, and hasToSave (lastTimeSave = frame/fps;
    # the first frame is 1, src/TimeManager.h:31)
    fps = 50.0
    if "wcsph" in name:
        spec = configs.wcsph_dam_32k()
    else:
        spec = configs.small_scene(sphg.DFSPH, n_side=8, spacing=0.05, jitter=0.01, seed=7)
    s = sphg.Scene(ctx, spec.params, spec.fluid, spec.boundary)
    last, frame, expect = 0.0, 1, []
    for k in range(steps):
        s.step(1)
        t = s.scalars().time
        if t - last >= 1.0 / fps:
            last = frame / fps
            expect.append((frame, t, s.get(sphg.FIELD_X), s.get(sphg.FIELD_V)))
            frame += 1
    s.close()
    assert len(expect) >= min_frames and [f.name for f in frames] == [f"dam_{k:04d}.json" for k, *_ in expect]
    for f, (k, t, x, v) in zip(frames, expect):
        with open(f) as fh:
            st = json.load(fh)
        assert st["time"] == t and st["numParticles"] == st["numActiveParticles"] == spec.n
        assert np.array_equal(np.array(st["positions"]), x.astype(np.float64))
        assert np.array_equal(np.array(st["velocities"]), v.astype(np.float64))


def test_pcisph_boundary_method_scene_json(ctx, tmp_path):
    """boundaryMethod 1 in a scene file (one Boundary entry, an inverted box,
    normalFct 0.5, tangFct 0.9) through the host layer ends in the same
    state, bit for bit, as the same scene through the C ABI."""
    with open(os.path.join(SCENES, "dam_break_dfsph_small.json")) as f:
        d = json.load(f)
    d["Configuration"]["boundaryMethod"] = 1
    b = d["Boundary"][0]
    b["normalFct"], b["tangFct"] = 0.5, 0.9
    b["box"][0]["inverted"] = True
    path = tmp_path / "pb.json"
    path.write_text(json.dumps(d))
    out = tmp_path / "final_pb.json"
    p = subprocess.run([RUN, "-f", str(path), "--steps", str(STEPS), "--final", str(out)],
                       capture_output=True, text=True, timeout=110)
    assert p.returncode == 0, p.stderr
    with open(out) as f:
        st = json.load(f)
    spec = configs.small_scene(sphg.DFSPH, n_side=8, spacing=0.05, jitter=0.01, seed=7)
    from dataclasses import replace
    params = replace(spec.params, boundary_method=sphg.BOUNDARY_PCISPH, pb_normal_fct=0.5, pb_tangential_fct=0.9)
    pts, nrm = sphg.sample_box_normals(b["box"][0]["min"], b["box"][0]["max"], spec.params.support_radius, True)
    s = sphg.Scene(ctx, params, spec.fluid, pts, boundary_normals=nrm)
    s.step(STEPS)
    xa, va = s.get(sphg.FIELD_X), s.get(sphg.FIELD_V)
    s.close()
    assert np.array_equal(np.array(st["positions"]), xa.astype(np.float64))
    assert np.array_equal(np.array(st["velocities"]), va.astype(np.float64))
