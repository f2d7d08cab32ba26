"""The C++ host layer (reference class surface + scene JSON) on the CPU:
scene parsing through the sphg_run front end (--info touches no GPU) and its
error paths.  The GPU run of the same scenes is in test_gpu_host_layer.py."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "sph_amd", "sphg_run")
SCENES = os.path.join(ROOT, "sph_amd", "scenes")


def _run(*args):
    if not os.path.exists(RUN):
        pytest.fail("sph_amd/sphg_run is not built (make -C sph_amd/host)")
    return subprocess.run([RUN, *args], capture_output=True, text=True, timeout=60)


@pytest.mark.parametrize("name,method,blocks,spheres,visc,st,adh", [
    ("dam_break_dfsph_small.json", 2, 1, 0, -1, -1, -1),
    ("dam_break_wcsph_32k.json", 0, 1, 0, 2, -1, -1),
    ("sphere_drop_pcisph.json", 1, 0, 1, 2, 3, 4),
])
def test_scene_info(name, method, blocks, spheres, visc, st, adh):
    p = _run("-f", os.path.join(SCENES, name), "--info")
    assert p.returncode == 0, p.stderr
    info = json.loads(p.stdout)
    assert info["simulationMethod"] == method and info["boundaryMethod"] == 2
    assert info["particleRadius"] == 0.025
    assert (info["fluidBlocks"], info["fluidSpheres"]) == (blocks, spheres)
    assert (info["viscosityMethod"], info["surfaceTensionMethod"], info["adhesionMethod"]) == (visc, st, adh)


def _variant(tmp_path, edit):
    with open(os.path.join(SCENES, "dam_break_dfsph_small.json")) as f:
        d = json.load(f)
    edit(d)
    p = tmp_path / "scene.json"
    p.write_text(json.dumps(d))
    return str(p)


def test_missing_required_key_is_an_error(tmp_path):
    # every Configuration key the method reads is required (src/SceneLoader.cpp:38-80)
    path = _variant(tmp_path, lambda d: d["Configuration"].pop("etaV"))
    p = _run("-f", path, "--info")
    assert p.returncode == 1 and "etaV" in p.stderr


def test_unknown_keys_are_ignored(tmp_path):
    """As in the reference's nlohmann-based loader, keys it does not know are
    ignored at every level, so a scene written for another front end (or
    with this one's additive keys) still parses."""
    def edit(d):
        d["Unknown"] = {"x": 1}
        d["Configuration"]["someFutureKey"] = [1, 2, 3]
        d["Fluid"]["comment"] = "ignored"
        d["Boundary"][0]["note"] = {"nested": True}
    path = _variant(tmp_path, edit)
    p = _run("-f", path, "--info")
    assert p.returncode == 0, p.stderr
    info = json.loads(p.stdout)
    ref = json.loads(_run("-f", os.path.join(SCENES, "dam_break_dfsph_small.json"), "--info").stdout)
    assert info == ref


def test_wrong_type_is_an_error(tmp_path):
    path = _variant(tmp_path, lambda d: d["Configuration"].__setitem__("particleRadius", "big"))
    p = _run("-f", path, "--info")
    assert p.returncode == 1 and "particleRadius" in p.stderr


def test_device_other_than_gpu_is_refused(tmp_path):
    path = _variant(tmp_path, lambda d: d["Configuration"].__setitem__("device", "cpu"))
    p = _run("-f", path)
    assert p.returncode == 1 and "GPU only" in p.stderr


def test_unsupported_boundary_method_is_refused(tmp_path):
    path = _variant(tmp_path, lambda d: d["Configuration"].__setitem__("boundaryMethod", 3))
    p = _run("-f", path)
    assert p.returncode == 1 and "boundaryMethod" in p.stderr


def test_pcisph_boundary_method_parses(tmp_path):
    """boundaryMethod 1: each Boundary entry needs normalFct and tangFct, and
    each box and sphere its inverted flag (src/SceneLoader.cpp:161-190)."""
    def edit(d, drop=None):
        d["Configuration"]["boundaryMethod"] = 1
        for b in d["Boundary"]:
            b["normalFct"], b["tangFct"] = 0.5, 0.9
            for box in b["box"]:
                box["inverted"] = True
            if drop:
                del b[drop]
    p = _run("-f", _variant(tmp_path, edit), "--info")
    assert p.returncode == 0, p.stderr
    assert json.loads(p.stdout)["boundaryMethod"] == 1
    for key in ("normalFct", "tangFct"):
        p = _run("-f", _variant(tmp_path, lambda d, k=key: edit(d, k)), "--info")
        assert p.returncode == 1 and key in p.stderr, (key, p.stderr)


def test_cube_boundary_method_parses(tmp_path):
    path = _variant(tmp_path, lambda d: d["Configuration"].__setitem__("boundaryMethod", 0))
    p = _run("-f", path, "--info")
    assert p.returncode == 0, p.stderr
    assert json.loads(p.stdout)["boundaryMethod"] == 0


def test_emitters_parse(tmp_path):
    """The emitter keys of src/SceneLoader.cpp:123-139; every key is required."""
    em = {"position": [0.5, 0.3, 0.4], "rotation": [1, 0, 0, 0], "velocity": 1.0, "numParticles": 10,
          "startTime": 0.0, "type": 0, "width": 0.2, "height": 0.2, "spacing": 1.0}

    def edit(d):
        d["Fluid"]["FluidModel"][0]["emitters"] = [em, dict(em, type=1, numParticles=7)]
    p = _run("-f", _variant(tmp_path, edit), "--info")
    assert p.returncode == 0, p.stderr
    info = json.loads(p.stdout)
    assert (info["emitters"], info["emitterParticles"]) == (2, 17)
    for key in em:
        def drop(d, key=key):
            d["Fluid"]["FluidModel"][0]["emitters"] = [{k: v for k, v in em.items() if k != key}]
        p = _run("-f", _variant(tmp_path, drop), "--info")
        assert p.returncode == 1 and key in p.stderr, (key, p.stderr)
    def bad_rot(d):
        d["Fluid"]["FluidModel"][0]["emitters"] = [dict(em, rotation=[1, 0, 0])]
    p = _run("-f", _variant(tmp_path, bad_rot), "--info")
    assert p.returncode == 1 and "rotation" in p.stderr


def test_usage_and_extension():
    assert _run().returncode == 2
    assert _run("-f", "scene.txt").returncode == 2


@pytest.mark.parametrize("method", [0, 1, 2])
def test_viscosity_methods_parse(tmp_path, method):
    def edit(d):
        d["Fluid"]["viscosityMethod"] = method
        d["Fluid"]["FluidModel"][0].update({"viscosity": 0.05, "boundaryViscosity": 0.02})
    p = _run("-f", _variant(tmp_path, edit), "--info")
    assert p.returncode == 0, p.stderr
    assert json.loads(p.stdout)["viscosityMethod"] == method
