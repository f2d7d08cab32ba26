"""CPU-side checks of libsphg: every function declared in include/sphg.h is
exported, struct layouts agree with the Python mirror, and the scene
set-up path (samplers, generators, argument validation) matches the
oracle's restatement of the reference bit for bit."""
import ctypes as C
import json
import os
import re
import subprocess

import numpy as np
import pytest

from sph_amd import configs, sphg
import oracle.pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "closed_forms.json")))


def _declared():
    text = open(os.path.join(ROOT, "include", "sphg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sphg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = sphg.load()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", sphg.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (sphg_[a-z0-9_]+)", out))
    assert set(names) <= exported


def test_struct_layouts_match_header(tmp_path):
    """Every field of the ctypes mirrors sits at the offset the C compiler
    gives it in include/sphg.h (a probe compiled with gcc)."""
    structs = {"sphg_scene_desc": sphg.SceneDesc, "sphg_scene_scalars": sphg.SceneScalars,
               "sphg_step_stats": sphg.StepStats, "sphg_emitter_desc": sphg.EmitterDesc,
               "sphg_tune_result": sphg.TuneResult}
    lines = ['#include "sphg.h"', "#include <stdio.h>", "#include <stddef.h>", "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines) + "\n")
    inc = os.path.join(ROOT, "include")
    subprocess.run(["gcc", "-I", inc, str(src), "-o", str(tmp_path / "probe")], check=True)
    out = subprocess.run([str(tmp_path / "probe")], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in structs.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_version_and_errors_without_device_calls():
    lib = sphg.load()
    assert lib.sphg_version().startswith(b"sphg")
    # argument validation happens before any device call
    assert lib.sphg_scene_create(None, None, None) == -1
    assert b"NULL" in lib.sphg_last_error()


@pytest.mark.parametrize("key", sorted(GOLD["box_shell_counts"]))
def test_box_shell_counts_and_bitexact(key):
    dims, h = key.split("@h")
    L = tuple(float(v) for v in dims.split("x"))
    h = float(h)
    a = sphg.sample_box((0.0, 0.0, 0.0), L, h)
    b = po.sample_cube((0.0, 0.0, 0.0), L, h)
    assert len(a) == GOLD["box_shell_counts"][key]
    assert np.array_equal(a, b.astype(np.float32))


def test_sphere_sampler_bitexact():
    for R, pr in ((0.5, 0.025), (0.1, 0.005), (1.3, 0.01)):
        a = sphg.sample_sphere((0.1, -0.2, 0.3), R, pr)
        b = po.sample_sphere((0.1, -0.2, 0.3), R, pr)
        assert len(a) == len(b) > 0
        assert np.array_equal(a, b.astype(np.float32))
        assert np.allclose(np.linalg.norm(a - np.array([0.1, -0.2, 0.3]), axis=1), R, rtol=1e-6)


def test_fluid_block_lattice_and_jitter():
    mn, mx, r = (0.025, 0.025, 0.025), (1.625, 1.625, 1.625), 0.025
    a = sphg.fluid_block(mn, mx, r)
    b = po.fluid_block(mn, mx, r)
    assert len(a) == 32 ** 3
    assert np.array_equal(a, b.astype(np.float32))
    j1 = sphg.fluid_block(mn, mx, r, 0.01, 1)
    j2 = sphg.fluid_block(mn, mx, r, 0.01, 1)
    j3 = sphg.fluid_block(mn, mx, r, 0.01, 2)
    assert np.array_equal(j1, j2) and not np.array_equal(j1, j3)
    d = np.abs(j1.astype(np.float64) - b)
    assert d.max() <= 0.01 * 2 * r + 2e-7 and d.max() > 0.009 * 2 * r   # + FP32 rounding of the position


def test_splitmix_matches_library_jitter():
    """The Python SplitMix64 (used for scene seeds) and the library's
    jitter draw the same stream."""
    mn, mx, r = (0.0, 0.0, 0.0), (0.1, 0.1, 0.1), 0.025
    base = sphg.fluid_block(mn, mx, r).astype(np.float64)
    jit = sphg.fluid_block(mn, mx, r, 0.01, 42).astype(np.float64)
    rng = configs.SplitMix64(42)
    exp = base.copy()
    for i in range(len(base)):
        for a in range(3):
            exp[i, a] = float(np.float32(base[i, a] + (2 * rng.uniform() - 1) * 0.01 * 2 * r))
    assert np.allclose(jit, exp, atol=1e-7)


def test_workload_sizes():
    c1 = configs.wcsph_dam_32k()
    assert (c1.n, c1.nb) == (32768, 10474)
    c2 = configs.dfsph_dam_1m(n_side=20)
    assert c2.n == 8000
    assert np.all(c2.fluid.min(0) > 0) and np.all(c2.fluid.max(0) < np.array([0.8, 0.43, 0.43]))


def test_desc_validation_rejects_unsupported():
    lib = sphg.load()
    p = sphg.SceneParams(method=sphg.DFSPH)
    d = p.to_desc()
    d.boundary_method = 0
    h = C.c_void_p()
    # a NULL context is rejected before the description is looked at
    assert lib.sphg_scene_create(None, C.byref(d), C.byref(h)) == -1


def test_bench_valu_ceiling_from_profile():
    """bench.py's roofline.valu: the dominant kernel's VALU instructions per
    launch from the committed SQ counter summary, issued at 1024 SIMDs x
    2.4 GHz / 4 cycles, as a fraction of a given launch time."""
    import bench
    v = bench.valu_ceiling("df_err_density", 100e-6)
    assert v is not None and v["instr_per_launch"] > 0
    assert abs(v["issue_us"] - v["instr_per_launch"] * 4 / (1024 * 2.4e9) * 1e6) < 0.01
    assert abs(v["frac"] - v["issue_us"] / 100.0) < 1e-3
    assert 0 < v["busy_measured"] < 1
    assert bench.valu_ceiling("no_such_kernel", 1e-6) is None
