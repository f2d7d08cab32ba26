"""Block-size / scan-chunk library variants give the default build's
results bit for bit (the reductions add per-chunk partials whose grouping
changes with the block size, so this is a check, not a tautology): each
built variant steps a small DFSPH dam break in a process of its own (a
process loads one libsphg) and its final positions and iteration count are
compared with the default library's."""
import os

import pytest

from sph_amd import tune_layout

pytestmark = pytest.mark.gpu

VARIANTS = [(128, 1024), (512, 1024), (256, 256), (256, 4096)]


def test_layout_variants_match_default():
    built = [v for v in VARIANTS if os.path.exists(tune_layout.lib_path(*v))]
    if not built:
        pytest.skip("no layout variant is built (make -C sph_amd/csrc BLOCK=... CELL_CHUNK=...)")
    r = tune_layout.tune("dfsph_small", [tune_layout.DEFAULT] + built, timeout=120.0, log=lambda s: None)
    rows = {(x["block"], x["scan_chunk"]): x for x in r["variants"]}
    ref = rows[tune_layout.DEFAULT]
    for v in built:
        assert rows[v]["x_max_diff_vs_default"] == 0.0, (v, rows[v])
        assert rows[v]["density_iterations"] == ref["density_iterations"], (v, rows[v])
