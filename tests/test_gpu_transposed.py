"""The transposed Fourier-space pass (large 2-D arrays: W is transposed,
the column pass runs on contiguous columns, and W is transposed back) must
compute the same bits as the row-major pass.  The layout is fixed per shape
at its first use, so each side runs in a child process with the layout
forced (SPX_TR=1, or SPX_TR=0 with SPX_NO_T4=1), on non-square shapes so a
swapped index shows, for a single phaser with shrinkwrap, for a batch with
per-reconstruction algorithms and an idle slot, and for graph replays of a
batch against its eager iterations.  The shapes include lengths with radix-7
stages (which never take the 4-column tiled layout: their lines have no
multiple of 16 threads, or 4 of them do not fit in LDS)."""
import json
import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]

CHILD = r"""
import json, sys
root, out = sys.argv[1], sys.argv[2]
shapes = json.loads(sys.argv[3])
sys.path[:0] = [root, root + "/tests"]
import numpy as np
from libspimage_amd import spimage as sp
from _helpers import GpuPhaser, random_case
res = {}
# a single phaser, HIO with an area shrinkwrap every 5 iterations
shape = tuple(shapes["single"])
obj, amp, meas, sup, g0 = random_case(shape, seed=21)
bmap, amap = sp.smap([(0, 2.0)]), sp.smap([(0, 0.1)])
sarr = sp.sp_support_array_init(sp.sp_support_area_alloc(bmap, amap), 5)
with GpuPhaser(sp, "hio", amp, meas, sup, g0, beta=0.9, sup_array=sarr) as g:
    assert g.iterate(12) == 0
    res["layout_single"] = sp.spx_phaser_tiled(g.ph)
    res["model_single"] = g.get("model")
    res["ef_single"] = np.array(g.efourier())
# a batch: per-reconstruction algorithms and betas, one idle slot
bshape = tuple(shapes["batch"])
pats = [random_case(bshape, seed=30 + r) for r in range(4)]
b = sp.Batch(bshape, 4)
for r, (_, a, m, s_, _) in enumerate(pats):
    b.set_data(r, a, m, s_)
b.init_model(sp.SpModelRandomPhases, 5)
b.set_algorithm(sp.sp_phasing_hio_alloc(sp.smap([(0, 0.9)]), 0))
b.set_rec_algorithms(np.array([sp.SpHIO, sp.SpRAAR, sp.SpER, sp.SpHIO], np.int32),
                     np.array([0.9, 0.7, 0.0, 0.5], np.float32))
b.iterate(6)
b.set_active(np.array([1, 1, 0, 1], np.int32))
b.iterate(6)
res["model_batch"] = b.model()
res["ef_batch"] = b.efourier()[0]
b.close()
# graph replays (captured after the transposed pass's set-up) against eager
hio = sp.sp_phasing_hio_alloc(sp.smap([(0, 0.9)]), 0)
def graph_batch(graphs):
    g = sp.Batch(bshape, 3)
    for r, (_, a, m, s_, _) in enumerate(pats[:3]):
        g.set_data(r, a, m, s_)
    g.init_model(sp.SpModelRandomPhases, 9)
    g.set_algorithm(hio)
    g.use_graphs(graphs)
    for chunk in (5, 9, 4):
        g.iterate(chunk)
        g.synchronize()
    m_ = g.model()
    g.close()
    return m_
res["graph_equal"] = np.array(np.array_equal(graph_batch(True), graph_batch(False)))
np.savez(out, **res)
"""


def _child(tmp_path, name, shapes, **env_extra):
    env = dict(os.environ)
    for k in ("SPX_FFT_E", "SPX_NO_TUNE", "SPX_T4", "SPX_NO_T4", "SPX_COL_WIDE", "SPX_NT", "SPX_TR",
              "SPX_TUNE_FILE"):
        env.pop(k, None)
    env.update(env_extra)
    out = tmp_path / f"{name}.npz"
    print(f"child process {name} {shapes}: {env_extra}", flush=True)
    r = subprocess.run([sys.executable, "-c", CHILD, str(ROOT), str(out), json.dumps(shapes)], env=env,
                       capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr[-3000:]
    return dict(np.load(out))


KEYS = ("model_single", "ef_single", "model_batch", "ef_batch")


@pytest.mark.parametrize("shapes", [{"single": [96, 160], "batch": [160, 96]},
                                    {"single": [112, 140], "batch": [140, 112]},
                                    {"single": [8, 7168], "batch": [8, 7168]}])
def test_transposed_pass_equals_row_major(tmp_path, shapes):
    # the same variants on both sides: forced elements per thread
    tr = _child(tmp_path, "transposed", shapes, SPX_TR="1", SPX_FFT_E="8")
    rm = _child(tmp_path, "row_major", shapes, SPX_TR="0", SPX_NO_T4="1", SPX_FFT_E="8")
    assert int(tr["layout_single"]) == 2 and int(rm["layout_single"]) == 0
    for key in KEYS:
        assert np.array_equal(tr[key], rm[key]), key
    assert bool(tr["graph_equal"]) and bool(rm["graph_equal"])
