"""Tuning records (SPX_TUNE_FILE): every choice the tuner times is appended
to the file, and a later process reads the file before its first choice and
makes the same choices without timing anything.  Each case runs a child
process, since a process reads the file once, before its first choice."""
import json
import os
import subprocess
import sys
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]

CHILD = r"""
import json, sys
root = sys.argv[1]
sys.path[:0] = [root, root + "/tests"]
from libspimage_amd import spimage as sp
from _helpers import GpuPhaser, random_case
out = {"variant": list(sp.fft_variant(256))}
obj, amp, meas, sup, g0 = random_case((256, 256), seed=1)
with GpuPhaser(sp, "hio", amp, meas, sup, g0) as g:
    out["tiled"] = sp.spx_phaser_tiled(g.ph)
    out["nt"] = sp.spx_phaser_nt_hints(g.ph)
print(json.dumps(out))
"""


def _child(path, **extra):
    env = dict(os.environ, SPX_TUNE_FILE=str(path))
    for k in ("SPX_FFT_E", "SPX_NO_TUNE", "SPX_T4", "SPX_NO_T4", "SPX_COL_WIDE", "SPX_NT"):
        env.pop(k, None)
    env.update(extra)
    print(f"child process, SPX_TUNE_FILE={path} {extra}", flush=True)
    r = subprocess.run([sys.executable, "-c", CHILD, str(ROOT)], env=env, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_records_are_read_before_the_first_choice(tmp_path):
    # mask 5 is not among the masks the tuner times, so only the file can set it
    f = tmp_path / "tune.txt"
    f.write_text("shape 256 256 1 8 4 0\n"
                 "shape 999 999 1 8 8 0\n"       # no plan for 999: skipped
                 "pair 256 256 0 1 0 5\n"
                 "not a record\n")
    before = f.read_text()
    out = _child(f)
    assert out == {"variant": [8, 4], "tiled": 1, "nt": 5}
    assert f.read_text() == before  # nothing was timed, so nothing was appended


def test_choices_are_recorded_and_replayed(tmp_path):
    f = tmp_path / "tune.txt"
    first = _child(f)
    lines = f.read_text().splitlines()
    shape = [ln.split() for ln in lines if ln.startswith("shape 256 256 1 ")]
    pair = [ln.split() for ln in lines if ln.startswith("pair 256 256 0 ")]
    assert len(shape) == 1 and len(pair) == 1, lines
    ex, ey, ez = map(int, shape[0][4:7])
    assert ex in (4, 8, 16) and ey in (4, 8, 16) and ez == 0
    assert [ex, ey] == first["variant"]
    tiled, wide, hints = map(int, pair[0][4:7])
    assert tiled == first["tiled"] and hints == first["nt"] and wide in (0, 1)
    again = _child(f)
    assert again == first
    assert f.read_text().splitlines() == lines


def test_forced_row_and_column_elements(tmp_path):
    # SPX_FFT_E=<row>,<column> forces the variants (no tuning, nothing recorded)
    f = tmp_path / "tune.txt"
    out = _child(f, SPX_FFT_E="16,4", SPX_NT="3")
    assert out["variant"] == [16, 4] and out["nt"] == 3
    assert not f.exists() or "shape 256 256 1" not in f.read_text()
