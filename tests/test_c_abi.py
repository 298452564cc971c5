"""The boundary as a C caller sees it: gcc builds against include/ and links
libspimage_hip.so the way the reference's C tests (tests/PhasingTests.c) and
Hawk link libspimage.

The struct layouts and enum values are compared with
tests/golden/reference_layout.json, which tests/golden/make_layout.py wrote by
compiling the same probe against the reference's own headers."""
import ctypes as C
import json
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
LIBDIR = ROOT / "libspimage_amd" / "_lib"


def reference_layout():
    return json.loads((GOLDEN / "reference_layout.json").read_text())


def _gcc(args, out):
    r = subprocess.run(["gcc", "-std=gnu99", "-Wall", "-Werror", *args, "-o", str(out)], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr
    return out


def test_header_layout_equals_reference(tmp_path):
    exe = _gcc([f"-I{ROOT / 'include'}", str(GOLDEN / "layout_probe.c")], tmp_path / "probe")
    ours = json.loads(subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout)
    ref = reference_layout()
    assert set(ours) == set(ref)
    diff = {k: (ours[k], ref[k]) for k in ref if ours[k] != ref[k]}
    assert not diff, diff


def test_ctypes_mirror_layout_equals_reference():
    """libspimage_amd/spimage.py declares the same structs for Python callers."""
    from libspimage_amd import spimage as sp
    ref = reference_layout()
    structs = {"Image": sp.Image, "Detector": sp.Detector, "sp_3matrix": sp.sp_3matrix, "sp_list": sp.sp_list,
               "sp_smap": sp.sp_smap, "SpPhasingAlgorithm": sp.SpPhasingAlgorithm,
               "SpSupportAlgorithm": sp.SpSupportAlgorithm, "SpSupportArray": sp.SpSupportArray,
               "SpPhaser": sp.SpPhaser}
    checked = 0
    for key, want in ref.items():
        if key.startswith("sizeof "):
            name = key.split()[1]
            if name in structs and name != "SpPhaser":
                assert C.sizeof(structs[name]) == want, key
                checked += 1
        elif "." in key:
            name, field = key.split(".")
            if name in structs:
                assert getattr(structs[name], field).offset == want, key
                checked += 1
        elif key.startswith("enum "):
            name = key.split()[1]
            if hasattr(sp, name):
                assert getattr(sp, name) == want, key
                checked += 1
    assert checked > 80


def _caller_defines():
    ref = reference_layout()
    pick = {"REF_SIZEOF_IMAGE": "sizeof Image", "REF_SIZEOF_SPSUPPORTARRAY": "sizeof SpSupportArray",
            "REF_SPSUPPORTARRAY_UPDATE_PERIOD": "SpSupportArray.update_period",
            "REF_SPPHASER_ALGORITHM": "SpPhaser.algorithm", "REF_SPPHASER_SUP_ALGORITHM": "SpPhaser.sup_algorithm",
            "REF_SPPHASER_ITERATION": "SpPhaser.iteration", "REF_SPPHASER_MODEL": "SpPhaser.model",
            "REF_SPPHASER_ENGINE": "SpPhaser.engine", "REF_SPPHASER_GP": "SpPhaser.gp"}
    return [f"-D{m}={ref[k]}" for m, k in pick.items()]


def build_c_caller(out_dir: Path) -> Path:
    lib = LIBDIR / "libspimage_hip.so"
    assert lib.exists(), "build the library first: python -m libspimage_amd.build"
    return _gcc([*_caller_defines(), f"-I{ROOT / 'include'}", str(ROOT / "tests" / "c" / "phasing_convergence.c"),
                 f"-L{LIBDIR}", "-lspimage_hip", f"-Wl,-rpath,{LIBDIR}", "-lm"], out_dir / "phasing_convergence")


def test_c_caller_compiles_links_and_loads(tmp_path):
    exe = build_c_caller(tmp_path)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "layout ok" in r.stdout, r.stderr


def test_c_caller_rejects_a_mismatched_layout(tmp_path):
    """The compile-time layout check is live: a wrong expected size fails the build."""
    src = ROOT / "tests" / "c" / "phasing_convergence.c"
    defs = [d if not d.startswith("-DREF_SIZEOF_IMAGE=") else "-DREF_SIZEOF_IMAGE=48" for d in _caller_defines()]
    r = subprocess.run(["gcc", "-std=gnu99", "-c", *defs, f"-I{ROOT / 'include'}", str(src), "-o",
                        str(tmp_path / "x.o")], capture_output=True, text=True)
    assert r.returncode != 0 and "Image differs from the reference" in r.stderr


@pytest.mark.gpu
def test_c_caller_convergence_on_gpu(tmp_path, sp):
    exe = build_c_caller(tmp_path)
    r = subprocess.run([str(exe), "run"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    assert "converged" in r.stdout


def test_launch_parameters_match_reference_rule():
    """src/cuda_util.c:39-60: block 64 doubled while size > 2^31-1 blocks;
    grid = ceil(size / block).  No device is needed."""
    from libspimage_amd import spimage as sp
    g, b = C.c_int(), C.c_int()
    for size, want in ((1, (1, 64)), (64, (1, 64)), (65, (2, 64)), (1 << 20, (1 << 14, 64)),
                       ((2**31 - 1) * 64, (2**31 - 1, 64)), ((2**31 - 1) * 64 + 1, (1 << 30, 128))):
        sp.sp_cuda_launch_parameters(size, C.byref(g), C.byref(b))
        assert (g.value, b.value) == want, size
