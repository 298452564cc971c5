"""Writes tests/golden/reference_layout.json: the sizes, field offsets and
enum values of the boundary's types as the reference's own headers define
them on this platform (x86-64, gcc).

The probe tests/golden/layout_probe.c is compiled with -DPROBE_REFERENCE
against the reference's include/ directory (read in place, nothing copied)
and run; the binary goes to oracle/_ref/, which git ignores.  Run from the
repository root in a checkout that has the reference:

    python tests/golden/make_layout.py [/path/to/reference]
"""
from __future__ import annotations

import json
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]


def main(ref: Path) -> None:
    inc = ref / "include"
    out_dir = ROOT / "oracle" / "_ref"
    out_dir.mkdir(parents=True, exist_ok=True)
    exe = out_dir / "layout_probe_reference"
    subprocess.run(["gcc", "-std=gnu99", "-w", "-DPROBE_REFERENCE", f"-I{inc}", f"-I{inc / 'spimage'}",
                    str(HERE / "layout_probe.c"), "-o", str(exe)], check=True)
    layout = json.loads(subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout)
    dest = HERE / "reference_layout.json"
    dest.write_text(json.dumps(layout, indent=1, sort_keys=True) + "\n")
    print(f"{dest}: {len(layout)} entries")


if __name__ == "__main__":
    main(Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference"))
