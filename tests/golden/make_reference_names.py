"""Writes reference_julia_names.json: the names Mavi.jl's sources define
(functions, short-form methods, structs, abstract types, modules, constants
and top-level globals, qualified or not), read as text from the Julia
sources under the reference checkout.  The static checks of the Julia glue
resolve every reference name the glue uses against this list.

    python tests/golden/make_reference_names.py [/path/to/Mavi.jl]
"""
import json
import re
import sys
from pathlib import Path

DEFS = [
    r"^\s*(?:@\w+\s+)*function\s+(?:[\w.]+\.)?([A-Za-z_]\w*!?)",
    r"^\s*(?:@inline\s+)?(?:[\w.]+\.)?([A-Za-z_]\w*!?)\s*(?:\{[^}]*\})?\([^=]*\)\s*(?:where\s+.*)?=[^=]",
    r"^\s*(?:@\w+\s+)*(?:mutable\s+)?struct\s+([A-Za-z_]\w*)",
    r"^\s*abstract\s+type\s+([A-Za-z_]\w*)",
    r"^\s*(?:bare)?module\s+([A-Za-z_]\w*)",
    r"^\s*const\s+([A-Za-z_]\w*)\s*=",
    r"^([A-Za-z_]\w*)\s*=[^=]",
]


def names(root: Path) -> list:
    out = set()
    for f in sorted((root / "src").rglob("*.jl")):
        for line in f.read_text(errors="replace").splitlines():
            for pat in DEFS:
                m = re.match(pat, line)
                if m:
                    out.add(m.group(1))
    return sorted(out)


if __name__ == "__main__":
    root = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    dest = Path(__file__).resolve().parent / "reference_julia_names.json"
    dest.write_text(json.dumps({"source": "Mavi.jl src/**/*.jl", "names": names(root)}, indent=0) + "\n")
    print(f"wrote {dest}")
