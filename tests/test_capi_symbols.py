"""The C-ABI library loads and exports every symbol include/mavi_hip.h declares
(no compute call: this runs without a GPU)."""
import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(mavi_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_header_symbols():
    from mavi_hip.build import library_path

    lib = ctypes.CDLL(str(library_path()))
    syms = declared_symbols()
    assert "mavi_step" in syms and "mavi_bin" in syms
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_header():
    from mavi_hip._lib import SIGNATURES

    extra_known = set()
    missing = [s for s in declared_symbols() if s not in SIGNATURES and s not in extra_known]
    # every C entry point has a ctypes signature, or is bound by a sibling module
    from mavi_hip import _lib

    bound = set(SIGNATURES)
    for mod in ("batch", "learned"):
        try:
            m = __import__(f"mavi_hip.{mod}", fromlist=["SIGNATURES"])
            bound |= set(getattr(m, "SIGNATURES", {}))
        except ImportError:
            pass
    missing = [s for s in declared_symbols() if s not in bound]
    assert not missing, missing


def test_abi_version_callable():
    from mavi_hip.build import library_path

    lib = ctypes.CDLL(str(library_path()))
    v = re.search(r"#define MAVI_ABI_VERSION (\d+)", (ROOT / "include" / "mavi_hip.h").read_text())
    assert lib.mavi_abi_version() == int(v.group(1))


def test_descriptor_layout_matches_header(tmp_path):
    """The ctypes mirror of mavi_sys_desc has the C layout (gcc on the header)."""
    import subprocess

    from mavi_hip._lib import SysDesc

    fields = [f for f, _ in SysDesc._fields_]
    src = tmp_path / "layout.c"
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "mavi_hip.h"\nint main(void) {\n'
                   '  printf("%zu\\n", sizeof(mavi_sys_desc));\n'
                   + "".join(f'  printf("%zu\\n", offsetof(mavi_sys_desc, {f}));\n' for f in fields)
                   + "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    out = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert out[0] == ctypes.sizeof(SysDesc)
    assert out[1:] == [getattr(SysDesc, f).offset for f in fields]


def test_host_layout_helpers():
    import numpy as np
    from mavi_hip.configs import as_points
    from mavi_hip.init_states import rectangular_grid

    r = 2 ** (1 / 6) / 2
    pos, g = rectangular_grid(32, 32, 0.4, r)
    assert pos.shape == (1024, 2)
    assert g.length == g.height
    assert abs(g.length - 77.2 * r) < 1e-12  # 32*2r + 0.4 r * 33
    assert abs(pos[0, 0] - 1.4 * r) < 1e-15 and abs(pos[1, 0] - pos[0, 0] - 2.4 * r) < 1e-14
    m = np.arange(6.0).reshape(2, 3)
    assert as_points(m).shape == (3, 2)
