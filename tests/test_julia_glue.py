"""Static checks of the Julia glue mavi.jl_amd/julia/MaviHIP.jl (Julia is not
installed here, so the file is checked as text):

- every block it opens is closed, and the module closes last;
- every ccall names a function of include/mavi_hip.h with the same arity,
  argument kinds and return kind, and passes as many values as it declares;
- its isbits mirrors of mavi_sys_desc, mavi_wall and mavi_ring_source have
  the C field order, offsets and size (gcc offsetof on the header);
- every name it takes from Mavi (qualified, or one of the names it uses
  unqualified through `using`) is defined in the reference sources
  (tests/golden/reference_julia_names.json)."""
import json
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
GLUE = ROOT / "mavi.jl_amd" / "julia" / "MaviHIP.jl"
HEADER = ROOT / "include" / "mavi_hip.h"

OPENERS = {"function", "macro", "module", "baremodule", "struct", "if", "for", "while", "let", "begin", "quote",
           "do", "try"}
IDENT = re.compile(r"[A-Za-z_][A-Za-z_0-9!]*")


def tokens(src: str):
    """(kind, text, line) for identifiers/keywords and brackets, skipping
    comments, strings (with $(...) interpolation), chars and symbols."""
    i, n, line = 0, len(src), 1
    out = []

    def skip_string(i, quote):
        # i points just past the opening quote(s); returns the index past the closing one
        nonlocal line
        while i < n:
            if src.startswith(quote, i):
                return i + len(quote)
            c = src[i]
            if c == "\\":
                i += 2
                continue
            if c == "\n":
                line += 1
            if c == "$" and i + 1 < n and src[i + 1] == "(":
                depth, i = 1, i + 2
                while i < n and depth:
                    if src[i] == "(":
                        depth += 1
                    elif src[i] == ")":
                        depth -= 1
                    elif src[i] == '"':
                        i = skip_string(i + 1, '"') - 1
                    i += 1
                continue
            i += 1
        raise AssertionError(f"unterminated string at line {line}")

    while i < n:
        c = src[i]
        if c == "\n":
            line += 1
            i += 1
        elif src.startswith("#=", i):
            depth, i = 1, i + 2
            while i < n and depth:
                if src.startswith("#=", i):
                    depth, i = depth + 1, i + 2
                elif src.startswith("=#", i):
                    depth, i = depth - 1, i + 2
                else:
                    line += src[i] == "\n"
                    i += 1
        elif c == "#":
            while i < n and src[i] != "\n":
                i += 1
        elif src.startswith('"""', i):
            i = skip_string(i + 3, '"""')
        elif c == '"':
            i = skip_string(i + 1, '"')
        elif c == "'":
            prev = src[i - 1] if i else " "
            if prev.isalnum() or prev in "_)]}'":  # adjoint
                i += 1
            else:
                m = re.match(r"'(\\.[^']*|[^\\'])'", src[i:])
                assert m, f"bad char literal at line {line}"
                i += m.end()
        elif c == ":" and i + 1 < n and (src[i + 1].isalpha() or src[i + 1] == "_") and \
                (i == 0 or not (src[i - 1].isalnum() or src[i - 1] in "_)]}:")):
            m = IDENT.match(src, i + 1)
            out.append(("sym", m.group(0), line))
            i = m.end()
        elif c.isalpha() or c == "_":
            m = IDENT.match(src, i)
            out.append(("id", m.group(0), line))
            i = m.end()
        elif c in "([{":
            out.append(("open", c, line))
            i += 1
        elif c in ")]}":
            out.append(("close", c, line))
            i += 1
        else:
            i += 1
    return out


def block_structure(src: str):
    """Checks that blocks balance; returns the (keyword, open line, close
    line) list in closing order."""
    toks = tokens(src)
    stack = []  # (keyword, bracket depth, line)
    brackets = []
    closed = []
    prev = None
    for k, (kind, text, line) in enumerate(toks):
        if kind == "open":
            brackets.append(text)
        elif kind == "close":
            assert brackets, f"unbalanced {text} at line {line}"
            want = {")": "(", "]": "[", "}": "{"}[text]
            assert brackets[-1] == want, f"{text} closes {brackets[-1]} at line {line}"
            brackets.pop()
        elif kind == "id":
            depth = len(brackets)
            in_brackets = depth > (stack[-1][1] if stack else 0)
            if text == "end":
                if stack and stack[-1][1] == depth:
                    kw, _, l0 = stack.pop()
                    closed.append((kw, l0, line))
                else:
                    assert "[" in brackets, f"stray end at line {line}"
            elif text == "type" and prev in ("abstract", "primitive"):
                stack.append((prev + " type", depth, line))
            elif text == "struct" and prev == "mutable":
                stack.append(("mutable struct", depth, line))
            elif text in OPENERS:
                if text in ("for", "if") and in_brackets:
                    pass  # comprehension / generator clause
                elif text == "begin" and in_brackets and brackets[-1] == "[":
                    pass  # x[begin]
                else:
                    stack.append((text, depth, line))
            prev = text
            continue
        prev = None
    assert not brackets, f"unclosed brackets {brackets}"
    assert not stack, f"unclosed blocks: {[(kw, l) for kw, _, l in stack]}"
    return closed


def split_top(s: str):
    """Split on top-level commas (outside brackets and strings)."""
    parts, depth, cur, i = [], 0, "", 0
    while i < len(s):
        c = s[i]
        if c == '"':
            j = s.index('"', i + 1)
            cur += s[i:j + 1]
            i = j + 1
            continue
        if c in "([{":
            depth += 1
        elif c in ")]}":
            depth -= 1
        if c == "," and depth == 0:
            parts.append(cur.strip())
            cur = ""
        else:
            cur += c
        i += 1
    if cur.strip():
        parts.append(cur.strip())
    return parts


def balanced(src: str, i: int) -> int:
    """Index past the bracket that closes the one at src[i]."""
    depth = 0
    while True:
        c = src[i]
        if c == '"':
            i = src.index('"', i + 1)
        elif c in "([{":
            depth += 1
        elif c in ")]}":
            depth -= 1
            if depth == 0:
                return i + 1
        i += 1


def ccalls(src: str):
    src = re.sub(r"#.*", "", src)
    out = []
    for m in re.finditer(r"\bccall\(", src):
        end = balanced(src, m.end() - 1)
        args = split_top(src[m.end():end - 1])
        target = re.match(r"\(\s*:(\w+)\s*,\s*LIB\s*\)", args[0])
        assert target, f"ccall target {args[0]!r}"
        tup = args[2]
        assert tup.startswith("(") and tup.endswith(")"), tup
        types = split_top(tup[1:-1])
        out.append((target.group(1), args[1], types, args[3:]))
    return out


def header_prototypes():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    protos = {}
    for m in re.finditer(r"(const\s+char\s*\*|int)\s*(mavi_\w+)\s*\(([^)]*)\)\s*;", text):
        ret, name, params = m.groups()
        params = [p.strip() for p in params.split(",")] if params.strip() not in ("", "void") else []
        protos[name] = ("ptr" if "*" in ret else "i32", params)
    return protos


C_SCALAR = {"int": "i32", "int32_t": "i32", "int64_t": "i64", "uint64_t": "u64", "double": "f64", "float": "f32"}
JL_SCALAR = {"Cint": "i32", "Int32": "i32", "Int64": "i64", "UInt64": "u64", "Float64": "f64", "Float32": "f32",
             "Cfloat": "f32", "Cdouble": "f64"}
# pointee of a Julia Ptr{T} / Ref{T} -> the C pointee it stands for
JL_POINTEE = {"Float64": "double", "Int64": "int64_t", "Int32": "int32_t", "UInt16": "uint16_t", "Float32": "float",
              "UInt64": "uint64_t", "SysDesc": "mavi_sys_desc", "MaviRingSource": "mavi_ring_source",
              "MaviWall": "mavi_wall"}


def c_param_kind(p: str):
    p = re.sub(r"\bconst\b", "", p).strip()
    if "*" in p:
        base = p.split("*")[0].strip()
        stars = p.count("*")
        return ("ptr", base, stars)
    base = p.rsplit(None, 1)[0].strip()
    return (C_SCALAR[base], base, 0)


def jl_kind(t: str):
    t = t.strip()
    m = re.fullmatch(r"(Ptr|Ref)\{(.+)\}", t)
    if m:
        inner = m.group(2)
        if inner == "Cvoid":
            return ("ptr", None, 1)
        m2 = re.fullmatch(r"Ptr\{(.+)\}", inner)
        if m2:
            return ("ptr", None, 2)
        return ("ptr", JL_POINTEE.get(inner), 1)
    if t == "Cstring":
        return ("ptr", "char", 1)
    return (JL_SCALAR[t], t, 0)


def test_glue_blocks_close():
    src = GLUE.read_text()
    closed = block_structure(src)
    # the module is the first block opened and the last one closed
    assert closed[-1][0] == "module"
    assert re.search(r"^end # module\s*$", src, flags=re.M)
    # the source-record loop (a Sink branch with `continue`) closes inside its function
    fn = [c for c in closed if c[0] == "function"]
    assert len(fn) >= 30


def test_block_checker_catches_a_missing_end():
    """The checker itself: the shape of the missing `end` an earlier version
    of the glue had (an `if` ... `continue` with no `end`) fails it."""
    good = "module M\nfunction f(x)\n    for s in x\n        if s > 1\n            continue\n        end\n" \
           "        y = [i for i in 1:3 if i > 1]\n        z = x[end]\n    end\nend\nend # module\n"
    block_structure(good)
    bad = good.replace("            continue\n        end\n", "            continue\n")
    with pytest.raises(AssertionError):
        block_structure(bad)


def test_ccalls_match_header():
    protos = header_prototypes()
    calls = ccalls(GLUE.read_text())
    assert len(calls) >= 25
    seen = set()
    for name, ret, types, values in calls:
        assert name in protos, f"{name} is not declared in mavi_hip.h"
        cret, params = protos[name]
        assert (ret == "Cstring") == (cret == "ptr") and ret in ("Cint", "Cstring"), (name, ret)
        assert len(types) == len(params), f"{name}: {len(types)} argument types, C has {len(params)}"
        assert len(values) == len(types), f"{name}: {len(values)} values for {len(types)} types"
        for t, p in zip(types, params):
            jk, ck = jl_kind(t), c_param_kind(p)
            assert jk[0] == ck[0], f"{name}: {t} vs {p}"
            if jk[0] == "ptr":
                assert jk[2] == ck[2] or (jk[2] == 1 and jk[1] is None), f"{name}: {t} vs {p}"
                if jk[1] is not None:
                    assert jk[1] == ck[1], f"{name}: {t} vs {p}"
        seen.add(name)
    # the entry points the glue's step, batch and rings paths need
    for need in ("mavi_sys_create", "mavi_step", "mavi_rings_step", "mavi_rings_info", "mavi_batch_add_replica",
                 "mavi_batch_run", "mavi_batch_download", "mavi_sys_noise_set", "mavi_rings_var_get"):
        assert need in seen, need


JL_SIZE = {"Int64": 8, "UInt64": 8, "Float64": 8, "Int32": 4, "Cint": 4, "UInt32": 4}


def jl_struct_layout(src: str, name: str):
    m = re.search(rf"^struct {name}\b[^\n]*\n(.*?)^end", src, flags=re.S | re.M)
    assert m, name
    fields, off, maxal = [], 0, 1
    for line in m.group(1).splitlines():
        line = line.split("#")[0].strip()
        if not line:
            continue
        fname, ftype = [x.strip() for x in line.split("::")]
        t = re.fullmatch(r"NTuple\{(\d+),\s*(\w+)\}", ftype)
        if t:
            size, al = int(t.group(1)) * JL_SIZE[t.group(2)], JL_SIZE[t.group(2)]
        elif ftype.startswith("Ptr{"):
            size = al = 8
        else:
            size = al = JL_SIZE[ftype]
        off = (off + al - 1) // al * al
        fields.append((fname, off))
        off += size
        maxal = max(maxal, al)
    return fields, (off + maxal - 1) // maxal * maxal


@pytest.mark.parametrize("jl,c", [("SysDesc", "mavi_sys_desc"), ("MaviWall", "mavi_wall"),
                                  ("MaviRingSource", "mavi_ring_source")])
def test_glue_struct_layouts(tmp_path, jl, c):
    fields, size = jl_struct_layout(GLUE.read_text(), jl)
    src = tmp_path / "layout.c"
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "mavi_hip.h"\nint main(void) {\n'
                   f'  printf("%zu\\n", sizeof({c}));\n'
                   + "".join(f'  printf("%zu\\n", offsetof({c}, {f}));\n' for f, _ in fields)
                   + "  return 0;\n}\n")
    exe = tmp_path / "layout"
    r = subprocess.run(["gcc", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr  # a field name the C struct lacks fails here
    out = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert out[0] == size, f"{jl}: Julia size {size}, C {out[0]}"
    assert out[1:] == [o for _, o in fields], f"{jl}: offsets {[o for _, o in fields]} vs C {out[1:]}"
    # every C field is mirrored (no field left out at the end)
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    body = re.search(r"typedef struct \{((?:(?!typedef struct).)*?)\}\s*" + c + ";", text, flags=re.S).group(1)
    c_fields = [re.findall(r"(\w+)\s*(?:\[\d+\])?\s*$", decl)[0]
                for stmt in body.split(";") if stmt.strip() for decl in stmt.split(",")]
    assert [f for f, _ in fields] == c_fields


# names the glue takes unqualified through `using Mavi...`
UNQUALIFIED = ["StandardSys", "Sequencial", "RectangleCfg", "PeriodicWalls", "RigidWalls", "LenJonesCfg",
               "HarmTruncCfg", "SzaboCfg", "RunTumbleCfg", "SpaceCfg", "get_forces", "particle_radius", "System"]
MODULES = {"Mavi", "Integration", "Experiments", "Rings", "Configs", "Systems", "ChunksMod", "States",
           "NeighborsMod", "Sources", "MaviSerder"}


def test_glue_reference_names_exist():
    names = set(json.loads((ROOT / "tests" / "golden" / "reference_julia_names.json").read_text())["names"])
    src = re.sub(r"#.*", "", GLUE.read_text())
    src = re.sub(r'"""(.*?)"""', "", src, flags=re.S)
    missing = []
    for m in re.finditer(r"\b((?:Mavi|Integration|Experiments|Rings|Configs)(?:\.[A-Za-z_]\w*!?)+)", src):
        parts = m.group(1).split(".")
        for p in parts[1:-1]:
            if p not in MODULES:
                missing.append(m.group(1))
        if parts[-1] not in names and parts[-1] not in MODULES:
            missing.append(m.group(1))
    for u in UNQUALIFIED:
        if u not in names:
            missing.append(u)
    assert not missing, sorted(set(missing))
