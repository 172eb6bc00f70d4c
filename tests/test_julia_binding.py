"""The Julia backend (julia/src/backends/mi355x.jl) against the C ABI, on the
CPU: Julia is not in this image, so the file is checked as text.

- every ccall names a function include/fomo_mi355x.h declares, with as many
  argument types as the declaration has parameters, each of a matching kind
  (Int32 / int32_t, Ptr{Float32} / float*, Ptr{Int32} / int32_t*, handles),
  and the declared return type;
- FomoModelDesc has fomo_model_desc's fields, in order, with their widths;
- the methods that the unchanged run_shot! / run_shots! / init_medium call
  (shot_manager.jl:46-72,145-222, init.jl:93-97) exist for MI355XBackend, and
  Wavefield(nx, nz, ::MI355XBackend) builds the reference's Wavefield type
  from HipMatrix <: AbstractMatrix{Float32} (Structures.jl:18-31);
- block keywords and `end`s balance, and brackets nest."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL = os.path.join(ROOT, "julia", "src", "backends", "mi355x.jl")
HDR = os.path.join(ROOT, "include", "fomo_mi355x.h")


def _strip_jl(text: str) -> str:
    text = re.sub(r'"""(.|\n)*?"""', '""', text)
    text = re.sub(r'"(\\.|[^"\\])*"', '""', text)
    return re.sub(r"#.*", "", text)


def _split_top(s: str) -> list:
    """Splits s at commas outside brackets."""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "([{":
            depth += 1
        elif ch in ")]}":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def _matching(s: str, i: int) -> int:
    """Index of the bracket closing the one at s[i]."""
    depth = 0
    for k in range(i, len(s)):
        if s[k] in "([{":
            depth += 1
        elif s[k] in ")]}":
            depth -= 1
            if depth == 0:
                return k
    raise ValueError("unbalanced")


def header_decls() -> dict:
    text = re.sub(r"/\*.*?\*/", "", open(HDR).read(), flags=re.S)
    decls = {}
    for m in re.finditer(r"([A-Za-z_][\w\s\*]*?)\b(fomo_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", text):
        ret = " ".join(m.group(1).split())
        params = [p.strip() for p in m.group(3).split(",")] if m.group(3).strip() not in ("", "void") else []
        decls[m.group(2)] = (ret, params)
    return decls


def julia_ccalls() -> list:
    src = _strip_jl(open(JL).read())
    calls = []
    for m in re.finditer(r"\bccall\(", src):
        close = _matching(src, m.end() - 1)
        args = _split_top(src[m.end():close])
        name = re.match(r"\(:(fomo_[a-z0-9_]+),\s*LIBFOMO\)", args[0])
        assert name, f"ccall target is not (:fomo_*, LIBFOMO): {args[0]}"
        types = args[2]
        assert types.startswith("(") and types.endswith(")"), types
        tlist = _split_top(types[1:-1])
        calls.append((name.group(1), args[1], tlist, args[3:]))
    return calls


def _kind_c(param: str) -> str:
    p = " ".join(param.replace("const", "").split())
    if "**" in p.replace(" ", "") or re.search(r"\*\s*const\s*\*", param):
        return "ptr-to-handle"
    if p.startswith("float") and "*" in p:
        return "float*"
    if p.startswith("int32_t") and "*" in p:
        return "int32*"
    if p.startswith("fomo_model_desc"):
        return "desc*"
    if "*" in p:
        return "handle"
    if p.startswith("int32_t"):
        return "int32"
    if p.startswith("int64_t"):
        return "int64"
    raise AssertionError(param)


def _kind_jl(t: str) -> str:
    t = t.replace(" ", "")
    return {"Int32": "int32", "Int64": "int64", "Ptr{Float32}": "float*", "Ptr{Int32}": "int32*",
            "Ptr{Cvoid}": "handle", "Ref{Ptr{Cvoid}}": "ptr-to-handle",
            "Ptr{Ptr{Cvoid}}": "ptr-to-handle", "Ref{FomoModelDesc}": "desc*"}[t]


def test_every_ccall_matches_its_declaration():
    decls = header_decls()
    assert len(decls) >= 50
    calls = julia_ccalls()
    assert len(calls) >= 30
    for name, ret, types, args in calls:
        assert name in decls, f"{name} is not declared in include/fomo_mi355x.h"
        cret, params = decls[name]
        assert len(types) == len(params), f"{name}: {len(types)} argument types, header has {len(params)}"
        assert len(args) == len(types), f"{name}: {len(args)} arguments for {len(types)} types"
        for t, p in zip(types, params):
            assert _kind_jl(t) == _kind_c(p), f"{name}: Julia {t} for C `{p}`"
        want_ret = {"int32_t": "Int32", "const char*": "Cstring", "int64_t": "Int64"}[cret.replace(" *", "*")]
        assert ret.strip() == want_ret, f"{name}: returns {ret}, header {cret}"


def test_the_calls_the_run_path_needs_are_bound():
    names = {c[0] for c in julia_ccalls()}
    for need in ("fomo_ctx_create", "fomo_ctx_last_error", "fomo_ctx_synchronize",
                 "fomo_model_create", "fomo_model_destroy", "fomo_wavefield_create",
                 "fomo_wavefield_destroy", "fomo_wavefield_reset", "fomo_wavefield_download",
                 "fomo_wavefield_upload", "fomo_shot_create", "fomo_shot_destroy",
                 "fomo_shot_gather", "fomo_time_step", "fomo_run_time_loop",
                 "fomo_batch_create", "fomo_batch_run", "fomo_batch_destroy",
                 "fomo_backup_boundary", "fomo_inject_source", "fomo_update_velocity",
                 "fomo_apply_habc_velocity", "fomo_update_stress", "fomo_apply_habc_stress",
                 "fomo_apply_free_surface", "fomo_record_receivers", "fomo_frames_create",
                 "fomo_run_time_loop_frames", "fomo_frames_download", "fomo_frames_destroy"):
        assert need in names, need


def test_model_desc_matches_the_header_struct():
    text = re.sub(r"/\*.*?\*/", "", open(HDR).read(), flags=re.S)
    body = re.search(r"typedef struct fomo_model_desc \{(.*?)\} fomo_model_desc;", text, re.S).group(1)
    cfields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        ctype, names = decl.split(None, 1)
        for n in names.split(","):
            cfields.append((n.strip(), {"int32_t": "Int32", "float": "Float32"}[ctype]))
    src = _strip_jl(open(JL).read())
    body = re.search(r"struct FomoModelDesc\n(.*?)\nend", src, re.S).group(1)
    jfields = [tuple(x.strip() for x in line.split("::")) for line in body.strip().splitlines()]
    assert jfields == cfields


def test_methods_of_the_unchanged_api_exist_for_the_backend():
    src = _strip_jl(open(JL).read())
    assert re.search(r"struct MI355XBackend <: AbstractBackend end", src)
    assert re.search(r"mutable struct HipMatrix <: AbstractMatrix\{Float32\}", src)
    for name in ("reset!", "run_time_loop!", "run_shots!", "time_step!", "_create_source",
                 "_create_receivers", "_get_gather", "to_device", "synchronize", "Wavefield",
                 "backup_boundary!", "inject_source!", "update_velocity!", "apply_habc_velocity!",
                 "update_stress!", "apply_habc_stress!", "apply_free_surface!", "record_receivers!"):
        pat = re.escape(name) + r"\([^)]*::MI355XBackend"
        assert re.search(pat, src), f"no {name} method for ::MI355XBackend"
    # the unchanged run_shot! takes W::Wavefield: the backend's Wavefield
    # constructor returns the reference's parametric type, not a new one
    ctor = re.search(r"function Wavefield\(nx::Int, nz::Int, ::MI355XBackend\)(.*?)\nend", src, re.S)
    assert ctor and re.search(r"return Wavefield\((f\(\d\),\s*){9}f\(\d\)\)", ctor.group(1))
    assert "HipMatrix(d, Int32(id)" in ctor.group(1)
    # indices come from the reference's own _coord_to_index
    assert src.count("_coord_to_index(") >= 2 and "function _coord_to_index" not in src


def test_blocks_and_brackets_balance():
    src = _strip_jl(open(JL).read())
    src = re.sub(r"\[[^\[\]]*\bend\b[^\[\]]*\]", "[]", src)  # a[end] is not a block end
    # `for`/`if`/`while` open a block at the start of a statement; inside a
    # comprehension or generator they do not
    openers = len(re.findall(r"\b(function|try|begin|let|struct|do|quote|macro|module)\b", src))
    openers += len(re.findall(r"^\s*(for|if|while)\b", src, re.M))
    ends = len(re.findall(r"\bend\b", src))
    assert openers == ends, (openers, ends)
    depth = 0
    for ch in src:
        depth += ch in "([{"
        depth -= ch in ")]}"
        assert depth >= 0
    assert depth == 0
