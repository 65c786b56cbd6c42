"""The bindings of the integration guide at the repository root, checked against include/ars.h (the
Rust shim cannot be compiled here: no Rust toolchain).  Every function the
shim's `extern "C"` blocks declare exists in the header with the same
argument count, argument types and return type; every constant the shim
defines or the text quotes has the header's value; and the Python stub's
methods exist on the package's Codec."""
import importlib.util
import pathlib
import re

ROOT = pathlib.Path(__file__).resolve().parents[1]

RUST_TO_C = {
    "c_int": "int",
    "u32": "uint32_t",
    "u64": "uint64_t",
    "usize": "size_t",
    "*const u8": "const uint8_t *",
    "*mut u8": "uint8_t *",
    "*mut u64": "uint64_t *",
    "*mut c_void": "void *",
    "*const c_void": "const void *",
    "*mut ArsCtx": "ars_ctx *",
    "*mut *mut ArsCtx": "ars_ctx **",
}


def c_type(decl):
    """'const uint8_t *positions' -> 'const uint8_t *' (the parameter name dropped)."""
    decl = " ".join(decl.split())
    m = re.match(r"^(.*?)\s*\b[A-Za-z_]\w*$", decl)
    t = m.group(1) if m and m.group(1).strip() else decl
    stars = t.count("*")
    base = " ".join(t.replace("*", " ").split())
    return f"{base} {'*' * stars}" if stars else base


def header_functions():
    text = (ROOT / "include" / "ars.h").read_text()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    text = re.sub(r"^\s*#.*$", " ", text, flags=re.M)
    out = {}
    for stmt in text.split(";"):
        stmt = re.split(r"[{}]", stmt)[-1].strip()
        m = re.match(r"^([A-Za-z_][\w\s\*]*?)\b(ars_\w+)\s*\(([^()]*)\)$", stmt, flags=re.S)
        if not m:
            continue
        ret, name, args = m.group(1), m.group(2), m.group(3).strip()
        params = [] if args in ("", "void") else [c_type(a) for a in args.split(",")]
        out[name] = (c_type(ret + " x"), params)
    return out


def header_constants():
    text = (ROOT / "include" / "ars.h").read_text()
    return {m.group(1): int(m.group(2)) for m in re.finditer(r"#define\s+(ARS_\w+)\s+\(?(-?\d+)u?\)?", text)}


def rust_blocks():
    md = (ROOT / "INTEGRATION.md").read_text()
    return re.findall(r"```rust\n(.*?)```", md, flags=re.S)


def rust_externs():
    decls = {}
    for block in rust_blocks():
        for ext in re.findall(r'extern "C"\s*\{(.*?)\n\}', block, flags=re.S):
            for m in re.finditer(r"fn\s+(\w+)\s*\(([^)]*)\)\s*(?:->\s*([\w\s\*]+?))?\s*;", ext, flags=re.S):
                name, args, ret = m.group(1), m.group(2), (m.group(3) or "()").strip()
                params = []
                for a in filter(None, (x.strip() for x in args.split(","))):
                    params.append(" ".join(a.split(":", 1)[1].split()))
                decls[name] = (ret, params)
    return decls


def test_header_parser_sees_the_api():
    fns = header_functions()
    assert fns["ars_create"] == ("int", ["int", "ars_ctx **"])
    assert fns["ars_last_error"] == ("const char *", [])
    assert len(fns) >= 20


def test_rust_shim_declarations_match_the_header():
    fns = header_functions()
    decls = rust_externs()
    assert {"ars_create", "ars_destroy", "ars_encode", "ars_recover", "ars_encode_host",
            "ars_recover_host", "ars_decode"} <= set(decls)
    for name, (ret, params) in decls.items():
        assert name in fns, f"{name} is not in include/ars.h"
        c_ret, c_params = fns[name]
        assert RUST_TO_C[ret] == c_ret, (name, ret, c_ret)
        assert len(params) == len(c_params), (name, params, c_params)
        for i, (r, c) in enumerate(zip(params, c_params)):
            assert RUST_TO_C[r] == c, (name, i, r, c)


def test_quoted_constants_match_the_header():
    consts = header_constants()
    seen = 0
    for block in rust_blocks():
        for m in re.finditer(r"const\s+(ARS_\w+)\s*:\s*\w+\s*=\s*(\d+)\s*;", block):
            assert consts[m.group(1)] == int(m.group(2)), m.group(1)
            seen += 1
    md = (ROOT / "INTEGRATION.md").read_text()
    for m in re.finditer(r"`(ARS_\w+)` \(= (-?\d+)\)", md):
        assert consts[m.group(1)] == int(m.group(2)), m.group(1)
        seen += 1
    assert seen >= 3


def test_python_stub_methods_exist():
    spec = importlib.util.spec_from_file_location(
        "ars_amd_decls", ROOT / "additive-fft-reed-solomon_amd" / "__init__.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    md = (ROOT / "INTEGRATION.md").read_text()
    stub = re.search(r"```python\n(.*?)```", md, flags=re.S).group(1)
    for meth in set(re.findall(r"codec\.(\w+)\(", stub)):
        assert hasattr(mod.Codec, meth), meth


def load_pkg():
    spec = importlib.util.spec_from_file_location(
        "ars_amd_bind", ROOT / "additive-fft-reed-solomon_amd" / "__init__.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_ctypes_signatures_match_the_header():
    """Every ctypes signature the package installs has the header's arity,
    return type and scalar argument widths (any pointer binds as a pointer)."""
    import ctypes
    mod = load_pkg()
    scalar = {"int": ctypes.c_int, "uint32_t": ctypes.c_uint32, "size_t": ctypes.c_size_t,
              "uint64_t": ctypes.c_uint64, "int32_t": ctypes.c_int32, "const char *": ctypes.c_char_p}
    fns = header_functions()
    for name, (res, args) in mod.SIGNATURES.items():
        c_ret, c_params = fns[name]
        assert len(args) == len(c_params), (name, len(args), c_params)
        if c_ret in scalar:
            assert ctypes.sizeof(res) == ctypes.sizeof(scalar[c_ret]), (name, "return")
        for i, (a, c) in enumerate(zip(args, c_params)):
            if c.endswith("*"):
                assert a is ctypes.c_void_p or hasattr(a, "_type_") and ctypes.sizeof(a) == 8, (name, i, a, c)
            else:
                assert ctypes.sizeof(a) == ctypes.sizeof(scalar[c]), (name, i, a, c)
                assert (a in (ctypes.c_int, ctypes.c_int32)) == (c in ("int", "int32_t")), (name, i, "signedness")


def test_ctypes_structs_match_the_c_layout(tmp_path):
    """sizeof and every field offset of the package's ctypes structures equal
    the C compiler's for the header's structs."""
    import ctypes
    import subprocess
    mod = load_pkg()
    structs = {"ars_job": mod.Job, "ars_batch_info": mod.BatchInfo, "ars_tune_result": mod.TuneResult,
               "ars_server_config": mod.ServerConfig, "ars_request": mod.Request,
               "ars_server_stats": mod.ServerStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "ars.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'    printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'    printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    r = subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", str(ROOT / "include"), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        s, f, v = line.split()
        got[(s, f)] = int(v)
    for cname, py in structs.items():
        assert got[(cname, "sizeof")] == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)
