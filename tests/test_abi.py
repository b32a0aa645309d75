"""CPU checks of the C-ABI library: it loads and exports every entry point
declared in include/gv.h.  No compute calls (there is no GPU here)."""
import ctypes
import os
import re

import gv_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(gv_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(gv_amd.LIB_PATH)
    decl = declared_symbols()
    assert len(decl) >= 20
    missing = [n for n in sorted(decl) if not hasattr(L, n)]
    assert not missing, missing
    assert set(gv_amd.EXPORTED) <= decl


def test_python_binding_loads_and_config_defaults():
    cfg = gv_amd.default_config()
    assert cfg.block_slots == 64
    assert cfg.pr_thread_max <= cfg.pr_group_max <= cfg.pr_wave_max
    assert 0 < cfg.compact_threshold < 1
    cfg2 = gv_amd.default_config(block_slots=128)
    assert cfg2.block_slots == 128


def test_struct_layout_matches_header():
    # gv_config: u32, (pad), u64, f64, f32, 3 x u32, 3 x int, (pad) -> 56 bytes
    assert ctypes.sizeof(gv_amd.GvConfig) == 56
    assert ctypes.sizeof(gv_amd.GvBatchStats) == 48


def test_cpp_facade_compiles_and_links(tmp_path):
    import shutil
    import subprocess
    if not shutil.which("g++"):
        import pytest
        pytest.skip("g++ not available")
    src = tmp_path / "use_gv.cpp"
    src.write_text(
        '#include "gv.hpp"\n'
        "int main() {\n"
        "  std::vector<uint32_t> s{0, 1}, d{1, 2};\n"
        "  gv::GraphVine g(3, s, d);\n"
        "  std::vector<float> x;\n"
        "  gv::static_pagerank(&g, 0.85f, 1e-6f, 10, x);\n"
        "  std::vector<uint64_t> t;\n"
        "  gv::static_tc(&g, t);\n"
        "  return (int)gv::bc(&g, {0}).size() == 3 ? 0 : 1;\n"
        "}\n")
    lib_dir = os.path.dirname(gv_amd.LIB_PATH)
    r = subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src),
                        "-L", lib_dir, "-lgv", "-o", str(tmp_path / "use_gv")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_struct_layouts_match_c(tmp_path):
    """sizeof and every field offset of the public structs, compiled by the
    host C compiler from the headers, equal the ctypes mirrors."""
    import shutil
    import subprocess
    import pytest
    from gv_amd import serve
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    pairs = [("gv_config", gv_amd.GvConfig), ("gv_batch_stats", gv_amd.GvBatchStats),
             ("gv_pagerank_stats", gv_amd.GvPagerankStats),
             ("gv_pagerank_opts", gv_amd.GvPagerankOpts), ("gv_request", serve.GvRequest),
             ("gv_response", serve.GvResponse), ("gv_server_opts", serve.GvServerOpts),
             ("gv_server_stats", serve.GvServerStats)]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "gv_serve.h"',
             'int main(void) {']
    expect = []
    for cname, ct in pairs:
        lines.append(f'  printf("%zu\\n", sizeof({cname}));')
        expect.append(ctypes.sizeof(ct))
        for f, _ in ct._fields_:
            lines.append(f'  printf("%zu\\n", offsetof({cname}, {f}));')
            expect.append(getattr(ct, f).offset)
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    assert got == expect
