"""The C ABI library on the CPU: it loads, exports every function its headers
declare, its ctypes mirror has the C struct layouts, and its host-only
pieces (generators, text format, CLI argument checks) behave.  No call here
touches a GPU."""
from __future__ import annotations

import ctypes as C
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from graphs import grid, random_graph

HDR = ROOT / "include" / "dmolp_hip.h"


def _lib_path():
    from dmolp_amd import _abi
    return _abi.LIB_PATH


def declared_c_functions():
    text = re.sub(r"/\*.*?\*/", "", HDR.read_text(), flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?(?:int|void|char)\s*\*?\s*(dmolp_\w+)\s*\(",
                                 text, flags=re.M)))


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", "-C", str(_lib_path())],
                         check=True, capture_output=True, text=True).stdout
    return out


def test_library_exports_every_declared_function():
    names = declared_c_functions()
    assert len(names) >= 30, names
    syms = exported_symbols()
    defined = {ln.split()[-1] for ln in syms.splitlines() if " T " in ln}
    missing = [n for n in names if n not in defined]
    assert not missing, missing
    # The C++ entry points with the reference's signatures.
    for sig in ("run_phase1(int, int, char const*, int, int, bool, bool, Graph&, GhostNodes&)",
                "run_phase2(int, int, int, Graph&, GhostNodes&, int)",
                "run_phase2_ex(int, int, int, Graph&, GhostNodes&, int, int, double)"):
        assert sig in syms, sig


def test_ctypes_signatures_cover_the_header():
    from dmolp_amd import _abi
    names = declared_c_functions()
    missing = [n for n in names if n not in _abi.SIGNATURES]
    assert not missing, missing


def test_loads_without_gpu_and_answers_host_queries():
    from dmolp_amd import _abi
    lib = _abi.lib()
    assert lib.dmolp_abi_version() == int(re.search(r"DMOLP_ABI_VERSION (\d+)", HDR.read_text())[1])
    for code in range(0, -10, -1):
        s = lib.dmolp_status_string(code)
        assert s and len(s) >= 2
    assert lib.dmolp_last_error(None) is not None  # never NULL


STRUCTS = {
    "Tuning": "dmolp_tuning", "GenSpec": "dmolp_gen_spec", "Csr": "dmolp_csr",
    "Phase1Metrics": "dmolp_phase1_metrics", "Phase2Metrics": "dmolp_phase2_metrics",
    "Trace": "dmolp_trace", "Job": "dmolp_job", "Result": "dmolp_result",
}


def test_ctypes_struct_layouts_match_the_header(tmp_path):
    from dmolp_amd import _abi
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "dmolp_hip.h"', "int main(void){"]
    for py, c in STRUCTS.items():
        st = getattr(_abi, py)
        lines.append(f'printf("{py} size %zu\\n", sizeof({c}));')
        for f, _ in st._fields_:
            lines.append(f'printf("{py} {f} %zu\\n", offsetof({c}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for ln in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        a, b, v = ln.split()
        got[(a, b)] = int(v)
    for py in STRUCTS:
        st = getattr(_abi, py)
        assert got[(py, "size")] == C.sizeof(st), py
        for f, _ in st._fields_:
            assert got[(py, f)] == getattr(st, f).offset, (py, f)


def _check_csr(off, col, n):
    assert off[0] == 0 and len(off) == n + 1 and off[-1] == len(col)
    deg = np.diff(off)
    assert (deg >= 0).all()
    rows = np.repeat(np.arange(n), deg)
    assert (col >= 0).all() and (col < n).all()
    assert not (rows == col).any(), "self-loop"
    # sorted strictly ascending within each row
    same = rows[1:] == rows[:-1]
    assert (col[1:][same] > col[:-1][same]).all()
    # symmetric
    a = np.sort(rows.astype(np.int64) * n + col)
    b = np.sort(col.astype(np.int64) * n + rows)
    assert (a == b).all()


@pytest.mark.parametrize("spec_args", [("rmat", 10, 16, 1), ("rmat", 12, 8, 7), ("chunglu", 5000, 40000, 3)])
def test_host_generator_csr_is_valid_and_deterministic(spec_args):
    from dmolp_amd import api
    kind = spec_args[0]
    if kind == "rmat":
        spec = lambda: api.rmat_spec(spec_args[1], spec_args[2], seed=spec_args[3])
        n = 1 << spec_args[1]
    else:
        spec = lambda: api.chung_lu_spec(spec_args[1], spec_args[2], gamma=2.0, dmax=500.0,
                                         seed=spec_args[3])
        n = spec_args[1]
    off, col = api.gen_host(spec())
    _check_csr(off, col, n)
    off2, col2 = api.gen_host(spec())
    assert (off == off2).all() and (col == col2).all()
    if kind == "rmat":
        # another seed gives another graph
        off3, col3 = api.gen_host(api.rmat_spec(spec_args[1], spec_args[2], seed=spec_args[3] + 1))
        assert len(col3) != len(col) or not (col3 == col).all()


def test_grid_generator_is_the_5_point_mesh():
    from dmolp_amd import api
    off, col = api.gen_host(api.grid_spec(7, 5))
    ref_off, ref_col = grid(7, 5)
    assert (off == ref_off).all() and (col == ref_col).all()


def test_text_format_round_trip(tmp_path):
    from dmolp_amd import api
    off, col = random_graph(300, 1500, seed=4)
    p = tmp_path / "g.txt"
    api.write_adjacency_text(p, off, col)
    lines = p.read_text().splitlines()
    assert int(lines[0]) == 300 and len(lines) == 301
    off2, col2 = api.read_adjacency_text(p)
    assert (off2 == off).all() and (col2 == col).all()


def test_text_reader_normalises_rows(tmp_path):
    """Rows are sorted and deduplicated, self-loops dropped; a last line
    without a newline is kept."""
    from dmolp_amd import api
    p = tmp_path / "g.txt"
    p.write_text("4\n3 1 1 0\n0 2 1\n1\n0")
    off, col = api.read_adjacency_text(p)
    assert list(off) == [0, 2, 4, 5, 6]
    assert list(col) == [1, 3, 0, 2, 1, 0]


def test_text_reader_vertex_count_is_the_line_count(tmp_path):
    """The header's V is replaced by the number of adjacency lines, as the
    reference's Allreduce of the line count does (init.cpp:96)."""
    from dmolp_amd import api
    p = tmp_path / "g.txt"
    p.write_text("7\n1\n0 2\n1\n")
    off, col = api.read_adjacency_text(p)
    assert list(off) == [0, 1, 3, 4] and list(col) == [1, 0, 2, 1]
    p.write_text("2\n1\n0 2\n1\n")
    off, col = api.read_adjacency_text(p)
    assert len(off) == 4


@pytest.mark.parametrize("body", [
    "2\n4294967297\n0\n",            # wraps to 1 in a 32-bit parser
    "2\n1\n99999999999999999999999\n",  # overflows 64 bits
    "2\n1\n2\n",                      # id outside [0, V)
])
def test_text_reader_rejects_ids_out_of_range(tmp_path, body):
    from dmolp_amd import api
    p = tmp_path / "g.txt"
    p.write_text(body)
    with pytest.raises(api.DmolpError) as e:
        api.read_adjacency_text(p)
    assert e.value.status == -9


def test_binary_csr_round_trip_and_checks(tmp_path):
    from dmolp_amd import api
    off, col = random_graph(500, 3000, seed=8)
    p = tmp_path / "g.csr"
    api.write_csr_binary(p, off, col)
    assert p.stat().st_size == 32 + 8 * len(off) + 4 * len(col)
    off2, col2 = api.read_csr_binary(p)
    assert (off2 == off).all() and (col2 == col).all()
    raw = bytearray(p.read_bytes())
    for bad in (raw[:-4],                                # truncated
                raw[:3] + b"X" + raw[4:],                # magic
                raw[:32 + 8] + (10 ** 9).to_bytes(8, "little") + raw[32 + 16:]):  # offset
        q = tmp_path / "bad.csr"
        q.write_bytes(bytes(bad))
        with pytest.raises(api.DmolpError) as e:
            api.read_csr_binary(q)
        assert e.value.status == -9
    col3 = col.copy()
    col3[7] = len(off)  # a column outside [0, n)
    api.write_csr_binary(tmp_path / "c.csr", off, col3)
    with pytest.raises(api.DmolpError):
        api.read_csr_binary(tmp_path / "c.csr")


def test_cached_loader_writes_then_reads_the_cache(tmp_path):
    import os
    from dmolp_amd import api
    off, col = random_graph(400, 2000, seed=9)
    p = tmp_path / "g.txt"
    api.write_adjacency_text(p, off, col)
    o1, c1, hit1 = api.load_adjacency_cached(p)
    assert not hit1 and (tmp_path / "g.txt.csr").exists()
    o2, c2, hit2 = api.load_adjacency_cached(p)
    assert hit2 and (o2 == off).all() and (c2 == col).all() and (o1 == off).all()
    # A text file newer than its cache is parsed again (and re-cached).
    off4, col4 = random_graph(300, 900, seed=10)
    api.write_adjacency_text(p, off4, col4)
    st = (tmp_path / "g.txt.csr").stat()
    os.utime(p, ns=(st.st_atime_ns, st.st_mtime_ns + 10 ** 9))
    o3, c3, hit3 = api.load_adjacency_cached(p)
    assert not hit3 and (o3 == off4).all() and (c3 == col4).all()
    # A damaged cache falls back to the text file.
    (tmp_path / "g.txt.csr").write_bytes(b"garbage")
    os.utime(p, ns=(st.st_atime_ns, st.st_mtime_ns - 10 ** 9))
    o5, c5, hit5 = api.load_adjacency_cached(p)
    assert not hit5 and (o5 == off4).all()


def test_cli_rejects_bad_arguments_before_touching_the_device():
    exe = ROOT / "dmolp_amd" / "hpc_partitioning"
    if not exe.exists():
        pytest.skip("CLI not built")
    for args in ([], ["-m", "bfs", "g.txt", "4"], ["-m", "sideways", "g.txt", "4", "10"],
                 ["-m", "bfs", "g.txt", "4", "0"], ["--bogus"]):
        r = subprocess.run([str(exe), *args], capture_output=True, text=True, timeout=30)
        assert r.returncode == 1, (args, r.stdout, r.stderr)
        assert "Usage" in r.stdout


def test_no_product_module_imports_the_oracle():
    """The product (package, library sources, bench timing path) never
    routes through the oracle."""
    for p in list((ROOT / "dmolp_amd").rglob("*.py")) + list((ROOT / "dmolp_amd").rglob("*.hip")) \
            + list((ROOT / "dmolp_amd").rglob("*.cpp")) + list((ROOT / "dmolp_amd").rglob("*.hpp")):
        text = p.read_text()
        assert "oracle" not in text.lower() or "ref_cpu" not in text, p
        assert "libref_cpu" not in text and "oracle_py" not in text, p
