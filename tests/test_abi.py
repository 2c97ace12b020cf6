"""CPU: the C-ABI library loads and exports every entry point that
include/hkg.h declares; host-only entry points behave (no device calls)."""
from __future__ import annotations

import os
import re
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "hkg.h")


def declared() -> list[str]:
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(hkg_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    for must in ["hkg_open_dir", "hkg_knn", "hkg_knn_device", "hkg_close", "hkg_build", "hkg_write_dir",
                 "hkg_server_start", "hkg_submit", "hkg_poll", "hkg_server_stop", "hkg_inspect_dir", "hkg_autotune"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    from hercules_amd import hkg
    L = hkg.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", hkg.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (hkg_\w+)$", out, flags=re.M))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing
    # the Python binding covers the whole ABI, with no extra names
    assert set(hkg.SIGNATURES) == set(declared())
    for n in declared():
        assert getattr(L, n) is not None


def test_no_torch_or_hip_types_in_header():
    code = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    assert not re.search(r"\bhip\w*", code) and "torch" not in code
    assert "#include <hip" not in code


def test_status_strings():
    from hercules_amd import hkg
    for s in range(9):
        assert hkg.lib().hkg_status_str(s).decode() == hkg.STATUS[s]


def test_inspect_dir_errors(tmp_path):
    from hercules_amd import hkg
    with pytest.raises(hkg.HkgError) as e:
        hkg.inspect_dir(str(tmp_path))
    assert e.value.status == 3  # missing root.idx -> i/o error
    # truncated root.idx -> malformed
    (tmp_path / "root.idx").write_bytes(b"\x01\x00\x00")
    with pytest.raises(hkg.HkgError) as e:
        hkg.inspect_dir(str(tmp_path))
    assert e.value.status == 4


def test_query_opts_layout_and_defaults(tmp_path):
    """The ctypes mirror of hkg_query_opts has the C layout (a C program
    compiled against include/hkg.h prints the offsets), and
    hkg_default_query_opts sets the reference CLI's defaults
    (hercules.c:48-51, :70-72): exact search, 80-leaf budget, 0.25 / 0.9."""
    import ctypes as C

    from hercules_amd import hkg
    fields = ["path", "batch", "profile", "cand_cap", "bf16_bound_pass", "epsilon", "delta", "nprobes",
              "approx_leaves", "eapca_threshold", "sax_threshold", "batched_filter", "bf16_query_tiles"]
    sfields = ["device_ms", "bf16_pairs", "stage_ms", "stage_launches", "mode_leaves_first", "mode_sequential"]
    tfields = [f for f, _ in hkg.Tuning._fields_]
    prog = tmp_path / "layout.c"
    prog.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "hkg.h"\nint main(void){\n'
                    + "".join(f'printf("%zu\\n", offsetof(hkg_query_opts, {f}));\n' for f in fields)
                    + 'printf("%zu\\n", sizeof(hkg_query_opts));\n'
                    + "".join(f'printf("%zu\\n", offsetof(hkg_query_stats, {f}));\n' for f in sfields)
                    + 'printf("%zu\\n", sizeof(hkg_query_stats));\n'
                    + "".join(f'printf("%zu\\n", offsetof(hkg_tuning, {f}));\n' for f in tfields)
                    + 'printf("%zu\\n", sizeof(hkg_tuning));\nreturn 0;}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(REPO, "include"), str(prog), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [getattr(hkg.QueryOpts, f).offset for f in fields] + [C.sizeof(hkg.QueryOpts)]
    want += [getattr(hkg.QueryStats, f).offset for f in sfields] + [C.sizeof(hkg.QueryStats)]
    want += [getattr(hkg.Tuning, f).offset for f in tfields] + [C.sizeof(hkg.Tuning)]
    assert got == want
    qo = hkg.query_opts()
    assert qo.epsilon == 0.0 and qo.nprobes == 0 and qo.delta == 1.0 and qo.approx_leaves == 80
    assert abs(qo.eapca_threshold - 0.25) < 1e-7 and abs(qo.sax_threshold - 0.9) < 1e-7
    with pytest.raises(TypeError):
        hkg.query_opts(epsilonn=1.0)
