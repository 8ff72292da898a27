"""The C-ABI library loads and exports every function include/*.h declares (CPU only:
no compute calls)."""
import glob
import os
import re

import pytest

from rusty_h264_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(rh264_\w+)\s*\(", txt, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_functions():
    names = _declared()
    assert {"rh264_init", "rh264_build_planes", "rh264_full_search", "rh264_subpel_refine",
            "rh264_plain_search", "rh264_block_metric", "rh264_predict_blocks"} <= names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    missing = [n for n in sorted(_declared()) if not hasattr(lib, n)]
    assert not missing, missing
    # every declared function has a ctypes prototype (so nothing is called untyped)
    assert not (_declared() - set(_lib.PROTOTYPES)), sorted(_declared() - set(_lib.PROTOTYPES))


def test_lambda_table_host_side_matches_truncation():
    # rh264_lambda_table is pure host code: (int64)(lambda*b), mb16.rs:2563
    from rusty_h264_amd import accel
    lam = accel.lambda_me(26)
    lr = accel.lambda_table(lam)
    assert [int(v) for v in lr[:10]] == [int(lam * b) for b in range(10)]


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(RuntimeError):
        _lib.load()


# C name of every ctypes mirror field, struct by struct (the mirrors' own names where
# they differ from the header's)
_LAYOUTS = {
    "rh264_search_params": (_lib.SearchParams, {}),
    "rh264_diamond": (_lib.Diamond, {"lam": "lambda", "lam_fp": "lambda_fp"}),
    "rh264_stream": (_lib.Stream, {}),
    "rh264_seq": (_lib.Seq, {}),
}


def test_struct_layouts_match_the_header(tmp_path):
    """gcc compiles a probe against include/rh264_me.h that prints sizeof and every
    field's offsetof; the ctypes mirrors must agree field for field (a field added or
    retyped on one side only would shift every later one)."""
    import ctypes as C
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("no gcc")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "rh264_me.h"', "int main(void) {"]
    for cname, (cls, rename) in _LAYOUTS.items():
        lines.append(f'    printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            cf = rename.get(f, f)
            lines.append(f'    printf("{cname} {f} %zu\\n", offsetof({cname}, {cf}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-o", str(exe),
                    str(src)], check=True)
    got = {}
    for ln in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        s, f, v = ln.split()
        got[(s, f)] = int(v)
    for cname, (cls, _) in _LAYOUTS.items():
        assert got[(cname, "sizeof")] == C.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert got[(cname, f)] == getattr(cls, f).offset, (cname, f)
