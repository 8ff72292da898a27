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
