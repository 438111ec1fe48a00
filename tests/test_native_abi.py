"""The C-ABI library loads and exports every entry point ``include/tenet_amd.h`` declares
(no compute call: this runs without a GPU)."""

import re
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_library_exports_every_declared_symbol():
    from tenet import _native

    header = (ROOT / "include" / "tenet_amd.h").read_text()
    declared = set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(tn_\w+)\s*\(", header, re.M))
    assert declared >= {"tn_recouple", "tn_gemm_grouped", "tn_dot", "tn_axpby"}
    lib = _native.load_library()
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert set(_native._SIGNATURES) == declared
    assert lib.tn_version() >= 1
