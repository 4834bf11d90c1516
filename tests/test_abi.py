"""The C ABI library loads and exports every function include/*.h declares
(no GPU calls)."""
import ctypes as C
import glob
import os
import re

import vvm_amd as V

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(vvm_\w+)\s*\(", src, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_the_abi():
    names = declared_functions()
    for must in ("vvm_dev_create", "vvm_dev_destroy", "vvm_dev_upload", "vvm_dev_download", "vvm_dev_step",
                 "vvm_dev_poisson", "vvm_dev_op", "vvm_dev_last_error", "vvm_init_case"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(V.LIB_PATH)
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_sizes_match_header():
    """The ctypes mirrors have the C layout (sizes computed by the compiler)."""
    lib = C.CDLL(V.LIB_PATH)
    assert lib.vvm_field_len(10, 7) == 70 and lib.vvm_hist_len(10, 7) == 140
    assert C.sizeof(V.VvmHostView) == 8 * (len(V.FIELDS2D) + len(V.HISTS) + len(V.ROWS) + len(V.PROFILES)) \
        + 8 * 3 + 16 + 8
