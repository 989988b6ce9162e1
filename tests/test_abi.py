"""The C-ABI library loads and exports every entry point include/cv2_md.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "cv2_md.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cv2_[a-z_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("cv2_create", "cv2_destroy", "cv2_set_box", "cv2_set_lj", "cv2_upload_atoms",
              "cv2_build_neighbors", "cv2_compute_forces", "cv2_run_vv", "cv2_run_langevin",
              "cv2_get_pairs", "cv2_download", "cv2_set_ghosts"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from cv2md import device
    lib = ctypes.CDLL(device.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_header():
    from cv2md import device
    L = device.lib()
    for s in declared_symbols():
        assert getattr(L, s).restype is not None or s == "cv2_last_error"
