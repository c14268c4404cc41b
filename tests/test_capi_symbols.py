"""The C-ABI library loads without a GPU and exports every function that
include/*.h declares (no compute calls here)."""
import ctypes
import os
import re

from conftest import ROOT

# headers of the product library; hpm_types/philox/fmath hold inline code only
HEADERS = ["hpm_capi.h", "hpm_io.h", "hpm_scene.h", "hpm_prior.h", "hpm_pipeline.h", "hpm_batch.h", "hpm_unet.h",
           "hpm_fusion.h", "hpm_acmm.h"]


def declared_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b(hpm_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_headers_declare_functions():
    for h in HEADERS:
        assert declared_functions(os.path.join(ROOT, "include", h)), h


def test_library_exports_every_declared_symbol(hpm):
    lib = ctypes.CDLL(hpm.LIB_PATH)
    missing = []
    for h in HEADERS:
        for name in declared_functions(os.path.join(ROOT, "include", h)):
            if not hasattr(lib, name):
                missing.append(f"{h}:{name}")
    assert not missing, missing


def test_bindings_cover_every_declared_symbol(hpm):
    declared = set()
    for h in HEADERS:
        declared |= set(declared_functions(os.path.join(ROOT, "include", h)))
    assert declared <= set(hpm.SIGNATURES), sorted(declared - set(hpm.SIGNATURES))
    hpm.lib()  # binds every signature


def test_params_struct_matches_header(hpm):
    # hpm_params: 22 four-byte fields; hpm_camera: 25 four-byte fields
    assert ctypes.sizeof(hpm.HpmParams) == 88
    assert ctypes.sizeof(hpm.HpmCamera) == 100
