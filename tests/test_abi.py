"""CPU checks of the C ABI: the built library loads and exports every entry
point declared in include/neb_hip.h, with ctypes signatures for each."""
import ctypes
import os
import re

import pytest

from neb_amd import _hip as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "neb_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(neb_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_bound_surface():
    decl = declared_symbols()
    assert decl, "no declarations parsed"
    assert sorted(H.SIGNATURES) == decl


def test_library_exports_every_declared_symbol():
    if not os.path.exists(H.LIB_PATH):
        pytest.fail("libnebhip.so is not built (run python -m neb_amd.build)")
    lib = ctypes.CDLL(H.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name
    L = H.load()
    assert L.neb_abi_version() == H.ABI_VERSION


def test_argument_errors_raise_without_touching_the_gpu():
    H.load()
    # unknown precision / dimension are rejected before any launch
    with pytest.raises(ValueError):
        H.call("neb_jet_layer_forward", 7, 2, 0, 8, 8, 8, 16, 8, 16, 8, 0, 1.0, 16, 16, 8, None)
    with pytest.raises(ValueError):
        H.call("neb_jet_layer_forward", 0, 4, 0, 8, 8, 8, 16, 8, 16, 8, 0, 1.0, 16, 16, 8, None)
    with pytest.raises(ValueError):
        H.call("neb_cross", 0, 0, None, None, None, None, None, None, 0, 0, 4, 4, None, 0, 1.0, 0.0, 16, 4, 16,
               0, None)
    # split fields: bad dimension, split Gram: null lo plane, reverse: fields
    # dtype that does not fit the precision
    with pytest.raises(ValueError):
        H.call("neb_jet_output_fields_split", 7, 8, 8, 8, 16, 8, 16, 8, 16, 16, 16, 16, 8, 0, None)
    # centering: a lo plane needs bf16 planes, a row stride below the width
    with pytest.raises(ValueError):
        H.call("neb_center_columns", H.F32, 16, 16, 8, 8, 8, 16, None)
    with pytest.raises(ValueError):
        H.call("neb_center_columns", H.BF16, 16, 16, 8, 8, 4, 16, None)
    with pytest.raises(ValueError):
        H.call("neb_gram_split", 16, None, 8, 8, 8, 1.0, 0.0, 16, 8, 16, 1 << 20, None)
    with pytest.raises(ValueError):
        H.call("neb_jet_output_backward", H.F32, 2, 0, 8, 8, 8, 16, 16, 8, H.BF16, 0, 16, 16, 8, 16, 8, 1.0, 16, 8,
               None)
    # inverse Cholesky factor: T missing, T aliasing A, T's row stride below K
    with pytest.raises(ValueError):
        H.call("neb_spd_inverse_chol", 8, 16, 8, 32, 8, None, 8, 48, 64, 1 << 20, None)
    with pytest.raises(ValueError):
        H.call("neb_spd_inverse_chol", 8, 16, 8, 32, 8, 16, 8, 48, 64, 1 << 20, None)
    with pytest.raises(ValueError):
        H.call("neb_spd_inverse_chol", 8, 16, 8, 32, 8, 40, 4, 48, 64, 1 << 20, None)
    assert H.load().neb_cross_workspace_size(1024, 1024, 131072, 1) > 0


def test_cross_occupancy_knob_arguments():
    """neb_set_cross_occupancy takes a workgroup cap >= 0 (host-side setting only)."""
    from neb_amd import _hip as H
    lib = H.load()
    assert lib.neb_set_cross_occupancy(-1) == 4  # NEB_ERR_ARG
    for v in (1, 128, 100000, 0):
        assert lib.neb_set_cross_occupancy(v) == 0


def test_split_precision_entry_points_check_arguments_without_the_gpu():
    H.load()
    # unknown dimension, null lo plane with points to process, bad leading dimension
    with pytest.raises(ValueError):
        H.call("neb_jet_layer_forward_x3", 4, 0, 8, 8, 8, 16, 16, 8, 16, 16, 8, 0, 1.0, 16, None, 16, 16, 8, None)
    with pytest.raises(ValueError, match="null"):
        H.call("neb_jet_layer_forward_x3", 2, 0, 8, 8, 8, 16, None, 8, 16, 16, 8, 0, 1.0, 16, None, 16, 16, 8, None)
    with pytest.raises(ValueError):
        H.call("neb_jet_output_fields_x3", 3, 8, 8, 16, 16, 16, 8, 16, 16, 8, 16, 16, 16, 16, 8, 0, None)
    with pytest.raises(ValueError):
        H.call("neb_split_bf16", H.BF16, 16, 4, 4, 4, 16, 16, 4, None)
    with pytest.raises(ValueError, match="null"):
        H.call("neb_split_bf16", H.F32, 16, 4, 4, 4, None, 16, 4, None)
    # zero points: sizes are checked, nothing is launched, null buffers are accepted
    H.call("neb_jet_layer_forward_x3", 2, 0, 0, 8, 8, None, None, 8, None, None, 8, None, 1.0, None, None, None, None,
           8, None)
    H.call("neb_split_bf16", H.F32, None, 0, 4, 4, None, None, 4, None)


def test_missing_library_fails_loudly(tmp_path):
    """No CPU or PyTorch fallback: without libnebhip.so every entry point raises."""
    import subprocess
    import sys
    code = ("import neb_amd._hip as H; H.LIB_PATH = '/nonexistent/libnebhip.so'; H._lib = None\n"
            "try:\n    H.load()\nexcept RuntimeError as e:\n    print('raised', e)\n")
    p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT, timeout=120)
    assert p.returncode == 0 and "raised" in p.stdout and "not built" in p.stdout, p.stdout + p.stderr
