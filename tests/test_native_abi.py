"""The C-ABI library loads and exports every function include/pyprop_amd.h declares
(no device calls: runs without a GPU)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "pyprop_amd.h"
LIB = ROOT / "pyprop_amd" / "_lib" / "libpyprop_amd.so"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s+(pp_[a-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("pp_tensor_potential_multiply", "pp_tp_create", "pp_tp_multiply", "pp_overlap_apply",
                     "pp_gmres_solve", "pp_pamp_propagate", "pp_cayley_step", "pp_precond_solve"):
        assert required in names


@pytest.mark.skipif(not LIB.exists(), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(LIB))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert missing == []


def test_python_binding_covers_header():
    from pyprop_amd import _native
    bound = {n for n, _, _ in _native.SIGNATURES}
    assert set(declared_functions()) <= bound


@pytest.mark.skipif(not LIB.exists(), reason="library not built")
def dropin_functions():
    text = (ROOT / "include" / "pyprop_amd_dropin.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*void\s+(tensorpotentialmultiply_[a-z_]+_)\s*\(", text, flags=re.M)))


@pytest.mark.skipif(not LIB.exists(), reason="library not built")
def test_library_exports_the_dropin_symbols():
    """The generated kernels' names (generator.py:1428-1451) for the storage combinations
    of the helium and hydrogen examples."""
    names = dropin_functions()
    assert names == ["tensorpotentialmultiply_diag_bandnh_", "tensorpotentialmultiply_simp_bandnh_",
                     "tensorpotentialmultiply_simp_simp_bandnh_"]
    lib = ctypes.CDLL(str(LIB))
    assert all(hasattr(lib, n) for n in names)


@pytest.mark.skipif(not LIB.exists(), reason="library not built")
def test_version_and_error_strings():
    from pyprop_amd import _native
    lib = _native.load()
    assert b"gfx950" in lib.pp_version()
    # an invalid call reports through pp_last_error without touching a device
    rc = lib.pp_overlap_create(0, 0, None, None)
    assert rc == -1
    assert lib.pp_last_error()
