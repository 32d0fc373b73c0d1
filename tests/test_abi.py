"""The C-ABI library loads and exports every symbol include/opo.h declares;
the Python mirror binds all of them.  No compute calls (CPU-only)."""
import ctypes
import os
import re

from tests.conftest import ROOT


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "opo.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(opo_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("opo_create", "opo_upload", "opo_crystal_pass", "opo_run", "opo_download", "opo_destroy",
              "opo_batch_create", "opo_batch_run", "opo_strerror"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    import cuopo_amd as P
    L = ctypes.CDLL(P.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_python_mirror_binds_every_symbol():
    import cuopo_amd as P
    assert set(P.SIGNATURES) == set(declared_symbols())
    P.lib()  # binds argtypes for all


def test_strerror_codes():
    import cuopo_amd as P
    L = P.lib()
    assert L.opo_strerror(0) == b"success"
    assert L.opo_strerror(-1) == b"invalid argument"
    assert L.opo_strerror(-5) == b"non-finite intracavity power"
    assert L.opo_strerror(99) == b"unknown status"


def test_invalid_params_rejected_without_gpu():
    """Validation happens before any HIP call."""
    import cuopo_amd as P
    d = P.derive(P.Args(N=1000, NZ=4, NRT=2))
    p = P.params_from(d)
    try:
        P.Solver(p)
    except P.OpoError as e:
        assert e.status == P.ERR_INVALID
        assert "power of two" in str(e)
    else:
        raise AssertionError("N=1000 accepted")


def test_binding_checks_grid_and_field_lengths():
    """The binding refuses grids that are not exactly N long and fields
    shorter than N before anything reaches the C ABI (which reads N values)."""
    import numpy as np
    import pytest
    import cuopo_amd as P
    assert P._f64(np.arange(4), 4).dtype == np.float64
    with pytest.raises(ValueError):
        P._f64(np.zeros(3), 4)
    with pytest.raises(ValueError):
        P._f64(np.zeros(5), 4)
    assert P._f64(None, 4) is None
    with pytest.raises(ValueError):
        P._c128(np.zeros(3), 4)
