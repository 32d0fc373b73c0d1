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


def test_out_of_range_params_rejected_without_gpu():
    """Grid sizes outside [256, 2^18], slices or round trips below one, an
    unknown precision or regime, and a non-positive or non-finite step are
    refused by validation, before any HIP call."""
    import math
    import pytest
    import cuopo_amd as P
    base = P.params_from(P.derive(P.Args(N=1024, NZ=4, NRT=2)))
    bad = [("N", 128, "power of two"), ("N", 1 << 19, "power of two"), ("N", 3 << 10, "power of two"),
           ("NZ", 0, "NZ and NRT"), ("NRT", 0, "NZ and NRT"), ("precision", 7, "precision"),
           ("regime", 5, "regime"), ("Nrts", -1, "Nrts"), ("dz", 0.0, "dz"), ("dz", math.nan, "dz"),
           ("dk", math.inf, "dk")]
    for field, value, msg in bad:
        p = P.Params.from_buffer_copy(base)
        setattr(p, field, value)
        with pytest.raises(P.OpoError) as e:
            P.Solver(p)
        assert e.value.status == P.ERR_INVALID and msg in str(e.value), (field, value, str(e.value))
    # batches: at least one instance, and one shared equation count / regime / NZ / precision
    with pytest.raises(P.OpoError) as e:
        P.Batch([])
    assert "count must be >= 1" in str(e.value)
    for field, value in [("NZ", 5), ("three_eqs", 1 - base.three_eqs), ("regime", 1 - base.regime),
                         ("precision", 1 - base.precision)]:
        q = P.Params.from_buffer_copy(base)
        setattr(q, field, value)
        with pytest.raises(P.OpoError) as e:
            P.Batch([base, q])
        assert "must share" in str(e.value), field


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
