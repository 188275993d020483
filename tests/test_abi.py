"""The C-ABI library loads without a GPU and exports every declared symbol."""
import ctypes as ct
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "qclhip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qclhip_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_operator_set():
    names = declared()
    for required in ("qclhip_current_green", "qclhip_dyson_keldysh", "qclhip_contract",
                     "qclhip_fft_hilbert", "qclhip_lo_contraction", "qclhip_embedding_component",
                     "qclhip_boundary_flux", "qclhip_state_populations", "qclhip_seed_lesser"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from qclhip import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libqclhip.so not built (run __graft_entry__.build())")
    lib = ct.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers exactly the declared entry points
    assert sorted(_lib.declared_symbols()) == declared()
    _lib.load()
    assert lib.qclhip_max_basis() == 48


def test_missing_library_fails_loudly(tmp_path):
    from qclhip import _lib
    saved = _lib._lib
    try:
        _lib._lib = None
        with pytest.raises(ImportError):
            _lib.load(str(tmp_path / "absent.so"))
    finally:
        _lib._lib = saved


def test_error_codes_map_to_reference_exceptions():
    from qclhip import _lib
    with pytest.raises(_lib.SingularException) as e:
        _lib.check(7)
    assert e.value.block == 7
    for code, exc in ((_lib.ERR_ARGUMENT, _lib.ArgumentError), (_lib.ERR_DIMENSION, _lib.DimensionMismatch),
                      (_lib.ERR_DOMAIN, _lib.DomainError), (_lib.ERR_UNSUPPORTED, _lib.UnsupportedError)):
        with pytest.raises(exc):
            _lib.check(code)
