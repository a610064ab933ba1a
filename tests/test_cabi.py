"""The C-ABI library: loads without a GPU, exports every symbol that
include/wrfdc.h declares, carries the product's field table and rejects
unsupported configurations with the documented codes (checked before any
device allocation, so these run on the CPU)."""

import ctypes
import os
import re

import pytest

from wrfdc import _lib
from wrfdc.config import CConfig, DycoreConfig
from wrfdc.fields import FIELD_SPECS

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(REPO, "include", "wrfdc.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wrfdc_[a-z_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libwrfdc.so is not built (run __graft_entry__.build())")
    return _lib.load()


def test_exports_every_declared_symbol(lib):
    syms = header_symbols()
    assert len(syms) >= 12
    raw = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in syms if not hasattr(raw, s)]
    assert not missing, missing
    assert set(syms) == set(_lib.SIGNATURES), "ctypes signature table differs from the header"


def test_field_table_matches_registry(lib):
    assert lib.wrfdc_field_count() == len(FIELD_SPECS)
    for i, (name, kind) in enumerate(FIELD_SPECS):
        assert lib.wrfdc_field_name(i).decode() == name
        assert chr(lib.wrfdc_field_kind(i)) == kind
    assert lib.wrfdc_field_name(len(FIELD_SPECS)) is None
    assert lib.wrfdc_field_name(-1) is None


def test_config_struct_layout():
    # 14 ints then 25 doubles, no interior padding
    assert ctypes.sizeof(CConfig) == 14 * 4 + 25 * 8
    assert CConfig.dx.offset == 56


def _create(lib, **over):
    base = dict(nx=24, ny=20, nz=10, ip=32, nmoist=1, nscalar=0)
    base.update(over)
    cfg = DycoreConfig(**base).as_struct()
    ctx = ctypes.c_void_p()
    return lib.wrfdc_create(ctypes.byref(cfg), ctypes.byref(ctx)), ctx


@pytest.mark.parametrize(
    "over,code",
    [
        (dict(nx=8), _lib.WRFDC_EINVAL),
        (dict(nz=3), _lib.WRFDC_EINVAL),
        (dict(ip=20), _lib.WRFDC_EINVAL),
        (dict(nmoist=0), _lib.WRFDC_EINVAL),
        (dict(dt=0.0), _lib.WRFDC_EINVAL),
        (dict(spec_zone=2, spec_bdy_width=6), _lib.WRFDC_ENOTIMPL),
        (dict(top_lid=1), _lib.WRFDC_ENOTIMPL),
        (dict(spec_bdy_width=6), _lib.WRFDC_EINVAL),
        (dict(nx=4095, ny=2047, nz=127, ip=4096), _lib.WRFDC_EINVAL),  # 2^30 elements per 3-D field
    ],
)
def test_create_rejects_bad_configs(lib, over, code):
    rc, ctx = _create(lib, **over)
    assert rc == code
    assert not ctx.value


def test_create_failure_message_reaches_python(lib):
    """wrfdc_create has no context to hold its message: wrfdc_last_error(NULL)
    returns it (per thread), and raise_for puts it into the exception."""
    rc, ctx = _create(lib, nx=8)
    assert rc == _lib.WRFDC_EINVAL
    assert b"domain too small" in lib.wrfdc_last_error(None)
    with pytest.raises(ValueError, match="domain too small"):
        _lib.raise_for(rc, lib, None, "wrfdc_create")
    rc, ctx = _create(lib, top_lid=1)
    assert rc == _lib.WRFDC_ENOTIMPL
    with pytest.raises(NotImplementedError, match="wrfdc_create: .+"):
        _lib.raise_for(rc, lib, None, "wrfdc_create")


def test_error_mapping_uses_reference_exception_types(lib):
    with pytest.raises(ValueError):
        _lib.raise_for(_lib.WRFDC_EINVAL, lib, None, "x")
    with pytest.raises(NotImplementedError):
        _lib.raise_for(_lib.WRFDC_ENOTIMPL, lib, None, "x")
    with pytest.raises(FloatingPointError):
        _lib.raise_for(_lib.WRFDC_ENONFINITE, lib, None, "x")
    with pytest.raises(KeyError):
        _lib.raise_for(_lib.WRFDC_EUNBOUND, lib, None, "x")


def test_python_validation_matches_library():
    with pytest.raises(ValueError):
        DycoreConfig(nx=8, ny=20, nz=10, ip=16, nmoist=1, nscalar=0).validate()
    with pytest.raises(ValueError):
        DycoreConfig(nx=4095, ny=2047, nz=127, ip=4096, nmoist=1, nscalar=0).validate()
    DycoreConfig(nx=4095, ny=2047, nz=126, ip=4096, nmoist=1, nscalar=0).validate()
    with pytest.raises(ValueError):
        DycoreConfig(nx=24, ny=20, nz=10, ip=32, nmoist=1, nscalar=0, options={"rk_ord": 2}).validate()
    with pytest.raises(NotImplementedError):
        DycoreConfig(nx=24, ny=20, nz=10, ip=32, nmoist=1, nscalar=0, options={"diff_opt": 1}).validate()


def test_product_has_no_cpu_fallback(monkeypatch, tmp_path):
    """Without the shared object every entry point raises; nothing falls back."""
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(_lib, "_LIB", None)
    with pytest.raises(_lib.LibraryMissingError):
        _lib.load()
    from wrfdc.device import Domain

    with pytest.raises(_lib.LibraryMissingError):
        Domain(DycoreConfig(nx=24, ny=20, nz=10, ip=32, nmoist=1, nscalar=0))
