"""The C-ABI library loads and exports every symbol declared in include/pk_api.h, and the
argument checks that run before any HIP call reject bad input (CPU only: no kernel runs)."""
import ctypes as C
import os
import re

import pytest

from jax_powspec_amd import _build, _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pk_api.h")


def header_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(pk_\w+)\s*\(", text, flags=re.M)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_build.lib_path()):
        pytest.fail("libpowspec_amd.so is not built (run __graft_entry__.build())")
    return _lib.load()


def test_header_declares_the_expected_surface():
    syms = header_symbols()
    for s in ("pk_scatter", "pk_r2c", "pk_powspec_mesh", "pk_estimate", "pk_plan_create", "pk_last_error"):
        assert s in syms


def test_every_header_symbol_is_exported(lib):
    missing = [s for s in header_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_bindings_cover_the_header():
    assert sorted(n for n, _, _ in _lib.SIGNATURES) == header_symbols()


def test_version_and_supported_sizes(lib):
    assert lib.pk_version() == 1
    for n in (8, 12, 64, 96, 128, 192, 256, 384, 512, 768, 1024):
        assert lib.pk_supported_n(n) == 1, n
    for n in (0, 7, 10, 100, 2048):
        assert lib.pk_supported_n(n) == 0, n


def test_invalid_arguments_fail_before_any_launch(lib):
    nbytes = C.c_size_t()
    assert lib.pk_scatter_workspace_size(1, 10, 0, None, C.byref(nbytes)) == _lib.PK_E_INVALID
    assert b"out of range" in lib.pk_last_error()
    bad_tile = (C.c_int * 3)(1, 4, 4)
    assert lib.pk_scatter_workspace_size(32, 10, 0, C.cast(bad_tile, C.c_void_p), C.byref(nbytes)) == _lib.PK_E_INVALID
    big_tile = (C.c_int * 3)(64, 64, 64)
    assert lib.pk_scatter_workspace_size(64, 10, 0, C.cast(big_tile, C.c_void_p), C.byref(nbytes)) == _lib.PK_E_INVALID
    assert lib.pk_scatter_workspace_size(64, 1000, 2, None, C.byref(nbytes)) == _lib.PK_OK
    assert nbytes.value > 1000 * 16
    # bad stencil / strides are rejected by pk_scatter before touching the device
    assert lib.pk_scatter(None, 1, None, 1, None, 1, None, 1, 0, 0.0, 0.0, 0.0, 1.0, 16, 7, 0, None, None,
                          C.c_void_p(8), None, None, 0, None) == _lib.PK_E_INVALID
    assert lib.pk_scatter(None, 0, None, 1, None, 1, None, 1, 0, 0.0, 0.0, 0.0, 1.0, 16, 0, 0, None, None,
                          C.c_void_p(8), None, None, 0, None) == _lib.PK_E_INVALID
    assert lib.pk_plan_create(16, None) == _lib.PK_E_INVALID


def test_extension_entry_points_validate_before_any_launch(lib):
    nbytes = C.c_size_t()
    # mock generator
    assert lib.pk_gaussian_workspace_size(7, C.byref(nbytes)) == _lib.PK_E_INVALID
    assert lib.pk_gaussian_workspace_size(64, C.byref(nbytes)) == _lib.PK_OK
    assert nbytes.value >= (3 * 32 * 32 + 1) * 4
    assert lib.pk_populate_workspace_size(0, C.byref(nbytes)) == _lib.PK_E_INVALID
    assert lib.pk_populate_workspace_size(2048, C.byref(nbytes)) == _lib.PK_E_INVALID
    assert lib.pk_populate_workspace_size(12, C.byref(nbytes)) == _lib.PK_OK
    assert nbytes.value >= 12 ** 3 // 2 * 2
    assert lib.pk_populate_count(None, 16, 1.0, 1.0, 0.0, 1, None, 0, None, None) == _lib.PK_E_INVALID
    assert lib.pk_populate_count(C.c_void_p(256), 16, 1.0, -1.0, 0.0, 1, None, 0, None, None) == _lib.PK_E_INVALID
    assert lib.pk_populate_count(C.c_void_p(256), 16, 1.0, 1.0, 0.0, 1, None, 0, None, None) == _lib.PK_E_WORKSPACE
    assert lib.pk_populate_place(C.c_void_p(256), None, 0.0, 16, 1.0, 0.0, 1, None, 0, None, -1, None) == \
        _lib.PK_E_INVALID
    # plan-based entry points reject a null plan
    assert lib.pk_gaussian_field(None, 1.0, None, None, 2, 0, 1, None, None, 0, None) == _lib.PK_E_INVALID
    assert lib.pk_c2r(None, None, None, None, None) == _lib.PK_E_INVALID
    assert lib.pk_gaussian_mesh(None, 1.0, None, None, 2, 0, 1, 9, None, None, None, 0, None) == _lib.PK_E_INVALID
    assert lib.pk_xi_workspace_size(None, 4, C.byref(nbytes)) == _lib.PK_E_INVALID
    assert lib.pk_bispec_workspace_size(None, 4, C.byref(nbytes)) == _lib.PK_E_INVALID


def test_python_error_mapping(lib):
    with pytest.raises(ValueError):
        _lib.check("pk_scatter_workspace_size", lib.pk_scatter_workspace_size(1, 1, 0, None, None))


def test_pk_job_layout_matches_the_header(tmp_path):
    """ctypes' PkJob has the size and field offsets of struct pk_job as a C compiler lays it out."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    fields = [f for f, _ in _lib.PkJob._fields_]
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "pk_api.h"\nint main(void){\n'
                   '  printf("%zu\\n", sizeof(pk_job));\n'
                   + "".join(f'  printf("%zu\\n", offsetof(pk_job, {f}));\n' for f in fields) + "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert out[0] == C.sizeof(_lib.PkJob)
    assert out[1:] == [getattr(_lib.PkJob, f).offset for f in fields]
