"""The C-ABI library loads and exports every symbol include/*.h declares
(CPU only: no compute call is made here)."""
import ctypes
import subprocess

import pytest

from clbool_amd import _capi


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_capi.LIB_PATH)
    declared = _capi.declared_symbols()
    assert len(declared) >= 40
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, missing
    # the Python bindings cover the same set
    bound = {n for n, _, _ in _capi.SIGNATURES}
    assert set(declared) <= bound, sorted(set(declared) - bound)


def test_exported_symbols_are_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", _capi.LIB_PATH], capture_output=True, text=True).stdout
    names = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for n in _capi.declared_symbols():
        assert n in names, n


def test_bindings_load():
    L = _capi.load()
    assert L.clb_last_error() is not None


def test_kernel_registry_names():
    import clbool_amd as cb
    cb.kernel_lookup("prefix_sum", "scan_blelloch")
    cb.kernel_lookup("hash/hash_tb", "hash_numeric_tb")
    with pytest.raises(cb.ClboolError) as e:
        cb.kernel_lookup("prepare_positions", "some name")  # test/test_errors.cpp:25-36
    assert e.value.status_name == "CLBOOL_CREATE_KERNEL_ERROR"
    with pytest.raises(cb.ClboolError) as e:
        cb.kernel_lookup("no_such_program", "x")
    assert e.value.status_name == "CLBOOL_NO_SUCH_PROGRAM"


def test_invalid_platform_id_raises():
    import clbool_amd as cb
    # env.cpp:15-20 -- an unknown platform id is an initialisation error
    with pytest.raises(cb.ClboolError) as e:
        cb.create_controls(platform_id=7)
    assert e.value.status_name == "CLBOOL_INITIALIZATION_ERROR"
