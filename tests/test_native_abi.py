"""libqbx.so loads on a machine without a GPU and exports every symbol include/qbx.h
declares; compute entry points are not called here."""
import ctypes
import re
from pathlib import Path

import pytest

from qubo_lib_amd import _native

HEADER = Path(__file__).resolve().parents[1] / "include" / "qbx.h"
CPU_HEADER = HEADER.with_name("qbx_cpu.h")


def declared_functions(header=HEADER):
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qbx_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(_native.EXPORTED_SYMBOLS)


def test_library_exports_every_declared_symbol():
    lib = _native.load_library()
    for name in declared_functions():
        assert isinstance(getattr(lib, name), ctypes._CFuncPtr), name


def test_cpu_library_exports_every_declared_symbol():
    assert declared_functions(CPU_HEADER) == sorted(_native.CPU_EXPORTED_SYMBOLS)
    lib = _native.load_cpu_library()
    for name in declared_functions(CPU_HEADER):
        assert isinstance(getattr(lib, name), ctypes._CFuncPtr), name


def test_struct_layouts_match_header():
    # sizes of the ABI structs as laid out by the C compiler for include/qbx.h
    assert ctypes.sizeof(_native.Opts) == 16
    assert ctypes.sizeof(_native.Result) == 8 * 7 + 4 * 8  # 88 bytes
    assert ctypes.sizeof(_native.DevInfo) == 64 + 4 * 4 + 8 + 8


def test_no_cpu_fallback_without_gpu():
    import torch  # noqa: F401  (only to learn whether a GPU is present)
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(_native.NativeError):
        _native.Context(0)
    from qubo_lib_amd.solvers import QBF
    from qubo_lib_amd.problems import RandomQUBO
    with pytest.raises(_native.NativeError):
        QBF().solve(RandomQUBO.init_random_instance(seed=0, N=8, dtype="int32"), device="cuda")


def test_path_codes_match_header():
    text = HEADER.read_text()
    codes = {int(v): k for k, v in re.findall(r"#define QBX_PATH_([A-Z0-9_]+)\s+(\d+)", text)}
    assert {c: name.lower() for c, name in codes.items()} == _native.PATHS
    assert int(re.search(r"#define QBX_PAIR_BOUND\s+(\d+)", text).group(1)) == _native.PAIR_BOUND
