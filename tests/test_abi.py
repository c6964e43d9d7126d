"""The C ABI library loads and exports every entry point include/sdr_amd.h declares
(no compute call: this runs without a GPU), and fails loudly without a device."""
import re
from pathlib import Path

import numpy as np
import pytest

import sdr_amd
from sdr_amd import _abi as A
from sdr_amd import scenes

HEADER = Path(__file__).resolve().parent.parent / "include" / "sdr_amd.h"


def declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(?:int|const char\*)\s+(sdr_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = sdr_amd.load_library()
    names = declared()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib.dll, n)]
    assert not missing, missing


def test_python_bindings_cover_the_header():
    names = {"sdr_" + k for k in {**A.SIGNATURES, **A.DEVICE_ONLY}}
    assert set(declared()) <= names | {"sdr_last_error"}


def test_oracle_exports_the_same_surface():
    import oracle

    lib = oracle.load()
    for name in A.SIGNATURES:
        assert hasattr(lib.dll, "orc_" + name), name


def test_no_device_is_an_error_not_a_fallback():
    n, _ = sdr_amd.device_info()
    if n > 0:
        pytest.skip("a device is present")
    with pytest.raises(sdr_amd.AbiError) as e:
        scenes.make_solver(scenes.box_scene())
    assert e.value.code == A.SDR_E_NODEV


def test_invalid_arguments_are_rejected():
    lib = sdr_amd.load_library()
    with pytest.raises(sdr_amd.AbiError) as e:
        lib.solver_new(None, 0.1, None, None, None, None)
    assert e.value.code == A.SDR_E_INVALID
    with pytest.raises(ValueError):
        sdr_amd.project_capturing_with_source(4, 4, 4, 0.1, (0, 0, 0), np.zeros(5), np.zeros(5),
                                              np.zeros(5), np.zeros(64, np.uint8), 1.0, 1.0)
