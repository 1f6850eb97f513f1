"""The C-ABI library: it loads without a GPU and exports exactly what
include/diorama.h declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

import diorama_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "diorama.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dio_[a-z0-9_]+)\s*\(", text)))


def test_library_is_built():
    assert os.path.exists(diorama_amd.LIB_PATH), "run python -m diorama_amd.build"


def test_exports_every_declared_symbol():
    lib = ctypes.CDLL(diorama_amd.LIB_PATH)
    missing = [name for name in declared_functions() if not hasattr(lib, name)]
    assert not missing, missing


def test_python_mirror_binds_every_declared_symbol():
    assert sorted(diorama_amd.SIGNATURES) == declared_functions()


def test_library_loads_and_reports_without_gpu():
    lib = diorama_amd.lib()
    assert lib.dio_version().decode().startswith("diorama-mi355x")
    settings = diorama_amd.default_settings(3.5)
    # DEFAULT_SETTINGS of src/trainer.ts:65-82
    assert settings["extent"] == pytest.approx(3.5)
    assert settings["position_lr"] == pytest.approx([0.00016, 0.0000016])
    assert settings["sh_lr"] == pytest.approx(0.0025 / 20)
    assert settings["densify_from"] == 500 and settings["densify_until"] == 15000
    assert settings["densify_interval"] == 100 and settings["opacity_reset_interval"] == 3000
    assert settings["densify_grad_threshold"] == pytest.approx(0.0002)


def test_null_handles_are_errors_not_crashes():
    lib = diorama_amd.lib()
    assert lib.dio_prefix_sum_encode(None) == -1
    assert b"null" in lib.dio_last_error()
    assert lib.dio_trainer_step(None) == -1
