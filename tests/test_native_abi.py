"""The C ABI library loads on the host and exports every entry point the
header declares (no compute calls: those need a GPU)."""

import re
from pathlib import Path

import pytest

from pararealml_amd import _native as nat

HEADER = Path(__file__).resolve().parent.parent / "include" / "pml.h"


def declared_symbols():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pml_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "pml_fdm_solve" in syms and "pml_parareal_update" in syms
    assert set(syms) == set(nat.SIGNATURES), (
        set(syms) ^ set(nat.SIGNATURES)
    )


def test_library_exports_every_declared_symbol():
    if not nat.LIB_PATH.exists():
        pytest.fail("libpml.so is not built; run python -m pararealml_amd.build")
    lib = nat.load_library()
    for name in declared_symbols():
        assert hasattr(lib, name), name
    assert lib.pml_version() >= 100


def test_config_queries_need_no_gpu():
    lib = nat.load_library()
    n = lib.pml_fdm_configs(nat.PML_EQ["cahn_hilliard"], nat.PML_INT_RK4, nat.PML_F32)
    assert n >= 1
    assert lib.pml_fdm_config_name(nat.PML_EQ["cahn_hilliard"], nat.PML_INT_RK4, nat.PML_F32, 0)
    assert lib.pml_fdm_configs(99, 0, 0) == 0


def test_device_lib_refuses_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("GPU visible")
    with pytest.raises(RuntimeError):
        nat.device_lib()
