"""The C-ABI library loads without a GPU and exports every function that
include/*.h declares (no compute calls here)."""
import ctypes
import glob
import os
import re

import pytest

import mesobasic_amd
from mesobasic_amd import lbm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(lbm_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_abi():
    names = declared_functions()
    for must in ("lbm_create", "lbm_step", "lbm_get_fields", "lbm_destroy", "lbm_conservation"):
        assert must in names


@pytest.mark.parametrize("name", declared_functions())
def test_symbol_exported(name):
    lib = ctypes.CDLL(lbm.LIB_PATH)
    assert hasattr(lib, name), name


def test_abi_version_and_errors_without_gpu():
    lib = mesobasic_amd.load()
    assert lib.lbm_abi_version() == 2
    # host-only entry points work without a device
    import numpy as np
    r = mesobasic_amd.noise_rho((3, 4), 1)
    assert r.shape == (3, 4)
    with pytest.raises(mesobasic_amd.LbmError):
        mesobasic_amd.Lattice(19, (4, 4, 2), 1.0, -5.5, walls=True)  # walls need nz >= 3
