"""mesobasic_amd -- MI355X-native Shan-Chen lattice-Boltzmann step.

The product is the C ABI in include/mesobasic_lbm.h, implemented by HIP
kernels for gfx950 (mesobasic_amd/csrc) and built into
mesobasic_amd/libmesobasic_lbm.so; mesobasic_amd/host holds the C++ drop-in
model classes and driver.  This Python package binds the C ABI for tests,
benchmarks and scripting.
"""
from .lbm import (BF16N, BF16S, FP16N, FP16S, FP32, FP64, KERNEL_AUTO, KERNEL_FUSED, KERNEL_TWO_PASS,
                  FIELDS_CURRENT, FIELDS_REFERENCE, Lattice, LbmError, droplet_rho, load,
                  noise_rho)

__all__ = ["Lattice", "LbmError", "load", "noise_rho", "droplet_rho", "FP64", "FP32", "FP16S",
           "BF16S", "FP16N", "BF16N", "KERNEL_AUTO", "KERNEL_TWO_PASS", "KERNEL_FUSED", "FIELDS_REFERENCE",
           "FIELDS_CURRENT"]
