"""The C ABI library: loads without a GPU, exports every function that
include/kgpu.h declares, and its host-only helpers agree with the oracle.
CPU only (no device call)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import kgpu
import kgo

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    lib = kgpu.lib()
    declared = kgpu.exported_symbols()
    assert len(declared) > 40
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", kgpu.abi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    names = {l.split()[-1] for l in out.splitlines() if l.strip()}
    assert not [s for s in declared if s not in names]


def test_abi_version_and_error_text():
    lib = kgpu.lib()
    assert lib.kg_abi_version() == 1
    assert isinstance(lib.kg_last_error(), bytes)


def test_zipf_cdf_matches_oracle():
    """The host-built Zipf table the device generator reads is the oracle's."""
    n = 100_000
    a = np.zeros(n)
    kgpu.abi.check(kgpu.lib().kg_zipf_cdf(n, 1.1, a.ctypes.data))
    assert np.array_equal(a, kgo.zipf_cdf(n, 1.1))


def test_bloom_bytes_for_matches_oracle():
    for items, fpp in [(750000, 0.03), (750000, 0.01), (100000, 0.01), (1, 0.5), (0, 0.01),
                       (4_000_000, 0.01)]:
        assert kgpu.lib().kg_bloom_bytes_for(items, fpp) == kgo.lib().kgo_sbbf_bytes_for(items, fpp)


def test_no_torch_types_in_header():
    text = open(os.path.join(ROOT, "include", "kgpu.h")).read()
    for bad in ("torch", "at::", "Tensor", "hipStream_t", "std::"):
        assert bad not in text
