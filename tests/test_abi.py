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


def _c_decls():
    """{function name: parameter count} and {struct: field count} of kgpu.h,
    parsed here independently of the crate's generator."""
    import re
    text = open(os.path.join(ROOT, "include", "kgpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    fns = {}
    for m in re.finditer(r"\b(kg_\w+)\s*\(([^;{}]*?)\)\s*;", text, flags=re.S):
        args = " ".join(m.group(2).split())
        fns[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    structs = {}
    for m in re.finditer(r"typedef\s+struct\s+\w+\s*\{(.*?)\}\s*(\w+)\s*;", text, flags=re.S):
        structs[m.group(2)] = len([f for f in m.group(1).split(";") if f.strip()])
    return fns, structs


def test_rust_sys_crate_declares_the_header():
    """krishiv-gpu-sys (the crate a Rust build of the engine links) declares
    every function of include/kgpu.h with the same number of parameters and
    every struct with the same fields, and is what its generator writes from
    the header today (regenerate after changing the header)."""
    import importlib.util
    import re
    crate = os.path.join(ROOT, "krishiv-engine_amd", "rust", "krishiv-gpu-sys")
    rs = open(os.path.join(crate, "src", "lib.rs")).read()
    fns, structs = _c_decls()
    got = {}
    for m in re.finditer(r"pub fn (kg_\w+)\(([^)]*)\)", rs):
        args = m.group(2).strip()
        got[m.group(1)] = 0 if not args else args.count(",") + 1
    assert got == fns
    for name, nfields in structs.items():
        m = re.search(r"pub struct " + name + r" \{(.*?)\n\}", rs, flags=re.S)
        assert m, name
        assert m.group(1).count("pub ") == nfields, name
    spec = importlib.util.spec_from_file_location("gen_lib", os.path.join(crate, "gen_lib.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    assert gen.generate() == rs
