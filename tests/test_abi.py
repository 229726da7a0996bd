"""C-ABI checks that need no GPU: the library loads, exports every entry point
include/f3d.h declares, and rejects bad arguments before touching the device."""
import ctypes
import os
import re

import pytest

from conftest import ROOT
from fiber3d_amd import _lib as L

HEADER = os.path.join(ROOT, "include", "f3d.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(f3d_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding_table():
    assert _declared() == sorted(L.SIGNATURES)


def test_library_exports_all_symbols():
    lib = L.load()
    for name in _declared():
        assert hasattr(lib, name), name
    assert lib.f3d_abi_version() == 100


def test_bad_arguments_rejected_on_host():
    lib = L.load()
    null = None
    rc = lib.f3d_pairsum(7, L.F64, 0, 10, null, null, null, null, 3, 1, 10, null, 3, 1, null,
                         null, 1.0, 5e-3, 1e-5, null, 3, 1, null, 0, null)
    assert rc == L.EINVAL and b"unknown kernel" in lib.f3d_last_error()
    rc = lib.f3d_pairsum(L.OSEEN, 9, 0, 10, null, null, null, null, 3, 1, 10, null, 3, 1, null,
                         null, 1.0, 5e-3, 1e-5, null, 3, 1, null, 0, null)
    assert rc == L.EINVAL and b"precision" in lib.f3d_last_error()
    rc = lib.f3d_pairsum(L.OSEEN, L.F64, 0, 10, null, null, null, null, 3, 1, 10, null, 3, 1,
                         null, null, -1.0, 5e-3, 1e-5, null, 3, 1, null, 0, null)
    assert rc == L.EINVAL and b"eta" in lib.f3d_last_error()
    # lo given without hi
    rc = lib.f3d_pairsum(L.OSEEN, L.F64, 0, 10, null, null, null, null, 3, 1, 10, null, 3, 1,
                         ctypes.c_void_p(8), null, 1.0, 5e-3, 1e-5, null, 3, 1, null, 0, null)
    assert rc == L.EINVAL
    # the self-pair flag with exclusion ranges, and an unknown flag
    rc = lib.f3d_pairsum(L.OSEEN, L.F64, L.EXCLUDE_SELF, 10, null, null, null, null, 3, 1, 10,
                         null, 3, 1, ctypes.c_void_p(8), ctypes.c_void_p(8), 1.0, 5e-3, 1e-5,
                         null, 3, 1, null, 0, null)
    assert rc == L.EINVAL and b"exclusive" in lib.f3d_last_error()
    rc = lib.f3d_pairsum(L.OSEEN, L.F64, 64, 10, null, null, null, null, 3, 1, 10, null, 3, 1,
                         null, null, 1.0, 5e-3, 1e-5, null, 3, 1, null, 0, null)
    assert rc == L.EINVAL and b"flags" in lib.f3d_last_error()
    rc = lib.f3d_pairsum_noalloc(7, L.F64, 0, 10, null, null, null, null, 3, 1, 10, null, 3, 1,
                                 null, null, 1.0, 5e-3, 1e-5, null, 3, 1, null)
    assert rc != L.OK
    rc = lib.f3d_lu_factor_batched(L.F64, 4, 300, null, null, null, null, null, null, null,
                                   null, null, null, null, null, null)
    assert rc == L.EINVAL and b"nmax" in lib.f3d_last_error()
    rc = lib.f3d_lu_solve_batched(L.F32C, 4, 64, null, null, null, null, null, null, null,
                                  null, null, null, null, -1, 1e-12, null, null)
    assert rc == L.EINVAL
    rc = lib.f3d_block_gemv_batched(3, 0, null, null, null, null, null, null, null, null, 1.0,
                                    0.0, null)
    assert rc == L.EINVAL
    assert lib.f3d_pairsum_workspace_size(0, 5, 10, 10) == 0
    assert lib.f3d_pairsum_workspace_size(0, L.F32C, 65536, 65536) > 65536 * 32


def test_no_cpu_fallback_without_library(monkeypatch, tmp_path):
    monkeypatch.setattr(L, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(L, "_lib", None)
    with pytest.raises(L.F3DError):
        L.load()


def test_binding_arity_matches_header():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    for m in re.finditer(r"\b(f3d_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", text):
        name, args = m.group(1), m.group(2).strip()
        n = 0 if args in ("void", "") else len(args.split(","))
        assert n == len(L.SIGNATURES[name][1]), name


def test_product_package_does_not_import_the_oracle():
    """The oracle is the checker only: importing every module of the
    package in a fresh interpreter loads nothing from oracle/."""
    import subprocess
    import sys
    code = ("import sys, os, pkgutil, importlib, fiber3d_amd\n"
            "d = fiber3d_amd.__path__[0]\n"
            "for m in pkgutil.iter_modules(fiber3d_amd.__path__):\n"
            "    if os.path.exists(os.path.join(d, m.name + '.py')):\n"
            "        importlib.import_module('fiber3d_amd.' + m.name)\n"
            "bad = [k for k in sys.modules if k == 'oracle' or k.startswith('oracle.')]\n"
            "print(bad)\n"
            "sys.exit(1 if bad else 0)\n")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus 2 exits with a message before any
    GPU work."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "bench.py", "--gpus", "2"], cwd=root,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
