"""The launch-shape tuning table: its shape and ranges, how the library
installs it, and the library's rejection of shapes it cannot run (host-only
calls: no device work)."""
import json
import os

import pytest

from fiber3d_amd import _lib as L
from fiber3d_amd import autotune as AT


@pytest.fixture(autouse=True)
def _restore_table():
    yield
    # leave the process with the saved table installed, as at import
    AT.apply()


def _saved():
    assert os.path.exists(AT.CACHE), "fiber3d_amd/tuning.json is missing"
    with open(AT.CACHE) as fh:
        return json.load(fh)


def test_saved_table_is_valid_and_complete():
    """Every kernel and precision has a default and tuned size buckets over
    the whole tuning grid, the Oseen sum also the large fiber layout's
    bucket 18/18; the LU and GEMM sections cover both precisions and both
    layer widths."""
    t = AT.validate(_saved())
    keys = {f"{k}/{p}" for k in AT.KERNELS for p in AT.PRECS}
    assert set(t["pairsum"]) == keys
    assert set(t["pairsum_buckets"]) == keys
    for key in keys:
        sizes = {(int(a), int(b)) for a, b in (s.split("/") for s in t["pairsum_buckets"][key])}
        grid = {(AT.lg(nt), AT.lg(ns)) for nt in AT.GRID_NT for ns in AT.GRID_NS}
        assert sizes == (grid | {(18, 18)} if key.startswith("oseen/") else grid)
        assert t["pairsum"][key]["k"] == t["pairsum_buckets"][key]["16/16"]["k"]
    assert set(t["lu"]) == {"f64", "f32c"}
    assert set(t["gemm"]) == {"128", "256"}


def test_installed_plans_follow_the_table():
    """After apply(), a call of a tuned size uses its tuned K and the source
    split its wave target implies; an untuned size takes the nearest tuned
    size's shape; the 65536 x 65536 default holds beyond the buckets."""
    t = AT.apply()
    for key, buckets in t["pairsum_buckets"].items():
        kernel, prec = key.split("/")
        for size, v in buckets.items():
            a, b = (int(s) for s in size.split("/"))
            nt, ns = 1 << a, 1 << b
            k, chunks, chunk = AT.plan(kernel, prec, ns, nt)
            assert k == v["k"], (key, size)
            waves_per_split = -(-nt // (64 * k))
            tiles = -(-ns // 64)
            want = min(tiles, max(1, -(-v["waves"] // waves_per_split)))
            per = -(-tiles // want)
            assert chunk == 64 * per and chunks == -(-ns // chunk), (key, size)
        tuned = {tuple(int(s) for s in size.split("/")): v for size, v in buckets.items()}
        full = AT.expand_buckets(tuned)
        assert len(full) == (AT.LG_MAX - AT.LG_MIN + 1) ** 2
        for (a, b), v in list(full.items())[::37]:
            assert AT.plan(kernel, prec, 1 << b, 1 << a)[0] == v["k"]


def test_library_rejects_bad_shapes():
    lib = L.load()
    assert lib.f3d_pairsum_set_launch(L.OSEEN, L.F32C, 3, 8192) != 0
    assert lib.f3d_pairsum_set_launch(L.OSEEN, L.F32C, 2, 10) != 0
    assert lib.f3d_pairsum_set_launch_bucket(L.OSEEN, L.F32C, 7, 12, 2, 8192) != 0
    assert lib.f3d_pairsum_set_launch_bucket(L.OSEEN, L.F32C, 12, 23, 2, 8192) != 0
    assert lib.f3d_pairsum_set_launch_bucket(L.OSEEN, L.F64, 12, 12, 4, 1 << 21) != 0
    assert lib.f3d_pairsum_set_launch_bucket(9, L.F64, 12, 12, 4, 8192) != 0
    assert lib.f3d_lu_set_launch(L.F64, 5, 1) != 0
    assert lib.f3d_lu_set_launch(L.F32C, 9, 1) != 0
    assert lib.f3d_lu_set_launch(L.F32C, 4, 2) != 0
    assert lib.f3d_gemm_set_stream(192, 1) != 0
    assert lib.f3d_gemm_set_stream(128, 3) != 0
    assert lib.f3d_gemm_set_stream(128, -1) != 0
    bad = _saved()
    bad["pairsum"]["oseen/f32c"]["k"] = 3
    with pytest.raises(ValueError):
        AT.validate(bad)
    bad = _saved()
    bad["pairsum_buckets"]["oseen/f64"]["30/12"] = {"k": 2, "waves": 8192}
    with pytest.raises(ValueError):
        AT.validate(bad)
    bad = _saved()
    bad["lu"]["f64"]["factor_waves_per_cu"] = 8
    with pytest.raises(ValueError):
        AT.validate(bad)


def test_bucket_reset_returns_to_the_default():
    """k = 0 clears a bucket: the call falls back to the (kernel, precision)
    shape."""
    AT.set_launch("rotlet", "f32c", 1, 4096)
    AT.set_launch_bucket("rotlet", "f32c", 12, 12, 4, 8192)
    assert AT.plan("rotlet", "f32c", 4096, 4096)[0] == 4
    AT.set_launch_bucket("rotlet", "f32c", 12, 12, 0, 0)
    assert AT.plan("rotlet", "f32c", 4096, 4096)[0] == 1
