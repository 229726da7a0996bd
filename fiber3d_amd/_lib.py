"""ctypes binding of libf3d.so (declarations in include/f3d.h).

The library is the product: nothing here falls back to a CPU or PyTorch
implementation.  If libf3d.so is missing or a call fails, an exception is
raised.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libf3d.so")

# constants mirrored from include/f3d.h
OK = 0
EINVAL = -1
EHIP = -2
EWORKSPACE = -3
OSEEN, ROTLET, STRESSLET = 0, 1, 2
F64, F32C = 0, 1
ACCUMULATE = 1
CLAMP_NEAR = 2
EXCLUDE_SELF = 4

_c_p = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_u32 = ctypes.c_uint32
_f64 = ctypes.c_double
_sz = ctypes.c_size_t

# name -> (restype, argtypes)
SIGNATURES = {
    "f3d_last_error": (ctypes.c_char_p, []),
    "f3d_abi_version": (ctypes.c_int, []),
    "f3d_pairsum_set_launch": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _i32, _i32]),
    "f3d_pairsum_set_launch_bucket": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _i32, _i32,
                                                     _i32, _i32]),
    "f3d_pairsum_plan": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _i64, _i64, _c_p]),
    "f3d_lu_set_launch": (ctypes.c_int, [ctypes.c_int, _i32, _i32]),
    "f3d_gemm_set_stream": (ctypes.c_int, [_i32, _i32]),
    "f3d_pairsum_timer": (ctypes.c_int, [ctypes.c_int]),
    "f3d_pairsum_timer_read": (ctypes.c_int, [ctypes.POINTER(ctypes.c_double),
                                              ctypes.POINTER(ctypes.c_int32)]),
    "f3d_pairsum_workspace_size": (_sz, [ctypes.c_int, ctypes.c_int, _i64, _i64]),
    "f3d_pairsum": (ctypes.c_int, [
        ctypes.c_int, ctypes.c_int, _u32,
        _i64, _c_p, _c_p, _c_p, _c_p, _i64, _i64,
        _i64, _c_p, _i64, _i64,
        _c_p, _c_p,
        _f64, _f64, _f64,
        _c_p, _i64, _i64,
        _c_p, _sz, _c_p]),
    "f3d_pairsum_noalloc": (ctypes.c_int, [
        ctypes.c_int, ctypes.c_int, _u32,
        _i64, _c_p, _c_p, _c_p, _c_p, _i64, _i64,
        _i64, _c_p, _i64, _i64,
        _c_p, _c_p,
        _f64, _f64, _f64,
        _c_p, _i64, _i64, _c_p]),
    "f3d_periphery_workspace_size": (_sz, [ctypes.c_int, _i64]),
    "f3d_periphery_apply": (ctypes.c_int, [
        ctypes.c_int, _i64, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p,
        _i64, _i64, _f64, _f64, _f64, _c_p, _c_p, _sz, _c_p]),
    "f3d_oseen_matrix": (ctypes.c_int, [
        _i64, _c_p, _i64, _i64, _i64, _c_p, _i64, _i64, _f64, _f64, _f64, _i32, _c_p, _i64, _c_p]),
    "f3d_stresslet_matrix": (ctypes.c_int, [
        _i64, _c_p, _c_p, _i64, _i64, _f64, _f64, _f64, _c_p, _i64, _c_p]),
    "f3d_lu_factor_batched": (ctypes.c_int, [
        ctypes.c_int, _i64, _i32, _c_p, _c_p, _c_p, _c_p,
        _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p]),
    "f3d_lu_solve_batched": (ctypes.c_int, [
        ctypes.c_int, _i64, _i32, _c_p, _c_p, _c_p, _c_p,   # n, a_off, x_off, A
        _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p,           # lu64 lu32 rs cs perm x y
        _i32, _f64, _c_p, _c_p]),
    "f3d_gemm_bf16": (ctypes.c_int, [
        ctypes.c_int, _i64, _i64, _i64, _i32, _c_p, _i64, _i64, _c_p, _i64, _i64,
        _c_p, _i64, _i64, _c_p, _c_p, _c_p, ctypes.c_float, _c_p]),
    "f3d_fiber_assemble": (ctypes.c_int, [
        _i64, _i32, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p,
        _f64, _f64, _f64, _f64, _f64, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p]),
    "f3d_fiber_rhs": (ctypes.c_int, [
        _i64, _i32, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p,
        _f64, _f64, _f64, _f64, _f64, _c_p, _c_p]),
    "f3d_gmres_workspace_size": (_sz, [_i64, _i32]),
    "f3d_gmres_begin": (ctypes.c_int, [_i64, _i32, _c_p, _c_p, _c_p, _c_p]),
    "f3d_gmres_current": (ctypes.c_int, [_i64, _i32, _c_p, _c_p, _c_p, _c_p]),
    "f3d_gmres_arnoldi": (ctypes.c_int, [_i64, _i32, _c_p, _c_p, _c_p, _c_p]),
    "f3d_gmres_solve": (ctypes.c_int, [_i64, _i32, _i32, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p]),
    "f3d_dense_gemv_batched": (ctypes.c_int, [_i64, _i32, _c_p, _c_p, _c_p, _c_p]),
    "f3d_dense_gemv_batched_f32": (ctypes.c_int, [_i64, _i32, _c_p, _c_p, _c_p, _c_p]),
    "f3d_block_gemv_batched": (ctypes.c_int, [
        _i64, _i32, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _f64, _f64, _c_p]),
}

_lib = None


class F3DError(RuntimeError):
    pass


def load():
    """Load libf3d.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise F3DError(
            f"{LIB_PATH} not found: build it with `python -m fiber3d_amd.build` "
            "(there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    missing = [n for n in SIGNATURES if not hasattr(lib, n)]
    if missing:
        raise F3DError(f"{LIB_PATH} lacks symbols {missing}: rebuild it")
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc, what):
    if rc != OK:
        msg = load().f3d_last_error().decode(errors="replace")
        raise F3DError(f"{what} failed (code {rc}): {msg}")


def ptr(t):
    """Device pointer of a torch tensor (or None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())
