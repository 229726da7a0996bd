"""Drop-in replacements for the reference's Green's-function sums.

The functions named ``*_numba`` keep the names, signatures, argument meaning
and return convention of python/kernels/kernels.py: float64 NumPy arrays of 3N
values (flat or (N,3), AoS) in, a new flat float64 NumPy array out.  They run
on the GPU through libf3d.so; nothing here computes on the CPU.

``pairsum`` is the device-tensor twin used by the GPU-resident pipeline: torch
float64 CUDA tensors in and out, no host round trip.

A module-level precision (``set_precision``) selects FP64 (default, parity
path) or FP32-compensated evaluation for the drop-in names.
"""
import ctypes

import numpy as np
import torch

from . import _lib as L

_PREC = {"f64": L.F64, "f32c": L.F32C}
_precision = "f64"
_ws = {}


def set_precision(p):
    """Select 'f64' (default) or 'f32c' for the drop-in functions."""
    global _precision
    if p not in _PREC:
        raise ValueError(f"precision must be one of {sorted(_PREC)}")
    _precision = p


def device():
    if not torch.cuda.is_available():
        raise L.F3DError("fiber3d_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda")


_tuned = False


def _ensure_tuned():
    """Install the saved launch-shape table (fiber3d_amd/tuning.json) once."""
    global _tuned
    if not _tuned:
        _tuned = True
        from . import autotune
        autotune.apply()


# Workspace buffers replaced by a larger one stay allocated for the life of
# the process: a recorded HIP graph keeps the device address it was recorded
# with, and freeing the buffer would hand that memory to other tensors.
_ws_retired = []


def workspace(nbytes, tag="default"):
    """A cached device byte buffer of at least nbytes (grown by >= 1.5x on
    demand; a replaced buffer is retired, never freed)."""
    key = (tag, torch.cuda.current_device())
    buf = _ws.get(key)
    if buf is None or buf.numel() < nbytes:
        size = max(int(nbytes), 256)
        if buf is not None:
            _ws_retired.append(buf)
            size = max(size, int(1.5 * buf.numel()))
        buf = torch.empty(size, dtype=torch.uint8, device=device())
        _ws[key] = buf
    return buf


def _stream():
    return ctypes_stream(torch.cuda.current_stream())


def ctypes_stream(s):
    return ctypes.c_void_p(s.cuda_stream)


def _strides3(t):
    """(sp, sc) of an (N,3)-shaped float64 tensor."""
    if t.dim() != 2 or t.shape[1] != 3:
        raise ValueError(f"expected an (N,3) tensor, got shape {tuple(t.shape)}")
    if t.dtype != torch.float64:
        raise TypeError("expected float64")
    return t.stride(0), t.stride(1)


def _as_common(ts):
    """Make a list of (N,3) tensors (None allowed) share one (sp, sc)."""
    ref = next(t for t in ts if t is not None)
    st = _strides3(ref)
    out = []
    for t in ts:
        if t is None:
            out.append(None)
        elif _strides3(t) != st:
            out.append(torch.empty_strided(ref.shape, ref.stride(), dtype=ref.dtype,
                                           device=ref.device).copy_(t))
        else:
            out.append(t)
    return out, st


def pairsum(kernel, src_pos, src_den, trg_pos, src_nrm=None, src_w=None, excl=None,
            eta=1.0, reg=5e-3, eps=1e-5, precision="f64", out=None, accumulate=False,
            clamp_near=False, exclude_self=False):
    """Device pair sum (libf3d f3d_pairsum) on torch CUDA tensors.

    kernel: 'oseen' | 'rotlet' | 'stresslet'.  src_pos, src_den, src_nrm,
    trg_pos, out: (N,3) float64 CUDA tensors with any strides (an SoA buffer
    (3,N) is passed as its transpose).  src_w: (Ns,) float64 weights.
    excl: (lo, hi) int32 (Nt,) tensors of excluded source ranges.
    exclude_self: drop the pair (source i, target i) (self-sums, j != i).
    Returns out (allocated AoS if None).
    """
    k = {"oseen": L.OSEEN, "rotlet": L.ROTLET, "stresslet": L.STRESSLET}[kernel]
    lib = L.load()
    _ensure_tuned()
    (sp_, sd_, sn_), (s_sp, s_sc) = _as_common([src_pos, src_den, src_nrm])
    nt = trg_pos.shape[0]
    ns = sp_.shape[0]
    t_sp, t_sc = _strides3(trg_pos)
    if out is None:
        out = torch.empty((nt, 3), dtype=torch.float64, device=trg_pos.device)
    o_sp, o_sc = _strides3(out)
    lo = hi = None
    if excl is not None:
        lo, hi = excl
        assert lo.dtype == torch.int32 and hi.dtype == torch.int32 and lo.numel() == nt
    if src_w is not None:
        assert src_w.dtype == torch.float64 and src_w.numel() == ns and src_w.is_contiguous()
    prec = _PREC[precision]
    need = lib.f3d_pairsum_workspace_size(k, prec, ns, nt)
    ws = workspace(need)
    flags = ((L.ACCUMULATE if accumulate else 0) | (L.CLAMP_NEAR if clamp_near else 0)
             | (L.EXCLUDE_SELF if exclude_self else 0))
    rc = lib.f3d_pairsum(k, prec, flags, ns, L.ptr(sp_), L.ptr(sn_), L.ptr(sd_), L.ptr(src_w),
                         s_sp, s_sc, nt, L.ptr(trg_pos), t_sp, t_sc, L.ptr(lo), L.ptr(hi),
                         float(eta), float(reg), float(eps), L.ptr(out), o_sp, o_sc,
                         L.ptr(ws), ws.numel(), _stream())
    L.check(rc, "f3d_pairsum")
    return out


def periphery_apply(pos, nrm, w, ex, ey, ez, q, add=None, eta=1.0, reg=5e-3, eps=1e-5,
                    precision="f64", out=None):
    """Matrix-free shell operator (f3d_periphery_apply) on (N,3) CUDA tensors."""
    lib = L.load()
    _ensure_tuned()
    (pos, nrm, ex, ey, ez, q, add), (sp, sc) = _as_common([pos, nrm, ex, ey, ez, q, add])
    n = pos.shape[0]
    if out is None:
        out = torch.empty_strided(pos.shape, pos.stride(), dtype=torch.float64, device=pos.device)
    elif _strides3(out) != (sp, sc):
        raise ValueError("out must share the layout of the inputs")
    prec = _PREC[precision]
    ws = workspace(lib.f3d_periphery_workspace_size(prec, n))
    rc = lib.f3d_periphery_apply(prec, n, L.ptr(pos), L.ptr(nrm), L.ptr(w), L.ptr(ex), L.ptr(ey),
                                 L.ptr(ez), L.ptr(q), L.ptr(add), sp, sc, float(eta), float(reg),
                                 float(eps), L.ptr(out), L.ptr(ws), ws.numel(), _stream())
    L.check(rc, "f3d_periphery_apply")
    return out


def kernel_timer(enable=True):
    """Start (clearing earlier records) or stop HIP-event timing of the pair
    kernel inside every later pairsum / periphery_apply call."""
    L.check(L.load().f3d_pairsum_timer(1 if enable else 0), "f3d_pairsum_timer")


def kernel_timer_read():
    """(summed milliseconds, launch count) of the pair-kernel launches timed
    since kernel_timer(True); waits for the last one."""
    ms = ctypes.c_double()
    n = ctypes.c_int32()
    L.check(L.load().f3d_pairsum_timer_read(ctypes.byref(ms), ctypes.byref(n)),
            "f3d_pairsum_timer_read")
    return ms.value, n.value


# ---------------------------------------------------------------------------
# Drop-in names (python/kernels/kernels.py)

def _to_dev(a, n=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1, 3))
    if n is not None and a.shape[0] != n:
        raise ValueError(f"expected {n} points, got {a.shape[0]}")
    return torch.from_numpy(a).to(device())


def _np_pairsum(kernel, r_source, r_target, density, normal, eta, reg, eps, self_sum=False):
    s = _to_dev(r_source)
    t = s if self_sum else _to_dev(r_target)
    d = _to_dev(density, s.shape[0])
    nm = _to_dev(normal, s.shape[0]) if normal is not None else None
    u = pairsum(kernel, s, d, t, src_nrm=nm, eta=eta, reg=reg, eps=eps,
                precision=_precision, exclude_self=self_sum)
    return u.cpu().numpy().ravel()


def oseen_kernel_source_target_numba(r_source, r_target, density, eta=1.0, reg=5e-3,
                                     epsilon_distance=1e-5):
    """Drop-in for kernels.oseen_kernel_source_target_numba (kernels/kernels.py:794-846)."""
    return _np_pairsum("oseen", r_source, r_target, density, None, eta, reg, epsilon_distance)


def rotlet_kernel_source_target_numba(r_source, r_target, density, eta=1.0, reg=5e-3,
                                      epsilon_distance=1e-5):
    """Drop-in for kernels.rotlet_kernel_source_target_numba (kernels/kernels.py:856-908)."""
    return _np_pairsum("rotlet", r_source, r_target, density, None, eta, reg, epsilon_distance)


def stresslet_kernel_source_target_numba(r_source, r_target, normal, density, eta=1.0,
                                         reg=5e-3, epsilon_distance=1e-5):
    """Drop-in for kernels.stresslet_kernel_source_target_numba (kernels/kernels.py:1186-1217)."""
    return _np_pairsum("stresslet", r_source, r_target, density, normal, eta, reg,
                       epsilon_distance)


def stresslet_kernel_times_normal_times_density_numba(r_vectors, normal, density, eta=1.0,
                                                      reg=5e-3, epsilon_distance=1e-5):
    """Drop-in for kernels.stresslet_kernel_times_normal_times_density_numba
    (kernels/kernels.py:1123-1183): the self-sum over j != i (the i == j pair
    is dropped by F3D_EXCLUDE_SELF, as the reference's loop skips it)."""
    return _np_pairsum("stresslet", r_vectors, r_vectors, density, normal, eta, reg,
                       epsilon_distance, self_sum=True)


def complementary_kernel_times_density_numba(r_vectors, normal, density):
    """Drop-in for kernels.complementary_kernel_times_density_numba
    (kernels/kernels.py:1267-1283): n_i * sum_j n_j . q_j, on the GPU."""
    n = _to_dev(normal)
    q = _to_dev(density, n.shape[0])
    return (n * (n * q).sum()).cpu().numpy().ravel()


def _points(r, fmt, what):
    """Device (N, 3) view parameters of an AoS ('r') or component-block
    ('xyz') point array: (tensor, point stride, component stride, N)."""
    a = np.ascontiguousarray(np.asarray(r, dtype=np.float64).ravel())
    if a.size % 3:
        raise ValueError(f"{what}: size {a.size} is not a multiple of 3")
    n = a.size // 3
    if fmt == "r":
        return torch.from_numpy(a).to(device()), 3, 1, n
    if fmt == "xyz":
        return torch.from_numpy(a).to(device()), 1, n, n
    raise ValueError("input_format is 'r' or 'xyz'")


def oseen_tensor_source_target(r_source, r_target, eta=1.0, reg=5e-3, epsilon_distance=1e-5,
                               input_format="r"):
    """Drop-in for kernels.oseen_tensor_source_target (kernels/kernels.py:119-211):
    the dense (3 Nt, 3 Ns) Oseen tensor, built on the GPU by f3d_oseen_matrix."""
    s, ssp, ssc, ns = _points(r_source, input_format, "r_source")
    t, tsp, tsc, nt = _points(r_target, input_format, "r_target")
    G = torch.empty((3 * nt, 3 * ns), dtype=torch.float64, device=device())
    rc = L.load().f3d_oseen_matrix(ns, L.ptr(s), ssp, ssc, nt, L.ptr(t), tsp, tsc, float(eta),
                                   float(reg), float(epsilon_distance),
                                   0 if input_format == "r" else 1, L.ptr(G), 3 * ns, _stream())
    L.check(rc, "f3d_oseen_matrix")
    return G.cpu().numpy()


def oseen_tensor(r_vectors, eta=1.0, reg=5e-3, epsilon_distance=1e-5, input_format="r"):
    """Drop-in for kernels.oseen_tensor (kernels/kernels.py:31-116): the points
    as both sources and targets (diagonal blocks regularised, not zero)."""
    return oseen_tensor_source_target(r_vectors, r_vectors, eta, reg, epsilon_distance,
                                      input_format)


def stresslet_kernel_times_normal_numba(r_vectors, normal, eta=1.0, reg=5e-3,
                                        epsilon_distance=1e-5):
    """Drop-in for kernels.stresslet_kernel_times_normal_numba
    (kernels/kernels.py:1063-1120): the dense (3N, 3N) stresslet contracted with
    the normals, built on the GPU by f3d_stresslet_matrix."""
    x = _to_dev(r_vectors)
    n = _to_dev(normal, x.shape[0])
    N = x.shape[0]
    M = torch.empty((3 * N, 3 * N), dtype=torch.float64, device=device())
    rc = L.load().f3d_stresslet_matrix(N, L.ptr(x), L.ptr(n), 3, 1, float(eta), float(reg),
                                       float(epsilon_distance), L.ptr(M), 3 * N, _stream())
    L.check(rc, "f3d_stresslet_matrix")
    return M.cpu().numpy()


def complementary_kernel(r_vectors, normal):
    """Drop-in for kernels.complementary_kernel (kernels/kernels.py:1286-1295):
    the rank-one (3N, 3N) matrix n n^T, formed on the GPU."""
    n = _to_dev(normal).reshape(-1)
    return torch.outer(n, n).cpu().numpy()
