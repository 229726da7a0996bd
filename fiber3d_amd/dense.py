"""Batched per-fiber dense operations on the GPU (libf3d dense kernels).

BlockSet describes a batch of uneven square blocks packed column-major in one
float64 device buffer.  BatchedLU factors all of them in one persistent launch
and solves all of them in one launch; it replaces
tstep_utils.build_block_diagonal_lu_preconditioner_fiber
(python/tstep/tstep_utils.py:1405-1419) and the lu_solve loop of tstep.P_inv
(python/tstep/tstep.py:985-988).
"""
import warnings

import numpy as np
import torch
from scipy.linalg import LinAlgWarning

from . import _lib as L
from .kernels import _stream, device


class BlockSet:
    """Sizes and offsets of nb square blocks n[b] x n[b] (column-major).

    a_off[b]: element offset of block b in the matrix buffer;
    x_off[b]: element offset of block b's rows in a stacked vector.
    """

    def __init__(self, sizes):
        sizes = np.asarray(sizes, dtype=np.int64)
        if sizes.ndim != 1 or sizes.size == 0 or sizes.min() < 1:
            raise ValueError("sizes must be a non-empty 1-D list of positive orders")
        self.sizes = sizes
        self.nb = int(sizes.size)
        self.nmax = int(sizes.max())
        self.x_off_h = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        self.a_off_h = np.concatenate([[0], np.cumsum(sizes * sizes)[:-1]]).astype(np.int64)
        self.nrows = int(sizes.sum())
        self.nelem = int((sizes * sizes).sum())
        dev = device()
        self.n = torch.from_numpy(sizes.astype(np.int32)).to(dev)
        self.a_off = torch.from_numpy(self.a_off_h).to(dev)
        self.x_off = torch.from_numpy(self.x_off_h).to(dev)
        # largest blocks first: the persistent workgroups then finish together
        self.order = torch.from_numpy(np.argsort(-sizes, kind="stable").astype(np.int32)).to(dev)

    def pack(self, blocks):
        """List of (n_b, n_b) host arrays -> packed column-major device buffer."""
        buf = np.empty(self.nelem, dtype=np.float64)
        for b, A in enumerate(blocks):
            n = self.sizes[b]
            buf[self.a_off_h[b]:self.a_off_h[b] + n * n] = np.asarray(A, np.float64).T.ravel()
        return torch.from_numpy(buf).to(device())

    def unpack(self, buf):
        h = buf.detach().cpu().numpy()
        return [h[self.a_off_h[b]:self.a_off_h[b] + n * n].reshape(n, n).T.copy()
                for b, n in enumerate(self.sizes)]


class BatchedLU:
    """LU factors of every block of a BlockSet, FP64 or FP32 + FP64 refinement."""

    def __init__(self, blocks: BlockSet, A, precision="f64", refine=8, tol=1e-12):
        self.bs = blocks
        self.A = A  # FP64 matrices, kept for refinement
        self.precision = precision
        self.refine = int(refine)
        self.tol = float(tol)
        dev = A.device
        prec = {"f64": L.F64, "f32c": L.F32C}[precision]
        self._prec = prec
        self.lu64 = self.lu32 = self.rs = self.cs = None
        if prec == L.F64:
            self.lu64 = torch.empty(blocks.nelem, dtype=torch.float64, device=dev)
        else:
            self.lu32 = torch.empty(blocks.nelem, dtype=torch.float32, device=dev)
            self.rs = torch.empty(blocks.nrows, dtype=torch.float64, device=dev)
            self.cs = torch.empty(blocks.nrows, dtype=torch.float64, device=dev)
        self.piv = torch.empty(blocks.nrows, dtype=torch.int32, device=dev)
        self.perm = torch.empty(blocks.nrows, dtype=torch.int32, device=dev)
        self.info = torch.empty(blocks.nb, dtype=torch.int32, device=dev)
        self.iters = torch.zeros(blocks.nb, dtype=torch.int32, device=dev)
        self.factor()

    def singular_blocks(self):
        """(block, 1-based index of the first exactly-zero pivot) for every
        singular block, from the info array of the last factorisation (one
        device read)."""
        info = self.info.cpu().numpy()
        return [(int(b), int(k)) for b, k in enumerate(info) if k > 0]

    def factor(self):
        bs = self.bs
        rc = L.load().f3d_lu_factor_batched(
            self._prec, bs.nb, bs.nmax, L.ptr(bs.n), L.ptr(bs.a_off), L.ptr(bs.x_off),
            L.ptr(bs.order), L.ptr(self.A), L.ptr(self.lu64), L.ptr(self.lu32), L.ptr(self.rs),
            L.ptr(self.cs), L.ptr(self.piv), L.ptr(self.perm), L.ptr(self.info), _stream())
        L.check(rc, "f3d_lu_factor_batched")

    def solve(self, x, out=None, refine=None):
        """Solve every block system; x is the stacked (nrows,) FP64 vector."""
        bs = self.bs
        if out is None:
            out = torch.empty_like(x)
        ref = self.refine if refine is None else int(refine)
        if self._prec == L.F64 and refine is None:
            ref = 0
        rc = L.load().f3d_lu_solve_batched(
            self._prec, bs.nb, bs.nmax, L.ptr(bs.n), L.ptr(bs.a_off), L.ptr(bs.x_off),
            L.ptr(self.A), L.ptr(self.lu64), L.ptr(self.lu32), L.ptr(self.rs), L.ptr(self.cs),
            L.ptr(self.perm), L.ptr(x), L.ptr(out), ref, self.tol, L.ptr(self.iters), _stream())
        L.check(rc, "f3d_lu_solve_batched")
        return out


class BlockGemv:
    """y_b = alpha M_b x_b + beta y_b over uneven column-major blocks m_b x k_b."""

    def __init__(self, m, k, x_off, y_off):
        m = np.asarray(m, np.int64)
        k = np.asarray(k, np.int64)
        self.nb = int(m.size)
        self.mmax = int(m.max())
        self.m_off_h = np.concatenate([[0], np.cumsum(m * k)[:-1]]).astype(np.int64)
        self.nelem = int((m * k).sum())
        self.m_h, self.k_h = m, k
        dev = device()
        self.m = torch.from_numpy(m.astype(np.int32)).to(dev)
        self.k = torch.from_numpy(k.astype(np.int32)).to(dev)
        self.m_off = torch.from_numpy(self.m_off_h).to(dev)
        self.x_off = torch.from_numpy(np.asarray(x_off, np.int64)).to(dev)
        self.y_off = torch.from_numpy(np.asarray(y_off, np.int64)).to(dev)

    def pack(self, mats):
        buf = np.empty(self.nelem, dtype=np.float64)
        for b, M in enumerate(mats):
            o = self.m_off_h[b]
            buf[o:o + M.size] = np.asarray(M, np.float64).T.ravel()
        return torch.from_numpy(buf).to(device())

    def __call__(self, M, x, y, alpha=1.0, beta=0.0):
        rc = L.load().f3d_block_gemv_batched(
            self.nb, self.mmax, L.ptr(self.m), L.ptr(self.k), L.ptr(self.m_off), L.ptr(self.x_off),
            L.ptr(self.y_off), L.ptr(M), L.ptr(x), L.ptr(y), float(alpha), float(beta), _stream())
        L.check(rc, "f3d_block_gemv_batched")
        return y


class DenseGemv:
    """y = M x for a stack of dense square matrices through
    f3d_dense_gemv_batched (FP64 entries) or f3d_dense_gemv_batched_f32 (FP32
    entries, FP64 arithmetic): row-major, one wave per row.  Unlike a BLAS
    GEMV this allocates nothing per call, so it can sit inside a recorded HIP
    graph."""

    def __init__(self, mats):
        """mats: (count, m, m) or (m, m) FP64 or FP32 CUDA tensor."""
        if mats.dim() == 2:
            mats = mats[None]
        count, m, m2 = mats.shape
        assert m == m2 and mats.dtype in (torch.float64, torch.float32)
        self.count, self.m = count, m
        self.M = mats.contiguous()
        self._fn = ("f3d_dense_gemv_batched" if mats.dtype == torch.float64
                    else "f3d_dense_gemv_batched_f32")

    def __call__(self, x, y):
        """x, y: contiguous (count * m,) FP64 tensors (y must not alias x)."""
        assert x.numel() == y.numel() == self.count * self.m
        assert x.dtype == y.dtype == torch.float64
        assert x.is_contiguous() and y.is_contiguous() and x.data_ptr() != y.data_ptr()
        rc = getattr(L.load(), self._fn)(self.count, self.m, L.ptr(self.M), L.ptr(x),
                                         L.ptr(y), _stream())
        L.check(rc, self._fn)
        return y


def build_block_diagonal_lu_preconditioner_fiber(fibers, A_fibers_blocks):
    """Drop-in for tstep_utils.build_block_diagonal_lu_preconditioner_fiber
    (tstep/tstep_utils.py:1405-1419): (LU_all, P_all), per fiber the pair that
    scipy.linalg.lu_factor(A, check_finite=False) returns (packed LU, 0-based
    row interchanges), from one batched FP64 factorisation on the GPU with
    LAPACK's partial pivoting.  The reference's lu_solve((LU, P), x) calls
    apply them unchanged; fiber3d_amd.dense.BatchedLU keeps the factors on the
    device instead."""
    mats = [np.asarray(A, dtype=np.float64) for A in A_fibers_blocks]
    if not mats:
        return [], []
    bs = BlockSet([A.shape[0] for A in mats])
    lu = BatchedLU(bs, bs.pack(mats), precision="f64")
    for b, k in lu.singular_blocks():
        # scipy.linalg.lu_factor's warning for an exactly-zero pivot
        warnings.warn(f"Diagonal number {k} is exactly zero. Singular matrix (fiber {b}).",
                      LinAlgWarning, stacklevel=2)
    LU_all = bs.unpack(lu.lu64)
    piv = lu.piv.cpu().numpy()
    P_all = [piv[o:o + n].astype(np.int32) for o, n in zip(bs.x_off_h, bs.sizes)]
    return LU_all, P_all
