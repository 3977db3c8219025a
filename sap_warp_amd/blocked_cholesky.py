"""Blocked Cholesky of large symmetric positive definite matrices on the GPU.

Drop-in for the reference's ``BlockCholeskySolverBatched`` (``sim/blocked_cholesky.py:725-901``:
``factorize(A, num_active_equations)`` then ``solve(rhs, result)``), backed by the native
``sapx_chol_*`` entry points.  The reference factors each matrix in one workgroup; here a
matrix's block columns are factored and their Schur-complement updates applied by the whole
GPU (64 x 64 blocks, float32 with MFMA updates), which is what a single world of thousands of
dofs needs.  Solves take float64 right-hand sides and add float64-residual refinement sweeps
against a kept copy of the input.  Tensors are torch CUDA tensors on the current stream.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as nat


class BlockCholeskySolverBatched:
    """``blocked_cholesky.py:725-901`` on one MI355X.

    ``block_size`` is fixed at 64 (the MFMA tile of the updates); ``refine`` is the number of
    float64-residual sweeps each :meth:`solve` adds.
    """

    def __init__(self, max_num_equations: int, batch_size: int, block_size: int = 64, device="cuda:0",
                 refine: int = 1):
        if block_size != 64:
            raise ValueError("the native blocked Cholesky uses 64 x 64 blocks")
        self.lib = nat.load()
        self.max_num_equations = int(max_num_equations)
        self.batch_size = int(batch_size)
        self.block_size = 64
        self.device = device
        self.refine = int(refine)
        self.n = 0
        h = ctypes.c_void_p()
        nat.check(self.lib.sapx_chol_create(self.max_num_equations, self.batch_size, ctypes.byref(h)),
                  "sapx_chol_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.lib.sapx_chol_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _stream():
        import torch

        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def factorize(self, A, num_active_equations: int):
        """Factor ``A`` (batch, >= n, >= n) float32 CUDA, rows contiguous; the leading
        ``num_active_equations`` square is used (``factorize``, ``blocked_cholesky.py:781-809``)."""
        import torch

        if not (isinstance(A, torch.Tensor) and A.is_cuda and A.dtype == torch.float32):
            raise TypeError("A must be a float32 CUDA tensor")
        if A.dim() == 2:
            A = A.unsqueeze(0)
        if A.shape[0] != self.batch_size or A.stride(2) != 1:
            raise ValueError("A must be (batch_size, n, n) with contiguous rows")
        n = int(num_active_equations)
        if A.shape[1] < n or A.shape[2] < n:
            raise ValueError("A is smaller than num_active_equations")
        # a size-1 batch dimension may carry any stride
        a_stride = int(A.stride(0)) if A.shape[0] > 1 else int(A.stride(1)) * int(A.shape[1])
        nat.check(self.lib.sapx_chol_factorize(self._h, ctypes.c_void_p(A.data_ptr()), n, int(A.stride(1)),
                                               a_stride, self._stream()), "sapx_chol_factorize")
        self.n = n
        self._A = A  # keep the input alive until the enqueued copy has run

    def solve(self, rhs, result, refine: int | None = None):
        """``result = A^-1 rhs`` (``solve``, ``blocked_cholesky.py:811-823``): float64 CUDA tensors
        of batch_size x n elements each (any shape with that many contiguous elements)."""
        import torch

        for t in (rhs, result):
            if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
                raise TypeError("rhs and result must be contiguous float64 CUDA tensors")
            if t.numel() != self.batch_size * self.n:
                raise ValueError("rhs / result must hold batch_size x n values")
        if rhs.data_ptr() == result.data_ptr():
            raise ValueError("rhs and result must not alias")
        nat.check(self.lib.sapx_chol_solve(self._h, ctypes.c_void_p(rhs.data_ptr()), ctypes.c_void_p(result.data_ptr()),
                                           self.refine if refine is None else int(refine), self._stream()),
                  "sapx_chol_solve")

    def info(self) -> np.ndarray:
        """Per-matrix status of the last factorisation (0, or 1 + the first non-positive pivot's
        column); synchronises.  Raises if a solve's wait timed out."""
        out = np.zeros(self.batch_size, dtype=np.int32)
        nat.check(self.lib.sapx_chol_info(self._h, out.ctypes.data_as(ctypes.c_void_p)), "sapx_chol_info")
        return out
