"""Operator containers of the reference (``type(sq_op)``, ``type(cc_mat)``)
as device FP64 buffers.

``SqOp`` holds E0, fhh, fph, fpp and the six gam blocks of every channel in
the flat column-major layout of include/imsrg_hip.h, plus the herm flag
(basic_IMSRG.f90:65-76).  ``CCMat`` is the cross-coupled storage
(cross_coupled.f90:11-15) and ``CCWork`` the WCC workspace
(allocate_CC_wkspc, cross_coupled.f90:1075-1096).  torch provides the
device memory and the stream; all arithmetic is in libimsrg_hip.so.
"""
from __future__ import annotations

import numpy as np
import torch

from .basis import Basis, ROW_CLS, COL_CLS


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


class SqOp:
    def __init__(self, basis: Basis, data: torch.Tensor, herm: int = 1):
        if data.dtype != torch.float64 or data.numel() != basis.oplen or not data.is_cuda:
            raise ValueError("SqOp needs a CUDA float64 tensor of basis.oplen elements")
        self.basis = basis
        self.data = data.contiguous()
        self.herm = int(herm)

    # allocate_blocks / duplicate_sq_op (basic_IMSRG.f90:521-685, 2797-2860): zero operator
    @classmethod
    def zeros(cls, basis: Basis, herm: int = 1, device="cuda") -> "SqOp":
        return cls(basis, torch.zeros(basis.oplen, dtype=torch.float64, device=device), herm)

    @classmethod
    def from_numpy(cls, basis: Basis, arr: np.ndarray, herm: int = 1, device="cuda") -> "SqOp":
        t = torch.from_numpy(np.ascontiguousarray(arr, np.float64)).to(device)
        return cls(basis, t, herm)

    def numpy(self) -> np.ndarray:
        return self.data.detach().cpu().numpy()

    @property
    def ptr(self) -> int:
        return self.data.data_ptr()

    @property
    def E0(self) -> float:
        return float(self.data[0].item())

    def duplicate(self, herm: int | None = None) -> "SqOp":
        return SqOp.zeros(self.basis, self.herm if herm is None else herm, self.data.device)

    def clone(self) -> "SqOp":
        return SqOp(self.basis, self.data.clone(), self.herm)

    # vectorize / repackage, basic_IMSRG.f90:3929-4053
    def vectorize(self) -> torch.Tensor:
        """the reference's vectorize stream (basis.neq doubles, device)"""
        from . import _lib
        v = torch.empty(self.basis.neq, dtype=torch.float64, device=self.data.device)
        _lib.check(_lib.lib.imsrg_vectorize(self.basis.handle, self.ptr, v.data_ptr(), _stream()))
        return v

    @classmethod
    def repackage(cls, basis: Basis, vec: torch.Tensor, herm: int = 1) -> "SqOp":
        from . import _lib
        if vec.dtype != torch.float64 or not vec.is_cuda or vec.numel() != basis.neq:
            raise ValueError(f"repackage needs a CUDA float64 vector of {basis.neq} elements")
        op = cls.zeros(basis, herm, vec.device)
        _lib.check(_lib.lib.imsrg_repackage(basis.handle, vec.contiguous().data_ptr(), int(herm), op.ptr,
                                            _stream()))
        return op

    def write_bin(self, path: str) -> None:
        """the operator file of write_omega_checkpoint (basic_IMSRG.f90:3215-3301)"""
        from . import _lib
        _lib.check(_lib.lib.imsrg_write_op_bin(self.basis.handle, self.ptr, str(path).encode(), _stream()))

    @classmethod
    def read_bin(cls, basis: Basis, path: str, herm: int = 1) -> "SqOp":
        """read_omega_checkpoint's read + repackage (basic_IMSRG.f90:3304-3407)"""
        from . import _lib
        op = cls.zeros(basis, herm)
        _lib.check(_lib.lib.imsrg_read_op_bin(basis.handle, str(path).encode(), int(herm), op.ptr, _stream()))
        return op

    @staticmethod
    def omega_checkpoint_name(prefix: str, s: float) -> str:
        """write_omega_checkpoint's file name for Omega at flow parameter s
        (basic_IMSRG.f90:3239-3249)"""
        from . import _lib
        import ctypes as C
        buf = C.create_string_buffer(len(prefix) + 64)
        _lib.check(_lib.lib.imsrg_omega_checkpoint_name(str(prefix).encode(), float(s), buf, len(buf)))
        return buf.value.decode()

    def gam(self, q: int, g: int) -> torch.Tensor:
        """view of gam(g+1) of channel q as an (rows, cols) matrix"""
        n1, n2 = self.basis.block_shape(q, g)
        o = int(self.basis.goff[q, g])
        return self.data[o: o + n1 * n2].view(n2, n1).t()


class CCMat:
    """cross-coupled matrices CCX(q) (rlen x nph) of one operator"""

    def __init__(self, basis: Basis, herm: int = 1, device="cuda"):
        self.basis = basis
        self.herm = herm
        self.data = torch.zeros(max(basis.cclen, 1), dtype=torch.float64, device=device)

    @property
    def ptr(self) -> int:
        return self.data.data_ptr()


class CCWork:
    """WCC workspace: rlen x rlen per cross-coupled block"""

    def __init__(self, basis: Basis, device="cuda"):
        self.basis = basis
        self.data = torch.zeros(max(basis.wcclen, 1), dtype=torch.float64, device=device)

    @property
    def ptr(self) -> int:
        return self.data.data_ptr()


__all__ = ["SqOp", "CCMat", "CCWork", "ROW_CLS", "COL_CLS"]
