"""The CCSD vector function as a torch.autograd.Function.

forward  = ccgpu_rhs  (CCRHSN, cc/ccsd_rhs.F:20)
backward = ccgpu_lhtr (CC_LHTR, cc/cc_lhtr.F:20): the vector-Jacobian
           product lambda A is the left transformation, computed by the
           hand-written adjoint kernels, not by tracing.

Tensors are float64 in Dalton's layouts (t1 as (a,i) a-fastest, doubles
packed over ai <= bj).  Amplitudes on the context's GPU (torch device
tensors) stay on the device: the C ABI takes their addresses, the results
are device tensors, and nothing passes through host memory; CPU tensors go
through host arrays.  Example:

    t1 = torch.tensor(t1am, device="cuda", requires_grad=True)
    t2 = torch.tensor(t2am, device="cuda", requires_grad=True)
    om1, om2 = ccsd_residual(ctx, t1, t2)
    (om1 @ l1 + om2 @ l2).backward()      # t1.grad, t2.grad = lambda A
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import check, lib
from .cc import FP64, CCContext

_dp = C.POINTER(C.c_double)

# torch's HIP runtime (bundled with the torch wheel) and the library's (the
# system ROCm) are two runtimes in one process: torch's is initialised here,
# at import, before the library launches anything (a first torch
# initialisation that came after many library launches, cooperative ones
# among them, has found no device).
torch.cuda.is_available()


def _np(x: torch.Tensor) -> np.ndarray:
    return np.ascontiguousarray(x.detach().to("cpu", torch.float64).numpy())


def _dev(x: torch.Tensor) -> torch.Tensor:
    return x.detach().to(torch.float64).contiguous()


def _addr(x: torch.Tensor):
    return C.cast(C.c_void_p(x.data_ptr()), _dp)


def _pdiag(n2: int, device) -> torch.Tensor:
    """Positions ai(ai+3)/2 of the diagonal elements (ai, ai) of a packed
    doubles vector of n2 = nov(nov+1)/2 values."""
    nov = int((np.sqrt(8 * n2 + 1) - 1) // 2)
    ai = torch.arange(nov, device=device, dtype=torch.int64)
    return ai * (ai + 3) // 2


class CCSDResidual(torch.autograd.Function):
    @staticmethod
    def forward(fctx, t1, t2, cc: CCContext, precision: int):
        fctx.cc, fctx.precision = cc, precision
        if t1.is_cuda:
            a1, a2 = _dev(t1), _dev(t2)
            if a1.numel() != cc.nt1 or a2.numel() != cc.nt2:
                raise ValueError("amplitude sizes do not match NT1AMX/NT2AMX")
            om1 = torch.empty(cc.nt1, dtype=torch.float64, device=a1.device)
            om2 = torch.empty(cc.nt2, dtype=torch.float64, device=a1.device)
            torch.cuda.current_stream(a1.device).synchronize()  # the inputs are written
            check(lib().ccgpu_rhs(cc._h, _addr(a1), _addr(a2), _addr(om1), _addr(om2), precision))
            fctx.amps = (a1, a2)
            return om1, om2
        a1, a2 = _np(t1), _np(t2)
        om1, om2 = cc.rhs(a1, a2, precision)
        fctx.amps = (a1, a2)
        return torch.from_numpy(om1), torch.from_numpy(om2)

    @staticmethod
    def backward(fctx, g1, g2):
        cc = fctx.cc
        a1, a2 = fctx.amps
        if isinstance(a1, torch.Tensor):
            dev = a1.device
            z1 = torch.zeros(cc.nt1, dtype=torch.float64, device=dev) if g1 is None else _dev(g1).clone()
            z2 = torch.zeros(cc.nt2, dtype=torch.float64, device=dev) if g2 is None else _dev(g2).clone()
            # plain packed dot products: the C ABI's left vectors carry the
            # doubles diagonal halved, and its result the diagonal doubled
            d = _pdiag(cc.nt2, dev)
            z2[d] *= 2.0
            r1 = torch.empty_like(z1)
            r2 = torch.empty_like(z2)
            torch.cuda.current_stream(dev).synchronize()
            # the Jacobian is taken at the forward amplitudes
            check(lib().ccgpu_set_amplitudes(cc._h, _addr(a1), _addr(a2)))
            check(lib().ccgpu_lhtr(cc._h, fctx.precision, 1, _addr(z1), _addr(z2), _addr(r1), _addr(r2)))
            r2[d] *= 0.5
            return r1, r2, None, None
        cc.set_amplitudes(a1, a2)  # the Jacobian is taken at the forward amplitudes
        z1 = np.zeros(cc.nt1) if g1 is None else _np(g1)
        z2 = np.zeros(cc.nt2) if g2 is None else _np(g2)
        r1, r2 = cc.lhtr(z1, z2, fctx.precision, convention="natural")
        return torch.from_numpy(r1), torch.from_numpy(r2), None, None


def ccsd_residual(cc: CCContext, t1: torch.Tensor, t2: torch.Tensor, precision: int = FP64):
    """(Omega1, Omega2) with a backward pass through the left transformation."""
    return CCSDResidual.apply(t1, t2, cc, precision)
