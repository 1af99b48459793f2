"""The CCSD vector function as a torch.autograd.Function.

forward  = ccgpu_rhs  (CCRHSN, cc/ccsd_rhs.F:20)
backward = ccgpu_lhtr (CC_LHTR, cc/cc_lhtr.F:20): the vector-Jacobian
           product lambda A is the left transformation, computed by the
           hand-written adjoint kernels, not by tracing.

Tensors are float64 in Dalton's layouts (t1 as (a,i) a-fastest, doubles
packed over ai <= bj); they live on the host because the C ABI takes host
arrays, the device-resident state being held by the CCContext.  Example:

    t1 = torch.tensor(t1am, requires_grad=True)
    t2 = torch.tensor(t2am, requires_grad=True)
    om1, om2 = ccsd_residual(ctx, t1, t2)
    (om1 @ l1 + om2 @ l2).backward()      # t1.grad, t2.grad = lambda A
"""
from __future__ import annotations

import numpy as np
import torch

from .cc import FP64, CCContext


def _np(x: torch.Tensor) -> np.ndarray:
    return np.ascontiguousarray(x.detach().to("cpu", torch.float64).numpy())


class CCSDResidual(torch.autograd.Function):
    @staticmethod
    def forward(fctx, t1, t2, cc: CCContext, precision: int):
        a1, a2 = _np(t1), _np(t2)
        om1, om2 = cc.rhs(a1, a2, precision)
        fctx.cc, fctx.precision = cc, precision
        fctx.amps = (a1, a2)
        return torch.from_numpy(om1), torch.from_numpy(om2)

    @staticmethod
    def backward(fctx, g1, g2):
        cc = fctx.cc
        a1, a2 = fctx.amps
        cc.set_amplitudes(a1, a2)  # the Jacobian is taken at the forward amplitudes
        z1 = np.zeros(cc.nt1) if g1 is None else _np(g1)
        z2 = np.zeros(cc.nt2) if g2 is None else _np(g2)
        r1, r2 = cc.lhtr(z1, z2, fctx.precision, convention="natural")
        return torch.from_numpy(r1), torch.from_numpy(r2), None, None


def ccsd_residual(cc: CCContext, t1: torch.Tensor, t2: torch.Tensor, precision: int = FP64):
    """(Omega1, Omega2) with a backward pass through the left transformation."""
    return CCSDResidual.apply(t1, t2, cc, precision)
