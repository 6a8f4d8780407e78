"""torch.autograd.Function over the discrete adjoint.

``rk4_map(u0, cs2, adj)`` advances ``u0`` (a contiguous complex CUDA tensor
of shape (3, nz, ny, nx//2 + 1): the reference's three velocity fields) by
``adj.nsteps`` IF-RK4 steps with C_s^2 = ``cs2`` and returns u(T).  Its
backward calls the library's checkpointed adjoint, so any torch loss of
u(T) gives gradients with respect to u(0) and cs2 (and through it C_s).
torch's complex-gradient convention (dL/dRe + i dL/dIm per entry) is the
library's.  The tensors' memory is handed to the library by address; torch
is synchronised around each call, since the library runs on its own stream.
"""
from __future__ import annotations

import sys

import speccodelib_amd as _scl

if _scl._lib is not None and "torch" not in sys.modules:
    raise RuntimeError("import torch before libscl.so is loaded: torch bundles its own HIP runtime under the "
                       "same soname, and a process that loads libscl.so first cannot give torch the GPU")

import torch  # noqa: E402

from . import Adjoint  # noqa: E402


def _ptrs(t):
    return [t[q].data_ptr() for q in range(3)]


def _check_fields(t, adj: Adjoint, what):
    """A (3, nz, ny, nx//2+1) complex CUDA tensor of the plan's precision:
    the library reads and writes the plan's element size at each address, so
    a complex64 tensor handed to an FP64 plan would be overrun."""
    want = torch.complex64 if adj.plan.prec == _scl.SP else torch.complex128
    if t.dtype != want:
        raise ValueError(f"{what} has dtype {t.dtype}; this plan's precision needs {want}")
    if t.dim() != 4 or t.shape[0] != 3:
        raise ValueError(f"{what} must have shape (3, nz, ny, nx//2+1)")
    if tuple(t.shape[1:]) != adj.plan.shape:
        raise ValueError(f"{what} fields have shape {tuple(t.shape[1:])}, the plan expects {adj.plan.shape}")
    if not t.is_cuda:
        raise ValueError(f"{what} must be a CUDA tensor")


class RK4Map(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, cs2, adj: Adjoint):
        _check_fields(u0, adj, "u0")
        u0c = u0.detach().contiguous()
        uT = torch.empty_like(u0c)
        c = float(cs2)
        torch.cuda.synchronize(u0.device)
        adj.forward_device(c, _ptrs(u0c), _ptrs(uT))
        adj._fwd_token = object()
        ctx.adj, ctx.token, ctx.cs2 = adj, adj._fwd_token, c
        ctx.save_for_backward(u0c)
        ctx.cs2_meta = (cs2.dtype, cs2.device, cs2.shape) if torch.is_tensor(cs2) else None
        return uT

    @staticmethod
    def backward(ctx, gT):
        adj = ctx.adj
        (u0c,) = ctx.saved_tensors
        if adj._fwd_token is not ctx.token:
            # another forward ran on this adjoint since: restore this one's checkpoints
            tmp = torch.empty_like(u0c)
            torch.cuda.synchronize(u0c.device)
            adj.forward_device(ctx.cs2, _ptrs(u0c), _ptrs(tmp))
            adj._fwd_token = ctx.token
        _check_fields(gT, adj, "the gradient of u(T)")
        gTc = gT.contiguous()
        g0 = torch.empty_like(u0c)
        torch.cuda.synchronize(u0c.device)
        dcs2 = adj.backward_device(_ptrs(gTc), _ptrs(g0))
        gc = None
        if ctx.cs2_meta is not None:
            dt, dev, shape = ctx.cs2_meta
            gc = torch.full(shape, dcs2, dtype=dt, device=dev)
        return g0, gc, None


def rk4_map(u0, cs2, adj: Adjoint):
    return RK4Map.apply(u0, cs2, adj)
