"""Reverse mode of the advect + project step (periodic one-box domain,
density 1, fixed dt) over the C ABI, and a torch.autograd.Function around it
for gradient checks.

  A = Adjoint((n0, n1, n2), prob_hi)
  vel, s, iters = A.forward(vel0, s0, nsteps, dt, every)
  vel0_bar, s0_bar = A.backward(vel_bar, s_bar)

Fields are float64 CUDA tensors: velocity (3, n2, n1, n0), scalar
(n2, n1, n0) (x fastest, as the library's domain-wide arrays).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import KsgfMLMGInfo, _ip, _ptr, _stream, check, default_options, load  # noqa: F401


class Adjoint:
    def __init__(self, n, prob_hi, limiter=1, rtol=1e-11, atol=1e-14, options=None):
        L = load()
        self.n = tuple(int(v) for v in n)
        nn = np.ascontiguousarray(self.n, dtype=np.int32)
        ph = np.ascontiguousarray(prob_hi, dtype=np.float64)
        self.options = options or default_options()
        h = C.c_void_p()
        check(L.ksgf_adjoint_create(nn.ctypes.data_as(_ip), ph.ctypes.data_as(C.POINTER(C.c_double)), int(limiter),
                                    rtol, atol, C.byref(self.options), C.byref(h)), "ksgf_adjoint_create")
        self.h = h
        self.nsteps = 0

    def _check(self, t, ncomp):
        import torch

        shape = (ncomp,) + self.n[::-1] if ncomp > 1 else self.n[::-1]
        if tuple(t.shape) != shape or t.dtype != torch.float64 or not t.is_contiguous() or not t.is_cuda:
            raise ValueError(f"expected a contiguous float64 CUDA tensor of shape {shape}")

    def forward(self, vel0, s0, nsteps, dt, every=5):
        import torch

        self._check(vel0, 3)
        self._check(s0, 1)
        vel = torch.empty_like(vel0)
        s = torch.empty_like(s0)
        iters = (C.c_int * max(nsteps, 1))()
        check(load().ksgf_adjoint_forward(self.h, _ptr(vel0), _ptr(s0), int(nsteps), float(dt), int(every), _ptr(vel),
                                          _ptr(s), iters, _stream()), "adjoint_forward")
        self.nsteps = nsteps
        return vel, s, list(iters)[:nsteps]

    def backward(self, vel_bar, s_bar):
        import torch

        self._check(vel_bar, 3)
        self._check(s_bar, 1)
        vb = torch.empty_like(vel_bar)
        sb = torch.empty_like(s_bar)
        check(load().ksgf_adjoint_backward(self.h, _ptr(vel_bar.contiguous()), _ptr(s_bar.contiguous()), _ptr(vb),
                                           _ptr(sb), _stream()), "adjoint_backward")
        return vb, sb

    def close(self):
        if self.h:
            load().ksgf_adjoint_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def advect_project_fn():
    """torch.autograd.Function: (vel0, s0) -> (vel_N, s_N) through an
    Adjoint object (passed as the first argument with nsteps, dt, every)."""
    import torch

    class AdvectProject(torch.autograd.Function):
        @staticmethod
        def forward(ctx, A, vel0, s0, nsteps, dt, every):
            vel, s, _ = A.forward(vel0.contiguous(), s0.contiguous(), nsteps, dt, every)
            ctx.A = A
            return vel, s

        @staticmethod
        def backward(ctx, vel_bar, s_bar):
            vb, sb = ctx.A.backward(vel_bar.contiguous(), s_bar.contiguous())
            return None, vb, sb, None, None, None

    return AdvectProject
