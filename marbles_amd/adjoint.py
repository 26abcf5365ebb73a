"""Gradients of a flow objective through the isothermal partial-bounce-back
step, by the engine's hand-written adjoint (C ABI mrb_iso_*), wrapped in a
torch.autograd.Function so that torch code can differentiate through it.

Parameters: per-cell solid fraction n_s (nz, ny, nx), relaxation time tau
(scalar), initial velocity u0 (3, nz, ny, nx); all float64 tensors on the
GPU.  The loss is the mean outlet-plane x-velocity plus 1e-3 times the final
kinetic energy.
"""
from __future__ import annotations

import ctypes
from ctypes import c_double, c_int, c_void_p

import numpy as np
import torch

from . import _lib
from ._lib import check


def _dev(t: torch.Tensor, shape) -> torch.Tensor:
    if t.device.type != "cuda" or t.dtype != torch.float64:
        raise ValueError("expected a float64 tensor on the GPU")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"expected shape {tuple(shape)}, got {tuple(t.shape)}")
    return t.contiguous()


class IsoSolver:
    """Isothermal D3Q27 box with inlet (x-lo), outflow (x-hi), periodic y/z."""

    def __init__(self, ncell, T0: float = 1.0 / 3.0, rho0: float = 1.0, u_in: float = 0.04):
        self.lib = _lib.load()
        self.ncell = tuple(int(n) for n in ncell)
        self.shape = tuple(reversed(self.ncell))
        h = c_void_p()
        check(self.lib.mrb_iso_create((c_int * 3)(*self.ncell), float(T0), float(rho0), float(u_in),
                                      ctypes.byref(h)))
        self.h = h

    def _stream(self):
        return c_void_p(torch.cuda.current_stream().cuda_stream)

    def forward(self, ns: torch.Tensor, tau: float, u0: torch.Tensor, nsteps: int) -> float:
        ns = _dev(ns, self.shape)
        u0 = _dev(u0, (3,) + self.shape)
        J = c_double()
        self._rec = None  # the engine's forward replaces a recording
        check(self.lib.mrb_iso_forward(self.h, c_void_p(ns.data_ptr()), float(tau), c_void_p(u0.data_ptr()),
                                       int(nsteps), ctypes.byref(J), self._stream()))
        return J.value

    def gradient(self, ns: torch.Tensor, tau: float, u0: torch.Tensor, nsteps: int, every: int = -1):
        """(J, dJ/dtau, dJ/dn_s, dJ/du0) with checkpoints every `every` steps
        (0: ceil(sqrt(nsteps)); negative: the longest segment whose states fit
        half of the free device memory, the whole run when they fit, so that
        nothing is recomputed).  Every interval gives the same gradient."""
        ns = _dev(ns, self.shape)
        u0 = _dev(u0, (3,) + self.shape)
        dns = torch.empty_like(ns)
        du0 = torch.empty_like(u0)
        J, dtau = c_double(), c_double()
        check(self.lib.mrb_iso_gradient(self.h, c_void_p(ns.data_ptr()), float(tau), c_void_p(u0.data_ptr()),
                                        int(nsteps), int(every), ctypes.byref(J), ctypes.byref(dtau),
                                        c_void_p(dns.data_ptr()), c_void_p(du0.data_ptr()), self._stream()))
        return J.value, dtau.value, dns, du0

    def record(self, ns: torch.Tensor, tau: float, u0: torch.Tensor, nsteps: int, every: int = -1):
        """Forward sweep keeping the checkpoints for reverse(); returns (J, token).
        The token names this recording: a later record() or forward() replaces it."""
        ns = _dev(ns, self.shape)
        u0 = _dev(u0, (3,) + self.shape)
        J = c_double()
        check(self.lib.mrb_iso_record(self.h, c_void_p(ns.data_ptr()), float(tau), c_void_p(u0.data_ptr()),
                                      int(nsteps), int(every), ctypes.byref(J), self._stream()))
        self._token = getattr(self, "_token", 0) + 1
        self._rec = (ns, u0)  # the engine reads them again in reverse()
        return J.value, self._token

    def record_interval(self) -> int:
        """Checkpoint interval of the last record() or gradient()."""
        e = ctypes.c_int()
        check(self.lib.mrb_iso_record_interval(self.h, ctypes.byref(e)))
        return e.value

    def reverse(self, token: int):
        """(dJ/dtau, dJ/dn_s, dJ/du0) of the recording `token`, or None when a
        later record() or forward() has replaced it."""
        if getattr(self, "_token", None) != token or getattr(self, "_rec", None) is None:
            return None
        ns, u0 = self._rec
        dns = torch.empty_like(ns)
        du0 = torch.empty_like(u0)
        dtau = c_double()
        check(self.lib.mrb_iso_reverse(self.h, ctypes.byref(dtau), c_void_p(dns.data_ptr()),
                                       c_void_p(du0.data_ptr()), self._stream()))
        self._rec = None
        return dtau.value, dns, du0

    def state(self) -> np.ndarray:
        """Final populations of the last forward run, shape (27, nz, ny, nx)."""
        out = np.empty((27,) + self.shape)
        check(self.lib.mrb_iso_state(self.h, out.ctypes.data_as(ctypes.POINTER(c_double))))
        return out

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.mrb_iso_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FlowLoss(torch.autograd.Function):
    """J(n_s, tau, u0).  The forward pass records the checkpointed forward
    sweep; backward runs the reverse sweep (or, if another evaluation on the
    same solver has replaced the recording, the whole gradient again)."""

    @staticmethod
    def forward(ctx, ns, tau, u0, solver: IsoSolver, nsteps: int, every: int = -1):
        if any(ctx.needs_input_grad[:3]):
            J, token = solver.record(ns.detach(), float(tau), u0.detach(), nsteps, every)
            ctx.solver, ctx.token, ctx.args = solver, token, (float(tau), nsteps, every)
            ctx.save_for_backward(ns.detach(), u0.detach())
        else:
            J = solver.forward(ns.detach(), float(tau), u0.detach(), nsteps)
        return torch.tensor(J, dtype=torch.float64, device=ns.device)

    @staticmethod
    def backward(ctx, gJ):
        ns, u0 = ctx.saved_tensors
        got = ctx.solver.reverse(ctx.token)
        if got is None:
            tau, nsteps, every = ctx.args
            _, dtau, dns, du0 = ctx.solver.gradient(ns, tau, u0, nsteps, every)
        else:
            dtau, dns, du0 = got
        dtau = torch.tensor(dtau, dtype=torch.float64, device=ns.device)
        return gJ * dns, (gJ * dtau).reshape(()), gJ * du0, None, None, None


def flow_loss(solver: IsoSolver, ns, tau, u0, nsteps: int, every: int = -1):
    return FlowLoss.apply(ns, tau, u0, solver, nsteps, every)
