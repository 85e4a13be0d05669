"""Differentiable direct N-body integration on the GPU.

* :class:`PairForce` — torch.autograd.Function over the hand-written HIP
  kernels g6a_forces_* (acc, jerk of the softened pair sum, the arithmetic of
  lib/g6lib/g6lib.c:165-183) and g6a_adjoint_* (their vector-Jacobian
  product: dL/dx, dL/dv, dL/dm from dL/dacc, dL/djerk).  float32 and float64
  tensors on the GPU; no CPU fallback.
* :func:`hermite_integrate` — fixed-step 4th-order Hermite, the scheme of
  src/amuse_hermite/interface.cc (predictor :266-281, time-symmetric
  corrector :536-556, fixed step as in its test mode :618-637), with the
  state checkpointed every `checkpoint` steps and the steps in between
  recomputed in the backward pass.
* :func:`lagrangian_radii` and :func:`radii_loss` — mass radii about the
  centre of mass (the quantity of src/amuse/datamodel/particle_attributes.py
  LagrangianRadii, here about the COM), linearly interpolated in the
  cumulative mass so they depend smoothly on masses and positions.
"""
from __future__ import annotations

import ctypes as C

import torch
from torch.utils.checkpoint import checkpoint

from .g6 import load_library

_lib = None


def _kernels():
    global _lib
    if _lib is None:
        L = load_library()
        vp = C.c_void_p
        for name, real in (("f32", C.c_float), ("f64", C.c_double)):
            f = getattr(L, "g6a_forces_" + name)
            f.restype = C.c_int
            f.argtypes = [C.c_int, vp, vp, vp, real, vp, vp, vp]
            g = getattr(L, "g6a_adjoint_" + name)
            g.restype = C.c_int
            g.argtypes = [C.c_int, vp, vp, vp, vp, vp, real, vp, vp, vp, vp]
        _lib = L
    return _lib


def _check(t: torch.Tensor, name: str) -> torch.Tensor:
    if not t.is_cuda:
        raise RuntimeError(f"{name} must be a GPU tensor (the pair kernels have no CPU path)")
    return t.contiguous()


def _suffix(dtype):
    if dtype == torch.float32:
        return "f32"
    if dtype == torch.float64:
        return "f64"
    raise TypeError(f"unsupported dtype {dtype}")


class PairForce(torch.autograd.Function):
    """acc, jerk = PairForce.apply(x, v, m, eps2)  with x, v (n, 3), m (n,)."""

    @staticmethod
    def forward(ctx, x, v, m, eps2: float):
        x, v, m = _check(x, "x"), _check(v, "v"), _check(m, "m")
        n = x.shape[0]
        acc = torch.empty_like(x)
        jerk = torch.empty_like(x)
        L = _kernels()
        stream = torch.cuda.current_stream(x.device).cuda_stream
        rc = getattr(L, "g6a_forces_" + _suffix(x.dtype))(
            n, x.data_ptr(), v.data_ptr(), m.data_ptr(), float(eps2), acc.data_ptr(),
            jerk.data_ptr(), stream)
        if rc != 0:
            raise RuntimeError(f"g6a_forces failed ({rc})")
        ctx.save_for_backward(x, v, m)
        ctx.eps2 = float(eps2)
        return acc, jerk

    @staticmethod
    def backward(ctx, ga, gj):
        x, v, m = ctx.saved_tensors
        ga = torch.zeros_like(x) if ga is None else ga.contiguous()
        gj = torch.zeros_like(x) if gj is None else gj.contiguous()
        n = x.shape[0]
        gx = torch.empty_like(x)
        gv = torch.empty_like(v)
        gm = torch.empty_like(m)
        L = _kernels()
        stream = torch.cuda.current_stream(x.device).cuda_stream
        rc = getattr(L, "g6a_adjoint_" + _suffix(x.dtype))(
            n, x.data_ptr(), v.data_ptr(), m.data_ptr(), ga.data_ptr(), gj.data_ptr(), ctx.eps2,
            gx.data_ptr(), gv.data_ptr(), gm.data_ptr(), stream)
        if rc != 0:
            raise RuntimeError(f"g6a_adjoint failed ({rc})")
        return gx, gv, gm, None


def pair_force(x, v, m, eps2):
    return PairForce.apply(x, v, m, eps2)


def hermite_step(x, v, a, j, m, dt: float, eps2: float):
    """One fixed step (src/amuse_hermite/interface.cc:266-281, 536-556)."""
    xp = x + v * dt + a * (dt * dt / 2) + j * (dt * dt * dt / 6)
    vp = v + a * dt + j * (dt * dt / 2)
    a1, j1 = pair_force(xp, vp, m, eps2)
    v1 = v + (a + a1) * (dt / 2) + (j - j1) * (dt * dt / 12)
    x1 = x + (v + v1) * (dt / 2) + (a - a1) * (dt * dt / 12)
    return x1, v1, a1, j1


def _segment(nsteps: int, dt: float, eps2: float):
    def run(x, v, a, j, m):
        for _ in range(nsteps):
            x, v, a, j = hermite_step(x, v, a, j, m, dt, eps2)
        return x, v, a, j
    return run


def hermite_integrate(x0, v0, m, dt: float, nsteps: int, eps2: float, checkpoint_every: int = 16):
    """Final (x, v) after nsteps fixed steps; the autograd graph keeps only the
    state every `checkpoint_every` steps (0: keep every step)."""
    a, j = pair_force(x0, v0, m, eps2)
    x, v = x0, v0
    done = 0
    while done < nsteps:
        k = min(checkpoint_every if checkpoint_every > 0 else nsteps, nsteps - done)
        seg = _segment(k, dt, eps2)
        if checkpoint_every > 0 and torch.is_grad_enabled():
            x, v, a, j = checkpoint(seg, x, v, a, j, m, use_reentrant=False)
        else:
            x, v, a, j = seg(x, v, a, j, m)
        done += k
    return x, v


def lagrangian_radii(x, m, fractions=(0.1, 0.5, 0.9)):
    """Mass radii about the centre of mass, interpolated in cumulative mass."""
    M = m.sum()
    com = (m[:, None] * x).sum(0) / M
    r = torch.linalg.vector_norm(x - com, dim=1)
    order = torch.argsort(r.detach())
    rs = r[order]
    cm = torch.cumsum(m[order], 0) / M
    out = []
    for q in fractions:
        qt = torch.as_tensor(q, dtype=x.dtype, device=x.device)
        k = int(torch.searchsorted(cm.detach(), qt).item())
        k = min(max(k, 1), r.numel() - 1)
        c0, c1 = cm[k - 1], cm[k]
        w = (qt - c0) / (c1 - c0)
        out.append(rs[k - 1] + w * (rs[k] - rs[k - 1]))
    return torch.stack(out)


def radii_loss(x_final, m, targets):
    """Squared distance of the final 10/50/90 % mass radii from targets."""
    return ((lagrangian_radii(x_final, m) - targets) ** 2).sum()
