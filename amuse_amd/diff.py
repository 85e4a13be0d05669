"""Differentiable direct N-body integration on the GPU.

* :class:`PairForce` — torch.autograd.Function over the hand-written HIP
  kernels g6a_forces_* (acc, jerk of the softened pair sum, the arithmetic of
  lib/g6lib/g6lib.c:165-183) and g6a_adjoint_* (their vector-Jacobian
  product: dL/dx, dL/dv, dL/dm from dL/dacc, dL/djerk).  float32 and float64
  tensors on the GPU; no CPU fallback.
* :class:`HermiteStep` — one fixed 4th-order Hermite step, the scheme of
  src/amuse_hermite/interface.cc (predictor :266-281, time-symmetric
  corrector :536-556, fixed step as in its test mode :618-637): predictor,
  pair sum and corrector as three kernels (g6a_step_predict_*, g6a_forces_*,
  g6a_step_correct_*), and in the backward pass the transposes of the two
  linear maps around the pair-sum adjoint (g6a_step_correct_t_*,
  g6a_adjoint_*, g6a_step_predict_t_*).
* :func:`hermite_integrate` — HermiteStep repeated, with the state
  checkpointed every `checkpoint` steps and the steps in between recomputed
  in the backward pass.
* :func:`lagrangian_radii` and :func:`radii_loss` — mass radii about the
  centre of mass (the quantity of src/amuse/datamodel/particle_attributes.py
  LagrangianRadii, here about the COM), linearly interpolated in the
  cumulative mass so they depend smoothly on masses and positions.  The
  nodes sit at the middle of each particle's cumulative-mass interval: with
  equal masses the fractions 0.1 / 0.5 / 0.9 would otherwise fall exactly on
  a node for many N (0.5 for every even N), a kink where the gradient in the
  masses is one-sided.
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
        d = C.c_double
        for name, real in (("f32", C.c_float), ("f64", C.c_double)):
            f = getattr(L, "g6a_forces_" + name)
            f.restype = C.c_int
            f.argtypes = [C.c_int, vp, vp, vp, real, vp, vp, vp]
            g = getattr(L, "g6a_adjoint_" + name)
            g.restype = C.c_int
            g.argtypes = [C.c_int, vp, vp, vp, vp, vp, real, vp, vp, vp, vp]
            for fn, args in (("g6a_step_predict_", [C.c_int, vp, vp, vp, vp, d, vp, vp, vp]),
                             ("g6a_step_correct_", [C.c_int, vp, vp, vp, vp, vp, vp, d, vp, vp,
                                                    vp]),
                             ("g6a_step_correct_t_", [C.c_int, vp, vp, vp, vp, d, vp, vp, vp, vp,
                                                      vp, vp, vp]),
                             ("g6a_step_predict_t_", [C.c_int, vp, vp, d, vp, vp, vp, vp, vp])):
                h = getattr(L, fn + name)
                h.restype = C.c_int
                h.argtypes = args
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


def _ptr(t):
    return None if t is None else t.data_ptr()


def _call(name, dtype, *args):
    rc = getattr(_kernels(), name + _suffix(dtype))(*args)
    if rc != 0:
        raise RuntimeError(f"{name}{_suffix(dtype)} failed ({rc})")


class HermiteStep(torch.autograd.Function):
    """x1, v1, a1, j1 = HermiteStep.apply(x, v, a, j, m, dt, eps2): one fixed
    step (src/amuse_hermite/interface.cc:266-281, 536-556) in three kernels;
    the backward pass runs the transposes of the predictor and corrector
    around the pair-sum adjoint."""

    @staticmethod
    def forward(ctx, x, v, a, j, m, dt: float, eps2: float):
        x, v, a, j, m = (_check(t, n) for t, n in ((x, "x"), (v, "v"), (a, "a"), (j, "j"),
                                                    (m, "m")))
        n = x.shape[0]
        stream = torch.cuda.current_stream(x.device).cuda_stream
        xp, vp = torch.empty_like(x), torch.empty_like(v)
        _call("g6a_step_predict_", x.dtype, n, x.data_ptr(), v.data_ptr(), a.data_ptr(),
              j.data_ptr(), float(dt), xp.data_ptr(), vp.data_ptr(), stream)
        a1, j1 = torch.empty_like(x), torch.empty_like(x)
        _call("g6a_forces_", x.dtype, n, xp.data_ptr(), vp.data_ptr(), m.data_ptr(), float(eps2),
              a1.data_ptr(), j1.data_ptr(), stream)
        x1, v1 = torch.empty_like(x), torch.empty_like(v)
        _call("g6a_step_correct_", x.dtype, n, x.data_ptr(), v.data_ptr(), a.data_ptr(),
              j.data_ptr(), a1.data_ptr(), j1.data_ptr(), float(dt), x1.data_ptr(),
              v1.data_ptr(), stream)
        ctx.save_for_backward(xp, vp, m)
        ctx.dt, ctx.eps2 = float(dt), float(eps2)
        return x1, v1, a1, j1

    @staticmethod
    def backward(ctx, gx1, gv1, ga1, gj1):
        xp, vp, m = ctx.saved_tensors
        n = xp.shape[0]
        stream = torch.cuda.current_stream(xp.device).cuda_stream
        gx1 = torch.zeros_like(xp) if gx1 is None else gx1.contiguous()
        gv1 = torch.zeros_like(xp) if gv1 is None else gv1.contiguous()
        ga1 = None if ga1 is None else ga1.contiguous()
        gj1 = None if gj1 is None else gj1.contiguous()
        gx, gv, ga, gj = (torch.empty_like(xp) for _ in range(4))
        ga1t, gj1t = torch.empty_like(xp), torch.empty_like(xp)
        _call("g6a_step_correct_t_", xp.dtype, n, gx1.data_ptr(), gv1.data_ptr(), _ptr(ga1),
              _ptr(gj1), ctx.dt, gx.data_ptr(), gv.data_ptr(), ga.data_ptr(), gj.data_ptr(),
              ga1t.data_ptr(), gj1t.data_ptr(), stream)
        gxp, gvp, gm = torch.empty_like(xp), torch.empty_like(vp), torch.empty_like(m)
        _call("g6a_adjoint_", xp.dtype, n, xp.data_ptr(), vp.data_ptr(), m.data_ptr(),
              ga1t.data_ptr(), gj1t.data_ptr(), ctx.eps2, gxp.data_ptr(), gvp.data_ptr(),
              gm.data_ptr(), stream)
        _call("g6a_step_predict_t_", xp.dtype, n, gxp.data_ptr(), gvp.data_ptr(), ctx.dt,
              gx.data_ptr(), gv.data_ptr(), ga.data_ptr(), gj.data_ptr(), stream)
        return gx, gv, ga, gj, gm, None, None


def hermite_step(x, v, a, j, m, dt: float, eps2: float):
    """One fixed step through the fused kernels (HermiteStep)."""
    return HermiteStep.apply(x, v, a, j, m, dt, eps2)


def hermite_step_reference(x, v, a, j, m, dt: float, eps2: float):
    """The same step as torch operations around PairForce (the reference the
    fused step is tested against)."""
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
    """Mass radii about the centre of mass, interpolated in cumulative mass
    between nodes at the middle of each particle's cumulative-mass
    interval."""
    M = m.sum()
    com = (m[:, None] * x).sum(0) / M
    r = torch.linalg.vector_norm(x - com, dim=1)
    order = torch.argsort(r.detach())
    rs = r[order]
    ms = m[order]
    cm = (torch.cumsum(ms, 0) - 0.5 * ms) / M
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
