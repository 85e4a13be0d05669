"""Differentiable integration: the adjoint kernel and the checkpointed
Hermite integrator (amuse_amd/diff.py).

* forward kernels: FP64 within 1e-15 sum|term| of the 80-bit oracle, FP32
  within 2e-6 sum|term|, at n = 2000 and 16384 (every lanes-per-i form the
  sizes select, one and two i per lane);
* adjoint kernel: equal to torch autograd through a plain float64 torch
  restatement of the pair sum (1e-11 relative), for x, v and m, at
  n = 96 .. 16384 (64 down to 8 lanes per i), and dL/dx, dL/dv summing to
  zero over the particles (translation and Galilean invariance);
* whole pipeline (integrate + Lagrangian-radius loss) in FP64: gradients
  match central finite differences along random directions;
* the fused Hermite step (predictor, pair sum and corrector kernels, with
  the transposed maps in the backward pass) equals the same step as torch
  operations around the pair-sum Function, values and gradients, in FP64;
* checkpointed and plain gradients are bit-identical;
* FP32 gradients agree with the FP64 reference gradient to the stated level.
"""
import numpy as np
import pytest

import oracle
from amuse_amd import ic

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def tensors(n, seed, dtype=torch.float64, eps_scale=1.0):
    m, x, v = ic.plummer(n, seed)
    rng = np.random.Generator(np.random.PCG64(seed))
    m = m * rng.uniform(0.5, 1.5, n)
    dev = torch.device("cuda")
    return (torch.tensor(x, dtype=dtype, device=dev), torch.tensor(v, dtype=dtype, device=dev),
            torch.tensor(m, dtype=dtype, device=dev))


def _pair_rows(x, v, m, eps2, i0, i1):
    """acc, jerk of particles i0..i1-1 as plain torch operations."""
    r = x[None, :, :] - x[i0:i1, None, :]
    w = v[None, :, :] - v[i0:i1, None, :]
    s = (r * r).sum(-1) + eps2
    self_pair = torch.arange(x.shape[0], device=x.device)[None, :] == \
        torch.arange(i0, i1, device=x.device)[:, None]
    y = torch.where(self_pair, torch.zeros_like(s), torch.rsqrt(s))
    y3, y5 = y ** 3, y ** 5
    rv = (r * w).sum(-1)
    acc = (m[None, :, None] * r * y3[..., None]).sum(1)
    jerk = (m[None, :, None] * (w * y3[..., None] - 3 * rv[..., None] * r * y5[..., None])).sum(1)
    return acc, jerk


def torch_pair_force(x, v, m, eps2, rows=2048):
    """The FP64 torch restatement of the pair sum; above `rows` particles in
    checkpointed blocks of rows (the saved tensors of one block at a time)."""
    from torch.utils.checkpoint import checkpoint
    n = x.shape[0]
    if n <= rows:
        return _pair_rows(x, v, m, eps2, 0, n)
    parts = [checkpoint(_pair_rows, x, v, m, eps2, i0, min(n, i0 + rows), use_reentrant=False)
             for i0 in range(0, n, rows)]
    return torch.cat([p[0] for p in parts]), torch.cat([p[1] for p in parts])


def test_fused_step_matches_torch_step():
    """HermiteStep against hermite_step_reference (torch operations around
    PairForce) on the same FP64 inputs: outputs within 1e-13 of their scale,
    vector-Jacobian products (all four cotangents, and x1 alone) within
    1e-11."""
    from amuse_amd.diff import hermite_step, hermite_step_reference, pair_force
    x, v, m = tensors(1536, 5)
    a, j = pair_force(x, v, m, 1e-4)
    rng = torch.Generator(device="cuda").manual_seed(3)
    dt = 1.0 / 256
    ins = [t.clone().requires_grad_(True) for t in (x, v, a, j, m)]
    ref = [t.clone().requires_grad_(True) for t in (x, v, a, j, m)]
    out = hermite_step(*ins, dt, 1e-4)
    oref = hermite_step_reference(*ref, dt, 1e-4)
    for o, r in zip(out, oref):
        assert (o - r).abs().max() <= 1e-13 * r.abs().max()
    cot = [torch.randn(o.shape, dtype=o.dtype, device=o.device, generator=rng) for o in out]
    for which in (slice(0, 4), slice(0, 1)):
        g = torch.autograd.grad(out[which], ins, cot[which], retain_graph=True)
        gr = torch.autograd.grad(oref[which], ref, cot[which], retain_graph=True)
        for name, u, w in zip(("x", "v", "a", "j", "m"), g, gr):
            err = ((u - w).abs().max() / w.abs().max()).item()
            print("fused step grad", name, which, err, flush=True)
            assert err <= 1e-11, (name, which, err)


@pytest.mark.parametrize("n", [2000, 16384])
def test_forward_kernels_vs_oracle(n):
    """FP64 against the oracle on the same inputs; FP32 against the oracle on
    the float-rounded inputs (the kernel's arithmetic, not input rounding).
    n = 2000 runs 64 lanes per i; n = 16384 16 (FP64) and 32 with two i per
    lane (FP32), checked on 512 sampled i."""
    from amuse_amd.diff import pair_force
    from conftest import par_oracle
    x, v, m = tensors(n, 1)
    ids = np.arange(n, dtype=np.int32)
    sel = np.arange(n) if n <= 2000 else np.arange(0, n, n // 512)
    for dtype, tols in ((torch.float64, (1e-15, 4e-15)), (torch.float32, (2e-6, 1e-5))):
        xx, vv, mm = x.to(dtype), v.to(dtype), m.to(dtype)
        a, j = pair_force(xx, vv, mm, 1e-4)
        xn, vn, mn = (t.double().cpu().numpy() for t in (xx, vv, mm))
        ea, ej, _, sa, sj, _ = par_oracle(oracle.force_ld, ids[sel], xn[sel], vn[sel], 1e-4, ids,
                                          mn, xn, vn)
        for val, ex, s, tol, name in ((a[sel], ea, sa, tols[0], "acc"), (j[sel], ej, sj, tols[1], "jerk")):
            d = np.abs(val.double().cpu().numpy() - ex).max(1)
            print(dtype, name, "max |d|/sum|term|", (d / s).max(), flush=True)
            assert (d <= tol * s).all(), (dtype, name, (d / s).max())
        # every particle: Newton's third law, sum m a = sum m jerk = 0
        for val, name in ((a, "acc"), (j, "jerk")):
            q = val.double().cpu().numpy()
            r = np.linalg.norm((mn[:, None] * q).sum(0)) / (mn * np.linalg.norm(q, axis=1)).sum()
            print(dtype, name, "|sum m q|/sum|m q|", r, flush=True)
            # (measured: FP64 <= 7e-17, FP32 <= 5.6e-9)
            assert r <= (1e-14 if dtype == torch.float64 else 1e-6), (dtype, name, r)


@pytest.mark.parametrize("n", [96, 4096, 8192, 16384])
def test_adjoint_matches_autograd_reference(n):
    """Against torch autograd of the FP64 restatement: n = 96, 4096, 8192
    and 16384 run the adjoint with 64, 32, 16 and 8 lanes per i."""
    from amuse_amd.diff import pair_force
    x, v, m = tensors(n, 2)
    gen = torch.Generator(device="cuda").manual_seed(3)
    ga = torch.randn(x.shape, dtype=x.dtype, device="cuda", generator=gen)
    gj = torch.randn(x.shape, dtype=x.dtype, device="cuda", generator=gen)
    grads = []
    for fn in (pair_force, torch_pair_force):
        xx, vv, mm = (t.clone().requires_grad_(True) for t in (x, v, m))
        a, j = fn(xx, vv, mm, 1e-3)
        ((a * ga).sum() + (j * gj).sum()).backward()
        grads.append((xx.grad, vv.grad, mm.grad))
    for g, r in zip(*grads):
        rel = (g - r).abs().max() / r.abs().max()
        assert rel < 1e-11, rel.item()
    # translation and Galilean invariance of the pair sum: shifting every x
    # (or every v) by the same vector changes no force, so the kernel's
    # dL/dx and dL/dv sum to zero over the particles up to rounding
    for g, name in ((grads[0][0], "x"), (grads[0][1], "v")):
        r = (torch.linalg.vector_norm(g.sum(0)) / torch.linalg.vector_norm(g, dim=1).sum()).item()
        print(f"n={n} |sum dL/d{name}|/sum|dL/d{name}| {r:.2e}", flush=True)
        assert r < 1e-15, (name, r)   # measured <= 4.4e-17 for n = 96 .. 16384


def loss_fn(x0, v0, m, nsteps, dt, eps2, ck, targets=None):
    """Radius loss after the run; targets (0.9 x the initial radii of the
    unperturbed state) are constants of the loss, as in the configuration."""
    from amuse_amd.diff import hermite_integrate, lagrangian_radii, radii_loss
    if targets is None:
        with torch.no_grad():
            targets = 0.9 * lagrangian_radii(x0, m)
    xf, _ = hermite_integrate(x0, v0, m, dt, nsteps, eps2, checkpoint_every=ck)
    return radii_loss(xf, m, targets)


def test_pipeline_gradient_vs_finite_differences():
    from amuse_amd.diff import lagrangian_radii
    x, v, m = tensors(64, 4)
    with torch.no_grad():
        tg = 0.9 * lagrangian_radii(x, m)
    args = (16, 1.0 / 64, 1e-3, 4, tg)
    xx, vv, mm = (t.clone().requires_grad_(True) for t in (x, v, m))
    L = loss_fn(xx, vv, mm, *args)
    L.backward()
    gen = torch.Generator(device="cuda").manual_seed(5)
    for which in range(3):
        base = [x, v, m][which]
        g = [xx.grad, vv.grad, mm.grad][which]
        d = torch.randn(base.shape, dtype=base.dtype, device="cuda", generator=gen)
        h = 1e-6
        with torch.no_grad():
            p = [x.clone(), v.clone(), m.clone()]
            q = [x.clone(), v.clone(), m.clone()]
            p[which] += h * d
            q[which] -= h * d
            fd = (loss_fn(*p, *args) - loss_fn(*q, *args)) / (2 * h)
        an = (g * d).sum()
        print("wrt", which, "analytic", an.item(), "finite difference", fd.item(), flush=True)
        assert abs(an - fd) <= 1e-6 * abs(fd) + 1e-12, (which, an.item(), fd.item())


def test_checkpointing_is_exact():
    x, v, m = tensors(128, 6)
    res = []
    for ck in (0, 4):
        xx, vv, mm = (t.clone().requires_grad_(True) for t in (x, v, m))
        loss_fn(xx, vv, mm, 16, 1.0 / 64, 1e-3, ck).backward()
        res.append((xx.grad, vv.grad, mm.grad))
    for a, b in zip(*res):
        assert torch.equal(a, b)


def test_fp32_gradient_vs_fp64():
    x, v, m = tensors(1024, 7)
    grads = {}
    for dtype in (torch.float64, torch.float32):
        xx, vv, mm = (t.to(dtype).clone().requires_grad_(True) for t in (x, v, m))
        loss_fn(xx, vv, mm, 32, 1.0 / 256, 1e-4, 16).backward()
        grads[dtype] = [g.double() for g in (xx.grad, vv.grad, mm.grad)]
    for g32, g64 in zip(grads[torch.float32], grads[torch.float64]):
        rel = torch.linalg.vector_norm(g32 - g64) / torch.linalg.vector_norm(g64)
        assert rel < 1e-3, rel.item()
