"""Differentiable Fellegi-Sunter likelihood over a pattern histogram (FP64).

The reference only fits the model by EM (faster/estimation.py:98-158). This
module adds the observed-data log-likelihood as a ``torch.autograd.Function``
whose forward and hand-written backward are HIP kernels (fh_fs_loglik_fwd /
fh_fs_loglik_bwd), its closed-form Hessian (fh_fs_loglik_hess), observed-
information standard errors, and gradient-based fits.

Parameters are logits ``theta = [a, beta_0..beta_{K-1}, alpha_0..alpha_{K-1}]``
with ``lambda = sigmoid(a)``, ``pi_0[k] = softmax(beta_k)`` (non-match) and
``pi_1[k] = softmax(alpha_k)`` (match), the order of the EM parameters
``[lambda, pi_0, pi_1]``. Softmax logits are shift-redundant, so standard
errors and fits use the *free* coordinates: every logit except level 0 of
each variable and class, which is held at 0.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check, ptr, stream


def _levels_arr(levels):
    return (ctypes.c_int32 * len(levels))(*[int(x) for x in levels])


def _dev_counts(counts):
    c = torch.as_tensor(np.asarray(counts, dtype=np.int64) if not isinstance(counts, torch.Tensor) else counts)
    return c.to(device=_lib.device(), dtype=torch.int64).contiguous()


class FSLogLik(torch.autograd.Function):
    """LL(theta) = sum_p N_p log(lambda P_1(p) + (1 - lambda) P_0(p))."""

    @staticmethod
    def forward(ctx, theta, counts, levels):
        L = _lib.load()
        th = theta.detach().to(device=_lib.device(), dtype=torch.float64).contiguous()
        out = torch.empty(1, dtype=torch.float64, device=_lib.device())
        check(L.fh_fs_loglik_fwd(len(levels), _levels_arr(levels), ptr(counts), ptr(th), ptr(out), ptr(None),
                                 stream()), "fh_fs_loglik_fwd")
        ctx.save_for_backward(th)
        ctx.counts = counts
        ctx.levels = levels
        ctx.in_device = theta.device
        return out[0].to(theta.device)

    @staticmethod
    def backward(ctx, grad_out):
        L = _lib.load()
        (th,) = ctx.saved_tensors
        g = grad_out.detach().reshape(1).to(device=_lib.device(), dtype=torch.float64).contiguous()
        grad = torch.empty_like(th)
        check(L.fh_fs_loglik_bwd(len(ctx.levels), _levels_arr(ctx.levels), ptr(ctx.counts), ptr(th), ptr(g),
                                 ptr(grad), stream()), "fh_fs_loglik_bwd")
        return grad.to(ctx.in_device), None, None


def fs_loglik(theta: torch.Tensor, counts, levels) -> torch.Tensor:
    """Differentiable log-likelihood (FP64 scalar tensor)."""
    return FSLogLik.apply(theta, _dev_counts(counts), tuple(int(x) for x in levels))


def fs_posteriors(theta: torch.Tensor, counts, levels) -> torch.Tensor:
    """Posterior match probability of every pattern under ``theta``."""
    L = _lib.load()
    c = _dev_counts(counts)
    th = theta.detach().to(device=_lib.device(), dtype=torch.float64).contiguous()
    out = torch.empty(1, dtype=torch.float64, device=_lib.device())
    post = torch.empty(c.numel(), dtype=torch.float64, device=_lib.device())
    check(L.fh_fs_loglik_fwd(len(levels), _levels_arr(levels), ptr(c), ptr(th), ptr(out), ptr(post), stream()),
          "fh_fs_loglik_fwd")
    return post


def fs_hessian(theta: torch.Tensor, counts, levels) -> torch.Tensor:
    """Closed-form Hessian d2 LL / d theta2 (dim x dim, FP64)."""
    L = _lib.load()
    c = _dev_counts(counts)
    th = theta.detach().to(device=_lib.device(), dtype=torch.float64).contiguous()
    dim = th.numel()
    h = torch.empty(dim * dim, dtype=torch.float64, device=_lib.device())
    check(L.fh_fs_loglik_hess(len(levels), _levels_arr(levels), ptr(c), ptr(th), ptr(h), stream()),
          "fh_fs_loglik_hess")
    return h.reshape(dim, dim)


# ---------------------------------------------------------------------------
# Parameterisation helpers
# ---------------------------------------------------------------------------

def free_index(levels) -> np.ndarray:
    """Positions of the free logits (level 0 of every variable and class is fixed at 0)."""
    n_lv = int(np.sum(levels))
    keep = [0]
    for cls in range(2):
        off = 0
        for L in levels:
            keep.extend(1 + cls * n_lv + off + l for l in range(1, int(L)))
            off += int(L)
    return np.array(keep, dtype=np.int64)


def params_to_logits(lam, Pi, levels) -> torch.Tensor:
    """(lambda, [pi_0, pi_1]) -> theta with level-0 logits at 0."""
    th = [np.log(lam) - np.log1p(-lam)]
    for cls in range(2):
        for k, L in enumerate(levels):
            p = np.asarray(Pi[cls][k], dtype=np.float64)
            th.extend((np.log(p) - np.log(p[0])).tolist())
    return torch.tensor(th, dtype=torch.float64)


def logits_to_params(theta, levels):
    t = theta.detach().cpu().numpy().astype(np.float64)
    lam = 1.0 / (1.0 + np.exp(-t[0]))
    n_lv = int(np.sum(levels))
    Pi = [[], []]
    for cls in range(2):
        off = 0
        for L in levels:
            z = t[1 + cls * n_lv + off:1 + cls * n_lv + off + int(L)]
            e = np.exp(z - z.max())
            Pi[cls].append(e / e.sum())
            off += int(L)
    return float(lam), Pi


def observed_information(theta, counts, levels) -> torch.Tensor:
    """-Hessian over the free logits."""
    H = fs_hessian(theta, counts, levels).cpu()
    idx = torch.as_tensor(free_index(levels))
    return -H[idx][:, idx]


def standard_errors(theta, counts, levels):
    """Observed-information standard errors of the free logits and, by the
    delta method, of lambda and every pi (dict of numpy arrays)."""
    info = observed_information(theta, counts, levels).numpy()
    cov = np.linalg.inv(info)
    se_free = np.sqrt(np.diag(cov))
    idx = free_index(levels)
    th = theta.detach().cpu().numpy().astype(np.float64)

    def natural(x_free):
        t = th.copy()
        t[idx] = x_free
        lam, Pi = logits_to_params(torch.as_tensor(t), levels)
        return np.concatenate([[lam], np.concatenate(Pi[0]), np.concatenate(Pi[1])])

    # Jacobian of the natural parameters w.r.t. the free logits (analytic).
    lam, Pi = logits_to_params(theta, levels)
    n_nat = 1 + 2 * int(np.sum(levels))
    J = np.zeros((n_nat, len(idx)))
    J[0, 0] = lam * (1 - lam)
    col = 1
    row_base = 1
    for cls in range(2):
        off = 0
        for k, L in enumerate(levels):
            p = Pi[cls][k]
            for l in range(1, int(L)):
                # d pi_m / d z_l = pi_m (delta_ml - pi_l)
                for m in range(int(L)):
                    J[row_base + cls * int(np.sum(levels)) + off + m, col] = p[m] * ((m == l) - p[l])
                col += 1
            off += int(L)
    cov_nat = J @ cov @ J.T
    return {"free": se_free, "natural": np.sqrt(np.maximum(np.diag(cov_nat), 0.0)), "cov_free": cov,
            "natural_values": natural(th[idx])}


# ---------------------------------------------------------------------------
# Fits
# ---------------------------------------------------------------------------

def fit(counts, levels, theta0, method="lbfgs", max_iter=500, tol=1e-14):
    """Maximises LL over the free logits from ``theta0``.

    'lbfgs': torch L-BFGS (strong-Wolfe line search) on LL / N, then Newton
    polishing steps with the closed-form Hessian; 'newton': Newton steps only
    (from a good start); 'gd': gradient descent on -LL / N with
    Barzilai-Borwein step lengths, stopping when the largest gradient entry
    is below ``tol``.
    Returns (theta, LL, iterations)."""
    c = _dev_counts(counts)
    lv = tuple(int(x) for x in levels)
    idx = torch.as_tensor(free_index(levels))
    base = theta0.detach().to(torch.float64).cpu().clone()
    N = float(c.sum().item())
    x = base[idx].clone().requires_grad_(True)

    def full(xf):
        t = base.clone()
        t = t.index_put((idx,), xf)
        return t

    def objective(xf):
        return -fs_loglik(full(xf), c, lv) / N

    iters = 0
    if method == "lbfgs":
        opt = torch.optim.LBFGS([x], lr=1.0, max_iter=max_iter, history_size=50, tolerance_grad=1e-13,
                                tolerance_change=1e-16, line_search_fn="strong_wolfe")

        def closure():
            opt.zero_grad()
            f = objective(x)
            f.backward()
            return f

        opt.step(closure)
        iters = opt.state[opt._params[0]].get("n_iter", 0)
    elif method == "gd":
        # Gradient descent: the direction is always the negative gradient; the
        # step length is the Barzilai-Borwein one, s's / s'y from the last step
        # and gradient change, kept only when a nonmonotone Armijo test
        # against the largest of the last 10 objective values passes, halved
        # otherwise (Raydan, SIAM J. Optim. 7 (1997) 26-33).
        f = objective(x)
        g, = torch.autograd.grad(f, x)
        recent = [f.item()]
        step = 1.0 / max(1.0, float(g.abs().max()))
        for it in range(max_iter):
            gg = float(g.dot(g))
            ref = max(recent[-10:])
            while True:
                xn = (x - step * g).detach().requires_grad_(True)
                fn = objective(xn)
                if fn.item() <= ref - 1e-4 * step * gg or step < 1e-30:
                    break
                step *= 0.5
            gn, = torch.autograd.grad(fn, xn)
            sv = (xn - x).detach()
            yv = (gn - g).detach()
            sy = float(sv.dot(yv))
            step = float(sv.dot(sv)) / sy if sy > 0 else 1.0
            step = min(max(step, 1e-12), 1e12)
            x, f, g = xn, fn, gn
            recent.append(f.item())
            iters = it + 1
            if float(g.abs().max()) < tol:
                break
    # Newton polishing with the closed-form Hessian (also the whole 'newton' method).
    if method in ("lbfgs", "newton"):
        for _ in range(20 if method == "lbfgs" else max_iter):
            xd = x.detach().clone().requires_grad_(True)
            f = -objective(xd) * N
            g, = torch.autograd.grad(f, xd)
            H = fs_hessian(full(xd.detach()), c, lv).cpu()[idx][:, idx]
            try:
                dx = torch.linalg.solve(H, -g.detach().cpu())
            except RuntimeError:  # singular: a parameter on the boundary
                break
            xn = (xd.detach() + dx)
            if float(fs_loglik(full(xn), c, lv)) < float(f) - 1e-9 * abs(float(f)):
                break
            x = xn.requires_grad_(True)
            iters += 1
            if float(dx.abs().max()) < 1e-13:
                break
    theta = full(x.detach())
    return theta, float(fs_loglik(theta, c, lv)), iters
