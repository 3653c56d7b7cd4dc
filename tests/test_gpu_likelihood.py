"""Differentiable FS likelihood over the 768-pattern histogram of the
1,000,000 x 1,000,000 synthetic linkage with three string variables, an exact
house number and a numeric birth day, each with a missing level (levels
[4, 4, 4, 3, 4]): hand-written backward vs central finite differences and an
FP64 torch autograd reference, closed-form Hessian vs autograd,
observed-information standard errors vs finite differences, and L-BFGS,
Newton and plain gradient-descent fits reaching the EM optimum within 1e-8
relative log-likelihood."""
import itertools

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

LEVELS = [4, 4, 4, 3, 4]


def _torch_ll(theta, counts, levels):
    """FP64 autograd reference of the same likelihood."""
    t = theta
    n_lv = sum(levels)
    G = torch.tensor(list(itertools.product(*(range(L) for L in levels))))
    la, l1a = torch.nn.functional.logsigmoid(t[0]), torch.nn.functional.logsigmoid(-t[0])
    s0 = torch.zeros(len(G), dtype=torch.float64) + l1a
    s1 = torch.zeros(len(G), dtype=torch.float64) + la
    off = 0
    for k, L in enumerate(levels):
        lp0 = torch.log_softmax(t[1 + off:1 + off + L], 0)
        lp1 = torch.log_softmax(t[1 + n_lv + off:1 + n_lv + off + L], 0)
        s0 = s0 + lp0[G[:, k]]
        s1 = s1 + lp1[G[:, k]]
        off += L
    lm = torch.logsumexp(torch.stack([s0, s1]), 0)
    return (torch.as_tensor(counts, dtype=torch.float64) * lm).sum()


@pytest.fixture(scope="module")
def em_optimum(config5_fit):
    """The histogram and EM's optimum on it (tolerance 1e-12)."""
    import faster

    _, _, comp = config5_fit
    counts = comp.Counts
    assert comp.Var_Levels == LEVELS
    np.random.seed(0)
    est = faster.Estimation(3, 1, counts, Levels=LEVELS)
    est.fit(Tolerance=1e-12, Max_Iter=100000)
    return counts, est


@pytest.fixture(scope="module")
def setup(em_optimum):
    from faster.likelihood import params_to_logits

    counts, est = em_optimum
    rng = np.random.default_rng(1)
    theta = params_to_logits(est.Lambda, est.Pi, LEVELS) + torch.as_tensor(rng.normal(0, 0.05, 1 + 2 * sum(LEVELS)))
    return counts, theta


def test_forward_and_gradient_vs_autograd(setup):
    from faster.likelihood import fs_loglik

    counts, theta = setup
    t1 = theta.clone().requires_grad_(True)
    ll = fs_loglik(t1, counts, LEVELS)
    ll.backward()
    t2 = theta.clone().requires_grad_(True)
    ref = _torch_ll(t2, counts, LEVELS)
    ref.backward()
    assert float(ll) == pytest.approx(float(ref), rel=1e-12)
    np.testing.assert_allclose(t1.grad.numpy(), t2.grad.numpy(), rtol=1e-9, atol=1e-9 * float(np.sum(counts)))


def test_gradient_vs_central_differences(setup):
    from faster.likelihood import fs_loglik

    counts, theta = setup
    N = float(np.sum(counts))
    t = theta.clone().requires_grad_(True)
    (fs_loglik(t, counts, LEVELS) / N).backward()
    g = t.grad.numpy()
    h = 1e-6
    fd = np.zeros_like(g)
    for i in range(len(g)):
        e = torch.zeros_like(theta)
        e[i] = h
        fd[i] = (float(fs_loglik(theta + e, counts, LEVELS)) - float(fs_loglik(theta - e, counts, LEVELS))) / (2 * h * N)
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-9)


def test_hessian_vs_autograd(setup):
    from faster.likelihood import fs_hessian

    counts, theta = setup
    H = fs_hessian(theta, counts, LEVELS).cpu().numpy()
    Href = torch.autograd.functional.hessian(lambda t: _torch_ll(t, counts, LEVELS), theta.clone()).numpy()
    scale = np.abs(Href).max()
    np.testing.assert_allclose(H, Href, rtol=1e-8, atol=1e-10 * scale)
    assert np.allclose(H, H.T)


def test_standard_errors_vs_finite_differences(setup):
    """Observed-information SEs from the closed-form Hessian match SEs from a
    Hessian built by central differences of the hand-written gradient, to 1e-6
    relative. The differences use h = 1e-4 and h = 5e-5 with Richardson
    extrapolation: with gradient terms of order N = 1e12, h = 1e-6 would
    leave a rounding error of eps * N / h ~ 1e2 against Hessian entries of
    order lambda * N ~ 1e8 for the match-class logits."""
    from faster.likelihood import fit, free_index, fs_loglik, standard_errors

    counts, theta = setup
    th, _, _ = fit(counts, LEVELS, theta, method="lbfgs")
    se = standard_errors(th, counts, LEVELS)
    idx = free_index(LEVELS)
    n = len(idx)

    def fd_hessian(h):
        Hfd = np.zeros((n, n))
        for a, i in enumerate(idx):
            grads = []
            for sgn in (1, -1):
                t = th.clone()
                t[i] += sgn * h
                t = t.requires_grad_(True)
                fs_loglik(t, counts, LEVELS).backward()
                grads.append(t.grad.numpy()[idx])
            Hfd[a] = (grads[0] - grads[1]) / (2 * h)
        return 0.5 * (Hfd + Hfd.T)

    Hfd = (4.0 * fd_hessian(5e-5) - fd_hessian(1e-4)) / 3.0
    se_fd = np.sqrt(np.diag(np.linalg.inv(-Hfd)))
    np.testing.assert_allclose(se["free"], se_fd, rtol=1e-6)
    assert np.all(np.isfinite(se["natural"]))


@pytest.mark.parametrize("method", ["lbfgs", "newton", "gd"])
def test_fit_reaches_em_optimum(em_optimum, method):
    from faster.likelihood import fit, fs_loglik, params_to_logits

    counts, est = em_optimum
    th_em = params_to_logits(est.Lambda, est.Pi, LEVELS)
    ll_em = float(fs_loglik(th_em, counts, LEVELS))
    rng = np.random.default_rng(2)
    start = th_em + torch.as_tensor(rng.normal(0, 0.05 if method != "newton" else 0.01, th_em.numel()))
    if method == "gd":
        th, ll, it = fit(counts, LEVELS, start, method="gd", max_iter=50000, tol=1e-13)
    else:
        th, ll, it = fit(counts, LEVELS, start, method=method, max_iter=500)
    print(f"{method}: {it} iterations, LL {ll!r} vs EM {ll_em!r}", flush=True)
    assert ll >= ll_em - 1e-8 * abs(ll_em), (ll, ll_em, it)
