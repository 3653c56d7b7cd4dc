"""fh_em_fit_noalloc (the allocation-free EM entry point) against fh_em_fit,
both called through the C ABI on the same inputs: identical parameters,
pattern probabilities, iteration counts and convergence flags, on two calls
in a row (the second with a smaller histogram, so stale scratch from the
first call would show), on the call that reads nothing back, and the size
limit rejected with an error."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _histogram(levels, seed, n_total=10 ** 9):
    """A two-class mixture histogram over itertools.product order."""
    pu_all, pm_all = np.ones(1), np.ones(1)
    rng = np.random.default_rng(seed)
    for L in levels:
        pm = np.sort(rng.dirichlet(np.arange(1, L + 1) * 3.0))
        pu = -np.sort(-rng.dirichlet(np.arange(1, L + 1)[::-1] * 3.0))
        pu_all = np.kron(pu_all, pu)
        pm_all = np.kron(pm_all, pm)
    mix = 0.01 * pm_all + 0.99 * pu_all
    return np.round(mix * n_total).astype(np.int64)


def _init(levels, seed):
    rng = np.random.default_rng(seed + 100)
    pi0 = [-np.sort(-rng.dirichlet(np.arange(1, L * 50 + 1, 50))) for L in levels]
    pi1 = [np.sort(rng.dirichlet(np.arange(1, L * 50 + 1, 50))) for L in levels]
    return np.concatenate([[0.1], *pi0, *pi1])


def _run(fn_name, levels, counts, params0, tol, max_iter, read_back=True):
    from faster import _lib

    L = _lib.load()
    dev = torch.device("cuda", 0)
    c = torch.as_tensor(counts, dtype=torch.int64, device=dev)
    par = torch.as_tensor(params0, dtype=torch.float64, device=dev).clone()
    ksi = torch.empty(len(counts), dtype=torch.float64, device=dev)
    lv = (ctypes.c_int32 * len(levels))(*levels)
    it, conv = ctypes.c_int(-1), ctypes.c_int(-1)
    fn = getattr(L, fn_name)
    st = fn(len(levels), lv, _lib.ptr(c), _lib.ptr(par), ctypes.c_double(tol), int(max_iter), _lib.ptr(ksi),
            ctypes.byref(it) if read_back else None, ctypes.byref(conv) if read_back else None, _lib.stream())
    _lib.check(st, fn_name)
    torch.cuda.synchronize()
    return par.cpu().numpy(), ksi.cpu().numpy(), it.value, conv.value


def test_em_noalloc_matches_em_fit():
    cases = [([4, 4, 4, 3, 4], 1, 1e-10), ([3, 3, 2], 2, 1e-7)]  # 768 patterns, then 18
    for levels, seed, tol in cases:
        counts = _histogram(levels, seed)
        p0 = _init(levels, seed)
        ref = _run("fh_em_fit", levels, counts, p0, tol, 3000)
        got = _run("fh_em_fit_noalloc", levels, counts, p0, tol, 3000)
        print(f"levels {levels}: {ref[2]} iterations, converged {ref[3]}", flush=True)
        assert got[2] == ref[2] and got[3] == ref[3]
        np.testing.assert_array_equal(got[0], ref[0])
        np.testing.assert_array_equal(got[1], ref[1])


def test_em_noalloc_without_readback():
    levels = [3, 3, 3, 2]
    counts = _histogram(levels, 5)
    p0 = _init(levels, 5)
    ref = _run("fh_em_fit", levels, counts, p0, 1e-8, 500)
    got = _run("fh_em_fit_noalloc", levels, counts, p0, 1e-8, 500, read_back=False)
    np.testing.assert_array_equal(got[0], ref[0])
    np.testing.assert_array_equal(got[1], ref[1])


def test_em_noalloc_rejects_beyond_limit():
    from faster import _lib

    L = _lib.load()
    levels = [4] * 9  # 262,144 patterns > 65,536
    lv = (ctypes.c_int32 * len(levels))(*levels)
    st = L.fh_em_fit_noalloc(len(levels), lv, None, None, ctypes.c_double(1e-6), 10, None, None, None,
                             _lib.stream())
    assert st == 2  # FH_ERR_ARG, before any device work
    assert b"scratch limit" in L.fh_last_error()
