"""gpk_mll_forward_noalloc: the exact-GP MLL and gradient with the library's fixed
device scratch in place of a caller workspace.  Its results must equal those of
gpk_mll_forward + gpk_mll_backward on the same inputs (same kernels, same
deterministic reductions), on two calls in a row with the second one smaller
(the scratch keeps the first call's contents), in both precisions; sizes above
its limit are refused."""

import ctypes

import numpy as np
import pytest
import torch

from gpjax_amd import _lib

pytestmark = pytest.mark.gpu

NOALLOC_MAX_N = 1024


def _inputs(cuda, n, d, seed):
    rng = np.random.default_rng(seed)
    X = torch.tensor(rng.uniform(-2, 2, (n, d)), device=cuda)
    y = torch.tensor(np.sin(rng.uniform(-2, 2, n)) + 0.1 * rng.standard_normal(n), device=cuda)
    ls = np.exp(rng.uniform(np.log(0.5), np.log(2.0), d))
    theta = torch.tensor(np.concatenate([ls, [1.2, 0.2, 0.3]]), device=cuda)
    return X, y, theta


def _with_workspace(cuda, kind, prec, X, y, theta, n, d):
    nbytes = _lib.lib().gpk_mll_workspace_bytes(n, d, prec)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=cuda)
    res = torch.empty(_lib.RESULT_LEN, dtype=torch.float64, device=cuda)
    grad = torch.empty(d + 3, dtype=torch.float64, device=cuda)
    st = _lib.stream_ptr(cuda)
    _lib.call("gpk_mll_forward", kind, prec, _lib.ptr(X), _lib.ptr(y), n, d, _lib.ptr(theta), d,
              1e-6, -1, _lib.ptr(ws), nbytes, _lib.ptr(res), st)
    _lib.call("gpk_mll_backward", kind, prec, _lib.ptr(X), n, d, _lib.ptr(theta), d, 1.0,
              _lib.ptr(ws), nbytes, _lib.ptr(grad), st)
    torch.cuda.synchronize(cuda)
    return res.cpu().numpy(), grad.cpu().numpy()


def _noalloc(cuda, kind, prec, X, y, theta, n, d):
    res = torch.empty(_lib.RESULT_LEN, dtype=torch.float64, device=cuda)
    grad = torch.empty(d + 3, dtype=torch.float64, device=cuda)
    _lib.call("gpk_mll_forward_noalloc", kind, prec, _lib.ptr(X), _lib.ptr(y), n, d,
              _lib.ptr(theta), d, 1e-6, -1, _lib.ptr(res), _lib.ptr(grad), _lib.stream_ptr(cuda))
    torch.cuda.synchronize(cuda)
    return res.cpu().numpy(), grad.cpu().numpy()


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_equals_workspace_entry_point(cuda, prec):
    for n, d, kind, seed in [(1000, 3, 0, 1), (300, 3, 3, 2)]:   # second call smaller
        X, y, theta = _inputs(cuda, n, d, seed)
        print(f"noalloc prec={prec} n={n} d={d} kind={kind}", flush=True)
        r_ws, g_ws = _with_workspace(cuda, kind, prec, X, y, theta, n, d)
        r_na, g_na = _noalloc(cuda, kind, prec, X, y, theta, n, d)
        assert r_ws[_lib.RES_INFO] == 0 and r_na[_lib.RES_INFO] == 0
        for slot in (_lib.RES_MLL, _lib.RES_LOGDET, _lib.RES_QUAD):
            assert abs(r_na[slot] - r_ws[slot]) <= 1e-13 * abs(r_ws[slot]), (slot, r_na, r_ws)
        assert np.linalg.norm(g_na - g_ws) <= 1e-13 * np.linalg.norm(g_ws), (g_na, g_ws)


def test_noalloc_forward_only_and_limits(cuda):
    n, d = 200, 2
    X, y, theta = _inputs(cuda, n, d, 5)
    res = torch.empty(_lib.RESULT_LEN, dtype=torch.float64, device=cuda)
    _lib.call("gpk_mll_forward_noalloc", 1, 0, _lib.ptr(X), _lib.ptr(y), n, d, _lib.ptr(theta),
              d, 1e-6, -1, _lib.ptr(res), None, _lib.stream_ptr(cuda))
    r_ws, _ = _with_workspace(cuda, 1, 0, X, y, theta, n, d)
    assert abs(res[0].item() - r_ws[0]) <= 1e-13 * abs(r_ws[0])
    big = NOALLOC_MAX_N + 1
    rc = _lib.lib().gpk_mll_forward_noalloc(0, 0, ctypes.c_void_p(_lib.ptr(X)),
                                            ctypes.c_void_p(_lib.ptr(y)), big, d,
                                            ctypes.c_void_p(_lib.ptr(theta)), d, 1e-6, -1,
                                            ctypes.c_void_p(_lib.ptr(res)), None, None)
    assert rc == 1 and "1024" in _lib.lib().gpk_last_error().decode()
