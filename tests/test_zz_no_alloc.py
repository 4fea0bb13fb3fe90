"""The allocation-free weight-gradient entry point
(cia_mlp_linear_backward_weight_noalloc) against the workspace entry point
(cia_mlp_linear_backward_weight), through the C ABI.

Two calls in a row, the second with a smaller input, then a call that adds into
the previous result: each equals the workspace entry point bit for bit (same
split rule, same kernels) and the float64 product of the bf16-exact inputs to
FP32 accumulation accuracy (2e-6 relative).  A shape beyond the static scratch
is rejected before any launch.
"""

import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _bf(a):
    t = torch.tensor(np.asarray(a, dtype=np.float32)).to(torch.bfloat16)
    return t.to(torch.float64).numpy()


def _p(t):
    return C.c_void_p(None) if t is None else C.c_void_p(t.data_ptr())


def _dev(a, dtype):
    return torch.tensor(np.asarray(a), dtype=dtype, device="cuda").contiguous()


def _case(M, N, K, seed):
    rng = np.random.default_rng(seed)
    dz = _bf(rng.standard_normal((M, N)))
    x = _bf(rng.standard_normal((M, K)))
    return dz, x


def _both(lib, st, dz, x, with_bias, accumulate, start=None):
    M, N = dz.shape
    K = x.shape[1]
    dZ, X = _dev(dz, torch.bfloat16), _dev(x, torch.bfloat16)
    out = []
    for variant in ("workspace", "noalloc"):
        dW = torch.zeros((N, K), dtype=torch.float32, device="cuda")
        db = torch.zeros(N, dtype=torch.float32, device="cuda") if with_bias else None
        if start is not None:
            dW.copy_(torch.as_tensor(start[0], device="cuda"))
            if with_bias:
                db.copy_(torch.as_tensor(start[1], device="cuda"))
        if variant == "workspace":
            n_ws = lib.cia_mlp_backward_weight_workspace(N, K)
            ws = torch.empty(n_ws, dtype=torch.float32, device="cuda")
            rc = lib.cia_mlp_linear_backward_weight(st, M, N, K, _p(dZ), _p(X), _p(dW), _p(db), _p(ws),
                                                    n_ws, accumulate)
        else:
            rc = lib.cia_mlp_linear_backward_weight_noalloc(st, M, N, K, _p(dZ), _p(X), _p(dW), _p(db),
                                                            accumulate)
        assert rc == 0, variant
        torch.cuda.synchronize()
        out.append((dW.cpu().numpy(), None if db is None else db.cpu().numpy()))
    return out


def test_noalloc_weight_gradient_equals_workspace_entry_point(gpu):
    from concrete_impact_ai_amd import _lib

    lib = _lib.lib()
    st = torch.cuda.current_stream().cuda_stream
    # first call: a 256 x 256 layer over 4096 rows (the largest layer of the surrogate)
    for (M, N, K, bias), seed in (((4096, 256, 256, True), 1), ((300, 12, 256, False), 2),
                                   ((77, 256, 13, True), 3)):
        dz, x = _case(M, N, K, seed)
        (w_ws, b_ws), (w_na, b_na) = _both(lib, st, dz, x, bias, 0)
        assert np.array_equal(w_ws, w_na), (M, N, K)
        ref = dz.T @ x
        assert np.max(np.abs(w_na - ref)) <= 2e-6 * np.max(np.abs(ref)) * max(1.0, np.sqrt(M / 64))
        if bias:
            assert np.array_equal(b_ws, b_na)
            rb = dz.sum(axis=0)
            assert np.max(np.abs(b_na - rb)) <= 2e-6 * M * np.max(np.abs(dz))
    # accumulate = 1 adds into an existing gradient
    dz, x = _case(200, 64, 96, 4)
    start = (np.full((64, 96), 0.5, np.float32), np.full(64, -0.25, np.float32))
    (w_ws, b_ws), (w_na, b_na) = _both(lib, st, dz, x, True, 1, start)
    assert np.array_equal(w_ws, w_na) and np.array_equal(b_ws, b_na)
    assert np.max(np.abs(w_na - (0.5 + dz.T @ x))) <= 1e-5 * np.max(np.abs(dz.T @ x))


def test_noalloc_rejects_shapes_beyond_the_scratch(gpu):
    from concrete_impact_ai_amd import _lib

    lib = _lib.lib()
    st = torch.cuda.current_stream().cuda_stream
    dZ = torch.zeros((64, 2048), dtype=torch.bfloat16, device="cuda")
    X = torch.zeros((64, 2048), dtype=torch.bfloat16, device="cuda")
    dW = torch.zeros((2048, 2048), dtype=torch.float32, device="cuda")
    rc = lib.cia_mlp_linear_backward_weight_noalloc(st, 64, 2048, 2048, _p(dZ), _p(X), _p(dW),
                                                    _p(None), 0)
    assert rc == _lib.CIA_ERR_INVALID_ARGUMENT
    assert "exceeds the static scratch" in _lib.last_error()
    assert float(dW.abs().max()) == 0.0  # nothing ran
