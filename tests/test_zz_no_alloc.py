"""bm_mlp_grad_noalloc (the gradient entry point that allocates nothing, its
buffers in a fixed per-GPU scratch array) against bm_mlp_grad on the same
inputs: identical gradients (the same kernels in the same order) and the
mean loss within FP32 summation rounding, on two calls in a row, the second
with a smaller batch; sizes beyond the scratch are refused."""
import ctypes as C

import numpy as np
import pytest

from brian_amd import _lib
from brian_amd import learned as L

pytestmark = pytest.mark.gpu


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def test_grad_noalloc_matches_grad():
    import torch
    rng = np.random.default_rng(12)
    m = L.MLP(64, 256, 4, seed=13)
    for B in (1024, 256):
        print(f"case B={B}", flush=True)
        X = rng.standard_normal((B, 64)).astype(np.float32)
        U = rng.random((B, 4)).astype(np.float32)
        loss_ref, g_ref = m.grad(X, U)
        Xd = torch.from_numpy(X).cuda()
        Ud = torch.from_numpy(U).cuda()
        loss_d = torch.full((1,), float("nan"), device="cuda")
        g_d = torch.full((m.nparams,), float("nan"), device="cuda")
        torch.cuda.synchronize()
        _lib.call("bm_mlp_grad_noalloc", m.h, _ptr(Xd), _ptr(Ud), B, _ptr(loss_d), _ptr(g_d))
        g = g_d.cpu().numpy()
        loss = float(loss_d.cpu().numpy()[0])
        np.testing.assert_array_equal(g, g_ref)
        # FP32 device mean against the host's FP64 mean of the same per-row losses
        assert abs(loss - loss_ref) <= 1e-6 * (1 + abs(loss_ref)), (loss, loss_ref)
    Xd = torch.zeros((2048, 64), device="cuda")
    Ud = torch.zeros((2048, 4), device="cuda")
    with pytest.raises(_lib.BrianMIError):
        _lib.call("bm_mlp_grad_noalloc", m.h, _ptr(Xd), _ptr(Ud), 2048, _ptr(loss_d), _ptr(g_d))
    with pytest.raises(_lib.BrianMIError):
        _lib.call("bm_mlp_grad_noalloc", m.h, _ptr(Xd), _ptr(Ud), 300, _ptr(loss_d), _ptr(g_d))
    m.close()
