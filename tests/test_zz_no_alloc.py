"""The allocation-free loss/gradient entry point (pelemp_mlp_loss_grad_noalloc)
gives what pelemp_mlp_loss_grad gives, bit for bit, on two calls in a row
(the second with a smaller batch, so it runs on a workspace that still holds
the first call's contents), and refuses a batch beyond its workspace."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lib():
    from pelemp_amd import _capi as C

    if C.device_count() < 1:
        pytest.skip("no HIP device")
    return C


def _case(C, n, seed):
    from pelemp_amd import surrogate as S

    rng = np.random.default_rng(seed)
    rr = S.input_ranges(S.DEFAULT_HULL)
    x = np.stack([rng.uniform(lo, hi, n) for lo, hi in rr], axis=1).astype(np.float32)
    t = rng.normal(0, 1, (n, 3)).astype(np.float32)
    w = (rng.uniform(0, 1, n) < 0.9).astype(np.float32)
    norm = S.make_norm(S.DEFAULT_HULL, rng.normal(0, 1, 3), rng.uniform(0.5, 2.0, 3))
    p = (S.init_params(seed) + rng.normal(0, 0.05, C.SURR_NPARAM)).astype(np.float32)
    bufs = []
    for a in (p, x, t, w):
        b = C.DeviceBuffer(a.nbytes)
        b.upload(np.ascontiguousarray(a))
        bufs.append(b)
    return p, norm, bufs


def test_noalloc_loss_grad_matches_allocating_call():
    C = _lib()
    from pelemp_amd import surrogate as S

    cases = [(C.MLP_NOALLOC_MAX_BATCH, 11, True), (700, 12, False)]
    for n, seed, quant in cases:
        print(f"noalloc loss_grad: n={n} seed={seed} quant={quant}", flush=True)
        p, norm, (pd, xd, td, wd) = _case(C, n, seed)
        loss_ref, g_ref = S.loss_grad(p, norm, xd, td, wd, n, quant=quant)
        gd = C.DeviceBuffer(4 * C.SURR_NPARAM)
        loss = S.loss_grad_noalloc(pd, norm, xd, td, wd, n, gd, quant=quant)
        g = gd.download(np.empty(C.SURR_NPARAM, np.float32))
        assert np.isfinite(loss_ref) and loss_ref > 0.
        assert loss == loss_ref, (n, loss, loss_ref)
        assert np.array_equal(g, g_ref), (n, np.max(np.abs(g - g_ref)))


def test_noalloc_rejects_a_batch_beyond_the_workspace():
    C = _lib()
    p, norm, (pd, xd, td, wd) = _case(C, 8, 13)
    gd = C.DeviceBuffer(4 * C.SURR_NPARAM)
    loss = ctypes.c_double()
    # the size check comes before any device work: the buffers are never read
    st = C.lib().pelemp_mlp_loss_grad_noalloc(pd.ptr, ctypes.byref(norm), 0, C.MLP_NOALLOC_MAX_BATCH + 1, xd.ptr,
                                              td.ptr, wd.ptr, gd.ptr, ctypes.byref(loss), None)
    assert st == C.E_ARG
    assert b"workspace" in C.lib().pelemp_last_error()
