"""The allocation-free network evaluation, gm_mlp_predict_noalloc_dev,
against gm_mlp_predict_dev (the same kernels with library-allocated
activation buffers), through the C ABI.

Each output row depends only on its own input row and the weights, and both
entry points run the same GEMM and epilogue kernels, so the outputs must be
equal bit for bit whatever the chunking.  Two calls in a row, the second with
fewer points, so the second reads a scratch that still holds the first call's
activations.  A network wider than the scratch is refused.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def capi():
    from goma_amd import capi as M

    M.lib()
    assert torch.cuda.is_available()
    return M


def _predict(L, fn, h, x):
    y = torch.full((x.shape[0], 5), float("nan"), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    rc = getattr(L, fn)(h, x.shape[0], x.data_ptr(), y.data_ptr())
    return rc, y


def test_noalloc_predict_matches_predict(capi):
    L = capi.lib()
    net = capi.Mlp(4, 5, hidden=4, width=256, seed=7, in_shift=[0.5] * 4, in_scale=[2.0] * 4)
    rng = np.random.default_rng(3)
    for npts in (5000, 777):  # 3 chunks (the last partial), then 1 partial chunk
        print(f"[case] noalloc predict npts {npts}", flush=True)
        x = torch.from_numpy(rng.random((npts, 4), dtype=np.float32)).cuda()
        rc, ref = _predict(L, "gm_mlp_predict_dev", net.h, x)
        assert rc == 0
        rc, y = _predict(L, "gm_mlp_predict_noalloc_dev", net.h, x)
        assert rc == 0
        assert torch.isfinite(y).all()
        assert torch.equal(y, ref)
    # no points: nothing to do
    assert L.gm_mlp_predict_noalloc_dev(net.h, 0, None, None) == 0
    net.close()


def test_noalloc_predict_refuses_a_wider_network(capi):
    L = capi.lib()
    net = capi.Mlp(3, 4, hidden=2, width=576, seed=1)
    x = torch.zeros((128, 3), dtype=torch.float32, device="cuda")
    y = torch.zeros((128, 4), dtype=torch.float32, device="cuda")
    assert L.gm_mlp_predict_noalloc_dev(net.h, 128, x.data_ptr(), y.data_ptr()) == capi.GM_ERR_ARG
    assert b"exceeds" in L.gm_last_error()
    net.close()
