"""neb_gram_split_noalloc (the split Gram without a workspace: partial tiles
in the library's file-scope device scratch) against neb_gram_split, through
the C ABI: bit for bit on two calls in a row, the second with a smaller
input, plus the accumulating form and the size limit."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # CPU-only host: the gpu tier runs on the MI355X box
    pytest.skip("needs a GPU", allow_module_level=True)

from neb_amd import _hip as H  # noqa: E402

DEV = "cuda"


def _planes(rows, K, seed):
    ld = (K + 7) // 8 * 8
    rng = np.random.default_rng(seed)
    x = torch.tensor(rng.normal(size=(rows, K)), dtype=torch.float32, device=DEV)
    hi = torch.zeros((rows, ld), dtype=torch.bfloat16, device=DEV)
    lo = torch.zeros_like(hi)
    s = torch.cuda.current_stream().cuda_stream
    H.call("neb_split_bf16", H.F32, x.data_ptr(), rows, K, K, hi.data_ptr(), lo.data_ptr(), ld, s)
    return hi, lo, ld


def _with_workspace(hi, lo, rows, K, ld, w, beta, G):
    need = H.load().neb_cross_workspace_size(K, K, rows, 1)
    ws = torch.empty(need, dtype=torch.uint8, device=DEV)
    H.call("neb_gram_split", hi.data_ptr(), lo.data_ptr(), rows, K, ld, w, beta, G.data_ptr(), K, ws.data_ptr(),
           need, torch.cuda.current_stream().cuda_stream)


def _no_alloc(hi, lo, rows, K, ld, w, beta, G):
    H.call("neb_gram_split_noalloc", hi.data_ptr(), lo.data_ptr(), rows, K, ld, w, beta, G.data_ptr(), K,
           torch.cuda.current_stream().cuda_stream)


@pytest.mark.parametrize("first,second", [((131072, 1024), (20000, 300)), ((65536, 512), (1000, 40))])
def test_noalloc_gram_matches_workspace_entry_point(first, second):
    for (rows, K), seed in ((first, 1), (second, 2)):
        hi, lo, ld = _planes(rows, K, seed)
        w = 1.0 / rows
        ref = torch.empty((K, K), dtype=torch.float64, device=DEV)
        out = torch.full((K, K), float("nan"), dtype=torch.float64, device=DEV)
        _with_workspace(hi, lo, rows, K, ld, w, 0.0, ref)
        _no_alloc(hi, lo, rows, K, ld, w, 0.0, out)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), (rows, K, float((out - ref).abs().max()))
        # accumulating form: G = G + w Phi^T Phi
        _with_workspace(hi, lo, rows, K, ld, w, 1.0, ref)
        _no_alloc(hi, lo, rows, K, ld, w, 1.0, out)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), (rows, K)


def test_noalloc_gram_rejects_k_beyond_scratch():
    G = torch.zeros((8, 8), dtype=torch.float64, device=DEV)
    with pytest.raises(ValueError, match="maximum"):
        H.call("neb_gram_split_noalloc", None, None, 0, 8192, 8192, 1.0, 0.0, G.data_ptr(), 8192,
               torch.cuda.current_stream().cuda_stream)
    # K = 3968 (31 tiles) fits at any row count: zero rows, beta = 0 writes zeros
    Gk = torch.full((3968, 3968), 1.0, dtype=torch.float64, device=DEV)
    H.call("neb_gram_split_noalloc", None, None, 0, 3968, 3968, 1.0, 0.0, Gk.data_ptr(), 3968,
           torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert float(Gk.abs().max()) == 0.0
