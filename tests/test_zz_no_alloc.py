"""The allocation-free projection (psdp_project_f32_noalloc) against
psdp_project_f32: same kernels, tile sizes and order of work, so the results
are bitwise equal.  Two calls in a row, the second on a smaller matrix (the
scratch array keeps the first call's contents), then the size limit."""
import numpy as np
import pytest

from oracle import psdp_oracle as O

torch = pytest.importorskip("torch")
import psd_projection_amd as P  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


@pytest.mark.parametrize("method,sizes", [("mixed", (1600, 300)), ("mixed_mx", (1664, 129)), ("composite_FP32", (1900, 77)),
                                          ("composite_FP32_emulated", (1400, 512)),
                                          ("composite_FP16", (1000, 999))])
def test_noalloc_equals_project(handle, method, sizes):
    print(f"noalloc {method} sizes {sizes}", flush=True)
    Xs = []
    for i, n in enumerate(sizes):
        X, Q, lam = O.spectral_input(n, 70 + i, "normal")
        Xs.append((torch.from_numpy(X.astype(np.float32)).to(DEV), Q, lam))
    got = []
    for X, _, _ in Xs:  # two allocation-free calls in a row, the second smaller
        handle.set_seed(21)
        got.append(P.project_noalloc(handle, X, method=method))
    torch.cuda.synchronize()
    for (X, Q, lam), g in zip(Xs, got):
        handle.set_seed(21)
        ref = P.project(handle, X, method=method)
        assert torch.equal(g, ref)
        err = O.rel_fro_error(g.cpu().numpy(), O.closed_form_projection(Q, lam))
        print(f"  n={X.shape[0]} rel err {err:.3e}", flush=True)
        assert err <= (1e-3 if method in ("mixed", "mixed_mx", "composite_FP16") else 2e-5) * \
            (5 if method == "composite_FP16" else 1)


def test_noalloc_rejects_sizes_over_its_scratch(handle):
    X = torch.zeros(4096, 4096, device=DEV)
    with pytest.raises(P.PsdpError) as e:
        P.project_noalloc(handle, X, method="mixed")
    assert e.value.status == P.ERR_INVALID_ARG
    # the handle keeps working afterwards, with and without the scratch array
    Y = torch.eye(64, device=DEV) * 0.5
    assert torch.allclose(P.project_noalloc(handle, Y, method="mixed"), Y, atol=1e-3)
    assert torch.allclose(P.project(handle, Y, method="mixed"), Y, atol=1e-3)
