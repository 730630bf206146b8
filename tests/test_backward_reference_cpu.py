"""CPU checks of the autograd reference used to pin the backward kernels:
its forward equals the oracle's optimal thresholding at lambda = 1 (the oracle
is itself pinned to the reference's golden vectors), and its gradient passes
finite differences."""

import numpy as np
import pytest
import torch

from tests.torch_shrink_ref import optimal_shrinkage_ref


def _volume(shape=(6, 6, 5, 14), seed=0):
    rng = np.random.default_rng(seed)
    X, Y, Z, T = shape
    sp = rng.standard_normal((X * Y * Z, 2))
    tc = rng.standard_normal((2, T))
    sig = 50.0 + (sp @ tc).reshape(X, Y, Z, T) * 5.0
    return sig + rng.standard_normal(shape)


@pytest.mark.parametrize("loss", ["fro", "nuc", "ope"])
def test_reference_forward_matches_oracle(loss):
    from oracle import cpu_ref as O

    data = _volume()
    T = data.shape[-1]
    mpm = O.mp_median(T / 27)
    out = optimal_shrinkage_ref(torch.from_numpy(data), torch.ones(4 * 4 * 3, dtype=torch.float64),
                                (3, 3, 3), (2, 2, 2), mpm, loss)
    ps, po = (3, 3, 3, T), (2, 2, 2, T)
    ref, _, _, _ = O.denoise(data, ps, po, lambda X, i: O.optimal_patch(X, O.SHRINKERS[loss], mpm))
    # the oracle (like the reference) accumulates the weights in float32
    np.testing.assert_allclose(out.numpy(), ref, rtol=1e-6, atol=1e-9)


def test_reference_gradient_finite_differences():
    from oracle import cpu_ref as O

    data = _volume((5, 5, 4, 12), seed=1)
    mpm = O.mp_median(12 / 27)
    x = torch.from_numpy(data).requires_grad_(True)
    lam = torch.full((3 * 3 * 2,), 1.1, dtype=torch.float64, requires_grad=True)
    gout = torch.from_numpy(np.random.default_rng(2).standard_normal(data.shape))

    def f(x, lam):
        return (optimal_shrinkage_ref(x, lam, (3, 3, 3), (2, 2, 2), mpm, "fro") * gout).sum()

    assert torch.autograd.gradcheck(f, (x, lam), eps=1e-6, atol=1e-5, rtol=1e-4, fast_mode=True)
