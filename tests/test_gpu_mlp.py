"""The planner network's HIP kernels (BF16 MFMA, FP32 accumulation) against
the CPU oracle: forward outputs and every gradient, plus a training run."""
import numpy as np
import pytest

import pyoracle as O
from gpsm_amd import _lib
from gpsm_amd.planner import MLP

pytestmark = pytest.mark.gpu

# Tolerances.  The device and the oracle round the same operands to BF16;
# they differ in FP32 accumulation order, which can flip the BF16 rounding of
# a hidden activation or delta by one unit in the last place (2^-8 relative).
FWD_RTOL, FWD_ATOL = 2e-2, 2e-2
GRAD_REL = 3e-2  # max |g_dev - g_ref| / max |g_ref|, per parameter block


def params(seed):
    rng = np.random.default_rng(seed)
    F, H = 32, 128
    p = [rng.normal(0, 1 / np.sqrt(F), H * F), rng.normal(0, 0.1, H), rng.normal(0, 1 / np.sqrt(H), H * H),
         rng.normal(0, 0.1, H), rng.normal(0, 1 / np.sqrt(H), H), rng.normal(0, 0.1, 1)]
    return np.concatenate(p).astype(np.float32)


BLOCKS = {"W1": (0, 4096), "b1": (4096, 4224), "W2": (4224, 20608), "b2": (20608, 20736), "W3": (20736, 20864),
          "b3": (20864, 20865)}


def test_param_roundtrip(ctx):
    m = MLP(ctx, seed=1)
    p = params(0)
    m.set(p)
    assert np.array_equal(m.get(), p)
    assert _lib.lib().gpsm_mlp_num_params() == _lib.N_PARAMS == len(p)
    m.free()


@pytest.mark.parametrize("B", [1, 31, 32, 33, 100, 1000])
def test_forward_matches_oracle(ctx, B):
    m = MLP(ctx)
    p = params(B)
    m.set(p)
    x = np.random.default_rng(B).normal(size=(B, 32)).astype(np.float32)
    y = m.forward(x)
    ref = O.mlp_forward(p, x, bf16=True)
    assert np.allclose(y, ref, rtol=FWD_RTOL, atol=FWD_ATOL), np.abs(y - ref).max()
    m.free()


@pytest.mark.parametrize("B", [32, 77, 256, 5000])
def test_gradients_match_oracle(ctx, B):
    m = MLP(ctx)
    p = params(100 + B)
    m.set(p)
    rng = np.random.default_rng(B)
    x = rng.normal(size=(B, 32)).astype(np.float32)
    t = rng.normal(size=B).astype(np.float32)
    loss, g = m.grad(x, t)
    rloss, rg = O.mlp_grad(p, x, t, bf16=True)
    assert abs(loss - rloss) <= 1e-2 * max(1.0, abs(rloss))
    for name, (a, b) in BLOCKS.items():
        ref = rg[a:b]
        err = np.abs(g[a:b] - ref).max() / max(np.abs(ref).max(), 1e-6)
        assert err < GRAD_REL, (name, err)
    m.free()


def test_training_fits_a_smooth_target(ctx):
    """Adam on the device drives the loss of a learnable target down."""
    m = MLP(ctx, seed=3)
    rng = np.random.default_rng(5)
    x = rng.normal(size=(4096, 32)).astype(np.float32)
    w = rng.normal(size=32).astype(np.float32)
    t = (np.tanh(x @ w / 4) * 3).astype(np.float32)
    first = None
    for epoch in range(30):
        idx = rng.permutation(len(x))
        for i in range(0, len(x), 256):
            b = idx[i:i + 256]
            loss = m.train_step(x[b], t[b], 1e-3)
            first = loss if first is None else first
    y = m.forward(x)
    final = 0.5 * np.mean((y - t) ** 2)
    assert final < 0.05 * first, (first, final)
    m.free()


def test_input_standardisation_is_applied_before_rounding(ctx):
    m = MLP(ctx)
    p = params(7)
    m.set(p)
    rng = np.random.default_rng(7)
    mu = rng.normal(size=32).astype(np.float32) * 10
    isd = (1.0 / (0.01 + rng.random(32))).astype(np.float32)
    m.set_norm(mu, isd)
    got_mu, got_isd = m.get_norm()
    assert np.array_equal(got_mu, mu) and np.array_equal(got_isd, isd)
    x = (mu + rng.normal(size=(64, 32)).astype(np.float32) / isd).astype(np.float32)
    y = m.forward(x)
    ref = O.mlp_forward(p, ((x - mu) * isd).astype(np.float32), bf16=True)
    assert np.allclose(y, ref, rtol=FWD_RTOL, atol=FWD_ATOL)
    m.free()
