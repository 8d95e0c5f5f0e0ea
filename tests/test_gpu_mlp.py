"""Warm-start MLP: the device FP32 kernels (matrix-core GEMMs, fused
epilogues, MSE, Adam) against the FP64 restatement in oracle/oracle_mlp.c
(itself pinned by finite differences), on the network the warm start trains
(scft_amd.warmstart.DIMS: 133 -> 4 x 256 tanh -> 136), on the configuration-4
shape 131 -> 4 x 256 tanh -> 128 and on small ragged shapes."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

def _warm_dims():
    import importlib.util
    import os
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "scft_amd",
                        "warmstart.py")
    spec = importlib.util.spec_from_file_location("scft_warmstart_dims", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return list(mod.DIMS)


SHAPES = [_warm_dims(), [131, 256, 256, 256, 256, 128], [5, 70, 3], [64, 64, 64]]


def _S():
    import scft_amd
    return scft_amd


@pytest.mark.parametrize("dims", SHAPES)
@pytest.mark.parametrize("batch", [37, 256])
def test_forward_matches_oracle(dims, batch):
    S = _S()
    m = S.Mlp(dims, seed=3)
    p = m.params().astype(np.float64)
    X = np.random.default_rng(1).standard_normal((batch, dims[0])).astype(np.float32)
    Y = m.forward(X).astype(np.float64)
    ref = O.mlp_forward(dims, p, X.astype(np.float64))
    err = np.abs(Y - ref).max() / max(1.0, np.abs(ref).max())
    print(f"dims {dims} batch {batch}: forward max rel err {err:.2e}")
    assert err <= 2e-5
    m.close()


@pytest.mark.parametrize("dims", SHAPES)
def test_gradient_matches_oracle(dims):
    S = _S()
    m = S.Mlp(dims, seed=4)
    p = m.params().astype(np.float64)
    rng = np.random.default_rng(2)
    X = rng.standard_normal((45, dims[0])).astype(np.float32)
    T = rng.standard_normal((45, dims[-1])).astype(np.float32)
    loss, g = m.loss_grad(X, T)
    lref, gref = O.mlp_loss_grad(dims, p, X.astype(np.float64), T.astype(np.float64))
    assert abs(loss - lref) <= 1e-5 * abs(lref)
    off = 0
    for l in range(len(dims) - 1):  # per layer: weights then biases
        for n in (dims[l + 1] * dims[l], dims[l + 1]):
            a, b = g[off:off + n].astype(np.float64), gref[off:off + n]
            err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)
            print(f"dims {dims} layer {l} block {n}: grad max rel err {err:.2e}")
            assert err <= 1e-4
            off += n
    assert off == m.n_params
    m.close()


def test_adam_step_matches_oracle():
    S = _S()
    dims = [5, 70, 3]
    m = S.Mlp(dims, seed=5)
    p0 = m.params().astype(np.float64)
    rng = np.random.default_rng(6)
    X = rng.standard_normal((20, 5)).astype(np.float32)
    T = rng.standard_normal((20, 3)).astype(np.float32)
    pr, mr, vr = p0.copy(), np.zeros_like(p0), np.zeros_like(p0)
    for t in (1, 2, 3):
        pcur = m.params().astype(np.float64)
        _, g = m.loss_grad(X, T)
        O.adam_step(pr, g.astype(np.float64), mr, vr, t, lr=1e-2)
        m.adam_step(1e-2)
        # same gradient fed to both: compare the parameter trajectories
        assert np.abs(m.params() - pr).max() <= 1e-5 * max(1.0, np.abs(pr).max())
        pr = m.params().astype(np.float64)  # re-sync to keep the comparison per step
    m.close()


def test_training_fits_a_smooth_map():
    """50 epochs of mini-batch Adam drive the loss of a smooth synthetic
    regression down by more than an order of magnitude."""
    S = _S()
    rng = np.random.default_rng(7)
    A = rng.standard_normal((8, 16)) / np.sqrt(8)
    X = rng.standard_normal((2048, 8)).astype(np.float32)
    Y = np.tanh(X @ A).astype(np.float32)
    m = S.Mlp([8, 64, 64, 16], seed=8)
    hist = m.train(X, Y, epochs=50, batch=256, lr=1e-3, seed=9)
    print("epoch losses:", hist[[0, 9, 24, 49]])
    assert hist[-1] < 0.1 * hist[0]
    lfin, _ = m.loss_grad(X, Y, want_grad=False)
    assert lfin < 0.1 * hist[0]
    m.close()
