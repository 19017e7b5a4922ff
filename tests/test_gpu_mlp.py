"""The surrogate network's device kernels (MFMA products, forward,
backward, Adam) against the CPU checker oracle/mlp_oracle.py."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu

DIMS = (3, 256, 256, 256, 65)


@pytest.fixture(scope="module")
def V():
    import inovesa_amd as V
    return V


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("M,N,K", [(16, 16, 16), (37, 53, 29), (256, 256, 256),
                                   (256, 65, 256), (65, 256, 256), (256, 3, 256),
                                   (1, 1, 1)])
def test_gemm_against_cpu(V, M, N, K):
    rng = np.random.default_rng(M * 1000 + N * 10 + K)
    A = rng.standard_normal((M, K)).astype(np.float32)
    B = rng.standard_normal((K, N)).astype(np.float32)
    C = V.gemm_bf16(A, B)
    ref = mo.bf16(A) @ mo.bf16(B)     # exact products of the BF16 operands
    scale = np.abs(mo.bf16(A)) @ np.abs(mo.bf16(B))
    assert np.all(np.abs(C - ref) <= 1e-6 * scale + 1e-30)


def test_gemm_transposed_operands(V):
    """The strides the backward products use: A^T and B^T views."""
    import ctypes as C_
    from inovesa_amd._lib import call
    rng = np.random.default_rng(7)
    M, N, K = 40, 24, 72
    At = np.ascontiguousarray(rng.standard_normal((K, M)).astype(np.float32))
    Bt = np.ascontiguousarray(rng.standard_normal((N, K)).astype(np.float32))
    C = np.empty((M, N), np.float32)
    p = lambda a: a.ctypes.data_as(C_.c_void_p)
    call("vfps_gemm_bf16", V.vfps.Device.default().h, M, N, K, p(At), At.size,
         1, M, p(Bt), Bt.size, 1, K, p(C))
    ref = mo.bf16(At).T @ mo.bf16(Bt).T
    scale = np.abs(mo.bf16(At)).T @ np.abs(mo.bf16(Bt)).T
    assert np.all(np.abs(C - ref) <= 1e-6 * scale + 1e-30)


def _data(B, seed=1):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (B, 3)).astype(np.float32)
    t = rng.standard_normal((B, 64)).astype(np.float32)
    lab = (rng.random(B) > 0.5).astype(np.float32)
    return x, t, lab


def test_forward_matches_oracle(V):
    net = V.MLP(DIMS, max_batch=256, seed=3)
    x, _, _ = _data(200)
    y = net.forward(x)
    ref, _ = mo.forward(mo.split(net.params(), DIMS), x, emulate=True)
    assert _rel(y, ref) <= 2e-3
    assert np.abs(y - ref).max() <= 2e-2 * np.abs(ref).max()


def test_gradients_match_oracle(V):
    net = V.MLP(DIMS, max_batch=256, seed=4)
    x, t, lab = _data(64, seed=2)
    loss = net.loss_grad(x, t, lab, 0.5)
    g = net.grads()
    rloss, rg = mo.loss_and_grads(mo.split(net.params(), DIMS), x, t, lab,
                                  0.5, emulate=True)
    assert loss == pytest.approx(rloss, rel=1e-3)
    gl = net.layers(g)
    for (dW, db), (rW, rb) in zip(gl, rg):
        assert _rel(dW, rW) <= 2e-2
        assert _rel(db, rb) <= 2e-2
        cos = float((dW.ravel() @ rW.ravel())
                    / (np.linalg.norm(dW) * np.linalg.norm(rW)))
        assert cos >= 0.9998
    # and against the unrounded FP64 gradients: BF16 operands only
    _, fg = mo.loss_and_grads(mo.split(net.params(), DIMS), x, t, lab, 0.5)
    assert _rel(g, mo.flatten(fg)) <= 5e-2


def test_adam_steps_match_closed_form(V):
    net = V.MLP(DIMS, max_batch=64, seed=5)
    x, t, lab = _data(32, seed=3)
    p0 = net.params().astype(np.float64)
    net.loss_grad(x, t, lab)
    g1 = net.grads().astype(np.float64)
    net.adam_step(lr=1e-3)
    p1 = net.params().astype(np.float64)
    ref1, m, v = mo.adam(p0, g1, np.zeros_like(p0), np.zeros_like(p0), 1, 1e-3)
    assert np.abs(p1 - ref1).max() <= 1e-6 * max(1.0, np.abs(p0).max())
    net.loss_grad(x, t, lab)
    g2 = net.grads().astype(np.float64)
    net.adam_step(lr=1e-3)
    p2 = net.params().astype(np.float64)
    ref2, _, _ = mo.adam(p1, g2, m, v, 2, 1e-3)
    assert np.abs(p2 - ref2).max() <= 1e-6 * max(1.0, np.abs(p0).max())


def test_training_reduces_loss(V):
    """A smooth synthetic map (3 inputs -> 64 outputs and a threshold
    label) is learned: the loss falls by more than 10x."""
    rng = np.random.default_rng(9)
    N = 2048
    x = rng.uniform(-1, 1, (N, 3)).astype(np.float32)
    k = np.linspace(0, 1, 64)
    t = (np.sin(3 * x[:, :1] + 2 * k) * (1 + x[:, 1:2])
         + 0.5 * x[:, 2:3] * k).astype(np.float32)
    lab = (x[:, 0] + 0.5 * x[:, 1] > 0).astype(np.float32)
    net = V.MLP(DIMS, max_batch=256, seed=11)
    losses = net.train(x, t, lab, batch=256, epochs=40, lr=2e-3)
    print("\nepoch losses:", " ".join(f"{v:.4f}" for v in losses[::5]))
    assert losses[-1] < losses[0] / 10
    y = net.forward(x[:256])
    acc = float(((y[:, 64] > 0) == (lab[:256] > 0.5)).mean())
    assert acc > 0.9
