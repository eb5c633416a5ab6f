"""BF16 MFMA perceptron kernels (include/nrg_mlp.h) against the FP64
reference (oracle/mlp_fp64.py), through the C-ABI.

Bars: the fused inference kernel (hidden width 256: activations in registers,
each hidden layer's K rows in the order its predecessor's accumulators hold
them, so its FP32 sums run in another order than the training GEMMs') and the
training-path forward agree to 2e-3 norm-wise; both agree with the FP64
reference that rounds to bf16 where the kernels do to 5e-3 (norm-wise); the
gradient of every layer agrees with that reference to 2e-2 norm-wise (bf16
rounding of activations and back-propagated gradients, FP32 accumulation,
split-K atomics), with cosine similarity >= 0.999.
"""
import numpy as np
import pytest

from oracle import mlp_fp64 as M

pytestmark = pytest.mark.gpu

DIMS = (12, 256, 4, 10)


def _net(seed=0):
    from nrg_amd.surrogate import MLP, init_params
    net = MLP(DIMS[0], DIMS[3], DIMS[1], DIMS[2])
    p = init_params(DIMS[0], DIMS[3], DIMS[1], DIMS[2], seed=seed)
    # non-zero biases so the bias paths are exercised
    p = p + np.random.default_rng(seed + 1).normal(0, 0.05, p.size).astype(np.float32) * (p == 0)
    net.set_params(p)
    return net, p


def _report(name, **kw):
    print("[mlp] %s " % name + " ".join("%s=%s" % (k, v) for k, v in kw.items()), flush=True)


def test_params_round_trip():
    net, p = _net(3)
    assert np.array_equal(net.get_params(), p)
    assert net.num_params == 12 * 256 + 256 + 3 * (256 * 256 + 256) + 256 * 10 + 10
    net.close()


@pytest.mark.parametrize("n", [1, 77, 300, 4099])
def test_forward_matches_reference(n):
    import torch
    net, p = _net(5)
    rng = np.random.default_rng(n)
    x = rng.normal(size=(n, DIMS[0])).astype(np.float32)
    xt = torch.as_tensor(x, device="cuda:0")
    y_inf = net.forward(xt).cpu().numpy().astype(np.float64)
    t = torch.zeros((n, DIMS[3]), dtype=torch.float32, device="cuda:0")
    _, _, y_tr = net.grad(xt, t)
    y_tr = y_tr.cpu().numpy().astype(np.float64)
    y_ref, _ = M.forward(p.astype(np.float64), DIMS, x.astype(np.float64), bf16=True)
    y_fp64, _ = M.forward(p.astype(np.float64), DIMS, x.astype(np.float64), bf16=False)
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    _report("forward", n=n, inf_vs_train=rel(y_inf, y_tr), inf_vs_ref=rel(y_inf, y_ref),
            inf_vs_fp64=rel(y_inf, y_fp64))
    assert rel(y_inf, y_tr) <= 2e-3
    assert rel(y_inf, y_ref) <= 5e-3
    assert rel(y_inf, y_fp64) <= 5e-2
    net.close()


def test_gradients_match_reference():
    import torch
    net, p = _net(7)
    n = 1000
    rng = np.random.default_rng(11)
    x = rng.normal(size=(n, DIMS[0])).astype(np.float32)
    t = rng.normal(size=(n, DIMS[3])).astype(np.float32)
    w = rng.uniform(0.5, 2.0, DIMS[3]).astype(np.float32)
    g, loss, _ = net.grad(torch.as_tensor(x, device="cuda:0"), torch.as_tensor(t, device="cuda:0"),
                          torch.as_tensor(w, device="cuda:0"))
    g = g.cpu().numpy().astype(np.float64)
    l_ref, g_ref, _ = M.loss_and_grad(p.astype(np.float64), DIMS, x.astype(np.float64),
                                      t.astype(np.float64), w.astype(np.float64), bf16=True)
    assert float(loss.item()) == pytest.approx(l_ref, rel=5e-3)
    # per parameter block (W_l, b_l)
    fan = [DIMS[0]] + [DIMS[1]] * DIMS[2] + [DIMS[3]]
    q = 0
    for i in range(DIMS[2] + 1):
        for size, what in ((fan[i] * fan[i + 1], "W%d" % i), (fan[i + 1], "b%d" % i)):
            a, b = g[q:q + size], g_ref[q:q + size]
            q += size
            rel = np.linalg.norm(a - b) / np.linalg.norm(b)
            cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))
            _report("grad", block=what, rel=rel, cos=cos)
            assert rel <= 2e-2 and cos >= 0.999, what
    net.close()


def test_adam_training_reduces_loss():
    """Fit a smooth 12 -> 10 map; 400 Adam steps on the GPU kernels."""
    import torch
    net, _ = _net(13)
    rng = np.random.default_rng(17)
    A = rng.normal(size=(12, 10)) / 3.0
    x = rng.normal(size=(8192, 12)).astype(np.float32)
    y = np.tanh(x @ A).astype(np.float32)
    xt = torch.as_tensor(x, device="cuda:0")
    yt = torch.as_tensor(y, device="cuda:0")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda:0")
    first = None
    for k in range(400):
        idx = torch.as_tensor(rng.choice(8192, 1024, replace=False), device="cuda:0")
        loss.zero_()
        torch.cuda.synchronize()
        net.train_step(xt[idx].contiguous(), yt[idx].contiguous(), 1e-3, loss=loss)
        if first is None:
            first = float(loss.item())
    _, final, _ = net.grad(xt, yt)
    _report("adam", first=first, final=float(final.item()))
    assert float(final.item()) < 0.05 * first
    net.close()
