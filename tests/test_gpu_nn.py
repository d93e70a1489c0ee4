"""The warm-start network's kernels against CPU references.

* the bf16 MFMA GEMM (every operand form, split-K with atomic accumulation) against numpy
  float64 on the same bf16-rounded inputs;
* the whole network's loss and parameter gradients against an independent torch float32
  CPU autograd model built from the same parameters (tolerances cover the bf16 rounding of
  activations and weights between layers);
* training lowers the loss on a tiny fixed problem.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bf16_round(a):
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


@pytest.mark.parametrize("a_form,b_form", [(0, 0), (0, 1), (1, 0), (1, 1)])
# (24, 200, 500): a 64-row tile with 40 empty rows, and K slices of two stages each
@pytest.mark.parametrize("shape,k_split", [((200, 70, 100), 0), ((64, 64, 64), 0),
                                           ((37, 129, 300), 64), ((3, 32, 1000), 96),
                                           ((24, 200, 500), 128)])
def test_gemm_matches_numpy(a_form, b_form, shape, k_split):
    import boltzmann_amd
    M, N, K = shape
    rng = np.random.default_rng(M * 7 + N)
    A = bf16_round(rng.standard_normal((M, K)))
    B = bf16_round(rng.standard_normal((N, K)))
    a = A if a_form == 0 else np.ascontiguousarray(A.T)
    b = B if b_form == 0 else np.ascontiguousarray(B.T)
    C = boltzmann_amd.gemm_check("gpu", a_form, b_form, a, b, k_split)
    ref = A.astype(np.float64) @ B.astype(np.float64).T
    np.testing.assert_allclose(C, ref, rtol=1e-5, atol=1e-4 * np.sqrt(K))


def _bf16(t):
    import torch
    return t.to(torch.bfloat16).to(torch.float32)


def layer_spec(w, depth):
    """(cin, cout, taps) in parameter order: encoder pairs by level, bottleneck pair,
    decoder from the deepest level up, 1x1 output (channels w * min(2^l, 4))."""
    ch = [w * min(2 ** l, 4) for l in range(depth + 1)]
    dch = [ch[max(l - 1, 0)] for l in range(depth)] + [ch[depth]]
    spec = []
    for l in range(depth + 1):
        spec += [(16 if l == 0 else ch[l - 1], ch[l], 9), (ch[l], ch[l], 9)]
    for l in range(depth - 1, -1, -1):
        spec.append((dch[l + 1] + ch[l], dch[l], 9))
    spec.append((dch[0], 3, 1))
    return spec


def torch_forward(params, x, w, depth=3, emulate=True):
    """The network of bolt_nn.hip in torch (NCHW, float32).  With emulate, values are rounded
    to bf16 where the GPU kernels store bf16: weights and the input, every layer's stored
    activation, and (in the backward pass) every pre-activation gradient."""
    import torch
    import torch.nn.functional as F

    class GradRound(torch.autograd.Function):
        @staticmethod
        def forward(ctx, t):
            return t

        @staticmethod
        def backward(ctx, g):
            return _bf16(g)

    R = _bf16 if emulate else (lambda t: t)
    G = GradRound.apply if emulate else (lambda t: t)
    spec = layer_spec(w, depth)
    Ws, bs, off = [], [], 0
    for cin, cout, taps in spec:
        Wf = params[off:off + cout * taps * cin]
        off += cout * taps * cin
        b = params[off:off + cout]
        off += cout
        k = 3 if taps == 9 else 1
        # ours: W[co][tap*cin + ci], tap = kx*3 + ky  ->  torch [co][ci][kh=kx][kw=ky]
        Wt = Wf.reshape(cout, taps, cin).transpose(0, 2, 1).reshape(cout, cin, k, k)
        Ws.append(torch.tensor(np.ascontiguousarray(Wt), requires_grad=True))
        bs.append(torch.tensor(np.ascontiguousarray(b), requires_grad=True))
    assert off == params.size
    xt = np.zeros((x.shape[0], 16) + x.shape[1:3], np.float32)
    xt[:, :8] = x.transpose(0, 3, 1, 2)
    X = R(torch.tensor(xt))
    last = len(spec) - 1

    def c(i, t):
        y = G(F.conv2d(t, R(Ws[i]), bs[i], padding=1 if i < last else 0))
        return R(F.relu(y)) if i < last else y

    up = lambda t: F.interpolate(t, scale_factor=2, mode="nearest")
    pool = lambda t: R(F.avg_pool2d(t, 2))
    skips, t = [], X
    for l in range(depth):
        t = c(2 * l + 1, c(2 * l, t))
        skips.append(t)
        t = pool(t)
    t = c(2 * depth + 1, c(2 * depth, t))
    for k, l in enumerate(range(depth - 1, -1, -1)):
        t = c(2 * depth + 2 + k, torch.cat([up(t), skips[l]], 1))
    out = c(last, t)
    return out, Ws, bs, spec


# (depth, nx, ny, min cosine, max relative error) per layer.  The deepest levels of the
# depth-5 net hold 4 x 2 pixels, too few to average out where bf16 rounding of a near tie
# goes the other way than in the reference; a wrong gradient scores cos < 0.9.
@pytest.mark.parametrize("depth,nx,ny,cos_min,rel_max", [(3, 32, 16, 0.999, 0.05),
                                                         (5, 128, 64, 0.995, 0.1)])
def test_network_loss_and_gradients_match_torch(depth, nx, ny, cos_min, rel_max):
    import torch
    import boltzmann_amd
    B, w = 2, 16
    net = boltzmann_amd.Network("gpu", nx, ny, B, width=w, seed=3, depth=depth)
    params = net.get_params()
    assert params.size == sum(co * t * ci + co for ci, co, t in layer_spec(w, depth))
    rng = np.random.default_rng(0)
    x = rng.standard_normal((B, nx, ny, 8)).astype(np.float32)
    y = rng.standard_normal((B, nx, ny, 3)).astype(np.float32)
    loss, grads = net.loss_grad(x, y)

    out, Ws, bs, spec = torch_forward(params, x, w, depth)
    tl = torch.mean((out - torch.tensor(y.transpose(0, 3, 1, 2))) ** 2)
    tl.backward()
    assert abs(loss - tl.item()) / tl.item() < 2e-3, (loss, tl.item())
    off = 0
    report = []
    for i, (cin, cout, taps) in enumerate(spec):
        k = 3 if taps == 9 else 1
        gW = grads[off:off + cout * taps * cin].reshape(cout, taps, cin).transpose(0, 2, 1)
        gW = gW.reshape(cout, cin, k, k)
        off += cout * taps * cin
        gb = grads[off:off + cout]
        off += cout
        for ours, ref in ((gW, Ws[i].grad.numpy()), (gb, bs[i].grad.numpy())):
            ref = ref.ravel().astype(np.float64)
            ours = ours.ravel().astype(np.float64)
            cos = ours @ ref / (np.linalg.norm(ours) * np.linalg.norm(ref) + 1e-30)
            rel = np.linalg.norm(ours - ref) / (np.linalg.norm(ref) + 1e-30)
            report.append((i, round(cos, 5), round(rel, 4)))
    print("per-layer (cos, rel) of W then b:", report, flush=True)
    for i, cos, rel in report:
        assert cos > cos_min and rel < rel_max, report


def test_training_reduces_loss():
    import boltzmann_amd
    nx, ny, B = 32, 16, 4
    net = boltzmann_amd.Network("gpu", nx, ny, B, width=16, seed=1)
    rng = np.random.default_rng(1)
    n = 16
    x = rng.standard_normal((n, nx, ny, 8)).astype(np.float32)
    # a smooth target that depends on the input: a learnable map
    y = np.stack([np.tanh(x[..., 0]), 0.5 * x[..., 1], x[..., 0] * x[..., 2] * 0.1], -1).astype(np.float32)
    net.set_data(x, y)
    losses = net.train(30, lr=2e-3, seed=0)
    assert np.isfinite(losses).all()
    assert losses[-1] < 0.5 * losses[0], losses
    pred = net.predict(x[:5])
    assert pred.shape == (5, nx, ny, 3) and np.isfinite(pred).all()


def test_prediction_does_not_depend_on_batching():
    """Samples never mix (the 3x3 gathers stay inside each image), so predicting n samples in
    batches, the last one partial, must give each sample what predicting it alone gives: this
    checks the feature upload offsets and the zero padding of the unused batch slots.  A
    sample's pixels can sit at other rows of a GEMM tile in the two runs, so the bound allows
    a bf16 rounding flip in the activations; a wrong offset changes the outputs by O(1)."""
    import boltzmann_amd
    nx, ny, B = 32, 16, 4
    net = boltzmann_amd.Network("gpu", nx, ny, B, width=16, seed=3)
    x = np.random.default_rng(3).standard_normal((7, nx, ny, 8)).astype(np.float32)
    together = net.predict(x)
    assert together.shape == (7, nx, ny, 3) and np.isfinite(together).all()
    scale = np.abs(together).max()
    worst = 0.0
    for i in range(7):
        alone = net.predict(x[i:i + 1])[0]
        worst = max(worst, float(np.abs(alone - together[i]).max()))
    print("largest difference alone vs batched:", worst, "of", scale, flush=True)
    # the bound must be tight enough to tell samples apart
    apart = min(float(np.abs(together[i] - together[0]).max()) for i in range(1, 7))
    print("smallest difference between samples:", apart, flush=True)
    assert worst <= 2e-2 * scale < apart
