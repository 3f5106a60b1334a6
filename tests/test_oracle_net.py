"""Pins of the oracle's correction network: central finite differences of the
FP64 loss, torch conv3d on the CPU, and a closed-form identity network."""
import numpy as np
import pytest

P = 4 * 6928


def layer_w(params, l):
    return params[l * 6928: l * 6928 + 6912].reshape(27, 16, 16)


def random_params(rng, scale=0.2):
    p = np.zeros(P, np.float32)
    cin = [2, 16, 16, 16]
    cout = [16, 16, 16, 1]
    for l in range(4):
        w = np.zeros((27, 16, 16), np.float32)
        w[:, : cin[l], : cout[l]] = rng.normal(scale=scale / np.sqrt(cin[l] * 27 / 8), size=(27, cin[l], cout[l]))
        p[l * 6928: l * 6928 + 6912] = w.ravel()
        p[l * 6928 + 6912: l * 6928 + 6912 + cout[l]] = rng.normal(scale=0.05, size=cout[l])
    return p


def bf16_exact(a):
    """float32 values exactly representable in bf16 (truncate the mantissa)."""
    a = np.ascontiguousarray(a, np.float32)
    return (a.view(np.uint32) & 0xFFFF0000).view(np.float32)


def test_param_count(orc):
    assert orc.orc_net_param_count() == P


def test_identity_network_reproduces_the_centre(orc):
    """Centre taps routing channel 0 through every layer: Y = X0 channel 0
    cropped to the central 16^3 (valid, valid, same, same)."""
    p = np.zeros(P, np.float32)
    for l in range(4):
        layer_w(p, l)[13, 0, 0] = 1.0
    rng = np.random.default_rng(0)
    x0 = bf16_exact(np.abs(rng.normal(size=(1, 20, 20, 20, 2))))
    for mode in (0, 1):
        y = np.zeros(4096, np.float32)
        orc.orc_net_forward(p, x0.ravel(), 1, mode, y)
        assert np.array_equal(y.reshape(16, 16, 16), x0[0, 2:18, 2:18, 2:18, 0])


def torch_forward(params, x0):
    torch = pytest.importorskip("torch")
    F = torch.nn.functional
    x = torch.from_numpy(np.ascontiguousarray(x0, np.float64)).permute(0, 4, 1, 2, 3)  # N C Z Y X
    cin = [2, 16, 16, 16]
    cout = [16, 16, 16, 1]
    pads = [0, 0, 1, 1]
    for l in range(4):
        w = layer_w(params, l).astype(np.float64)[:, : cin[l], : cout[l]]  # [t][c][o], t = dx + 3dy + 9dz
        wt = torch.from_numpy(w.reshape(3, 3, 3, cin[l], cout[l]).transpose(4, 3, 0, 1, 2).copy())  # o c dz dy dx
        bias = torch.from_numpy(params[l * 6928 + 6912: l * 6928 + 6912 + cout[l]].astype(np.float64))
        x = F.conv3d(x, wt, bias, padding=pads[l])
        if l < 3:
            x = F.relu(x)
    return x[:, 0].numpy()


def test_forward_matches_torch_conv3d(orc):
    rng = np.random.default_rng(1)
    p = random_params(rng)
    x0 = rng.normal(size=(2, 20, 20, 20, 2)).astype(np.float32)
    y = np.zeros(2 * 4096, np.float32)
    orc.orc_net_forward(p, x0.ravel(), 2, 0, y)
    want = torch_forward(p, x0)
    assert np.abs(y.reshape(2, 16, 16, 16) - want).max() <= 1e-6 * np.abs(want).max()


def test_gradient_matches_central_differences(orc):
    """FP64 mode: analytic gradient vs (L(p + h) - L(p - h)) / 2h for a sample
    of weights and biases of every layer."""
    rng = np.random.default_rng(2)
    p = random_params(rng).astype(np.float64)
    x0 = rng.normal(size=(1, 20, 20, 20, 2)).astype(np.float32)
    tgt = rng.normal(size=(1, 4096)).astype(np.float32)
    grad = np.zeros(P)
    pf = p.astype(np.float32)
    orc.orc_net_loss_grad(pf, x0.ravel(), tgt.ravel(), 1, 0, grad)
    cin = [2, 16, 16, 16]
    cout = [16, 16, 16, 1]
    picks = []
    for l in range(4):
        for _ in range(4):
            t, c, o = rng.integers(27), rng.integers(cin[l]), rng.integers(cout[l])
            picks.append(l * 6928 + (t * 16 + c) * 16 + o)
        picks.append(l * 6928 + 6912 + rng.integers(cout[l]))
    for k in picks:
        h = 2e-5 * max(abs(pf[k]), 1e-2)  # small: ReLU kinks crossed by the step grow with h
        hi, lo = pf.copy(), pf.copy()
        hi[k] += np.float32(h)
        lo[k] -= np.float32(h)
        g = np.zeros(P)
        lp = orc.orc_net_loss_grad(hi, x0.ravel(), tgt.ravel(), 1, 0, g)
        lm = orc.orc_net_loss_grad(lo, x0.ravel(), tgt.ravel(), 1, 0, g)
        fd = (lp - lm) / (float(hi[k]) - float(lo[k]))
        assert abs(fd - grad[k]) <= 1e-3 * abs(fd) + 1e-4 * np.abs(grad).max(), (k, fd, grad[k])
    # padding entries (layer 1 channels >= 2, layer 4 outputs >= 1) get no gradient
    assert np.all(layer_w(grad, 0)[:, 2:, :] == 0) and np.all(layer_w(grad, 3)[:, :, 1:] == 0)
