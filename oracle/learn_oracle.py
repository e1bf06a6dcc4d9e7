"""CPU oracle of the learned level-2 correction network (test
infrastructure only: used by tests/, never by the library).

The network of gpu-to_amd/csrc/learn.hip restated with torch CPU
(float64) conv3d and autograd, which are library routines:

    e1 = relu(conv(x, W1, stride 1))      4 -> 16
    e2 = relu(conv(e1, W2, stride 2))    16 -> 32
    e3 = relu(conv(e2, W3, stride 2))    32 -> 64
    a2 = relu(conv(e3, W4, stride 1))    64 -> 32
    u2 = up(a2) + e2
    a1 = relu(conv(u2, W5, stride 1))    32 -> 16
    u1 = up(a1) + e1
    y  = conv(u1, W6, stride 1)          16 -> 3

3x3x3 kernels with padding 1; up() repeats every entry twice along each
axis and crops to the finer grid.  The flat parameter layout is per layer
W[cout][cin][3][3][3] followed by b[cout].  There is no reference
counterpart: the network is new in this build, so it is pinned by this
independent restatement, and the restatement by finite differences.
"""
import numpy as np

LAYERS = [(4, 16, 1, True), (16, 32, 2, True), (32, 64, 2, True), (64, 32, 1, True), (32, 16, 1, True),
          (16, 3, 1, False)]


def nparams():
    return sum(co * ci * 27 + co for ci, co, _, _ in LAYERS)


def unpack(flat, torch):
    out, off = [], 0
    for ci, co, _, _ in LAYERS:
        w = torch.tensor(flat[off:off + co * ci * 27].reshape(co, ci, 3, 3, 3), dtype=torch.float64)
        off += co * ci * 27
        b = torch.tensor(flat[off:off + co], dtype=torch.float64)
        off += co
        out.append((w, b))
    return out


def _up(d, shape):
    for ax in (2, 3, 4):
        d = d.repeat_interleave(2, dim=ax)
    return d[:, :, :shape[2], :shape[3], :shape[4]]


def forward(params, x, torch):
    """params: list of (w, b) tensors; x: tensor [1, 4, n0, n1, n2]."""
    F = torch.nn.functional
    (w1, b1), (w2, b2), (w3, b3), (w4, b4), (w5, b5), (w6, b6) = params
    e1 = F.relu(F.conv3d(x, w1, b1, stride=1, padding=1))
    e2 = F.relu(F.conv3d(e1, w2, b2, stride=2, padding=1))
    e3 = F.relu(F.conv3d(e2, w3, b3, stride=2, padding=1))
    a2 = F.relu(F.conv3d(e3, w4, b4, stride=1, padding=1))
    u2 = _up(a2, e2.shape) + e2
    a1 = F.relu(F.conv3d(u2, w5, b5, stride=1, padding=1))
    u1 = _up(a1, e1.shape) + e1
    return F.conv3d(u1, w6, b6, stride=1, padding=1)


def run(flat, x, gy=None):
    """y = N(x) and, with gy, the flat gradient of <gy, y> w.r.t. the params."""
    import torch
    params = unpack(np.asarray(flat, dtype=np.float64), torch)
    for w, b in params:
        w.requires_grad_(True)
        b.requires_grad_(True)
    xt = torch.tensor(np.asarray(x, dtype=np.float64)[None])
    y = forward(params, xt, torch)
    if gy is None:
        return y.detach().numpy()[0], None
    (y * torch.tensor(np.asarray(gy, dtype=np.float64)[None])).sum().backward()
    g = np.concatenate([np.concatenate([w.grad.numpy().ravel(), b.grad.numpy().ravel()]) for w, b in params])
    return y.detach().numpy()[0], g
