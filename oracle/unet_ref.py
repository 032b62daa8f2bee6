"""CPU reference of the learned-warm-start U-Net (torch, FP32 on the CPU).

TEST INFRASTRUCTURE ONLY: the checker the GPU network is compared against.

The same network as the device code: 3x3x3 convolutions with zero
padding, ReLU, 2x2x2 average pooling, nearest upsampling, skip connections
by channel concatenation ([upsampled, skip]), a 1x1x1 output convolution.
With bf16=True the values the device keeps in BF16 (input, weights, every
stored activation) are rounded to BF16 at the same points, so the forward
pass differs from the device only by FP32 summation order.

Weight layout of the device: W[tap][cout_pad][cin_pad] with tap = kx + 3 ky
+ 9 kz (k in 0..2, offset k - 1 along x, y, z); volumes are channels-last
v = x + nx (y + ny z).
"""
import numpy as np
import torch
import torch.nn.functional as F

# cin, cout, taps, level, relu (device layer table)
LAYERS = [(2, 16, 27, 0, 1), (16, 16, 27, 0, 1), (16, 32, 27, 1, 1),
          (32, 32, 27, 1, 1), (32, 64, 27, 2, 1), (64, 64, 27, 2, 1),
          (96, 32, 27, 1, 1), (32, 32, 27, 1, 1), (48, 16, 27, 0, 1),
          (16, 16, 27, 0, 1), (16, 1, 1, 0, 0)]


def pad16(c):
    return (c + 15) // 16 * 16


def bf(t):
    return t.to(torch.bfloat16).to(torch.float32)


def torch_weight(w_dev, layer):
    """Device weights (flat, [taps][coutp][cinp]) -> torch (co, ci, kz, ky, kx)."""
    cin, cout, taps = layer[0], layer[1], layer[2]
    cinp, coutp = pad16(cin), pad16(cout)
    W = np.asarray(w_dev, dtype=np.float32).reshape(taps, coutp, cinp)
    W = W[:, :cout, :cin]
    k = 3 if taps == 27 else 1
    W = W.reshape(k, k, k, cout, cin)           # [kz][ky][kx]
    return torch.from_numpy(np.ascontiguousarray(W.transpose(3, 4, 0, 1, 2)))


def to_ncdhw(x_cl, dims, C):
    nx, ny, nz = dims
    a = np.asarray(x_cl, dtype=np.float32).reshape(nz, ny, nx, C)
    return torch.from_numpy(np.ascontiguousarray(a.transpose(3, 0, 1, 2)))[None]


def forward(weights, biases, x, bf16=True, requires_grad=False):
    """weights/biases: lists of torch tensors (device layout converted);
    x: (1, 2, nz, ny, nx).  Returns the output (1, 1, nz, ny, nx)."""
    r = bf if bf16 else (lambda t: t)
    Ws = [r(w) for w in weights]

    def conv(i, t):
        pad = 1 if LAYERS[i][2] == 27 else 0
        y = F.conv3d(t, Ws[i], biases[i][:LAYERS[i][1]], padding=pad)
        return F.relu(y) if LAYERS[i][4] else y

    t = r(x)
    e0a = r(conv(0, t))
    e0 = r(conv(1, e0a))
    p0 = r(F.avg_pool3d(e0, 2))
    e1a = r(conv(2, p0))
    e1 = r(conv(3, e1a))
    p1 = r(F.avg_pool3d(e1, 2))
    b0 = r(conv(4, p1))
    b1 = r(conv(5, b0))
    u1 = F.interpolate(b1, scale_factor=2, mode="nearest")
    d1a = r(conv(6, torch.cat([u1, e1], 1)))
    d1 = r(conv(7, d1a))
    u0 = F.interpolate(d1, scale_factor=2, mode="nearest")
    d0a = r(conv(8, torch.cat([u0, e0], 1)))
    d0 = r(conv(9, d0a))
    return conv(10, d0)


def grads(w_dev, b_dev, x_cl, t_cl, dims):
    """Loss 0.5 mean (y - t)^2 and its gradients with respect to every
    layer's weights (device layout, zero in the padding) and biases."""
    Wt = [torch_weight(w, L).requires_grad_(True) for w, L in zip(w_dev, LAYERS)]
    Bt = [torch.from_numpy(np.asarray(b, dtype=np.float32)).requires_grad_(True)
          for b in b_dev]
    x = to_ncdhw(x_cl, dims, 2)
    t = to_ncdhw(t_cl, dims, 1)
    y = forward(Wt, Bt, x, bf16=True)
    loss = 0.5 * ((y - t) ** 2).mean()
    loss.backward()
    gw, gb = [], []
    for W, B, L in zip(Wt, Bt, LAYERS):
        cin, cout, taps = L[0], L[1], L[2]
        k = 3 if taps == 27 else 1
        g = W.grad.numpy().transpose(2, 3, 4, 0, 1).reshape(taps, cout, cin)
        full = np.zeros((taps, pad16(cout), pad16(cin)), dtype=np.float32)
        full[:, :cout, :cin] = g
        gw.append(full.ravel())
        gb.append(B.grad.numpy())
    return float(loss), y.detach()[0, 0].numpy().transpose(2, 1, 0), gw, gb
