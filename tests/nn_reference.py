"""CPU reference of the guess network (torch, FP32) for the kernel tests.

Mirrors the device network layer by layer: BF16-rounded weights, FP32
bias and accumulation, activations rounded to BF16 where the device stores
them, nearest upsampling before the decoder convolutions, skips added after
the ReLU, output = last convolution + monopole term.  Forward rounding is a
straight-through operation in the backward pass; with grad_rounding the
backward pass rounds the gradient to BF16 wherever the device stores one
(the loss gradient, each adjoint convolution's output, each pooled or
upsampled gradient and the summed gradient of a tensor with two uses).
The two passes then differ only where an FP32 sum taken in another order
lands on the other side of a BF16 rounding boundary.

Weight convention: W[t][ci][co] with tap t = 9 (dx + 1) + 3 (dy + 1) + (dz + 1)
reading input cell (i + dx, j + dy, k + dz); zero padding.
"""
import numpy as np
import torch
import torch.nn.functional as F

LAYERS = [(2, 8, 0), (8, 8, 0), (8, 12, 0), (12, 12, 0), (12, 16, 0), (16, 16, 0),
          (16, 24, 0), (24, 24, 0), (24, 16, 1), (16, 12, 1), (12, 8, 1), (8, 1, 0)]


def bf16(x, enabled=True):
    if not enabled:
        return x
    return x + (x.to(torch.bfloat16).to(torch.float32) - x).detach()


class _GradBF16(torch.autograd.Function):
    """Identity forward; the incoming gradient is rounded to BF16."""

    @staticmethod
    def forward(ctx, x):
        return x.view_as(x)

    @staticmethod
    def backward(ctx, g):
        return g.to(torch.bfloat16).to(g.dtype)


def gbf16(x, enabled=True):
    return _GradBF16.apply(x) if enabled else x


def torch_weight(w):
    """W[27][cin][cout] -> conv3d weight [cout][cin][3][3][3]."""
    w = torch.as_tensor(w)
    cin, cout = w.shape[1], w.shape[2]
    return w.reshape(3, 3, 3, cin, cout).permute(4, 3, 0, 1, 2).contiguous()


def mono_field(n, mom4, box_len=1.0, origin=(0.0, 0.0, 0.0)):
    """Softened monopole potential in units of G M / L, FP32 as the device
    computes it (eps = 1.5 / n)."""
    c = [np.float32((mom4[1 + a] - origin[a]) / box_len) for a in range(3)]
    h = np.float32(1.0) / np.float32(n)
    x = (np.arange(n, dtype=np.float32) + np.float32(0.5)) * h
    dx = (x - c[0])[:, None, None]
    dy = (x - c[1])[None, :, None]
    dz = (x - c[2])[None, None, :]
    eps2 = (np.float32(1.5) * h) ** 2
    return (np.float32(-1.0) / np.sqrt(dx * dx + dy * dy + dz * dz + eps2)).astype(np.float32)


def forward(params, feat, mono, rounding=True, dtype=torch.float32, grad_rounding=False):
    """params: [(W[27][cin][cout], b[cout])] (tensors with requires_grad as
    wanted); feat: [n][n][n][2]; mono: [n][n][n].  Returns the output [n^3]."""
    r = lambda t: bf16(t, rounding)
    q = lambda t: gbf16(t, grad_rounding)
    x = torch.as_tensor(feat, dtype=dtype).permute(3, 0, 1, 2)[None]
    m = torch.as_tensor(mono, dtype=dtype)

    def conv(l, inp, relu=True):
        w, b = params[l]
        wt = torch_weight(r(w) if rounding else w).to(dtype)
        y = F.conv3d(inp, wt, bias=b.to(dtype), padding=1)
        return F.relu(y) if relu else y

    up = lambda t: q(F.interpolate(t, scale_factor=2, mode="nearest"))
    pool = lambda t: q(r(F.avg_pool3d(t, 2)))
    e1a = q(r(conv(0, x)))
    e1 = q(r(conv(1, e1a)))
    e2a = q(r(conv(2, pool(e1))))
    e2 = q(r(conv(3, e2a)))
    e3a = q(r(conv(4, pool(e2))))
    e3 = q(r(conv(5, e3a)))
    e4a = q(r(conv(6, pool(e3))))
    e4 = q(r(conv(7, e4a)))
    d3 = q(r(r(conv(8, up(e4))) + e3))
    d2 = q(r(r(conv(9, up(d3))) + e2))
    d1 = q(r(r(conv(10, up(d2))) + e1))
    out = q(conv(11, d1, relu=False)[0, 0] + m)
    return out


def loss_and_grads(layer_params, feat, mono, target, rounding=True, dtype=torch.float32,
                   grad_rounding=False):
    """Mean squared error and its gradients [(dW, db)] by autograd."""
    ps = [(torch.tensor(np.asarray(w), dtype=dtype, requires_grad=True),
           torch.tensor(np.asarray(b), dtype=dtype, requires_grad=True)) for w, b in layer_params]
    out = forward(ps, feat, mono, rounding, dtype, grad_rounding)
    t = torch.as_tensor(target, dtype=dtype)
    loss = torch.mean((out - t) ** 2)
    loss.backward()
    grads = [(w.grad.numpy(), b.grad.numpy()) for w, b in ps]
    return float(loss), out.detach().numpy(), grads
