"""Plain PyTorch (CPU, FP32 or FP64) statement of the pressure network, the
reference against which the library's BF16 matrix-core kernels are checked.

It follows the parameter layout of include/stroemung_hip.h (stm_net_get_params):
per convolution W[cout][9][cin] with tap 3*(dh+1) + (dw+1) weighing the input
pixel at offset (dh, dw), then b[cout]; then the 1x1 head.  The library keeps
convolution weights in BF16 for the matrix cores and biases and the head in
FP32; `bf16_weights` rounds the same way so that the remaining differences are
the BF16 rounding of activations and gradients.
"""
import numpy as np
import torch
import torch.nn.functional as F

CONVS = [(32, 32), (32, 32), (32, 64), (64, 64), (64, 128), (128, 128), (128, 256),
         (256, 256), (384, 128), (128, 128), (192, 64), (64, 64), (96, 32), (32, 32)]


def layout():
    out, off = [], 0
    for k, (cin, cout) in enumerate(CONVS):
        out.append(("conv%d.w" % k, off, (cout, 9, cin)))
        off += cout * 9 * cin
        out.append(("conv%d.b" % k, off, (cout,)))
        off += cout
        off = (off + 7) // 8 * 8
    out.append(("head.w", off, (32,)))
    off += 32
    out.append(("head.b", off, (1,)))
    off += 8
    return out, off


def to_bf16(a):
    return torch.as_tensor(a).to(torch.bfloat16).to(torch.float32)


def unpack(params, dtype=torch.float32, bf16_weights=True, requires_grad=False):
    lay, n = layout()
    assert params.size == n, (params.size, n)
    P = {}
    for name, off, shape in lay:
        size = int(np.prod(shape))
        t = torch.tensor(params[off:off + size], dtype=torch.float64).reshape(shape)
        if bf16_weights and name.endswith(".w") and name.startswith("conv"):
            t = to_bf16(t)
        t = t.to(dtype)
        if requires_grad:
            t.requires_grad_(True)
        P[name] = t
    return P


def conv3x3(t, w, b):
    """t: (nb, cin, H, W), w: (cout, 9, cin) in the library's tap order,
    zero padding 1 (cross-correlation, as F.conv2d)."""
    cout, _, cin = w.shape
    return F.conv2d(t, w.reshape(cout, 3, 3, cin).permute(0, 3, 1, 2), b, padding=1)


class _Bf16Store(torch.autograd.Function):
    """A tensor stored in BF16: the value rounds going forward and the
    gradient arriving at it rounds going back."""

    @staticmethod
    def forward(ctx, t):
        return to_bf16(t).to(t.dtype)

    @staticmethod
    def backward(ctx, g):
        return to_bf16(g).to(g.dtype)


def forward(P, x, bf16_acts=False):
    """x: (nb, H, W, 2) tensor -> y: (nb, H, W).  bf16_acts: store every
    activation (and the gradient arriving at it) in BF16 where the library's
    kernels do -- conv outputs and the inputs of the convolutions."""
    S = _Bf16Store.apply if bf16_acts else (lambda t: t)
    a = x.permute(0, 3, 1, 2)
    a = torch.cat([a, torch.zeros(a.shape[0], 30, a.shape[2], a.shape[3], dtype=a.dtype)], 1)

    def conv(k, t):
        return S(F.relu(conv3x3(t, P["conv%d.w" % k], P["conv%d.b" % k])))

    def up(t):
        return t.repeat_interleave(2, dim=2).repeat_interleave(2, dim=3)

    e0 = conv(1, conv(0, a))
    e1 = conv(3, conv(2, S(F.max_pool2d(e0, 2))))
    e2 = conv(5, conv(4, S(F.max_pool2d(e1, 2))))
    e3 = conv(7, conv(6, S(F.max_pool2d(e2, 2))))
    d2 = conv(9, conv(8, S(torch.cat([up(e3), e2], 1))))
    d1 = conv(11, conv(10, S(torch.cat([up(d2), e1], 1))))
    d0 = conv(13, conv(12, S(torch.cat([up(d1), e0], 1))))
    return torch.einsum("nchw,c->nhw", d0, P["head.w"]) + P["head.b"][0]


def loss(y, x, t):
    """mean over samples of sum m (y - t)^2 / sum m t^2, m = x[..., 1] > 0."""
    m = (x[..., 1] > 0).to(y.dtype)
    num = (m * (y - t) ** 2).sum(dim=(1, 2))
    den = (m * t * t).sum(dim=(1, 2))
    return (num / den).mean()


def gradients(params, x, t, dtype=torch.float32, bf16_acts=False):
    """Loss and the flat gradient in the library's layout."""
    P = unpack(params, dtype=dtype, requires_grad=True)
    xx = to_bf16(x).to(dtype)
    tt = torch.as_tensor(t).to(dtype)
    L = loss(forward(P, xx, bf16_acts), xx, tt)
    L.backward()
    lay, n = layout()
    g = np.zeros(n, np.float64)
    for name, off, shape in lay:
        g[off:off + int(np.prod(shape))] = P[name].grad.detach().double().numpy().ravel()
    return float(L.detach()), g
