"""FP64 CPU restatement of the learned-microphysics network -- TEST
INFRASTRUCTURE ONLY (the checker of the GPU's forward and backward kernels;
the product never imports it).

Network (cress_amd/csrc/mlp.h): x[16] (9 normalised features, zero padded)
-> z1 = x W1 + b1, h1 = gelu(z1) -> z2, h2 -> z3, h3 -> y = h3 W4 + b4
(16 outputs, the first 7 used), gelu the tanh form
    gelu(z) = z/2 (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3)))
(Hendrycks & Gimpel 2016).  Loss: mean over the batch and the 7 outputs of
the squared error of the normalised outputs.  The parameter block is the
library's: W1[16][128] b1 W2[128][128] b2 W3[128][128] b3 W4[128][16] b4[16].
"""
import numpy as np

NIN, NOUT, H, INP, OUTP = 9, 7, 128, 16, 16
OW1 = 0
OB1 = OW1 + INP * H
OW2 = OB1 + H
OB2 = OW2 + H * H
OW3 = OB2 + H
OB3 = OW3 + H * H
OW4 = OB3 + H
OB4 = OW4 + H * OUTP
NPARAM = OB4 + OUTP
NNORM = 96
C = np.sqrt(2.0 / np.pi)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OW1:OB1].reshape(INP, H), p[OB1:OW2], p[OW2:OB2].reshape(H, H), p[OB2:OW3],
            p[OW3:OB3].reshape(H, H), p[OB3:OW4], p[OW4:OB4].reshape(H, OUTP), p[OB4:NPARAM])


def pack(w1, b1, w2, b2, w3, b3, w4, b4):
    return np.concatenate([w1.ravel(), b1, w2.ravel(), b2, w3.ravel(), b3, w4.ravel(), b4])


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(C * (z + 0.044715 * z ** 3)))


def dgelu(z):
    t = np.tanh(C * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * C * (1.0 + 3.0 * 0.044715 * z * z)


def bf16(a):
    """Round to the nearest BF16 value, ties to even (the GPU's f2bf)."""
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def normalise(inp, norm):
    """Raw features (n, 9) -> padded normalised (n, 16)."""
    inp = np.asarray(inp, np.float64)
    x = np.zeros((inp.shape[0], INP))
    x[:, :NIN] = (inp - norm[:NIN]) * norm[16:16 + NIN]
    return x


def forward(p, x):
    w1, b1, w2, b2, w3, b3, w4, b4 = unpack(p)
    z1 = x @ w1 + b1
    h1 = gelu(z1)
    z2 = h1 @ w2 + b2
    h2 = gelu(z2)
    z3 = h2 @ w3 + b3
    h3 = gelu(z3)
    y = h3 @ w4 + b4
    return y, (x, z1, h1, z2, h2, z3, h3)


def loss_grad(p, x, t):
    """Mean squared error over the batch and the 7 outputs, and its
    gradient with respect to the parameter block."""
    w1, b1, w2, b2, w3, b3, w4, b4 = unpack(p)
    y, (x, z1, h1, z2, h2, z3, h3) = forward(p, x)
    n = x.shape[0]
    e = y[:, :NOUT] - t
    loss = (e * e).sum() / (n * NOUT)
    dy = np.zeros_like(y)
    dy[:, :NOUT] = 2.0 * e / (n * NOUT)
    gw4 = h3.T @ dy
    gb4 = dy.sum(0)
    dz3 = (dy @ w4.T) * dgelu(z3)
    gw3 = h2.T @ dz3
    gb3 = dz3.sum(0)
    dz2 = (dz3 @ w3.T) * dgelu(z2)
    gw2 = h1.T @ dz2
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ w2.T) * dgelu(z1)
    gw1 = x.T @ dz1
    gb1 = dz1.sum(0)
    return loss, pack(gw1, gb1, gw2, gb2, gw3, gb3, gw4, gb4)


def predict(p, norm, inp):
    """Raw features (n, 9) -> raw tendencies (n, 7)."""
    y, _ = forward(p, normalise(inp, norm))
    return y[:, :NOUT] * norm[48:48 + NOUT] + norm[32:32 + NOUT]


def forward_bf16(p, x):
    """The forward pass with the GPU's roundings: weights, inputs and every
    hidden activation rounded to BF16, products summed in FP64 (the GPU
    sums in FP32), biases and GELU in FP64."""
    w1, b1, w2, b2, w3, b3, w4, b4 = unpack(p)
    h = bf16(x)
    for w, b in ((w1, b1), (w2, b2), (w3, b3)):
        h = bf16(gelu(h @ bf16(w) + b))
    return h @ bf16(w4) + b4
