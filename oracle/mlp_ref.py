"""Numpy restatement of the learned seed's network (test infrastructure only).

The learned seed has no counterpart in the reference; this module states the
network the GPU kernels compute so the tests can check them: a 16-64-64-4
perceptron with ReLU, loss per row
  (out0 - y0)^2 + w_pass BCE(sigmoid(out1), y1)
    + (out2 - y2)^2 + w_pass BCE(sigmoid(out3), y3)
averaged over the rows, a NaN target leaving its term out. With ``quant=True`` it mirrors the kernels' arithmetic
(every matrix-product operand rounded to BF16, round to nearest even; FP32
accumulation, biases and ReLU masks); with ``quant=False`` it is the exact
FP64 network, whose gradients the tests pin by central differences.
"""
import numpy as np

N_FEAT, N_HID, N_OUT = 16, 64, 4
N_PARAMS = N_FEAT * N_HID + N_HID + N_HID * N_HID + N_HID + N_HID * N_OUT + N_OUT


def bf16_bits(x):
    """float32 -> bf16 bit patterns (uint16), round to nearest even."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def bf16(x):
    return (bf16_bits(x).astype(np.uint32) << 16).view(np.float32)


def unpack(p):
    o = 0
    W1 = p[o:o + N_FEAT * N_HID].reshape(N_FEAT, N_HID); o += N_FEAT * N_HID
    b1 = p[o:o + N_HID]; o += N_HID
    W2 = p[o:o + N_HID * N_HID].reshape(N_HID, N_HID); o += N_HID * N_HID
    b2 = p[o:o + N_HID]; o += N_HID
    W3 = p[o:o + N_HID * N_OUT].reshape(N_HID, N_OUT); o += N_HID * N_OUT
    b3 = p[o:o + N_OUT]
    return W1, b1, W2, b2, W3, b3


class _Num:
    def __init__(self, quant):
        self.quant = quant
        self.dt = np.float32 if quant else np.float64

    def q(self, x):
        return bf16(x) if self.quant else np.asarray(x, np.float64)

    def mm(self, a, b):
        return (np.asarray(a, np.float64) @ np.asarray(b, np.float64)).astype(self.dt)


def forward(p, X, quant=True):
    """Outputs [n][4] for standardised features X (rounded to BF16 when quant)."""
    N = _Num(quant)
    W1, b1, W2, b2, W3, b3 = unpack(np.asarray(p, N.dt))
    Xq = N.q(X)
    H1 = np.maximum(N.mm(Xq, N.q(W1)) + b1, 0)
    H2 = np.maximum(N.mm(N.q(H1), N.q(W2)) + b2, 0)
    return N.mm(N.q(H2), N.q(W3)) + b3


def _terms(O, y, w_pass):
    """Per-element loss terms and their derivatives (zero where y is NaN)."""
    m = ~np.isnan(y)
    yz = np.where(m, y, 0.0)
    L = np.zeros_like(O, dtype=np.float64)
    G = np.zeros_like(O, dtype=np.float64)
    for c in range(N_OUT):
        o = O[:, c].astype(np.float64)
        if c % 2 == 0:
            d = o - yz[:, c]
            L[:, c], G[:, c] = d * d, 2 * d
        else:
            L[:, c] = w_pass * (np.maximum(o, 0) - o * yz[:, c] + np.log1p(np.exp(-np.abs(o))))
            G[:, c] = w_pass * (1 / (1 + np.exp(-o)) - yz[:, c])
    return np.where(m, L, 0.0), np.where(m, G, 0.0)


def loss(p, X, y, w_pass, quant=True):
    O = forward(p, X, quant)
    L, _ = _terms(O, y, w_pass)
    return np.sum(L) / X.shape[0]


def loss_and_grad(p, X, y, w_pass, quant=True):
    """Mean loss and its gradient with respect to the packed parameters."""
    N = _Num(quant)
    W1, b1, W2, b2, W3, b3 = unpack(np.asarray(p, N.dt))
    n = X.shape[0]
    Xq = N.q(X)
    H1 = np.maximum(N.mm(Xq, N.q(W1)) + b1, 0)
    H1q = N.q(H1)
    H2 = np.maximum(N.mm(H1q, N.q(W2)) + b2, 0)
    H2q = N.q(H2)
    O = N.mm(H2q, N.q(W3)) + b3
    Lt, Gt = _terms(O, y, w_pass)
    L = np.sum(Lt) / n
    dO = (Gt / n).astype(N.dt)
    dOq = N.q(dO)
    gb3 = dO.sum(0)
    gW3 = N.mm(H2q.T, dOq)
    dH2 = N.mm(dOq, N.q(W3).T) * (H2 > 0)
    gb2 = dH2.sum(0)
    dH2q = N.q(dH2)
    gW2 = N.mm(H1q.T, dH2q)
    dH1 = N.mm(dH2q, N.q(W2).T) * (H1 > 0)
    gb1 = dH1.sum(0)
    gW1 = N.mm(Xq.T, N.q(dH1))
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return L, g


def init_params(seed=0):
    """He-initialised parameters (zero biases), float32 [N_PARAMS]."""
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0, np.sqrt(2 / N_FEAT), (N_FEAT, N_HID))
    W2 = rng.normal(0, np.sqrt(2 / N_HID), (N_HID, N_HID))
    W3 = rng.normal(0, np.sqrt(1 / N_HID), (N_HID, N_OUT))
    z = np.zeros
    return np.concatenate([W1.ravel(), z(N_HID), W2.ravel(), z(N_HID), W3.ravel(), z(N_OUT)]).astype(np.float32)
