"""MLP surrogate oracle -- TEST INFRASTRUCTURE ONLY.

A numpy restatement of the chemistry surrogate network (11 -> 256 -> 256 ->
256 -> 8, GELU in its tanh form) in the packed parameter layout of
the library, with its loss (mean squared error over the 8 outputs) and
backward pass.  The network has no counterpart in the reference; this file
is the checker of the GPU training kernels: only tests/ may use it.

Arithmetic is FP64.  With bf=True every GEMM operand is rounded to BF16
(round to nearest even) where the GPU kernels round it: the layer inputs
(the standardised features, the stored hidden activations), the weights
and biases, and the output-gradient operands of the backward GEMMs.

Parameter layout (FP32 array, as in include/ostc.h): layer l is a
row-major [out_pad x ld] matrix, weights in [:dout, :din], the bias in
column kpad, zeros elsewhere.

Pinned in tests/test_oracle_mlp.py: GELU against its logistic form and the
erf form it approximates, its derivative against finite differences, backward against central finite
differences of the FP64 loss, the layout against a hand-built network, and
the BF16 rounding against torch's bfloat16 cast (ties to even, representable
values unchanged).
"""
import math

import numpy as np

# (din, kpad, ld, dout, out_pad)
LAYERS = [(11, 16, 24, 256, 256), (256, 256, 264, 256, 256), (256, 256, 264, 256, 256), (256, 256, 264, 8, 16)]
N_PARAMS = sum(op * ld for _, _, ld, _, op in LAYERS)


def bf16(a):
    """Round to BF16 (nearest, ties to even) via the FP32 value; returns FP64."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64).reshape(np.shape(a))


K_GELU = math.sqrt(2.0 / math.pi)


def gelu(x):
    """x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))) / 2"""
    return 0.5 * x * (1.0 + np.tanh(K_GELU * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    t = np.tanh(K_GELU * (x + 0.044715 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * K_GELU * (1.0 + 3 * 0.044715 * x * x)


def offsets():
    off, o = [], 0
    for _, _, ld, _, op in LAYERS:
        off.append(o)
        o += op * ld
    return off


def unpack(p):
    """Packed parameters -> [(W [dout x din], b [dout])] in FP64."""
    p = np.asarray(p, dtype=np.float64)
    out = []
    for (din, kpad, ld, dout, op), o in zip(LAYERS, offsets()):
        M = p[o:o + op * ld].reshape(op, ld)
        out.append((M[:dout, :din].copy(), M[:dout, kpad].copy()))
    return out


def pack(layers):
    p = np.zeros(N_PARAMS)
    for (din, kpad, ld, dout, op), o, (W, b) in zip(LAYERS, offsets(), layers):
        M = np.zeros((op, ld))
        M[:dout, :din] = W
        M[:dout, kpad] = b
        p[o:o + op * ld] = M.ravel()
    return p


def forward(p, x, bf=True):
    """Returns (y, cache) with y [B x 8]."""
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(p)
    a = r(x)
    cache = []
    for l, (W, b) in enumerate(layers):
        z = a @ r(W).T + r(b)
        cache.append((a, z))
        if l < len(layers) - 1:
            a = r(gelu(z))
        else:
            y = z
    return y, cache


def loss_and_grad(p, x, t, bf=True):
    """Mean squared error over the 8 outputs and its gradient (packed)."""
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(p)
    y, cache = forward(p, x, bf)
    B, nout = y.shape
    d = y - t
    loss = float((d * d).sum() / (B * nout))
    dz = 2.0 * d / (B * nout)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        a, _ = cache[l]
        g = r(dz)
        grads[l] = (g.T @ a, g.sum(axis=0))
        if l > 0:
            W, _ = layers[l]
            dz = (g @ r(W)) * gelu_grad(cache[l - 1][1])
    return loss, pack(grads)


def init_params(seed):
    """Same distribution as ostc_surrogate_init (not the same stream)."""
    rng = np.random.default_rng(seed)
    layers = []
    for l, (din, _, _, dout, _) in enumerate(LAYERS):
        a = math.sqrt((6.0 if l < len(LAYERS) - 1 else 3.0) / din)
        layers.append((rng.uniform(-a, a, (dout, din)), np.zeros(dout)))
    return pack(layers)
