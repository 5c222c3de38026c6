"""CPU oracle of the engine's learned correction gate (TEST INFRASTRUCTURE
ONLY: used by tests/ and never by the product).

The gate is the engine's own addition (no counterpart in the reference): a
32-64-32-16-1 MLP with ReLU hidden layers and a sigmoid output, trained with
the mean binary cross-entropy.  This module restates the device arithmetic of
care_amd/csrc/mlp_device.h and mlp.hip in float64 numpy:

* inputs and the three weight matrices are rounded to bfloat16 (round to
  nearest even), as the device does before its matrix instructions;
* hidden activations of layers 1 and 2 are rounded to bfloat16 after ReLU,
  layer 3 and the output layer stay in full precision;
* in the backward pass the gradients entering a matrix product (dH3, dH2,
  dH1) are rounded to bfloat16, bias gradients are summed before rounding.

With ``bf16=False`` every rounding is skipped, which gives the plain float64
network that the finite-difference test pins.

Parameter vector layout (CARE_MLP_PARAMS = 4737 floats): W1[64][32], b1[64],
W2[32][64], b2[32], W3[16][32], b3[16], w4[16], b4; matrices output-major.
"""
import numpy as np

N_IN, N_H1, N_H2, N_H3 = 32, 64, 32, 16
SIZES = [("W1", (N_H1, N_IN)), ("b1", (N_H1,)), ("W2", (N_H2, N_H1)), ("b2", (N_H2,)),
         ("W3", (N_H3, N_H2)), ("b3", (N_H3,)), ("w4", (N_H3,)), ("b4", (1,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)
assert N_PARAMS == 4737


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return ((u << 16).astype(np.uint32)).view(np.float32).astype(np.float64).reshape(a.shape)


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).reshape(-1) for name, _ in SIZES])


def init_params(seed):
    """He-uniform initial parameters (zero biases)."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in SIZES:
        if len(shape) == 2:
            lim = np.sqrt(6.0 / shape[1])
            d[name] = rng.uniform(-lim, lim, size=shape)
        elif name == "w4":
            lim = np.sqrt(6.0 / N_H3)
            d[name] = rng.uniform(-lim, lim, size=shape)
        else:
            d[name] = np.zeros(shape)
    return pack(d).astype(np.float32)


def forward(x, params, bf16_round=True):
    """Logits z and the intermediates of the forward pass."""
    r = bf16 if bf16_round else (lambda v: np.asarray(v, dtype=np.float64))
    p = unpack(params)
    X = r(np.asarray(x, dtype=np.float32))
    if bf16_round:
        W1, W2, W3 = bf16(p["W1"].astype(np.float32)), bf16(p["W2"].astype(np.float32)), bf16(p["W3"].astype(np.float32))
    else:
        W1, W2, W3 = p["W1"], p["W2"], p["W3"]
    a1 = X @ W1.T + p["b1"]
    H1 = r(np.maximum(a1, 0).astype(np.float32)) if bf16_round else np.maximum(a1, 0)
    a2 = H1 @ W2.T + p["b2"]
    H2 = r(np.maximum(a2, 0).astype(np.float32)) if bf16_round else np.maximum(a2, 0)
    H3 = np.maximum(H2 @ W3.T + p["b3"], 0)
    z = H3 @ p["w4"] + p["b4"][0]
    return z, (X, W1, W2, W3, H1, H2, H3)


def loss_and_grad(x, y, params, bf16_round=True):
    """Mean binary cross-entropy over the rows and its gradient (flat)."""
    r = bf16 if bf16_round else (lambda v: np.asarray(v, dtype=np.float64))
    p = unpack(params)
    z, (X, W1, W2, W3, H1, H2, H3) = forward(x, params, bf16_round)
    y = np.asarray(y, dtype=np.float64)
    n = len(y)
    loss = np.mean(np.maximum(z, 0) - y * z + np.log1p(np.exp(-np.abs(z))))
    dz = (1.0 / (1.0 + np.exp(-z)) - y) / n
    g = {}
    g["w4"] = H3.T @ dz
    g["b4"] = np.array([dz.sum()])
    d3 = np.outer(dz, p["w4"]) * (H3 > 0)
    g["b3"] = d3.sum(axis=0)
    d3r = r(d3.astype(np.float32)) if bf16_round else d3
    g["W3"] = d3r.T @ H2
    d2 = (d3r @ W3) * (H2 > 0)
    g["b2"] = d2.sum(axis=0)
    d2r = r(d2.astype(np.float32)) if bf16_round else d2
    g["W2"] = d2r.T @ H1
    d1 = (d2r @ W2) * (H1 > 0)
    g["b1"] = d1.sum(axis=0)
    d1r = r(d1.astype(np.float32)) if bf16_round else d1
    g["W1"] = d1r.T @ X
    return loss, pack(g)
