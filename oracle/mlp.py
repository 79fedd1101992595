"""FP64 restatement of the HH surrogate MLP -- TEST INFRASTRUCTURE ONLY.

The surrogate has no counterpart in the reference (which integrates HH by forward Euler only,
src/neuron/hodgkin_huxley.rs:47-92), so this module is pinned by mathematics instead: its
backward pass against central finite differences of its own forward pass (exact mode), and its
bf16 rounding against hand-worked round-to-nearest-even cases.  Only tests/ use it, as the
checker of the device MLP's forward and gradient kernels.

Network: x (5) -> relu(W1 x + b1) (64) -> relu(W2 a1 + b2) (64) -> W3 a2 + b3 (4), all on
normalised quantities: x_n = (x - in_mu) * in_isd, target_n = (target - out_mu) / out_sd.
Loss = mean over samples and the 4 outputs of (y - target_n)^2.

`emulate=True` reproduces the device's roundings: weights to bf16 (layers 1-2; the output
layer's as the sum of two bf16 words, hi + bf16(W3 - hi)), the layer-1 bias as two bf16 words,
the normalised inputs as two bf16 words (x = hi + bf16(x - hi), the layer-1 MFMA's operands),
the layer-1 and layer-2 activations to bf16, the input normalisation in FP32; everything else
is FP64.
`emulate="weights"` / `"act"` apply only the weight (and bias) / only the input and activation
roundings, "w1", "w2", "w3" only that layer's weight rounding.
The backward pass treats the roundings as the identity (as the device does): the layer-1
weight gradient uses the unsplit inputs.
"""
import numpy as np

N_PARAMS = 4804
OFF = {"w1": (0, (64, 5)), "b1": (320, (64,)), "w2": (384, (64, 64)), "b2": (4480, (64,)),
       "w3": (4544, (4, 64)), "b3": (4800, (4,))}
NORM_FIELDS = {"in_mu": (0, 5), "in_isd": (5, 5), "out_mu": (10, 4), "out_sd": (14, 4),
               "lo": (18, 5), "hi": (23, 5), "scale_eps": (28, 1)}


def round_bf16(x):
    """float32 -> nearest bf16 (ties to even), returned as float32."""
    b = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32)


def split_bf16(x):
    """x as the sum of two bf16 words, hi + bf16(x - hi) (the difference is exact in FP32)."""
    x = np.asarray(x, np.float32)
    hi = round_bf16(x)
    return hi.astype(np.float64) + round_bf16((x - hi).astype(np.float32))


def unpack(params):
    p = np.asarray(params)
    return {k: p[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in OFF.items()}


def pack(parts):
    p = np.zeros(N_PARAMS, np.float64)
    for k, (o, s) in OFF.items():
        p[o:o + int(np.prod(s))] = np.asarray(parts[k], np.float64).ravel()
    return p


def unpack_norm(norm):
    v = np.asarray(norm, np.float32)
    return {k: v[o:o + c] for k, (o, c) in NORM_FIELDS.items()}


def normalise_inputs(norm, x):
    nm = unpack_norm(norm)
    x = np.asarray(x, np.float32)
    return ((x - nm["in_mu"]).astype(np.float32) * nm["in_isd"]).astype(np.float32)


def normalise_targets(norm, t):
    nm = unpack_norm(norm)
    t = np.asarray(t, np.float32)
    return ((t - nm["out_mu"]).astype(np.float32) / nm["out_sd"]).astype(np.float32)


def inside(norm, x):
    """Training-envelope test of raw states x (count x 5): the fallback selector."""
    nm = unpack_norm(norm)
    x = np.asarray(x, np.float32)
    return np.all((x >= nm["lo"]) & (x <= nm["hi"]), axis=1)


def _weights(params, emulate):
    w = unpack(params)
    out = {}
    for k, v in w.items():
        if emulate and emulate != "act":
            v = np.asarray(v, np.float32)
            if k.startswith("w") and (emulate in (True, "weights") or emulate == k):
                hi = round_bf16(v)
                # the output layer carries a second bf16 word: W3 = hi + bf16(W3 - hi)
                v = (hi.astype(np.float64) + round_bf16(v - hi)) if k == "w3" else hi
            if k == "b1" and emulate in (True, "weights", "w1"):
                v = split_bf16(v)  # two bf16 words against constant-1 inputs
        out[k] = np.asarray(v, np.float64)
    return out


def forward(params, xn, emulate=True):
    """Normalised outputs (count x 4) and the activations, from normalised inputs xn."""
    w = _weights(params, emulate)
    xe = split_bf16(xn) if emulate in (True, "act") else np.asarray(xn, np.float64)
    xn = np.asarray(xn, np.float64)
    z1 = xe @ w["w1"].T + w["b1"]
    a1 = np.maximum(z1, 0.0)
    if emulate in (True, "act"):
        a1 = round_bf16(a1.astype(np.float32)).astype(np.float64)
    z2 = a1 @ w["w2"].T + w["b2"]
    a2 = np.maximum(z2, 0.0)
    if emulate in (True, "act"):
        a2 = round_bf16(a2.astype(np.float32)).astype(np.float64)
    y = a2 @ w["w3"].T + w["b3"]
    return y, (xn, a1, a2, w)


def predict(params, norm, x, emulate=True):
    """Raw increments (count x 4) for raw states x (count x 5)."""
    nm = unpack_norm(norm)
    y, _ = forward(params, normalise_inputs(norm, x), emulate)
    return y * nm["out_sd"].astype(np.float64) + nm["out_mu"].astype(np.float64)


def loss_and_grad(params, xn, tn, emulate=True):
    """Mean squared error over samples and outputs, and its gradient in parameter order."""
    y, (xn, a1, a2, w) = forward(params, xn, emulate)
    tn = np.asarray(tn, np.float64)
    e = y - tn
    count = e.size
    loss = float(np.sum(e * e) / count)
    dy = 2.0 * e / count
    g = {"w3": dy.T @ a2, "b3": dy.sum(0)}
    dz2 = (dy @ w["w3"]) * (a2 > 0)
    g["w2"] = dz2.T @ a1
    g["b2"] = dz2.sum(0)
    dz1 = (dz2 @ w["w2"]) * (a1 > 0)
    g["w1"] = dz1.T @ xn
    g["b1"] = dz1.sum(0)
    return loss, pack(g)


def loss_only(params, xn, tn, emulate=False):
    y, _ = forward(params, xn, emulate)
    e = y - np.asarray(tn, np.float64)
    return float(np.sum(e * e) / e.size)


def init_params(seed):
    """He-uniform weights (the device initialiser's bounds) and small random biases, so that
    every bias gradient is exercised."""
    rng = np.random.default_rng(seed)
    parts = {"w1": rng.uniform(-1, 1, (64, 5)) * np.sqrt(6 / 5),
             "w2": rng.uniform(-1, 1, (64, 64)) * np.sqrt(6 / 64),
             "w3": rng.uniform(-1, 1, (4, 64)) * np.sqrt(1 / 64),
             "b1": rng.uniform(-0.1, 0.1, 64), "b2": rng.uniform(-0.1, 0.1, 64),
             "b3": rng.uniform(-0.1, 0.1, 4)}
    return pack(parts).astype(np.float32)
