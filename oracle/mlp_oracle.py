"""CPU oracle of the near-contact closure network -- TEST INFRASTRUCTURE ONLY.

A float64 numpy restatement of the 4 -> 64 -> 64 -> 3 MLP of
lbcuda_amd/csrc/lbm_mlp.hip (tanh hidden layers, linear output), its mean
squared error and the analytic gradient by back-propagation.  The reference
has no such network (its near-contact model is the lubrication term of
kernels_fluid_PART.CUF:1463-1501); this oracle is pinned by finite
differences of its own loss in tests/test_mlp_oracle.py.

Parameter vector layout (4675 floats): W1[64][4] b1[64] W2[64][64] b2[64]
W3[3][64] b3[3], rows are outputs.
"""
import numpy as np

NI, NH, NO = 4, 64, 3
SIZES = [("W1", (NH, NI)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NO, NH)), ("b3", (NO,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, o = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = p[o:o + k].reshape(shape)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SIZES])


def init_params(seed=0):
    """Glorot-uniform weights, zero biases (float32 vector)."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in SIZES:
        if name.startswith("W"):
            lim = np.sqrt(6.0 / (shape[0] + shape[1]))
            d[name] = rng.uniform(-lim, lim, shape)
        else:
            d[name] = np.zeros(shape)
    return pack(d).astype(np.float32)


def forward(p, x):
    """y (n,3) and the hidden activations for x (n,4)."""
    w = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    a1 = np.tanh(x @ w["W1"].T + w["b1"])
    a2 = np.tanh(a1 @ w["W2"].T + w["b2"])
    y = a2 @ w["W3"].T + w["b3"]
    return y, (x, a1, a2)


def loss(p, x, t):
    y, _ = forward(p, x)
    return float(np.mean((y - np.asarray(t, dtype=np.float64)) ** 2))


def loss_grad(p, x, t):
    """Mean squared error over all n*3 outputs and its gradient."""
    w = unpack(p)
    y, (x, a1, a2) = forward(p, x)
    t = np.asarray(t, dtype=np.float64)
    n = x.shape[0]
    dy = 2.0 * (y - t) / (n * NO)
    g = {"W3": dy.T @ a2, "b3": dy.sum(0)}
    d2 = (dy @ w["W3"]) * (1.0 - a2 ** 2)
    g["W2"] = d2.T @ a1
    g["b2"] = d2.sum(0)
    d1 = (d2 @ w["W2"]) * (1.0 - a1 ** 2)
    g["W1"] = d1.T @ x
    g["b1"] = d1.sum(0)
    return float(np.mean((y - t) ** 2)), pack(g)
