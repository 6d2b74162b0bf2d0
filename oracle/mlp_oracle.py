"""FP64 CPU restatement of the chemistry-surrogate MLP (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module, as the
checker.  The surrogate has no counterpart in the reference: it is pinned by finite-difference gradient
checks of this restatement (tests/test_surrogate_oracle_cpu.py), and the GPU kernels are compared with it
with the GPU's BF16 operand rounding emulated (`emulate=True`).

Network: 12 -> 128 -> 128 -> 128 -> 128 -> 10 with a hidden activation on the four hidden layers, linear output:
GELU in its tanh form 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))) (Hendrycks and Gimpel, 2016; act="gelu", the
default) or hard swish z clamp(z/6 + 1/2, 0, 1) (Howard et al., "Searching for MobileNetV3", 2019; act="hswish").
GPU numerics emulated here: every GEMM operand (layer inputs, weights, back-propagated gradients) is
rounded to BF16 (round to nearest even), products are accumulated in FP32 on the GPU (FP64 here), and
pre-activations, activations and the loss are FP32 on the GPU (FP64 here).
"""
import numpy as np

SIZES = (12, 128, 128, 128, 128, 10)


def bf16(x):
    """Round to BF16 (nearest even) and return as float64."""
    x32 = np.ascontiguousarray(x, dtype=np.float32)
    u = x32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


GELU_S, GELU_A = np.sqrt(2.0 / np.pi), 0.044715


def hswish(z):
    """Hard swish."""
    return z * np.clip(z / 6.0 + 0.5, 0.0, 1.0)


def hswish_grad(z):
    """Its derivative: 0 below -3, 1 above 3, z/3 + 1/2 between (the kinks take the outer values)."""
    return np.where(z <= -3.0, 0.0, np.where(z >= 3.0, 1.0, z / 3.0 + 0.5))


def gelu(z):
    """GELU, tanh form."""
    return 0.5 * z * (1.0 + np.tanh(GELU_S * (z + GELU_A * z ** 3)))


def gelu_grad(z):
    t = np.tanh(GELU_S * (z + GELU_A * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GELU_S * (1.0 + 3.0 * GELU_A * z * z)


ACTS = {"gelu": (gelu, gelu_grad), "hswish": (hswish, hswish_grad)}


def act(z, kind="gelu"):
    return ACTS[kind][0](z)


def act_grad(z, kind="gelu"):
    return ACTS[kind][1](z)


def init_params(seed, sizes=SIZES):
    """He-type initialisation: W_l (K_l, N_l) ~ N(0, 2/K_l) (1/K_l for the output layer), b_l = 0."""
    rng = np.random.default_rng(seed)
    params = []
    for l in range(len(sizes) - 1):
        k, n = sizes[l], sizes[l + 1]
        scale = np.sqrt((2.0 if l < len(sizes) - 2 else 1.0) / k)
        params.append((rng.standard_normal((k, n)) * scale, np.zeros(n)))
    return params


def forward(params, X, emulate=True, kind="gelu"):
    """Returns (outputs, cache).  X: (B, 12)."""
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    H = r(X)
    cache = []
    L = len(params)
    for l, (W, b) in enumerate(params):
        Wr = r(W)
        Z = H @ Wr + b
        cache.append((H, Wr, Z))
        if l < L - 1:
            H = r(act(Z, kind))
    return Z, cache


def loss_and_grads(params, X, Y, emulate=True, kind="gelu"):
    """Mean squared error over batch and outputs, and its gradients [(dW_l, db_l)]."""
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    out, cache = forward(params, X, emulate, kind)
    B, N = out.shape
    loss = np.mean((out - Y) ** 2)
    G = r(2.0 * (out - Y) / (B * N))
    grads = [None] * len(params)
    for l in range(len(params) - 1, -1, -1):
        H, Wr, _ = cache[l]
        grads[l] = (H.T @ G, G.sum(axis=0))
        if l > 0:
            dH = G @ Wr.T
            G = r(dH * act_grad(cache[l - 1][2], kind))
    return loss, grads


def adam_step(params, grads, state, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step in place; state = dict(t, m, v) (created on first use)."""
    if not state:
        state.update(t=0, m=[(np.zeros_like(W), np.zeros_like(b)) for W, b in params],
                     v=[(np.zeros_like(W), np.zeros_like(b)) for W, b in params])
    state["t"] += 1
    t = state["t"]
    c1, c2 = 1 - beta1 ** t, 1 - beta2 ** t
    new = []
    for l, ((W, b), (gW, gb)) in enumerate(zip(params, grads)):
        out = []
        for q, (p, g) in enumerate(((W, gW), (b, gb))):
            m = state["m"][l][q] * beta1 + (1 - beta1) * g
            v = state["v"][l][q] * beta2 + (1 - beta2) * g * g
            state["m"][l] = tuple(m if qq == q else state["m"][l][qq] for qq in range(2))
            state["v"][l] = tuple(v if qq == q else state["v"][l][qq] for qq in range(2))
            out.append(p - lr * (m / c1) / (np.sqrt(v / c2) + eps))
        new.append(tuple(out))
    params[:] = new
    return params
