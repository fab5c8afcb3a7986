"""FP64 numpy restatement of the schedule selector (test infrastructure).

MLP 16 -> 64 -> 64 -> 4 with ReLU, parameter layout
W1[64][16] b1[64] W2[64][64] b2[64] W3[4][64] b3[4] (5,508 values), loss =
sum_i weight_i * softmax-cross-entropy(logits_i, label_i).  Used to check the
GPU forward/backward kernels; pinned itself by finite differences.
"""
import numpy as np

IN, H, OUT = 16, 64, 4
SIZES = [("W1", (H, IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (OUT, H)), ("b3", (OUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(w):
    out, o = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = np.asarray(w[o:o + k], dtype=np.float64).reshape(shape)
        o += k
    return out


def pack(p):
    return np.concatenate([p[name].reshape(-1) for name, _ in SIZES])


def forward(w, x):
    p = unpack(w)
    a1 = x @ p["W1"].T + p["b1"]
    h1 = np.maximum(a1, 0)
    a2 = h1 @ p["W2"].T + p["b2"]
    h2 = np.maximum(a2, 0)
    return h2 @ p["W3"].T + p["b3"], (a1, h1, a2, h2)


def loss_grad(w, x, y, sw=None):
    x = np.asarray(x, dtype=np.float64)
    n = x.shape[0]
    sw = np.ones(n) if sw is None else np.asarray(sw, dtype=np.float64)
    p = unpack(w)
    logits, (a1, h1, a2, h2) = forward(w, x)
    z = logits - logits.max(axis=1, keepdims=True)
    e = np.exp(z)
    prob = e / e.sum(axis=1, keepdims=True)
    loss = float(np.sum(sw * -np.log(prob[np.arange(n), y])))
    dl = prob.copy()
    dl[np.arange(n), y] -= 1.0
    dl *= sw[:, None]
    g = {"W3": dl.T @ h2, "b3": dl.sum(0)}
    d2 = (dl @ p["W3"]) * (a2 > 0)
    g["W2"] = d2.T @ h1
    g["b2"] = d2.sum(0)
    d1 = (d2 @ p["W2"]) * (a1 > 0)
    g["W1"] = d1.T @ x
    g["b1"] = d1.sum(0)
    return loss, pack(g)


def init(seed=0):
    rng = np.random.default_rng(seed)
    p = {"W1": rng.uniform(-1, 1, (H, IN)) * np.sqrt(6 / IN), "b1": rng.uniform(-0.1, 0.1, H),
         "W2": rng.uniform(-1, 1, (H, H)) * np.sqrt(6 / H), "b2": rng.uniform(-0.1, 0.1, H),
         "W3": rng.uniform(-1, 1, (OUT, H)) * np.sqrt(6 / H), "b3": rng.uniform(-0.1, 0.1, OUT)}
    return pack(p)


def bf16_round(w):
    """Round FP32 values to BF16 (nearest even), returned as FP32."""
    u = np.asarray(w, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)
