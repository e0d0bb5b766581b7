"""ORACLE (test infrastructure only; never imported by the product).

CPU restatement of the learned initialiser's network in numpy: forward pass,
loss, backward pass and the output-head decoding of cudimot_amd's
cdm_mlp_* entry points (cudimot_amd/csrc/mlp/mlp.hip).  There is no
reference counterpart (the reference has no learned initialiser).

bf16=True rounds at the same points as the GPU kernels (inputs, weights,
hidden activations and back-propagated deltas are BF16; products accumulate
in wider precision), so GPU and oracle differ only by accumulation order.
bf16=False is plain FP64 and is what the finite-difference check pins.

Head kinds follow cudimot_amd.capi: SIGMOID, SIGMOID_LOG, SIGMOID_RATIO, AXIS.
"""
import numpy as np
from scipy.special import erf

SIGMOID, SIGMOID_LOG, SIGMOID_RATIO, AXIS = range(4)


def round_bf16(x):
    """Round-to-nearest-even to bfloat16, returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def sigmoid(o):
    return 1.0 / (1.0 + np.exp(-o))


def unflatten(w, shapes):
    """Flat weights (per layer W[out][in] then b[out]) -> [(W, b), ...]."""
    out, k = [], 0
    for (o, i) in shapes:
        W = np.asarray(w[k:k + o * i], dtype=np.float64).reshape(o, i)
        k += o * i
        b = np.asarray(w[k:k + o], dtype=np.float64)
        k += o
        out.append((W, b))
    return out


def flatten(layers):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers])


def out_index(head):
    idx, k = [], 0
    for e in head:
        idx.append(k)
        k += 3 if e["kind"] == AXIS else 1
    return idx, k


def encode(head, params):
    """True parameters [n][P] -> targets [n][n_out]."""
    idx, n_out = out_index(head)
    p = np.asarray(params, dtype=np.float64)
    t = np.zeros((p.shape[0], n_out))
    for e, a in zip(head, idx):
        x = p[:, e["param"]]
        if e["kind"] == AXIS:
            th, ph = x, p[:, e["param2"]]
            t[:, a:a + 3] = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1)
            continue
        if e["kind"] == SIGMOID:
            u = (x - e["lo"]) / (e["hi"] - e["lo"])
        elif e["kind"] == SIGMOID_LOG:
            u = (np.log(x) - np.log(e["lo"])) / (np.log(e["hi"]) - np.log(e["lo"]))
        else:
            u = (x / p[:, e["param2"]] - e["lo"]) / (e["hi"] - e["lo"])
        t[:, a] = np.clip(u, 0.0, 1.0)
    return t


def decode(head, out, params):
    """Raw outputs [n][n_out] -> parameters (entries in order; others kept)."""
    idx, _ = out_index(head)
    p = np.array(params, dtype=np.float64)
    o = np.asarray(out, dtype=np.float64)
    for e, a in zip(head, idx):
        if e["kind"] == AXIS:
            v = o[:, a:a + 3]
            r = np.linalg.norm(v, axis=1)
            ok = r > 0
            p[ok, e["param"]] = np.arccos(np.clip(v[ok, 2] / r[ok], -1, 1))
            p[ok, e["param2"]] = np.arctan2(v[ok, 1], v[ok, 0])
            continue
        s = sigmoid(o[:, a])
        if e["kind"] == SIGMOID:
            x = e["lo"] + (e["hi"] - e["lo"]) * s
        elif e["kind"] == SIGMOID_LOG:
            x = np.exp(np.log(e["lo"]) + (np.log(e["hi"]) - np.log(e["lo"])) * s)
        else:
            x = (e["lo"] + (e["hi"] - e["lo"]) * s) * p[:, e["param2"]]
        p[:, e["param"]] = np.clip(x, e["vmin"], e["vmax"])
    return p


def forward(layers, x, bf16=False):
    """Returns raw outputs and the cache for backward."""
    rb = round_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    a = rb(x)
    acts, zs = [a], []
    for W, b in layers[:-1]:
        z = a @ rb(W).T + b
        zs.append(z)
        a = rb(gelu(z))
        acts.append(a)
    W, b = layers[-1]
    return a @ rb(W).T + b, (acts, zs)


def loss_and_dout(head, out, target):
    """Mean loss over the batch and d(loss)/d(out) (already / batch)."""
    idx, _ = out_index(head)
    n = out.shape[0]
    g = np.zeros_like(out)
    loss = np.zeros(n)
    for e, a in zip(head, idx):
        w = e["weight"]
        if e["kind"] == AXIS:
            v, t = out[:, a:a + 3], target[:, a:a + 3]
            q = np.sum(v * t, 1)
            rr = np.sum(v * v, 1) + 1e-6
            loss += w * (1.0 - q * q / rr)
            g[:, a:a + 3] = (-2.0 * w * q / rr)[:, None] * (t - (q / rr)[:, None] * v)
        else:
            s = sigmoid(out[:, a])
            tt = target[:, a]
            sc = np.clip(s, 1e-7, 1 - 1e-7)
            tc = np.clip(tt, 1e-7, 1 - 1e-7)
            loss += w * (tc * np.log(tc / sc) + (1 - tc) * np.log((1 - tc) / (1 - sc)))
            g[:, a] = w * (s - tt)
    return loss.mean(), g / n


def loss_grad(layers, head, x, params_true, bf16=False):
    """Mean loss and flat gradient (same layout as the weights)."""
    rb = round_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    out, (acts, zs) = forward(layers, x, bf16)
    loss, d = loss_and_dout(head, out, encode(head, params_true))
    d = rb(d)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        grads[l] = (d.T @ acts[l], d.sum(0))
        if l > 0:
            d = rb((d @ rb(W)) * gelu_grad(zs[l - 1]))
    return loss, flatten(grads)
