"""FP64 numpy restatement of the learned-guess network of gm_mlp.hip, with the
BF16 rounding points of the device path switchable, for gradient checks.

Flat parameter layout (as gm_mlp_create): for each of the 4 layers the weight
matrix W [in][out] row-major followed by the bias [out].  Dimensions
kin = 64 * ceil(4 nlambda / 64), 256, 256, 256, 64 (11 outputs used).
"""
import numpy as np

HIDDEN, NOUT, NUSED = 256, 64, 11
K_GELU = np.sqrt(2.0 / np.pi)


def bf16(x):
    """Round to bfloat16 (nearest even) and return as float64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(K_GELU * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    u = K_GELU * (x + 0.044715 * x ** 3)
    t = np.tanh(u)
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * K_GELU * (1.0 + 3 * 0.044715 * x * x)


def layout(nlambda):
    kin = ((4 * nlambda + 63) // 64) * 64
    dims = [kin, HIDDEN, HIDDEN, HIDDEN, NOUT]
    out, off = [], 0
    for l in range(4):
        wo = off
        off += dims[l] * dims[l + 1]
        out.append((dims[l], dims[l + 1], wo, off))
        off += dims[l + 1]
    return out, off


def unpack(params, nlambda):
    lay, _ = layout(nlambda)
    return [(params[wo:wo + i * o].reshape(i, o).astype(np.float64),
             params[bo:bo + o].astype(np.float64)) for i, o, wo, bo in lay]


def prep_input(obs, kin):
    """obs [b, 4, nl] -> [b, kin]: divided by max(I[0], I[nl-1]); I -> 2 (I - 1),
    Q, U, V -> 10 x."""
    b, _, nl = obs.shape
    obs = obs.astype(np.float32)
    ic = np.maximum(np.maximum(obs[:, 0, 0], obs[:, 0, nl - 1]), np.float32(1e-6))
    v = (obs / ic[:, None, None]).astype(np.float64)
    v[:, 0] = (v[:, 0] - 1.0) * 2.0
    v[:, 1:] *= 10.0
    x = np.zeros((b, kin))
    x[:, :4 * nl] = v.reshape(b, 4 * nl)
    return x


def prep_target(models):
    m = np.asarray(models, dtype=np.float64)
    g = np.radians(m[:, 5])
    ph = 2 * np.radians(m[:, 6])
    ic = m[:, 7] + m[:, 8]
    t = np.zeros((len(m), NOUT))
    t[:, 0] = (m[:, 1] - 1500.0) / 866.0
    t[:, 1] = np.cos(g)
    t[:, 2] = np.sin(g) * 2 - 1
    t[:, 3] = np.cos(ph)
    t[:, 4] = np.sin(ph)
    t[:, 5] = m[:, 2] / 1.732
    t[:, 6] = (m[:, 3] - 0.04) / 0.01155
    t[:, 7] = (m[:, 4] - 0.255) / 0.1414
    t[:, 8] = (m[:, 0] - 10.5) / 5.485
    t[:, 9] = (m[:, 7] / ic - 0.25) / 0.1
    t[:, 10] = (ic - 1.0) / 0.2
    return t.astype(np.float32).astype(np.float64)


def forward(layers, x, rounding):
    r = bf16 if rounding else (lambda v: v)
    a = r(x)
    zs, acts = [], [a]
    for l, (W, b) in enumerate(layers):
        z = a @ r(W) + b
        if l < 3:
            zs.append(z)
            a = r(gelu(z))
            acts.append(a)
        else:
            y = z
    return y, zs, acts


def loss_and_grads(params, nlambda, obs, models, rounding=True):
    """Mean squared error over the 11 used outputs and its gradient with the
    device path's rounding points (x, W, activations, dY and dZ in BF16)."""
    r = bf16 if rounding else (lambda v: v)
    layers = unpack(params, nlambda)
    kin = layout(nlambda)[0][0][0]
    x = prep_input(obs, kin)
    t = prep_target(models)
    y, zs, acts = forward(layers, x, rounding)
    b = len(x)
    d = np.zeros_like(y)
    d[:, :NUSED] = y[:, :NUSED] - t[:, :NUSED]
    loss = np.sum(d ** 2) / (b * NUSED)
    dcur = r(2.0 * d / (b * NUSED))
    grads = [None] * 4
    for l in range(3, -1, -1):
        W, _ = layers[l]
        grads[l] = (acts[l].T @ dcur, dcur.sum(axis=0))
        if l > 0:
            dcur = r((dcur @ r(W).T) * gelu_grad(zs[l - 1]))
    flat = np.zeros(layout(nlambda)[1])
    for (i, o, wo, bo), (gw, gb) in zip(layout(nlambda)[0], grads):
        flat[wo:wo + i * o] = gw.ravel()
        flat[bo:bo + o] = gb
    return loss, flat


def decode(y, base):
    """Network outputs [b, >=11] -> start models, as k_decode."""
    y = np.asarray(y, dtype=np.float64)
    m = np.tile(np.asarray(base, dtype=np.float64), (len(y), 1))
    sg = np.maximum((y[:, 2] + 1.0) * 0.5, 0.0)
    gm = np.degrees(np.arctan2(sg, y[:, 1]))
    az = 0.5 * np.degrees(np.arctan2(y[:, 4], y[:, 3]))
    az = np.where(az < 0, az + 180.0, az)
    ic = np.maximum(y[:, 10] * 0.2 + 1.0, 0.05)
    s0f = np.clip(y[:, 9] * 0.1 + 0.25, 0.01, 0.99)
    m[:, 0] = np.clip(y[:, 8] * 5.485 + 10.5, 1.0, 2500.0)
    m[:, 1] = np.clip(y[:, 0] * 866.0 + 1500.0, 0.0, 5000.0)
    m[:, 2] = np.clip(y[:, 5] * 1.732, -20.0, 20.0)
    m[:, 3] = np.clip(y[:, 6] * 0.01155 + 0.04, 0.005, 0.6)
    m[:, 4] = np.clip(y[:, 7] * 0.1414 + 0.255, 0.0001, 10.0)
    m[:, 5] = np.clip(gm, 0.0, 180.0)
    m[:, 6] = np.clip(az, 0.0, 180.0)
    m[:, 7] = s0f * ic
    m[:, 8] = (1.0 - s0f) * ic
    return m
