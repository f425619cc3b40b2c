"""CPU restatement of the surrogate network and of its node encoding.

TEST INFRASTRUCTURE ONLY: the checker for galacticus_amd/csrc/surrogate.hip.
Only the test suite imports this module.

The reference has no learned component; what is restated here is the
network the engine defines (include/gal_surrogate.h): 32 -> 256 -> 256 ->
256 -> 31, SiLU, a weighted mean-square loss on standardised increments plus
a log-error head, and Adam with a cosine learning-rate decay.  Two versions:

* ``forward_bf16`` / ``backward_bf16`` round to bf16 (nearest even) at the
  same points as the GPU kernels (inputs, weights, hidden pre-activations and
  activations, the output gradient, the hidden gradients) and accumulate in
  float32, so GPU and CPU differ only by summation order and the exp of
  SiLU;
* ``forward64`` / ``backward64`` are the exact float64 network, whose
  gradient is pinned against central finite differences.
"""
from __future__ import annotations

import math

import numpy as np

N_IN, N_HID, N_OUT, OUT_LD, HEAD = 32, 256, 31, 32, 30
OUT_CLIP = 20.0   # network outputs enter sinh() clipped to +-20
FAN_IN = (N_IN, N_HID, N_HID, N_HID)
FAN_OUT = (N_HID, N_HID, N_HID, N_OUT)


def param_count() -> int:
    return sum(a * b + b for a, b in zip(FAN_IN, FAN_OUT))


def unpack(p):
    """Parameter vector -> [(W (fan_in, fan_out), b (fan_out))] (views)."""
    out, o = [], 0
    for a, b in zip(FAN_IN, FAN_OUT):
        W = p[o:o + a * b].reshape(a, b)
        o += a * b
        out.append((W, p[o:o + b]))
        o += b
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers])


def bf16(a):
    """Round float32 values to bf16 (nearest, ties to even); returns float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).copy()
    nan = np.isnan(a)
    out[nan] = np.nan
    return out


def bf16_bits(a):
    """bf16 bit patterns (uint16) of float32 values."""
    return (bf16(a).view(np.uint32) >> 16).astype(np.uint16)


def from_bf16_bits(b):
    return (np.asarray(b, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def _silu32(z):
    return (z / (np.float32(1) + np.exp(-z))).astype(np.float32)


def _dsilu32(z):
    s = (np.float32(1) / (np.float32(1) + np.exp(-z))).astype(np.float32)
    return (s * (np.float32(1) + z * (np.float32(1) - s))).astype(np.float32)


def forward_bf16(p, x):
    """x: float32 values already on the bf16 grid (n, 32)."""
    layers = unpack(np.asarray(p, dtype=np.float32))
    cache = {"x": x, "z": [], "h": []}
    a = x
    for l in range(3):
        W, b = layers[l]
        z = (a @ bf16(W)).astype(np.float32) + b
        h = bf16(_silu32(z))
        cache["z"].append(bf16(z))
        cache["h"].append(h)
        a = h
    W, b = layers[3]
    out = (a @ bf16(W)).astype(np.float32) + b
    cache["out"] = out
    return out, cache


def loss_grad(out, target, w, sigma, head_mask, head_w):
    """Loss (mean over rows) and its gradient wrt the 31 outputs (float32)."""
    n = out.shape[0]
    w = np.asarray(w, dtype=np.float32)
    r = (out[:, :30] - target[:, :30]).astype(np.float32)
    wsum = np.float32(w.sum())
    mse = (w * r * r).sum(axis=1)
    hm = np.asarray(head_mask, dtype=bool)
    o_c = np.clip(out[:, :30], -OUT_CLIP, OUT_CLIP).astype(np.float32)
    err = np.abs(np.sinh(o_c) - np.sinh(target[:, :30].astype(np.float32))).astype(np.float32)
    e = (err[:, hm] * np.asarray(sigma, dtype=np.float32)[hm]).max(axis=1)
    e = np.where(target[:, 31] != 0, np.maximum(e, np.float32(1.0)), e)   # clipped samples
    rh = out[:, HEAD] - np.log(e + np.float32(1e-7)).astype(np.float32)
    loss = float((mse / wsum + np.float32(head_w) * rh * rh).sum() / n)
    g = np.zeros((n, OUT_LD), dtype=np.float32)
    g[:, :30] = np.float32(2.0) * w * r / wsum / np.float32(n)
    g[:, HEAD] = np.float32(2.0 * head_w) * rh / np.float32(n)
    return loss, g


def backward_bf16(p, cache, g_out):
    """Gradient of the parameters given dL/d(out) (float32, n x 32)."""
    layers = unpack(np.asarray(p, dtype=np.float32))
    grads = [None] * 4
    d = bf16(g_out)[:, :N_OUT]
    for l in range(3, -1, -1):
        a = cache["x"] if l == 0 else cache["h"][l - 1]
        grads[l] = ((a.T @ d).astype(np.float32), d.sum(axis=0, dtype=np.float32))
        if l == 0:
            break
        W = bf16(layers[l][0])
        d = bf16((d @ W.T).astype(np.float32) * _dsilu32(cache["z"][l - 1]))
    return pack(grads)


def forward64(p, x):
    layers = unpack(np.asarray(p, dtype=np.float64))
    cache = {"x": x, "z": [], "h": []}
    a = x
    for l in range(3):
        W, b = layers[l]
        z = a @ W + b
        cache["z"].append(z)
        a = z / (1.0 + np.exp(-z))
        cache["h"].append(a)
    W, b = layers[3]
    out = a @ W + b
    return out, cache


def residual_error(out, target, sigma, hm):
    """Worst error of the head variables' residuals (float64): targets are
    asinh((r - m)/sigma), so the residual error is sigma |sinh(o) - sinh(t)|."""
    o = np.clip(out[:, :30], -OUT_CLIP, OUT_CLIP)
    err = np.abs(np.sinh(o) - np.sinh(target[:, :30])) * np.asarray(sigma)
    return err[:, hm].max(axis=1)


def network_increment(out, norm):
    """Residual increments (asinh units of the state) from network outputs:
    m + sigma sinh(o), with the outputs clipped to +-OUT_CLIP."""
    o = np.clip(np.asarray(out, dtype=np.float64)[..., :30], -OUT_CLIP, OUT_CLIP)
    return np.asarray(norm, dtype=np.float64)[2, :30] + np.asarray(norm, dtype=np.float64)[3, :30] * np.sinh(o)


def loss64(out, target, w, sigma, head_mask, head_w, e_fixed=None):
    """Float64 loss; the head target ln(e) is a constant of the step (no
    gradient flows through it), so finite differences hold it at e_fixed."""
    n = out.shape[0]
    r = out[:, :30] - target[:, :30]
    mse = (w * r * r).sum(axis=1) / w.sum()
    hm = np.asarray(head_mask, dtype=bool)
    if e_fixed is None:
        e = residual_error(out, target, sigma, hm)
        e = np.where(target[:, 31] != 0, np.maximum(e, 1.0), e)          # clipped samples
    else:
        e = e_fixed
    rh = out[:, HEAD] - np.log(e + 1e-7)
    return float((mse + head_w * rh * rh).sum() / n), e


def backward64(p, cache, out, target, w, sigma, head_mask, head_w):
    layers = unpack(np.asarray(p, dtype=np.float64))
    n = out.shape[0]
    r = out[:, :30] - target[:, :30]
    hm = np.asarray(head_mask, dtype=bool)
    e = residual_error(out, target, sigma, hm)
    e = np.where(target[:, 31] != 0, np.maximum(e, 1.0), e)              # clipped samples
    rh = out[:, HEAD] - np.log(e + 1e-7)
    d = np.zeros((n, N_OUT))
    d[:, :30] = 2.0 * w * r / w.sum() / n
    d[:, HEAD] = 2.0 * head_w * rh / n
    grads = [None] * 4
    for l in range(3, -1, -1):
        a = cache["x"] if l == 0 else cache["h"][l - 1]
        grads[l] = (a.T @ d, d.sum(axis=0))
        if l == 0:
            break
        z = cache["z"][l - 1]
        s = 1.0 / (1.0 + np.exp(-z))
        d = (d @ layers[l][0].T) * (s * (1.0 + z * (1.0 - s)))
    return pack(grads), e


def adam(p, m, v, g, step, lr0, lr_min, beta1, beta2, eps, total_steps):
    """One Adam step (step counts from 0 before the update), float32 state."""
    t = step + 1
    frac = min(t, total_steps) / total_steps
    lr = np.float32(lr_min + 0.5 * (lr0 - lr_min) * (1.0 + math.cos(math.pi * frac)))
    bc1 = np.float32(1.0 - beta1 ** t)
    bc2 = np.float32(1.0 - beta2 ** t)
    b1, b2 = np.float32(beta1), np.float32(beta2)
    m = b1 * m + (np.float32(1) - b1) * g
    v = b2 * v + (np.float32(1) - b2) * g * g
    p = p - lr * (m / bc1) / (np.sqrt(v / bc2) + np.float32(eps))
    return p.astype(np.float32), m.astype(np.float32), v.astype(np.float32)


# ---- node encoding (restates the contract of include/gal_surrogate.h) -------
ROLE_MASS, ROLE_METAL, ROLE_ANGMOM, ROLE_ENERGY, ROLE_BH, ROLE_UNIT, ROLE_CONST, ROLE_SPIN, ROLE_FIRST_ORDER = range(9)
# ROLE_FIRST_ORDER: not fed to the network, advanced by the first-order step
# of the rate at the interval's start (outer radius, stripped gas, bound mass)
VAR_ROLES = [
    ROLE_MASS, ROLE_METAL, ROLE_ANGMOM, ROLE_MASS, ROLE_ANGMOM, ROLE_METAL, ROLE_MASS, ROLE_METAL,  # 0-7
    ROLE_FIRST_ORDER, ROLE_FIRST_ORDER, ROLE_FIRST_ORDER, ROLE_ENERGY,                            # 8-11
    ROLE_MASS, ROLE_MASS, ROLE_CONST, ROLE_METAL, ROLE_MASS, ROLE_METAL, ROLE_ANGMOM,             # 12-18
    ROLE_MASS, ROLE_MASS, ROLE_METAL, ROLE_MASS, ROLE_METAL, ROLE_ANGMOM,                         # 19-24
    ROLE_BH, ROLE_UNIT, ROLE_CONST, ROLE_SPIN, ROLE_FIRST_ORDER,                                  # 25-29
]
_UNTRAINED = (ROLE_CONST, ROLE_SPIN, ROLE_FIRST_ORDER)
INPUT_VARS = [0, 1, 2, 3, 4, 5, 6, 7, 12, 13, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26]
# every state variable but the angular momenta that can change sign, the
# radiated energy, the black-hole radial position and the spin angular momentum
NON_NEGATIVE = frozenset(v for v in range(30) if v not in (2, 4, 11, 27, 28))


def node_frame(mass, r_vir, v_vir, f_baryon, unit_energy):
    unit = f_baryon * mass
    return {ROLE_MASS: 1e-4 * unit, ROLE_METAL: 1e-6 * unit, ROLE_ANGMOM: 1e-4 * unit * r_vir * v_vir,
            ROLE_ENERGY: unit_energy * mass * v_vir * v_vir * 1e-3, ROLE_BH: 1.0, ROLE_UNIT: 1.0}


def _baseline(y0, rate, dt, sc):
    """First-order increment in asinh space: the Euler step taken in y for a
    growing quantity, exponential decay on the time scale y/|f| for a
    decaying positive one, the Euler step in u = asinh(y/s) otherwise."""
    du = np.zeros(30)
    for v in range(30):
        role = VAR_ROLES[v]
        if role == ROLE_ENERGY:
            du[v] = math.asinh(rate[v] * dt / sc[role])
        elif role not in _UNTRAINED:
            s = sc[role]
            if rate[v] >= 0.0:
                du[v] = math.asinh((y0[v] + rate[v] * dt) / s) - math.asinh(y0[v] / s)
            elif y0[v] > 0.0:
                du[v] = math.asinh(y0[v] * math.exp(rate[v] * dt / y0[v]) / s) - math.asinh(y0[v] / s)
            else:
                du[v] = rate[v] * dt / math.sqrt(s * s + y0[v] * y0[v])
    return du


def encode_raw(y0, y1, t0, t1, halo, node_par, f_baryon, unit_energy, v_suppress, a3_reion, rate=None):
    """Raw inputs (32) and raw targets (30) of one node.  halo: dict with
    mass, r_vir, v_vir, t_dyn, a (expansion factor); rate: dy/dt at (t0, y0)
    for the Euler baseline (zero if None)."""
    M, rv, vv = halo["mass"], halo["r_vir"], halo["v_vir"]
    sc = node_frame(M, rv, vv, f_baryon, unit_energy)
    dt = t1 - t0
    x = np.zeros(32)
    for j, v in enumerate(INPUT_VARS):
        x[j] = math.asinh(y0[v] / sc[VAR_ROLES[v]])
    dtp = max(dt, 1e-12)
    sat = node_par[4] != 0.0
    lam = y0[28] / (M * vv * rv)
    x[22] = math.log(dtp)
    x[23] = math.log10(M) - 11.0
    x[24] = 0.1 * t0
    x[25] = (0.0 if sat else node_par[2]) * dt / M
    x[26] = math.log(max(lam, 1e-4))
    x[27] = y0[8] / rv
    x[28] = vv / 200.0
    x[29] = 1.0 if (halo["a"] ** 3 > a3_reion and vv < v_suppress) else 0.0
    x[30] = math.log(dtp / halo["t_dyn"])
    t = np.zeros(30)
    if y1 is not None:
        du = _baseline(y0, np.zeros(30) if rate is None else rate, dt, sc)
        for v in range(30):
            role = VAR_ROLES[v]
            if role == ROLE_ENERGY:
                t[v] = math.asinh((y1[v] - y0[v]) / sc[role]) - du[v]
            elif role not in _UNTRAINED:
                t[v] = math.asinh(y1[v] / sc[role]) - math.asinh(y0[v] / sc[role]) - du[v]
    return x, t


def _sinh(x):
    """sinh with IEEE overflow to +-inf (as the device function)."""
    try:
        return math.sinh(x)
    except OverflowError:
        return math.copysign(math.inf, x)


def decode(y0, d, dt, halo, node_par, f_baryon, unit_energy, rate=None, negative_tolerance=0.0):
    """Inverse of the target map: state at t1 from y0 and raw residuals d.
    A non-negative quantity that decodes into [-negative_tolerance s, 0) is
    set to zero."""
    sc = node_frame(halo["mass"], halo["r_vir"], halo["v_vir"], f_baryon, unit_energy)
    du = _baseline(y0, np.zeros(30) if rate is None else rate, dt, sc)
    y1 = np.array(y0, dtype=np.float64)
    for v in range(30):
        role = VAR_ROLES[v]
        if role == ROLE_CONST:
            continue
        if role == ROLE_SPIN:
            y1[v] = y0[v] + (0.0 if node_par[4] != 0.0 else node_par[3]) * dt
        elif role == ROLE_FIRST_ORDER:
            y1[v] = y0[v] + (0.0 if rate is None else rate[v]) * dt
            if v in NON_NEGATIVE and -negative_tolerance <= y1[v] < 0.0:
                y1[v] = 0.0
        elif role == ROLE_ENERGY:
            y1[v] = y0[v] + sc[role] * _sinh(du[v] + d[v])
        else:
            y1[v] = sc[role] * _sinh(math.asinh(y0[v] / sc[role]) + du[v] + d[v])
            if v in NON_NEGATIVE and -negative_tolerance * sc[role] <= y1[v] < 0.0:
                y1[v] = 0.0
    return y1
