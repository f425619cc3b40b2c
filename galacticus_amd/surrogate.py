"""Host side of the learned node-interval surrogate (include/gal_surrogate.h).

* ``SurrogateMLP`` wraps the bf16 MFMA network of the HIP library: parameters,
  Adam, loss configuration, inference, gradients and hipGraph-replayed
  training steps.
* ``generate_pairs`` evolves node histories interval by interval with the
  FP64 batched integrator and keeps every (state, interval -> state) pair on
  the GPU: the training data and the reference for the surrogate's error.
* ``train_surrogate`` encodes the pairs, fits the normalisation, and trains.
* ``SurrogateEvolver`` is the surrogate+fallback drop-in beside
  ``NodeEvolverBatched.evolve_device`` (same arguments, same statuses).

torch supplies device memory and streams only; all arithmetic runs in the
library's kernels.  Statistics of the encoded data (means and standard
deviations of the network inputs and targets) are reductions done with torch
on the device when the data set is built.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
import time
from typing import Optional

import numpy as np

from . import _lib
from . import workloads as wl
from .evolver import NodeEvolverBatched

N_IN, N_HID, N_OUT, OUT_LD, HEAD = 32, 256, 31, 32, 30
FAN_IN = (N_IN, N_HID, N_HID, N_HID)
FAN_OUT = (N_HID, N_HID, N_HID, N_OUT)
V = wl.V

# Variables whose increments the network learns (loss weight 1); the rest are
# carried over (zero rate) or advanced analytically (spin angular momentum).
INPUT_VARS = [0, 1, 2, 3, 4, 5, 6, 7, 12, 13, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26]
TRAINED_VARS = sorted(INPUT_VARS + [V["HOT_ENERGY_RADIATED"]])
# The error head predicts the worst asinh-space error of these masses
# (stellar, hot gas and black hole); cold-gas errors are reported, and the
# baryon and metal budgets of the gate bound them.
HEAD_VARS = [V["DISK_MASS_STELLAR"], V["SPHEROID_MASS_STELLAR"], V["HOT_MASS"], V["BH_MASS"]]

FLOPS_FORWARD_PER_ROW = 2 * sum(a * b for a, b in zip(FAN_IN, FAN_OUT))
FLOPS_TRAIN_PER_ROW = 3 * FLOPS_FORWARD_PER_ROW


def _ptr(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _stream(stream):
    return ctypes.c_void_p(0 if stream is None else stream.cuda_stream)


def param_count() -> int:
    return sum(a * b + b for a, b in zip(FAN_IN, FAN_OUT))


def init_params(seed: int = 0) -> np.ndarray:
    """Uniform He initialisation (bound sqrt(3/fan_in)); zero biases."""
    rng = np.random.Generator(np.random.PCG64(seed))
    parts = []
    for a, b in zip(FAN_IN, FAN_OUT):
        lim = math.sqrt(3.0 / a)
        parts.append(rng.uniform(-lim, lim, a * b))
        parts.append(np.zeros(b))
    return np.concatenate(parts).astype(np.float32)


@dataclasses.dataclass
class GateConfig:
    """Acceptance rule of a surrogate step (gal_surrogate_gate)."""
    error_max: float = 0.01          # predicted asinh-space error of the head masses
    mass_tolerance: float = 5.0e-3
    metal_tolerance: float = 2.0e-2
    bh_slack: float = 20.0
    negative_tolerance: float = 0.1      # small negatives (in units of the asinh scale) become zero

    def c_struct(self):
        return _lib.Gate(math.log(self.error_max), self.mass_tolerance, self.metal_tolerance, self.bh_slack,
                         self.negative_tolerance)


class SurrogateMLP:
    """The 32-256-256-256-31 SiLU network on the GPU (bf16 weights and
    activations, fp32 accumulation and master weights, Adam)."""

    def __init__(self, batch: int = 8192, params: Optional[np.ndarray] = None, seed: int = 0):
        self.L = _lib.load()
        self.batch = int(batch)
        h = ctypes.c_void_p()
        _lib.check(self.L.gal_mlp_create(self.batch, ctypes.byref(h)), "gal_mlp_create")
        self.handle = h
        self.n_params = int(self.L.gal_mlp_param_count())
        assert self.n_params == param_count()
        self.set_params(init_params(seed) if params is None else params)
        self._dataset = None

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.shape != (self.n_params,):
            raise ValueError(f"expected {self.n_params} parameters")
        _lib.check(self.L.gal_mlp_set_params(self.handle, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                             self.n_params), "gal_mlp_set_params")

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.n_params, dtype=np.float32)
        _lib.check(self.L.gal_mlp_get_params(self.handle, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                             self.n_params), "gal_mlp_get_params")
        return p

    def set_adam(self, lr=1e-3, lr_min=1e-5, beta1=0.9, beta2=0.999, eps=1e-8, total_steps=1):
        _lib.check(self.L.gal_mlp_set_adam(self.handle, lr, lr_min, beta1, beta2, eps, int(total_steps)),
                   "gal_mlp_set_adam")

    def set_loss(self, var_weight, target_sigma, head_vars, head_weight=0.05):
        w = np.ascontiguousarray(var_weight, dtype=np.float32)
        s = np.ascontiguousarray(target_sigma, dtype=np.float32)
        hv = np.ascontiguousarray(head_vars, dtype=np.int32)
        fp = ctypes.POINTER(ctypes.c_float)
        _lib.check(self.L.gal_mlp_set_loss(self.handle, w.ctypes.data_as(fp), s.ctypes.data_as(fp),
                                           hv.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), head_weight),
                   "gal_mlp_set_loss")

    def forward(self, x, out=None, stream=None, layers=False):
        """Network outputs (n x 32 fp32; column 31 stays as allocated, zero).
        layers=True runs the per-layer GEMM path instead of the fused kernel."""
        import torch
        n = int(x.shape[0])
        if out is None:
            out = torch.zeros((n, OUT_LD), dtype=torch.float32, device=x.device)
        fn = self.L.gal_mlp_forward_layers_d if layers else self.L.gal_mlp_forward_d
        _lib.check(fn(self.handle, n, _ptr(x), _ptr(out), _stream(stream)),
                   "gal_mlp_forward_layers_d" if layers else "gal_mlp_forward_d")
        return out

    def gradient(self, x, target, stream=None):
        import torch
        n = int(x.shape[0])
        grad = torch.empty(self.n_params, dtype=torch.float32, device=x.device)
        loss = torch.empty(1, dtype=torch.float32, device=x.device)
        _lib.check(self.L.gal_mlp_gradient_d(self.handle, n, _ptr(x), _ptr(target), _ptr(grad), _ptr(loss),
                                             _stream(stream)), "gal_mlp_gradient_d")
        return grad, loss

    def set_dataset(self, x, target, perm):
        self._dataset = (x, target, perm)  # keep the device buffers alive
        _lib.check(self.L.gal_mlp_set_dataset(self.handle, int(x.shape[0]), _ptr(x), _ptr(target), _ptr(perm)),
                   "gal_mlp_set_dataset")

    def train_steps(self, n_steps, use_graph=True, stream=None):
        import torch
        dev = self._dataset[0].device
        trace = torch.empty(max(int(n_steps), 1), dtype=torch.float32, device=dev)
        _lib.check(self.L.gal_mlp_train_d(self.handle, int(n_steps), 1 if use_graph else 0, _ptr(trace),
                                          _stream(stream)), "gal_mlp_train_d")
        return trace[:n_steps]

    def close(self):
        if getattr(self, "handle", None):
            self.L.gal_mlp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- data -------------------------------------------------------------------
@dataclasses.dataclass
class PairSet:
    """(state, interval -> state) pairs, device tensors, SoA like the evolver."""
    y0: "object"        # (30, n) float64
    y1: "object"
    node_par: "object"  # (GAL_NPARAM_NODE, n)
    t0: "object"
    t1: "object"

    @property
    def n(self):
        return int(self.t0.shape[0])


def generate_tree_pairs(physics, table, max_pairs: int, seed: int = 0, trees_per_forest: int = 1024,
                        log_root_mass=(10.0, 12.0), z_start: float = 5.0, mass_resolution: float = 1e9,
                        tol_abs: float = 1e-6, tol_rel: float = 1e-3, device="cuda:0") -> PairSet:
    """Training pairs from merger-tree trajectories: Parkinson-Cole-Helly
    forests (roots log-uniform in 10^log_root_mass Msun at z=0, the given
    resolution, evolved from z_start in the reference's walk order by the
    FP64 integrator), one forest per seed from `seed` on, until max_pairs
    intervals of non-satellite halos are recorded (the states before and
    after each interval, its times and halo parameters)."""
    import torch
    from . import trees as tr
    cosmo = tr.cosmology(physics)
    t_end = float(wl.scale_factor_to_cosmic_time(1.0))
    t_beg = float(wl.scale_factor_to_cosmic_time(1.0 / (1.0 + z_start)))
    parts = {"y0": [], "y1": [], "np": [], "t0": [], "t1": []}
    total, k = 0, 0
    while total < max_pairs:
        rng = np.random.Generator(np.random.PCG64(seed + k))
        roots = 10.0 ** rng.uniform(log_root_mass[0], log_root_mass[1], trees_per_forest)
        forest = tr.Forest.build_pch(trees_per_forest, roots, t_end, cosmo, tr.pch_params(mass_resolution, t_beg),
                                     seed=seed + k)
        wp = tr.walk_params(t_end)
        sched = tr.Schedule(forest, cosmo, wp)
        ev = NodeEvolverBatched(forest.n_nodes, physics, table, odeToleranceAbsolute=tol_abs,
                                odeToleranceRelative=tol_rel)
        y = torch.from_numpy(forest.initial_state(cosmo, wp)).to(device)
        y0, y1, npar, t0, t1, _ = tr.record_forest(ev, sched, y)
        keep = (npar[4] == 0) & (t1 > t0)
        idx = torch.nonzero(keep).flatten()[: max_pairs - total]
        parts["y0"].append(y0[:, idx])
        parts["y1"].append(y1[:, idx])
        parts["np"].append(npar[:, idx])
        parts["t0"].append(t0[idx])
        parts["t1"].append(t1[idx])
        total += int(idx.numel())
        k += 1
        del ev, y, y0, y1
    cat = lambda key, dim: torch.cat(parts[key], dim=dim).contiguous()
    return PairSet(cat("y0", 1), cat("y1", 1), cat("np", 1), cat("t0", 0), cat("t1", 0))


def generate_pairs(evolver: NodeEvolverBatched, hist: wl.HaloHistories, max_pairs: int, seed: int = 0,
                   device="cuda:0") -> PairSet:
    """Evolve every history interval by interval to z=0 with the FP64
    integrator and record each interval (at most max_pairs, in order)."""
    import torch
    dev = torch.device(device)
    b = hist.batch
    n = b.n
    if n > evolver.n_max:
        raise ValueError("the evolver's n_max is smaller than the number of histories")
    rng = np.random.Generator(np.random.PCG64(seed + 7919))
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    y = d(b.y)
    npar = d(b.node_par)
    t = b.t_start.copy()
    h = d(b.h_step)
    status = torch.empty(n, dtype=torch.int32, device=dev)
    out = {k: [] for k in ("y0", "y1", "np", "t0", "t1")}
    total = 0
    while total < max_pairs:
        active = np.nonzero(t < hist.t_final - 1e-9)[0]
        if active.size == 0:
            break
        t_next = hist.intervals(t, rng)
        if active.size < n:
            idx = torch.from_numpy(active).to(dev)
            ya = y[:, idx].contiguous()
            na = npar[:, idx].contiguous()
            ha = h[idx].contiguous()
        else:
            idx, ya, na, ha = None, y, npar, h
        t0 = d(t[active])
        t1 = d(t_next[active])
        y_start = ya.clone()
        evolver.evolve_device(ya, na, t0, t1, ha, status=status[:active.size])
        if idx is not None:
            y[:, idx] = ya
            h[idx] = ha
        keep = min(active.size, max_pairs - total)
        out["y0"].append(y_start[:, :keep])
        out["y1"].append(ya[:, :keep].clone())
        out["np"].append(na[:, :keep].clone())
        out["t0"].append(t0[:keep])
        out["t1"].append(t1[:keep])
        total += keep
        t[active] = t_next[active]
    cat = lambda k, dim: torch.cat(out[k], dim=dim).contiguous()
    return PairSet(cat("y0", 1), cat("y1", 1), cat("np", 1), cat("t0", 0), cat("t1", 0))


def encode(evolver: NodeEvolverBatched, pairs: PairSet, norm=None, raw=False, with_target=True, stream=None):
    """Encode pairs: (x bf16 as int16 tensor or None, x_raw or None, target)."""
    import torch
    n, dev = pairs.n, pairs.t0.device
    x = None if raw else torch.empty((n, N_IN), dtype=torch.int16, device=dev)
    x_raw = torch.empty((n, N_IN), dtype=torch.float32, device=dev) if raw else None
    tg = torch.empty((n, OUT_LD), dtype=torch.float32, device=dev) if with_target else None
    _lib.check(evolver.L.gal_surrogate_encode_d(evolver.handle, n, n, _ptr(pairs.y0),
                                                _ptr(pairs.y1 if with_target else None), _ptr(pairs.node_par),
                                                _ptr(pairs.t0), _ptr(pairs.t1), _ptr(norm), _ptr(x), _ptr(x_raw),
                                                _ptr(tg), _stream(stream)), "gal_surrogate_encode_d")
    return x, x_raw, tg


def fit_normalization(x_raw, t_raw, sample=2_000_000):
    """Normalisation (4 x 32 fp32) from at most `sample` rows: input means
    and standard deviations; for each residual target its median m and a
    robust scale sigma, 1.4826 x the median absolute deviation of the rows
    that deviate from m at all (a component that is empty in most rows still
    gets the scale of its active rows).  The network learns
    asinh((r - m) / sigma): unit scale for typical rows, logarithmic tails."""
    import torch
    xs = x_raw[:sample].double()
    ts = t_raw[:sample].double()
    norm = torch.zeros((4, 32), dtype=torch.float32, device=x_raw.device)
    norm[0] = xs.mean(0).float()
    norm[1] = xs.std(0).float()
    norm[1] = torch.where(norm[1] > 1e-6, norm[1], torch.ones_like(norm[1]))
    med = ts.median(dim=0).values
    dev = (ts - med).abs()
    for v in range(32):
        if v not in TRAINED_VARS:
            norm[2, v], norm[3, v] = 0.0, 1.0
            continue
        active = dev[:, v][dev[:, v] > 1e-9]
        norm[2, v] = float(med[v])
        norm[3, v] = max(1.4826 * float(active.median()), 1e-6) if active.numel() else 1.0
    return norm.contiguous()


def clip_targets(target, clip=16.0):
    """Clip targets to +-clip (asinh units: residuals beyond sigma sinh(16) =
    4.4e6 sigma) and flag clipped rows in column 31 (their error counts as
    large for the error head)."""
    import torch
    big = (target[:, :30].abs() > clip).any(dim=1)
    target[:, :30].clamp_(-clip, clip)
    target[:, 31] = big.to(target.dtype)
    return float(big.float().mean())


@dataclasses.dataclass
class TrainReport:
    epochs: int
    steps: int
    batch: int
    seconds: float
    samples_per_s: float
    tflops: float
    loss_first: float
    loss_last: float
    losses: list
    clipped_fraction: float = 0.0


def train_surrogate(evolver: NodeEvolverBatched, pairs: PairSet, epochs: int = 20, batch: int = 8192,
                    lr: float = 2e-3, lr_min: float = 2e-5, seed: int = 0, head_weight: float = 0.05,
                    use_graph: bool = True, mlp: Optional[SurrogateMLP] = None):
    """Fit the network to the pairs; returns (mlp, norm, TrainReport)."""
    import torch
    dev = pairs.t0.device
    _, x_raw, t_raw = encode(evolver, pairs, raw=True)
    norm = fit_normalization(x_raw, t_raw)
    del x_raw, t_raw
    x, _, target = encode(evolver, pairs, norm=norm)
    clipped = clip_targets(target)
    n = pairs.n
    steps_per_epoch = n // batch
    if steps_per_epoch < 1:
        raise ValueError("fewer pairs than one batch")
    mlp = SurrogateMLP(batch, seed=seed) if mlp is None else mlp
    mlp.set_adam(lr=lr, lr_min=lr_min, total_steps=epochs * steps_per_epoch)
    w = np.zeros(30, dtype=np.float32)
    w[TRAINED_VARS] = 1.0
    hv = np.zeros(30, dtype=np.int32)
    hv[HEAD_VARS] = 1
    mlp.set_loss(w, norm[3, :30].cpu().numpy(), hv, head_weight)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    losses = []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    perm = torch.empty(n, dtype=torch.int32, device=dev)
    mlp.set_dataset(x, target, perm)   # fixed buffers: one captured step graph serves every epoch
    for _ in range(epochs):
        perm.copy_(torch.randperm(n, device=dev, generator=gen))
        losses.append(mlp.train_steps(steps_per_epoch, use_graph=use_graph))
    torch.cuda.synchronize()
    secs = time.perf_counter() - t_start
    trace = torch.cat(losses).cpu().numpy()
    steps = epochs * steps_per_epoch
    rep = TrainReport(epochs, steps, batch, secs, steps * batch / secs, steps * batch * FLOPS_TRAIN_PER_ROW / secs / 1e12,
                      float(trace[:10].mean()), float(trace[-10:].mean()),
                      [float(v) for v in trace[::max(1, steps // 50)]], clipped)
    return mlp, norm, rep


class SurrogateEvolver:
    """Surrogate + FP64 fallback with the arguments of
    NodeEvolverBatched.evolve_device; the integrator handle supplies physics,
    tolerances and the fallback."""

    def __init__(self, evolver: NodeEvolverBatched, mlp: SurrogateMLP, norm, gate: Optional[GateConfig] = None):
        self.evolver, self.mlp, self.norm = evolver, mlp, norm
        self.gate = gate or GateConfig()

    def evolve_device(self, y, node_par, t_start, t_end, h_step, scale=None, status=None, accepted=None,
                      stream=None, count_rejected=True):
        """Returns (status, accepted, n_rejected).  With count_rejected=False
        the call only enqueues work on the stream and n_rejected is None
        (count ``accepted == 0`` once the stream is done)."""
        import torch
        n = int(t_start.shape[0])
        ld = int(y.shape[1])
        if status is None:
            status = torch.empty(n, dtype=torch.int32, device=y.device)
        if accepted is None:
            accepted = torch.empty(n, dtype=torch.int32, device=y.device)
        if scale is None:
            scale = torch.empty_like(y)
        n_rej = ctypes.c_int(0)
        g = self.gate.c_struct()
        _lib.check(self.evolver.L.gal_surrogate_evolve_d(
            self.evolver.handle, self.mlp.handle, _ptr(self.norm), ctypes.byref(g), n, ld, _ptr(y), _ptr(node_par),
            _ptr(t_start), _ptr(t_end), _ptr(h_step), _ptr(scale), _ptr(status), _ptr(accepted),
            ctypes.byref(n_rej) if count_rejected else None,
            _stream(stream)), "gal_surrogate_evolve_d")
        return status, accepted, (n_rej.value if count_rejected else None)
