"""Host-side mirror of the reference's node-evolver interface, batched.

The reference evolves one node per call through
``mergerTreeNodeEvolverStandard%evolve`` (source/merger_trees/node_evolver/
standard.F90:385-755), configured by the parameter-XML keys of
``<mergerTreeNodeEvolver value="standard">`` (standard.F90:165-256).  The
batched evolver keeps those keys, their defaults and the per-node status codes,
and evolves a whole wave of nodes per call on the GPU through the C ABI of
include/gal_ode.h.  Device memory comes from torch tensors (plumbing only);
all arithmetic is in the HIP library.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import _lib
from . import params as gp
from .cooling import CoolingTable, synthetic_cie_table

STATUS_SUCCESS = 0
STATUS_FAILURE = -1
STATUS_INTERRUPT = 9
STATUS_UNDERFLOW = 15
STATUS_XCPU = 1025


def _ptr(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def values_agree(value1, value2, abs_tol=None, rel_tol=None):
    """Values_Agree_Double (source/numerical/comparison.F90:129-155)."""
    v1 = np.asarray(value1, dtype=np.float64)
    v2 = np.asarray(value2, dtype=np.float64)
    diff = np.abs(v1 - v2)
    out = np.zeros(np.broadcast(v1, v2).shape, dtype=bool)
    if abs_tol is not None:
        out |= diff <= abs_tol
    if rel_tol is not None:
        out |= diff <= 0.5 * np.abs(v1 + v2) * rel_tol
    out |= v1 == v2
    return out


class NodeEvolverBatched:
    """Batched ``mergerTreeNodeEvolver`` for the standard node ODE system.

    Parameters mirror the reference's XML keys: ``odeToleranceAbsolute``,
    ``odeToleranceRelative``, ``reuseODEStepSize``, ``enforceNonNegativity``;
    ``odeAlgorithm`` must be ``rungeKuttaCashKarp`` (the reference default).
    """

    def __init__(self, n_max: int, physics: Optional[np.ndarray] = None, table: Optional[CoolingTable] = None,
                 odeToleranceAbsolute: float = gp.EVOLVER_DEFAULTS["odeToleranceAbsolute"],
                 odeToleranceRelative: float = gp.EVOLVER_DEFAULTS["odeToleranceRelative"],
                 reuseODEStepSize: bool = True, enforceNonNegativity: bool = False,
                 scales_provided: bool = False, odeAlgorithm: str = "rungeKuttaCashKarp",
                 cost_order: bool = True):
        if odeAlgorithm != "rungeKuttaCashKarp":
            raise ValueError("odeAlgorithm must be rungeKuttaCashKarp for the batched evolver")
        self.L = _lib.load()
        self.n_max = int(n_max)
        self.odeToleranceAbsolute = float(odeToleranceAbsolute)
        self.odeToleranceRelative = float(odeToleranceRelative)
        self.flags = gp.solver_flags(dict(reuseODEStepSize=reuseODEStepSize,
                                          enforceNonNegativity=enforceNonNegativity), scales_provided,
                                        cost_order)
        h = ctypes.c_void_p()
        _lib.check(self.L.gal_ode_batch_create(self.n_max, self.odeToleranceAbsolute, self.odeToleranceRelative,
                                               self.flags, ctypes.byref(h)), "gal_ode_batch_create")
        self.handle = h
        self.physics = gp.physics_params() if physics is None else np.asarray(physics, dtype=np.float64)
        self.set_physics(self.physics)
        self.table = synthetic_cie_table() if table is None else table
        self.set_cooling_table(self.table)

    @classmethod
    def from_parameter_xml(cls, path_or_text: str, n_max: int, table: Optional[CoolingTable] = None):
        physics, ev = gp.parse_parameter_xml(path_or_text)
        return cls(n_max, physics, table, odeToleranceAbsolute=ev["odeToleranceAbsolute"],
                   odeToleranceRelative=ev["odeToleranceRelative"], reuseODEStepSize=ev["reuseODEStepSize"],
                   enforceNonNegativity=ev["enforceNonNegativity"], odeAlgorithm=ev["odeAlgorithm"])

    # -- configuration ---------------------------------------------------
    def set_physics(self, physics: np.ndarray):
        p = np.ascontiguousarray(physics, dtype=np.float64)
        if p.shape != (gp.NPARAM,):
            raise ValueError(f"physics parameters must have shape ({gp.NPARAM},)")
        _lib.check(self.L.gal_ode_batch_set_physics(self.handle, p.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                    gp.NPARAM), "gal_ode_batch_set_physics")
        self.physics = p

    def set_cooling_table(self, t: CoolingTable):
        arr = np.ascontiguousarray(t.ln_lambda, dtype=np.float64)
        _lib.check(self.L.gal_ode_batch_set_cooling_table(self.handle, arr.ctypes.data_as(
            ctypes.POINTER(ctypes.c_double)), t.n_t, t.ln_t_min, t.d_ln_t, t.n_z, t.ln_z_min, t.d_ln_z),
            "gal_ode_batch_set_cooling_table")
        self.table = t

    def set_launch(self, nodes_per_block: int, blocks_per_cu: int, rhs_waves: Optional[int] = None):
        """Launch shape: nodes per workgroup, workgroups per CU and, if given,
        waves per group of nodes (2 splits every RHS over two SIMDs)."""
        _lib.check(self.L.gal_ode_batch_set_launch(self.handle, nodes_per_block, blocks_per_cu),
                   "gal_ode_batch_set_launch")
        if rhs_waves is not None:
            _lib.check(self.L.gal_ode_batch_set_rhs_waves(self.handle, int(rhs_waves)), "gal_ode_batch_set_rhs_waves")

    def set_wall_limit(self, seconds: float):
        """Wall-time limit of each later evolve, in seconds from the start of
        its launch (<= 0: none); nodes not finished by then return status
        1025 (errorStatusXCPU) with their state as it came in."""
        _lib.check(self.L.gal_ode_batch_set_wall_limit(self.handle, float(seconds)), "gal_ode_batch_set_wall_limit")

    def set_graph_stall_limit(self, seconds: float):
        """Operation-graph executor of forests: lanes give up, and the call
        fails, when no operation has become ready for this long (<= 0: the
        default, 60 s)."""
        _lib.check(self.L.gal_ode_batch_set_graph_stall_limit(self.handle, float(seconds)),
                   "gal_ode_batch_set_graph_stall_limit")

    def launch(self):
        """(nodes per workgroup, workgroups per CU, waves per group of nodes)."""
        bn, bb, bw = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.check(self.L.gal_ode_batch_get_launch(self.handle, ctypes.byref(bn), ctypes.byref(bb), ctypes.byref(bw)),
                   "gal_ode_batch_get_launch")
        return bn.value, bb.value, bw.value

    # -- evolution ---------------------------------------------------------
    @staticmethod
    def _check_shapes(n, y, node_par, *per_node, scale=None):
        """The kernels address y[v*ld + i] and node_par[k*ld + i] with ld =
        y.shape[1]: reject arrays that would put those addresses out of bounds."""
        ld = int(y.shape[1])
        if tuple(y.shape) != (gp.NVAR, ld) or n > ld:
            raise ValueError(f"y must have shape ({gp.NVAR}, ld >= n)")
        if tuple(node_par.shape) != (gp.NPARAM_NODE, ld):
            raise ValueError(f"node_par must have shape ({gp.NPARAM_NODE}, {ld})")
        if scale is not None and tuple(scale.shape) != (gp.NVAR, ld):
            raise ValueError(f"scale must have shape ({gp.NVAR}, {ld})")
        for a in per_node:
            if a is not None and int(a.shape[0]) < n:
                raise ValueError(f"per-node arrays need at least {n} entries")

    def evolve_device(self, y, node_par, t_start, t_end, h_step, scale=None, status=None, counters=None,
                      stream=None, planned=False):
        """Evolve in place on device tensors (float64 SoA, y: (30, ld)).

        planned=True launches the evolve in the claim order that plan_device
        computed for the same y (gal_ode_batch_evolve_planned_d)."""
        import torch
        n = int(t_start.shape[0])
        ld = int(y.shape[1])
        self._check_shapes(n, y, node_par, t_end, h_step, status, scale=scale)
        if counters is not None and int(counters.numel()) < 3 * n:
            raise ValueError(f"counters need {3 * n} entries")
        if status is None:
            status = torch.empty(n, dtype=torch.int32, device=y.device)
        if scale is None:
            scale = torch.empty_like(y)
        s = ctypes.c_void_p(0 if stream is None else stream.cuda_stream)
        fn = self.L.gal_ode_batch_evolve_planned_d if planned else self.L.gal_ode_batch_evolve_d
        _lib.check(fn(self.handle, n, ld, _ptr(y), _ptr(node_par), _ptr(t_start), _ptr(t_end), _ptr(h_step),
                      _ptr(scale), _ptr(status), _ptr(counters), s),
                   "gal_ode_batch_evolve_planned_d" if planned else "gal_ode_batch_evolve_d")
        return status

    def plan_device(self, y, node_par, t_start, t_end, scale=None, stream=None):
        """Order the nodes of the next evolve_device(..., planned=True) of the
        same y for claiming (gal_ode_batch_plan_d), without launching it."""
        n = int(t_start.shape[0])
        ld = int(y.shape[1])
        self._check_shapes(n, y, node_par, t_end, scale=scale)
        s = ctypes.c_void_p(0 if stream is None else stream.cuda_stream)
        _lib.check(self.L.gal_ode_batch_plan_d(self.handle, n, ld, _ptr(y), _ptr(node_par), _ptr(t_start),
                                               _ptr(t_end), _ptr(scale), s), "gal_ode_batch_plan_d")

    def evolve(self, y, node_par, t_start, t_end, h_step, scale=None):
        """Host arrays in, host arrays out (staged through HBM by the library).

        Returns (y, h_step, scale, status, counters[3, n])."""
        y = np.array(y, dtype=np.float64, order="C", copy=True)
        n = y.shape[1]
        h = np.array(h_step, dtype=np.float64, copy=True)
        sc = np.zeros_like(y) if scale is None else np.array(scale, dtype=np.float64, order="C", copy=True)
        st = np.zeros(n, dtype=np.int32)
        cnt = np.zeros(3 * n, dtype=np.int64)
        dp = ctypes.POINTER(ctypes.c_double)
        npar = np.ascontiguousarray(node_par, dtype=np.float64)
        ts = np.ascontiguousarray(t_start, dtype=np.float64)
        te = np.ascontiguousarray(t_end, dtype=np.float64)
        self._check_shapes(n, y, npar, ts, te, h, scale=sc)
        _lib.check(self.L.gal_ode_batch_evolve_h(self.handle, n, y.ctypes.data_as(dp), npar.ctypes.data_as(dp),
                                                 ts.ctypes.data_as(dp), te.ctypes.data_as(dp),
                                                 h.ctypes.data_as(dp), sc.ctypes.data_as(dp),
                                                 st.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                                 ctypes.c_void_p(cnt.ctypes.data)), "gal_ode_batch_evolve_h")
        return y, h, sc, st, cnt.reshape(3, n)

    def rhs_device(self, t, y, node_par, dydt=None, stream=None):
        import torch
        if dydt is None:
            dydt = torch.empty_like(y)
        self._check_shapes(int(t.shape[0]), y, node_par, scale=dydt)
        s = ctypes.c_void_p(0 if stream is None else stream.cuda_stream)
        _lib.check(self.L.gal_ode_rhs_d(self.handle, int(t.shape[0]), int(y.shape[1]), _ptr(t), _ptr(y),
                                        _ptr(node_par), _ptr(dydt), s), "gal_ode_rhs_d")
        return dydt

    def autotune(self, y, node_par, t_start, t_end, h_step, stream=None):
        """Time the launch shapes on copies of the inputs; keeps the fastest."""
        import torch
        n = int(t_start.shape[0])
        self._check_shapes(n, y, node_par, t_end, h_step)
        bn, bb = ctypes.c_int(), ctypes.c_int()
        scale = torch.empty_like(y)
        status = torch.empty(n, dtype=torch.int32, device=y.device)
        s = ctypes.c_void_p(0 if stream is None else stream.cuda_stream)
        _lib.check(self.L.gal_ode_batch_autotune(self.handle, n, int(y.shape[1]), _ptr(y), _ptr(node_par),
                                                 _ptr(t_start), _ptr(t_end), _ptr(h_step), _ptr(scale),
                                                 _ptr(status), s, ctypes.byref(bn), ctypes.byref(bb)),
                   "gal_ode_batch_autotune")
        return self.launch()

    def isAccurate(self, valueNode, valueExpected):
        """standardIsAccurate (node_evolver/standard.F90:757-768)."""
        return values_agree(valueNode, valueExpected, self.odeToleranceAbsolute, self.odeToleranceRelative)

    def close(self):
        if getattr(self, "handle", None):
            self.L.gal_ode_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def hadjust_device(y, yerr, yp, scale, nonneg, h, ord_=4, eps_abs=1e-6, eps_rel=1e-3, a_y=1.0, a_dydt=0.0):
    """Apply the scaled controller to n cases on the GPU (SoA tensors (dim, n))."""
    import torch
    L = _lib.load()
    dim, n = int(y.shape[0]), int(y.shape[1])
    result = torch.empty(n, dtype=torch.int32, device=y.device)
    _lib.check(L.gal_ode_hadjust_d(n, dim, ord_, eps_abs, eps_rel, a_y, a_dydt, _ptr(y), _ptr(yerr), _ptr(yp),
                                   _ptr(scale), _ptr(nonneg), _ptr(h), _ptr(result), ctypes.c_void_p(0)),
               "gal_ode_hadjust_d")
    return result
