"""Host side of the tree path (include/gal_tree.h): forests, the reference's
evolution order as a schedule, and its execution on the GPU.

* ``Forest.build_pch`` builds merger trees with the Parkinson, Cole & Helly
  (2008) rates in the Cole et al. (2000) construction
  (merger_trees/construct/builder/Cole2000/_class.F90:360-915).
* ``Forest.from_xml`` reads the reference's fully-specified tree format
  (merger_trees/construct/fully_specified.F90:269-450), e.g.
  testSuite/parameters/reproducibility/closedBoxTree.xml.
* ``Schedule`` runs the reference's tree walk (merger_trees/evolver/standard.F90:291-635)
  on the host and records every node-evolution interval, promotion, halo
  merger and galaxy merger in the reference's order, with its wave.
* ``evolve_forest`` runs a schedule on the GPU.
"""
from __future__ import annotations

import ctypes
import dataclasses
import xml.etree.ElementTree as ET
from typing import Optional

import numpy as np

from . import _lib
from . import params as gp
from . import workloads as wl

OP_EVOLVE, OP_PROMOTE, OP_BECOME_SATELLITE, OP_MERGE_GALAXY = 0, 1, 2, 3
OP_NAMES = {OP_EVOLVE: "evolve", OP_PROMOTE: "promote", OP_BECOME_SATELLITE: "satellite",
            OP_MERGE_GALAXY: "merge"}


class Cosmology(ctypes.Structure):
    _fields_ = [("omega_matter", ctypes.c_double), ("omega_dark_energy", ctypes.c_double),
                ("omega_baryon", ctypes.c_double), ("hubble_constant", ctypes.c_double),
                ("sigma_8", ctypes.c_double), ("index_spectral", ctypes.c_double),
                ("temperature_cmb", ctypes.c_double)]


class PCHParams(ctypes.Structure):
    _fields_ = [("mass_resolution", ctypes.c_double), ("merge_probability", ctypes.c_double),
                ("accretion_limit", ctypes.c_double), ("g0", ctypes.c_double), ("gamma1", ctypes.c_double),
                ("gamma2", ctypes.c_double), ("epsilon1", ctypes.c_double), ("time_earliest", ctypes.c_double)]


class WalkParams(ctypes.Structure):
    _fields_ = [("time_end", ctypes.c_double), ("timestep_relative", ctypes.c_double),
                ("timestep_absolute", ctypes.c_double), ("timestep_host_relative", ctypes.c_double),
                ("timestep_host_absolute", ctypes.c_double), ("satellite_offset_absolute", ctypes.c_double),
                ("satellite_offset_relative", ctypes.c_double), ("merging_multiplier", ctypes.c_double),
                ("orbit_radius_ratio", ctypes.c_double), ("orbit_circularity", ctypes.c_double),
                ("spin_parameter", ctypes.c_double)]


class ScheduleOp(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("a", ctypes.c_int32), ("b", ctypes.c_int32), ("wave", ctypes.c_int32),
                ("t0", ctypes.c_double), ("t1", ctypes.c_double), ("node_par", ctypes.c_double * gp.NPARAM_NODE)]


OP_DTYPE = np.dtype([("kind", np.int32), ("a", np.int32), ("b", np.int32), ("wave", np.int32), ("t0", np.float64),
                     ("t1", np.float64), ("node_par", np.float64, (gp.NPARAM_NODE,))])
assert OP_DTYPE.itemsize == ctypes.sizeof(ScheduleOp)


def cosmology(physics=None, sigma_8=0.807, index_spectral=0.961, temperature_cmb=2.7255) -> Cosmology:
    """The builder's cosmology from the engine's physics parameters
    (power spectrum values of parameters/quickTest.xml:50-57)."""
    p = gp.physics_params() if physics is None else physics
    return Cosmology(p[gp.P["OmegaMatter"]], p[gp.P["OmegaDarkEnergy"]], p[gp.P["OmegaBaryon"]],
                     p[gp.P["HubbleConstant"]], sigma_8, index_spectral, temperature_cmb)


def pch_params(mass_resolution=1e9, time_earliest=0.5, merge_probability=0.1, accretion_limit=0.1,
               g0=0.57, gamma1=0.38, gamma2=-0.01, epsilon1=0.1) -> PCHParams:
    """Cole2000 and PCH defaults (Cole2000/_class.F90:187-260; PCH 2008)."""
    return PCHParams(mass_resolution, merge_probability, accretion_limit, g0, gamma1, gamma2, epsilon1,
                     time_earliest)


def walk_params(time_end, timestep_relative=0.1, timestep_absolute=1.0, timestep_host_relative=0.1,
                timestep_host_absolute=1.0, satellite_offset_absolute=1e-2, satellite_offset_relative=1e-3,
                merging_multiplier=0.75, orbit_radius_ratio=0.6, orbit_circularity=0.5,
                spin_parameter=0.035) -> WalkParams:
    """Defaults: simple timestep (evolve/timesteps/simple.F90:40-80), host
    timesteps (parameters/quickTest.xml:297-300), satellite offsets
    (evolve/timesteps/standard.F90:113), Jiang (2008) multiplier 0.75
    (parameters/quickTest.xml:207-210); the orbit is fixed (radius ratio and
    circularity) and every halo has the same spin parameter."""
    return WalkParams(time_end, timestep_relative, timestep_absolute, timestep_host_relative, timestep_host_absolute,
                      satellite_offset_absolute, satellite_offset_relative, merging_multiplier, orbit_radius_ratio,
                      orbit_circularity, spin_parameter)


def node_events(kind, node, target, y, h_step, major_ratio: float = 0.25):
    """Apply promote/merge events (GAL_OP_PROMOTE, GAL_OP_MERGE_GALAXY) to
    host arrays y [30, n] and h_step [n] in place on the GPU; every node may
    appear in at most one event."""
    L = _bind(_lib.load())
    kind = np.ascontiguousarray(kind, dtype=np.int32)
    node = np.ascontiguousarray(node, dtype=np.int32)
    target = np.ascontiguousarray(target, dtype=np.int32)
    if not (y.flags.c_contiguous and y.dtype == np.float64 and h_step.flags.c_contiguous and h_step.dtype == np.float64):
        raise ValueError("y and h_step must be C-contiguous float64 arrays")
    if y.shape != (gp.NVAR, h_step.shape[0]) or not (len(kind) == len(node) == len(target)):
        raise ValueError("inconsistent event or state shapes")
    i32p, dp = ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double)
    _lib.check(L.gal_node_events_h(len(kind), kind.ctypes.data_as(i32p), node.ctypes.data_as(i32p),
                                   target.ctypes.data_as(i32p), h_step.shape[0], y.ctypes.data_as(dp),
                                   h_step.ctypes.data_as(dp), float(major_ratio)), "gal_node_events_h")


def _bind(L):
    if getattr(L, "_trees_bound", False):
        return L
    vp, c_int, i64, dp, ip = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.POINTER(ctypes.c_double), \
        ctypes.POINTER(ctypes.c_int32)
    L.gal_forest_build_pch.argtypes = [ctypes.POINTER(Cosmology), ctypes.POINTER(PCHParams), c_int, dp,
                                       ctypes.c_double, ctypes.c_uint64, ctypes.POINTER(vp)]
    L.gal_forest_from_arrays.argtypes = [i64, c_int, ctypes.POINTER(i64), ip, ip, ip, dp, dp, ctypes.POINTER(vp)]
    L.gal_forest_size.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(c_int)]
    L.gal_forest_arrays.argtypes = [vp, ctypes.POINTER(i64), ip, ip, ip, dp, dp]
    L.gal_forest_destroy.argtypes = [vp]
    L.gal_cosmology_sigma.argtypes = [ctypes.POINTER(Cosmology), c_int, dp, dp]
    L.gal_cosmology_delta_critical.argtypes = [ctypes.POINTER(Cosmology), c_int, dp, dp]
    L.gal_schedule_build.argtypes = [vp, ctypes.POINTER(Cosmology), ctypes.POINTER(WalkParams), ctypes.POINTER(vp)]
    L.gal_schedule_size.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(c_int)]
    L.gal_schedule_ops.argtypes = [vp, vp]
    L.gal_schedule_node_gyr.argtypes = [vp, dp]
    L.gal_schedule_destroy.argtypes = [vp]
    L.gal_forest_initial_state.argtypes = [vp, ctypes.POINTER(Cosmology), ctypes.POINTER(WalkParams),
                                           ctypes.c_double, ctypes.c_double, c_int, dp]
    L.gal_tree_evolve_d.argtypes = [vp, vp, c_int, vp, vp, vp, ctypes.c_double, vp]
    L.gal_tree_evolve_waves_d.argtypes = [vp, vp, c_int, vp, vp, vp, ctypes.c_double, vp]
    L.gal_node_events_h.argtypes = [c_int, ip, ip, ip, c_int, dp, dp, ctypes.c_double]
    L.gal_tree_evolve_surrogate_d.argtypes = [vp, vp, vp, ctypes.POINTER(_lib.Gate), vp, c_int, vp, vp, vp,
                                              ctypes.c_double, ctypes.POINTER(ctypes.c_int64), vp]
    L.gal_tree_evolve_record_d.argtypes = [vp, vp, c_int, vp, vp, vp, ctypes.c_double, vp, vp, vp, vp, vp,
                                           ctypes.c_int64, ctypes.POINTER(ctypes.c_int64), vp]
    for name in ("gal_forest_build_pch", "gal_forest_from_arrays", "gal_forest_size", "gal_forest_arrays",
                 "gal_forest_destroy", "gal_cosmology_sigma", "gal_cosmology_delta_critical", "gal_schedule_build",
                 "gal_schedule_size", "gal_schedule_ops", "gal_schedule_node_gyr", "gal_schedule_destroy",
                 "gal_forest_initial_state",
                 "gal_tree_evolve_d", "gal_tree_evolve_waves_d", "gal_tree_evolve_surrogate_d", "gal_tree_evolve_record_d",
                 "gal_node_events_h"):
        getattr(L, name).restype = c_int
    L._trees_bound = True
    return L


def _leaves(el, path):
    kids = list(el)
    if not kids:
        return {path: float(el.text)}
    out = {}
    for k in kids:
        out.update(_leaves(k, path + "/" + k.tag))
    return out


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def sigma(cosmo: Cosmology, mass) -> np.ndarray:
    L = _bind(_lib.load())
    m = np.ascontiguousarray(mass, dtype=np.float64).ravel()
    out = np.zeros_like(m)
    _lib.check(L.gal_cosmology_sigma(ctypes.byref(cosmo), m.size, _dp(m), _dp(out)), "gal_cosmology_sigma")
    return out


def delta_critical(cosmo: Cosmology, time) -> np.ndarray:
    L = _bind(_lib.load())
    t = np.ascontiguousarray(time, dtype=np.float64).ravel()
    out = np.zeros_like(t)
    _lib.check(L.gal_cosmology_delta_critical(ctypes.byref(cosmo), t.size, _dp(t), _dp(out)),
               "gal_cosmology_delta_critical")
    return out


class Forest:
    """A set of merger trees in flat arrays (global node indices)."""

    def __init__(self, handle):
        self.L = _bind(_lib.load())
        self.handle = handle
        n, nt = ctypes.c_int64(), ctypes.c_int()
        self.L.gal_forest_size(handle, ctypes.byref(n), ctypes.byref(nt))
        self.n_nodes, self.n_trees = int(n.value), int(nt.value)
        self.tree_offset = np.zeros(self.n_trees + 1, dtype=np.int64)
        self.parent = np.zeros(self.n_nodes, dtype=np.int32)
        self.first_child = np.zeros(self.n_nodes, dtype=np.int32)
        self.sibling = np.zeros(self.n_nodes, dtype=np.int32)
        self.mass = np.zeros(self.n_nodes)
        self.time = np.zeros(self.n_nodes)
        self.L.gal_forest_arrays(handle, self.tree_offset.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                 _ip(self.parent), _ip(self.first_child), _ip(self.sibling), _dp(self.mass),
                                 _dp(self.time))

    @classmethod
    def build_pch(cls, n_trees, root_mass, time_root, cosmo: Optional[Cosmology] = None,
                  pch: Optional[PCHParams] = None, seed=0):
        L = _bind(_lib.load())
        cosmo = cosmology() if cosmo is None else cosmo
        pch = pch_params() if pch is None else pch
        m = np.ascontiguousarray(np.broadcast_to(root_mass, (n_trees,)), dtype=np.float64)
        h = ctypes.c_void_p()
        _lib.check(L.gal_forest_build_pch(ctypes.byref(cosmo), ctypes.byref(pch), int(n_trees), _dp(m),
                                          float(time_root), int(seed), ctypes.byref(h)), "gal_forest_build_pch")
        return cls(h)

    @classmethod
    def from_arrays(cls, tree_offset, parent, first_child, sibling, mass, time):
        L = _bind(_lib.load())
        to = np.ascontiguousarray(tree_offset, dtype=np.int64)
        arrs = [np.ascontiguousarray(a, dtype=np.int32) for a in (parent, first_child, sibling)]
        m, t = np.ascontiguousarray(mass, dtype=np.float64), np.ascontiguousarray(time, dtype=np.float64)
        h = ctypes.c_void_p()
        _lib.check(L.gal_forest_from_arrays(int(m.size), int(to.size - 1),
                                            to.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), _ip(arrs[0]),
                                            _ip(arrs[1]), _ip(arrs[2]), _dp(m), _dp(t), ctypes.byref(h)),
                   "gal_forest_from_arrays")
        return cls(h)

    @classmethod
    def from_xml(cls, texts):
        """Fully-specified trees (one XML document per tree, <tree><node>...):
        index, parent, firstChild, sibling, basic/time, basic/mass.  Returns
        (forest, components) where components[i] maps component paths
        (e.g. "disk/massGas") of node i to values."""
        parent, first_child, sibling, mass, time, comps, offsets = [], [], [], [], [], [], [0]
        for text in ([texts] if isinstance(texts, str) else texts):
            root = ET.fromstring(text)
            nodes = root.findall("node")
            idx = [int(n.findtext("index")) for n in nodes]
            # the root (parent -1) first, others in document order
            order = sorted(range(len(nodes)), key=lambda k: 0 if int(nodes[k].findtext("parent")) == -1 else 1)
            local = {idx[k]: j for j, k in enumerate(order)}
            base = offsets[-1]
            g = lambda ref: -1 if int(ref) == -1 else base + local[int(ref)]
            for k in order:
                n = nodes[k]
                parent.append(g(n.findtext("parent")))
                first_child.append(g(n.findtext("firstChild")))
                sibling.append(g(n.findtext("sibling")))
                mass.append(float(n.findtext("basic/mass")))
                time.append(float(n.findtext("basic/time")))
                c = {}
                for comp in n:
                    if comp.tag not in ("index", "parent", "firstChild", "sibling", "basic"):
                        c.update(_leaves(comp, comp.tag))
                comps.append(c)
            offsets.append(base + len(nodes))
        return cls.from_arrays(offsets, parent, first_child, sibling, mass, time), comps

    def roots(self):
        return self.tree_offset[:-1]

    def split(self, n_parts: int):
        """Split into n_parts forests of consecutive trees (balanced by node
        count).  Returns [(forest, first_node)]: node i of part k is node
        first_node + i of this forest.  Trees are independent, so evolving the
        parts separately applies exactly the same operations to every node."""
        n_parts = max(1, min(int(n_parts), self.n_trees))
        sizes = np.diff(self.tree_offset)
        cum = np.cumsum(sizes)
        cuts = [0]
        for k in range(1, n_parts):
            cuts.append(int(np.searchsorted(cum, k * cum[-1] / n_parts)) + 1)
        cuts.append(self.n_trees)
        cuts = sorted(set(min(max(c, 0), self.n_trees) for c in cuts))
        parts = []
        for a, b in zip(cuts[:-1], cuts[1:]):
            if b <= a:
                continue
            n0, n1 = int(self.tree_offset[a]), int(self.tree_offset[b])
            rel = lambda v: np.where(v >= 0, v - n0, -1).astype(np.int32)
            parts.append((Forest.from_arrays(self.tree_offset[a:b + 1] - n0, rel(self.parent[n0:n1]),
                                             rel(self.first_child[n0:n1]), rel(self.sibling[n0:n1]),
                                             self.mass[n0:n1], self.time[n0:n1]), n0))
        return parts

    def tree_of_node(self):
        """Index of the tree that holds each node (nodes of a tree are
        contiguous, tree_offset[t] <= node < tree_offset[t+1])."""
        return (np.searchsorted(self.tree_offset, np.arange(self.n_nodes), side="right") - 1).astype(np.int32)

    def initial_state(self, cosmo: Cosmology, walk: WalkParams, metal_fraction=1e-4, black_hole_seed=100.0):
        y = np.zeros((gp.NVAR, self.n_nodes))
        _lib.check(self.L.gal_forest_initial_state(self.handle, ctypes.byref(cosmo), ctypes.byref(walk),
                                                   metal_fraction, black_hole_seed, self.n_nodes, _dp(y)),
                   "gal_forest_initial_state")
        return y

    def close(self):
        if getattr(self, "handle", None):
            self.L.gal_forest_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Schedule:
    """The reference's evolution order for a forest (host walk) with waves."""

    def __init__(self, forest: Forest, cosmo: Cosmology, walk: WalkParams):
        self.L = _bind(_lib.load())
        h = ctypes.c_void_p()
        _lib.check(self.L.gal_schedule_build(forest.handle, ctypes.byref(cosmo), ctypes.byref(walk),
                                             ctypes.byref(h)), "gal_schedule_build")
        self.handle = h
        n, nw = ctypes.c_int64(), ctypes.c_int()
        self.L.gal_schedule_size(h, ctypes.byref(n), ctypes.byref(nw))
        self.n_ops, self.n_waves = int(n.value), int(nw.value)
        self._ops = None

    @property
    def ops(self):
        """The operations as a structured array (OP_DTYPE), copied out of the
        library on first use (88 B per operation: a large forest's schedule
        is run without it)."""
        if self._ops is None:
            if not getattr(self, "handle", None):
                raise ValueError("schedule closed")
            ops = np.zeros(self.n_ops, dtype=OP_DTYPE)
            if self.n_ops:
                _lib.check(self.L.gal_schedule_ops(self.handle, ctypes.c_void_p(ops.ctypes.data)),
                           "gal_schedule_ops")
            self._ops = ops
        return self._ops

    def node_gyr(self):
        """Node-Gyr the schedule evolves: its evolve intervals' lengths summed."""
        v = ctypes.c_double()
        _lib.check(self.L.gal_schedule_node_gyr(self.handle, ctypes.byref(v)), "gal_schedule_node_gyr")
        return float(v.value)

    def close(self):
        if getattr(self, "handle", None):
            self.L.gal_schedule_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def evolve_forests(evolvers, schedules, ys, streams, major_ratio=0.25, executor="graph"):
    """Evolve independent forests concurrently: forest k on evolver k and
    stream streams[k], each driven by its own host thread (the library call
    releases the interpreter lock), so one forest's waves overlap the
    others'.  Several forests may share a stream: their launches then
    interleave in that stream's order (the forests touch disjoint state, so
    the results do not depend on the interleaving).  Returns the per-node
    status tensors."""
    import threading
    out = [None] * len(schedules)
    err = []

    def run(k):
        try:
            out[k] = evolve_forest(evolvers[k], schedules[k], ys[k], major_ratio=major_ratio, stream=streams[k],
                                   executor=executor)
        except Exception as e:  # reported after every thread has joined
            err.append(e)
    threads = [threading.Thread(target=run, args=(k,)) for k in range(len(schedules))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if err:
        raise err[0]
    return out


def evolve_forest(evolver, schedule: Schedule, y, scale=None, status=None, major_ratio=0.25, stream=None,
                  surrogate=None, executor="graph"):
    """Run the schedule on device state y (torch float64, (30, n_nodes));
    returns the per-node status tensor.  executor: "graph" (one persistent
    launch; each operation waits only for the operations before it on its
    own nodes: gal_tree_evolve_d) or "waves" (wave after wave:
    gal_tree_evolve_waves_d); both give the same state bit for bit.
    surrogate: a SurrogateEvolver whose network and gate take the evolve
    operations of non-satellite halos (gal_tree_evolve_surrogate_d, wave
    after wave); its count of accepted intervals is left in
    surrogate.tree_accepted."""
    import torch
    n = int(y.shape[1])
    if status is None:
        status = torch.zeros(n, dtype=torch.int32, device=y.device)
    if scale is None:
        scale = torch.zeros_like(y)
    L = _bind(evolver.L)
    s = ctypes.c_void_p(0 if stream is None else stream.cuda_stream)
    if surrogate is None:
        if executor not in ("graph", "waves"):
            raise ValueError(f"executor must be 'graph' or 'waves', not {executor!r}")
        name = "gal_tree_evolve_d" if executor == "graph" else "gal_tree_evolve_waves_d"
        _lib.check(getattr(L, name)(evolver.handle, schedule.handle, n, ctypes.c_void_p(y.data_ptr()),
                                    ctypes.c_void_p(scale.data_ptr()), ctypes.c_void_p(status.data_ptr()),
                                    float(major_ratio), s), name)
        return status
    n_acc = ctypes.c_int64(0)
    g = surrogate.gate.c_struct()
    _lib.check(L.gal_tree_evolve_surrogate_d(evolver.handle, surrogate.mlp.handle,
                                             ctypes.c_void_p(surrogate.norm.data_ptr()), ctypes.byref(g),
                                             schedule.handle, n, ctypes.c_void_p(y.data_ptr()),
                                             ctypes.c_void_p(scale.data_ptr()), ctypes.c_void_p(status.data_ptr()),
                                             float(major_ratio), ctypes.byref(n_acc), s),
               "gal_tree_evolve_surrogate_d")
    surrogate.tree_accepted = int(n_acc.value)
    return status


def record_forest(evolver, schedule: Schedule, y, major_ratio=0.25):
    """Run the schedule like evolve_forest and return every evolve operation's
    trajectory piece: (y_before, y_after, node_par, t0, t1) as device tensors
    (30 x n, 30 x n, 5 x n, n, n), in execution order, and the status."""
    import torch
    n = int(y.shape[1])
    n_ev = int((schedule.ops["kind"] == OP_EVOLVE).sum())
    cap = max(n_ev, 1)
    dev = y.device
    status = torch.zeros(n, dtype=torch.int32, device=dev)
    scale = torch.zeros_like(y)
    y0 = torch.empty((gp.NVAR, cap), dtype=torch.float64, device=dev)
    y1 = torch.empty_like(y0)
    t0 = np.zeros(cap)
    t1 = np.zeros(cap)
    npar = np.zeros((gp.NPARAM_NODE, cap))
    n_rec = ctypes.c_int64(0)
    L = _bind(evolver.L)
    _lib.check(L.gal_tree_evolve_record_d(evolver.handle, schedule.handle, n, ctypes.c_void_p(y.data_ptr()),
                                          ctypes.c_void_p(scale.data_ptr()), ctypes.c_void_p(status.data_ptr()),
                                          float(major_ratio), ctypes.c_void_p(y0.data_ptr()),
                                          ctypes.c_void_p(y1.data_ptr()), _dp(t0), _dp(t1), _dp(npar), cap,
                                          ctypes.byref(n_rec), None), "gal_tree_evolve_record_d")
    k = int(n_rec.value)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    return y0[:, :k], y1[:, :k], d(npar[:, :k]), d(t0[:k]), d(t1[:k]), status
