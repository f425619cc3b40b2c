"""ctypes binding of the CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: the checker for the MI355X engine.  Only the test
suite, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module.  The product (``galacticus_amd``) never
imports it and fails loudly when its own HIP library is missing.

Arrays follow the engine's structure-of-arrays convention: ``y[v, i]`` is
variable ``v`` of node ``i`` (numpy float64, C order, shape (30, n)).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "liboracle.so")
_lib = None

NVAR = 30
NPARAM_NODE = 7


class OrcCooling(ctypes.Structure):
    _fields_ = [
        ("ln_lambda", ctypes.POINTER(ctypes.c_double)),
        ("n_t", ctypes.c_int),
        ("n_z", ctypes.c_int),
        ("ln_t_min", ctypes.c_double),
        ("d_ln_t", ctypes.c_double),
        ("ln_z_min", ctypes.c_double),
        ("d_ln_z", ctypes.c_double),
    ]


def build() -> str:
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def build_native(dest_dir: str) -> str:
    """Compile the oracle for the host it runs on (-march=native, the
    reference's -O3 and -fno-math-errno, OpenMP) into dest_dir; returns the
    path.  Used for the CPU baseline, where the prebuilt library (portable
    x86-64-v3) would understate a newer host."""
    os.makedirs(dest_dir, exist_ok=True)
    out = os.path.join(dest_dir, "liboracle_native.so")
    cmd = ["gcc", "-O3", "-march=native", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-math-errno", "-std=c11",
           "-shared", "-o", out, os.path.join(_HERE, "oracle.c"), "-lm"]
    subprocess.run(cmd, check=True, capture_output=True)
    return out


def use_library(path: str) -> None:
    """Load the oracle from another build of oracle.c (same ABI)."""
    global _LIB_PATH, _lib
    _LIB_PATH = path
    _lib = None
    lib()


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        lp = ctypes.POINTER(ctypes.c_int64)
        cp = ctypes.POINTER(OrcCooling)
        L.orc_evolve_batch.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, dp, dp, dp, dp, ip, lp, dp, cp,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_int]
        L.orc_evolve_batch.restype = ctypes.c_int
        L.orc_rhs_batch.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, dp, dp, dp, cp, ctypes.c_double]
        L.orc_rhs_batch.restype = ctypes.c_int
        L.orc_scales_batch.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, dp, dp, dp]
        L.orc_scales_batch.restype = None
        L.orc_hadjust.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, dp, dp, dp, dp, ip, dp]
        L.orc_hadjust.restype = ctypes.c_int
        L.orc_cooling_function.argtypes = [cp, ctypes.c_double, ctypes.c_double]
        L.orc_cooling_function.restype = ctypes.c_double
        L.orc_cooling_time.argtypes = [dp, cp, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.orc_cooling_time.restype = ctypes.c_double
        L.orc_halo_properties.argtypes = [dp, cp, dp, ctypes.c_double, dp, ctypes.c_double, dp]
        L.orc_halo_properties.restype = None
        L.orc_test_ode_solve.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, dp, ctypes.c_double,
                                         ctypes.c_double, dp, lp]
        L.orc_test_ode_solve.restype = ctypes.c_int
        L.orc_test_rkck_step.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, dp, dp]
        L.orc_test_rkck_step.restype = ctypes.c_int
        L.orc_tree_run.argtypes = [ctypes.c_int64, ip, ip, ip, dp, dp, dp, ctypes.c_int, ip, ctypes.c_int, dp, dp, ip,
                                   dp, cp, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_int]
        L.orc_tree_run.restype = ctypes.c_int
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _lp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def cooling_struct(table):
    """Wrap a CoolingTable (galacticus_amd.cooling) for the oracle; keeps a ref."""
    arr = np.ascontiguousarray(table.ln_lambda, dtype=np.float64)
    s = OrcCooling(_dp(arr), table.n_t, table.n_z, table.ln_t_min, table.d_ln_t, table.ln_z_min, table.d_ln_z)
    s._keep = arr
    return s


def evolve(y, node_par, t_start, t_end, h_step, params, table, tol_abs, tol_rel, flags,
           scale=None, root_tol_rel=1.0e-14, nthreads=1):
    """Evolve nodes on the CPU.  Returns (y, h_step, scale, status, counters)."""
    L = lib()
    y = np.array(y, dtype=np.float64, order="C", copy=True)
    n = y.shape[1]
    node_par = np.ascontiguousarray(node_par, dtype=np.float64)
    t_start = np.ascontiguousarray(t_start, dtype=np.float64)
    t_end = np.ascontiguousarray(t_end, dtype=np.float64)
    h = np.array(h_step, dtype=np.float64, copy=True)
    sc = np.zeros((NVAR, n)) if scale is None else np.array(scale, dtype=np.float64, copy=True)
    status = np.zeros(n, dtype=np.int32)
    counters = np.zeros(3 * n, dtype=np.int64)
    params = np.ascontiguousarray(params, dtype=np.float64)
    cs = cooling_struct(table)
    L.orc_evolve_batch(n, n, _dp(y), _dp(node_par), _dp(t_start), _dp(t_end), _dp(h), _dp(sc), _ip(status),
                       _lp(counters), _dp(params), ctypes.byref(cs), tol_abs, tol_rel, flags, root_tol_rel,
                       nthreads)
    return y, h, sc, status, counters.reshape(3, n)


def rhs(t, y, node_par, params, table, root_tol_rel=1.0e-14):
    L = lib()
    y = np.ascontiguousarray(y, dtype=np.float64)
    n = y.shape[1]
    d = np.zeros_like(y)
    t = np.ascontiguousarray(np.broadcast_to(t, (n,)), dtype=np.float64)
    cs = cooling_struct(table)
    L.orc_rhs_batch(n, n, _dp(t), _dp(y), _dp(np.ascontiguousarray(node_par, dtype=np.float64)), _dp(d),
                    _dp(np.ascontiguousarray(params, dtype=np.float64)), ctypes.byref(cs), root_tol_rel)
    return d


def scales(t, y, node_par, params):
    L = lib()
    y = np.ascontiguousarray(y, dtype=np.float64)
    n = y.shape[1]
    s = np.zeros_like(y)
    t = np.ascontiguousarray(np.broadcast_to(t, (n,)), dtype=np.float64)
    L.orc_scales_batch(n, n, _dp(t), _dp(y), _dp(np.ascontiguousarray(node_par, dtype=np.float64)), _dp(s),
                       _dp(np.ascontiguousarray(params, dtype=np.float64)))
    return s


def hadjust(ord_, eps_abs, eps_rel, a_y, a_dydt, y, yerr, yp, scale, nonneg, h):
    L = lib()
    y = np.ascontiguousarray(y, dtype=np.float64)
    dim = y.shape[0]
    hh = np.array([h], dtype=np.float64)
    nn = np.ascontiguousarray(nonneg, dtype=np.int32)
    r = L.orc_hadjust(dim, ord_, eps_abs, eps_rel, a_y, a_dydt, _dp(y),
                      _dp(np.ascontiguousarray(yerr, dtype=np.float64)),
                      _dp(np.ascontiguousarray(yp, dtype=np.float64)),
                      _dp(np.ascontiguousarray(scale, dtype=np.float64)), _ip(nn), _dp(hh))
    return r, float(hh[0])


def cooling_function(table, temperature, metallicity_solar):
    cs = cooling_struct(table)
    return lib().orc_cooling_function(ctypes.byref(cs), temperature, metallicity_solar)


def cooling_time(params, table, temperature, density, metal_fraction):
    cs = cooling_struct(table)
    p = np.ascontiguousarray(params, dtype=np.float64)
    return lib().orc_cooling_time(_dp(p), ctypes.byref(cs), temperature, density, metal_fraction)


def halo_properties(params, table, node_par_row, t, y_row, root_tol_rel=1.0e-14):
    cs = cooling_struct(table)
    out = np.zeros(8)
    p = np.ascontiguousarray(params, dtype=np.float64)
    npr = np.ascontiguousarray(node_par_row, dtype=np.float64)
    yr = np.ascontiguousarray(y_row, dtype=np.float64)
    lib().orc_halo_properties(_dp(p), ctypes.byref(cs), _dp(npr), t, _dp(yr), root_tol_rel, _dp(out))
    return dict(r_vir=out[0], v_vir=out[1], t_dyn=out[2], t_vir=out[3], a=out[4], r_cool=out[5],
                r_cool_rate=out[6], t_avail=out[7])


def test_ode_solve(which, x0, x1, y0, tol_abs, tol_rel):
    y = np.array(y0, dtype=np.float64, copy=True)
    h = np.zeros(1)
    cnt = np.zeros(3, dtype=np.int64)
    st = lib().orc_test_ode_solve(which, x0, x1, _dp(y), tol_abs, tol_rel, _dp(h), _lp(cnt))
    return st, y, float(h[0]), cnt


def test_rkck_step(which, t, h, y0):
    """One uncontrolled Cash-Karp step of a test system: (status, y, yerr)."""
    y = np.array(y0, dtype=np.float64, copy=True)
    yerr = np.zeros_like(y)
    st = lib().orc_test_rkck_step(which, t, h, _dp(y), _dp(yerr))
    return st, y, yerr


def tree_run(ops, y0, h0, tree_of_node, n_trees, params, table, tol_abs, tol_rel, flags, major_ratio=0.25,
             root_tol_rel=1.0e-14, nthreads=1):
    """Apply a schedule's operations (structured array with the fields of
    gal_schedule_op) on the CPU, one at a time per tree, trees in parallel.
    Returns (y, h, status, n_failed)."""
    L = lib()
    y = np.array(y0, dtype=np.float64, order="C", copy=True)
    ld = y.shape[1]
    h = np.array(h0, dtype=np.float64, copy=True)
    status = np.zeros(ld, dtype=np.int32)
    kind = np.ascontiguousarray(ops["kind"], dtype=np.int32)
    a = np.ascontiguousarray(ops["a"], dtype=np.int32)
    b = np.ascontiguousarray(ops["b"], dtype=np.int32)
    t0 = np.ascontiguousarray(ops["t0"], dtype=np.float64)
    t1 = np.ascontiguousarray(ops["t1"], dtype=np.float64)
    npar = np.ascontiguousarray(ops["node_par"], dtype=np.float64)
    tree = np.ascontiguousarray(tree_of_node, dtype=np.int32)
    cs = cooling_struct(table)
    bad = L.orc_tree_run(len(kind), _ip(kind), _ip(a), _ip(b), _dp(t0), _dp(t1), _dp(npar), ld, _ip(tree), n_trees,
                         _dp(y), _dp(h), _ip(status), _dp(np.ascontiguousarray(params, dtype=np.float64)),
                         ctypes.byref(cs), tol_abs, tol_rel, flags, root_tol_rel, major_ratio, nthreads)
    if bad < 0:
        raise RuntimeError(f"orc_tree_run returned {bad}")
    return y, h, status, bad
