"""CPU restatement of the tree path: the reference's tree walk and the
sequential execution of its operations.

TEST INFRASTRUCTURE ONLY: the checker for galacticus_amd/csrc/trees.cpp and
galacticus_amd/csrc/tree_exec.hip.  Only the test suite imports this module.

``walk`` follows mergerTreeEvolverStandard%evolve (merger_trees/evolver/
standard.F90:291-635): repeated depth-first walks (walkers/all_nodes/
_class.F90:101-222) that evolve every evolvable node (:723-760) as far as
timeEvolveTo (:762-1035) allows, with the standard timestepper (simple +
satellite, evolve/timesteps/standard.F90:136-200, simple.F90:124-155,
satellite.F90 timeEvolveTo), promotion (node_evolver/standard.F90:1241-1327),
halo mergers (single_level_hierarchy.F90:65-115) with the Jiang (2008)
merging time (satellites/merging/timescale/Jiang2008.F90) on a fixed orbit,
and satellite mergers (evolve/timesteps/satellite.F90:210-278).  It returns
the operation log: (kind, a, b, t0, t1, node_par).

``run_sequential`` applies a log to per-node states one operation at a time,
evolving each interval with the C oracle, as the reference does while it
walks.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import math

import numpy as np

# The C library's cube root (numpy's vectorised cbrt can differ in the last bit).
_libm = ctypes.CDLL(ctypes.util.find_library("m"))
_libm.cbrt.restype = ctypes.c_double
_libm.cbrt.argtypes = [ctypes.c_double]
cbrt = _libm.cbrt

EVOLVE, PROMOTE, SATELLITE, MERGE = 0, 1, 2, 3

_PARSEC = 3.0856775814913673e16
_MPC = 1.0e6 * _PARSEC
_GYR = 1.0e9 * 3.15581497635456e7
_MSUN = 1.98847e30
G = 6.67430e-11 * _MSUN / 1.0e6 / _MPC          # Mpc (km/s)^2 / Msun
MPC_KMS_GYR = _MPC / 1.0e3 / _GYR


class Cosmo:
    def __init__(self, om, ol, h0):
        self.om, self.ol, self.h0 = om, ol, h0
        self.h0_gyr = h0 / MPC_KMS_GYR

    def a(self, t):
        # libm cbrt and pow, as the engine's host code uses
        return cbrt(self.om / self.ol) * math.sinh(1.5 * math.sqrt(self.ol) * self.h0_gyr * t) ** (2.0 / 3.0)

    def e2(self, a):
        return self.om / (a * a * a) + self.ol

    def hubble_time(self, t):
        return 1.0 / (self.h0_gyr * math.sqrt(self.e2(self.a(t))))

    def virial(self, m, t):
        a = self.a(t)
        e2 = self.e2(a)
        x = self.om / (a * a * a) / e2 - 1.0
        dc = 18.0 * math.pi * math.pi + 82.0 * x - 39.0 * x * x
        rho_c = 3.0 * self.h0 * self.h0 * e2 / (8.0 * math.pi * G)
        r = cbrt(3.0 * m / (4.0 * math.pi * dc * rho_c))
        return r, math.sqrt(G * m / r)


def walk(parent, first_child, sibling, mass, time, cosmo: Cosmo, w: dict, trees=None):
    """Operation log of the reference walk over every tree of a forest
    (global indices; trees = list of (first, last+1), root first)."""
    n = len(parent)
    if trees is None:
        trees = [(0, n)]
    log = []
    for lo, hi in trees:
        log += _walk_tree(parent, first_child, sibling, mass, time, cosmo, w, lo, hi)
    return log


def _walk_tree(parent0, child0, sib0, mass, time0, cosmo, w, lo, hi):
    P = {i: parent0[i] for i in range(lo, hi)}
    C = {i: child0[i] for i in range(lo, hi)}
    S = {i: sib0[i] for i in range(lo, hi)}
    SAT = {i: -1 for i in range(lo, hi)}          # first satellite
    is_sat = {i: False for i in range(lo, hi)}
    T = {i: float(time0[i]) for i in range(lo, hi)}
    tm = {i: math.inf for i in range(lo, hi)}      # merging times
    root = lo
    final = T[root]
    lam = w["spin_parameter"]

    # the halo of each node until it reaches its parent
    par = {}
    for i in range(lo, hi):
        m, t = float(mass[i]), float(time0[i])
        r, v = cosmo.virial(m, t)
        j = lam * m * v * r
        mr = jr = 0.0
        p = P[i]
        if p >= 0 and C[p] == i and time0[p] > t:
            mp, tp = float(mass[p]), float(time0[p])
            rp, vp = cosmo.virial(mp, tp)
            mr = (mp - m) / (tp - t)
            jr = (lam * mp * vp * rp - j) / (tp - t)
        par[i] = [t, m, mr, jr, 0.0, 0.0, 0.0]

    def halo_mass(i, t):
        q = par[i]
        return q[1] + q[2] * (t - q[0])

    def is_primary(i):
        return P[i] >= 0 and C[P[i]] == i

    def deepest(i):
        while SAT[i] >= 0 or C[i] >= 0:
            i = SAT[i] if SAT[i] >= 0 else C[i]
        return i

    def walker_next(i):
        if i == root:
            return -1
        if S[i] >= 0:
            return deepest(S[i])
        if is_sat[i]:
            return deepest(C[P[i]]) if C[P[i]] >= 0 else P[i]
        return P[i]

    def satellites(h):
        out, k = [], SAT[h]
        while k >= 0:
            out.append(k)
            k = S[k]
        return out

    def add_satellite(h, s):
        S[s] = -1
        P[s] = h
        sats = satellites(h)
        if sats:
            S[sats[-1]] = s
        else:
            SAT[h] = s

    def drop_satellite(h, s):
        sats = satellites(h)
        sats.remove(s)
        SAT[h] = sats[0] if sats else -1
        for a, b in zip(sats, sats[1:] + [-1]):
            S[a] = b
        S[s] = -1

    def hand_over_satellites(frm, to):
        for s in satellites(frm):
            add_satellite(to, s)
        SAT[frm] = -1

    def can_evolve(i):
        return (P[i] >= 0 and C[i] < 0 and T[i] < w["time_end"]
                and (T[i] < final or (is_sat[i] and T[i] <= final)))

    def target_time(i):
        t = T[i]
        to = w["time_end"]
        simple = t + min(w["timestep_relative"] * cosmo.hubble_time(t), w["timestep_absolute"])
        sat_step, sat_task = math.inf, False
        if is_sat[i] and tm[i] < math.inf:
            off = min(w["satellite_offset_absolute"], w["satellite_offset_relative"] * tm[i])
            h = P[i]
            if T[h] < tm[i] - off and P[h] >= 0:
                sat_step = max(tm[i] - t - 0.5 * off, 0.0) + t
            else:
                sat_step = tm[i]
                sat_task = C[h] < 0
        step, task = (sat_step, sat_task) if sat_step <= simple else (simple, False)
        if step > to:
            task = False
        else:
            to = step
        if to == t:
            return to, task
        if not is_sat[i]:
            if P[i] >= 0 and T[P[i]] < to:
                to = T[P[i]]
        else:
            h = P[i]
            th = T[h] if P[h] >= 0 else max(T[h], t)
            if C[h] >= 0:
                lim = max(th, t)
            else:
                lim = max(th + min(w["timestep_host_relative"] * cosmo.hubble_time(th), w["timestep_host_absolute"]), t)
            to = min(to, lim)
        if to == t:
            return to, task
        if is_primary(i):
            k = S[i]
            while k >= 0:
                to = min(to, max(t, T[k]))
                k = S[k]
        if to == t:
            return to, task
        for s in satellites(i):
            if T[s] < to:
                to = max(T[s], t)
                if to == t:
                    break
        if to == t:
            return to, task
        for s in satellites(i):
            if tm[s] < math.inf and max(tm[s], t) < to:
                to = max(tm[s], t)
        return to, task

    def merging_time(host, sat, t):
        q = halo_mass(host, t) / halo_mass(sat, t)
        if not q > 0.0:
            return 0.0
        r, v = cosmo.virial(halo_mass(host, t), t)
        t_dyn = MPC_KMS_GYR * r / v
        return (w["merging_multiplier"] * t_dyn * math.sqrt(w["orbit_radius_ratio"])
                * ((0.94 * w["orbit_circularity"] ** 0.60 + 0.60) / 2.0 / 0.43) * q / math.log(1.0 + q))

    log = []
    while True:
        evolvable_seen = False
        n_log = len(log)
        i = deepest(root)
        while i >= 0:
            p_saved, s_saved = P[i], S[i]
            nxt, found, alive = -1, False, True
            if can_evolve(i):
                evolvable_seen = True
                moved = False
                to, task = target_time(i)
                if to > T[i]:
                    log.append((EVOLVE, i, -1, T[i], to, list(par[i])))
                    T[i] = to
                    moved = True
                if task:
                    h = P[i]
                    log.append((MERGE, i, h, T[i], T[i], None))
                    hand_over_satellites(i, h)
                    drop_satellite(h, i)
                    alive = False
                if alive and P[i] >= 0 and T[i] >= T[P[i]]:
                    p = P[i]
                    if not is_primary(i):
                        if not is_sat[i]:
                            k = C[p]
                            while S[k] != i:
                                k = S[k]
                            S[k] = S[i]
                            tmerge = T[i] + merging_time(C[p], i, T[i])
                            add_satellite(p, i)
                            hand_over_satellites(i, p)
                            is_sat[i] = True
                            tm[i] = tmerge
                            par[i] = [T[i], halo_mass(i, T[i]), 0.0, 0.0, 1.0, 0.0, 0.0]
                            log.append((SATELLITE, i, p, T[i], tmerge, None))
                    elif S[i] < 0:
                        log.append((PROMOTE, i, p, T[i], T[p], None))
                        h = p
                        while is_sat[h]:
                            h = P[h]
                        hand_over_satellites(i, h)
                        C[p] = -1
                        alive = False
                        nxt, found = p, True
                if not found and alive and moved:
                    nxt = i
            if nxt < 0:
                if alive:
                    nxt = walker_next(i)
                elif not found:
                    nxt = deepest(s_saved if s_saved >= 0 else p_saved) if (s_saved >= 0 or p_saved >= 0) else -1
            i = nxt
        if not evolvable_seen:
            break
        if len(log) == n_log:
            raise RuntimeError("tree walk deadlocked")
    return log


def run_sequential(log, y0, h0, evolve_one, merge_galaxy):
    """Apply an operation log to states y (30 x n) one operation at a time.
    evolve_one(y_col, node_par, t0, t1, h) -> (y_col, h, status);
    merge_galaxy(y, sat, host) mutates y."""
    y = np.array(y0, dtype=np.float64, copy=True)
    h = np.array(h0, dtype=np.float64, copy=True)
    status = np.zeros(y.shape[1], dtype=np.int32)
    for kind, a, b, t0, t1, npar in log:
        if kind == EVOLVE:
            y[:, a], h[a], status[a] = evolve_one(y[:, a], npar, t0, t1, h[a])
        elif kind == PROMOTE:
            y[:, b] = y[:, a]
            h[b] = h[a]
        elif kind == MERGE:
            merge_galaxy(y, a, b)
    return y, h, status


# State indices used by the mass movements.
_V = dict(HOT_MASS=0, HOT_ABUNDANCES=1, HOT_ANGULAR_MOMENTUM=2, HOT_OUTFLOWED_MASS=3,
          HOT_OUTFLOWED_ANGULAR_MOMENTUM=4, HOT_OUTFLOWED_ABUNDANCES=5, HOT_UNACCRETED_MASS=6,
          HOT_UNACCRETED_ABUNDANCES=7, HOT_STRIPPED_MASS=9, HOT_STRIPPED_ABUNDANCES=10, HOT_ENERGY_RADIATED=11,
          DISK_MASS_STELLAR=12, DISK_MASS_STELLAR_FORMED=13,
          DISK_ABUNDANCES_STELLAR=15, DISK_MASS_GAS=16, DISK_ABUNDANCES_GAS=17, DISK_ANGULAR_MOMENTUM=18,
          SPHEROID_MASS_STELLAR=19, SPHEROID_MASS_STELLAR_FORMED=20, SPHEROID_ABUNDANCES_STELLAR=21,
          SPHEROID_MASS_GAS=22, SPHEROID_ABUNDANCES_GAS=23, SPHEROID_ANGULAR_MOMENTUM=24, BH_MASS=25)


def merge_galaxy_simple(y, sat, host, major_ratio=0.25):
    """mergerMassMovementsSimple (satellites/merging/mass_movements/simple.F90:208-290)
    with spheroid destinations: the satellite's gas and stars join the host
    spheroid, the host disk joins them in a major merger; the satellite's
    circumgalactic reservoirs and black hole join the host's."""
    V = _V
    gal = lambda s: (y[V["DISK_MASS_GAS"], s] + y[V["DISK_MASS_STELLAR"], s] + y[V["SPHEROID_MASS_GAS"], s]
                     + y[V["SPHEROID_MASS_STELLAR"], s])
    ms, mh = gal(sat), gal(host)
    major = ms > 0 and mh > 0 and min(ms, mh) >= major_ratio * max(ms, mh)
    to_sph = [("DISK_MASS_GAS", "SPHEROID_MASS_GAS"), ("DISK_ABUNDANCES_GAS", "SPHEROID_ABUNDANCES_GAS"),
              ("DISK_MASS_STELLAR", "SPHEROID_MASS_STELLAR"), ("DISK_ABUNDANCES_STELLAR", "SPHEROID_ABUNDANCES_STELLAR"),
              ("DISK_MASS_STELLAR_FORMED", "SPHEROID_MASS_STELLAR_FORMED"),
              ("DISK_ANGULAR_MOMENTUM", "SPHEROID_ANGULAR_MOMENTUM")]
    same = ["SPHEROID_MASS_GAS", "SPHEROID_ABUNDANCES_GAS", "SPHEROID_MASS_STELLAR", "SPHEROID_ABUNDANCES_STELLAR",
            "SPHEROID_MASS_STELLAR_FORMED", "SPHEROID_ANGULAR_MOMENTUM", "HOT_MASS", "HOT_ABUNDANCES",
            "HOT_ANGULAR_MOMENTUM", "HOT_OUTFLOWED_MASS", "HOT_OUTFLOWED_ANGULAR_MOMENTUM", "HOT_OUTFLOWED_ABUNDANCES",
            "HOT_UNACCRETED_MASS", "HOT_UNACCRETED_ABUNDANCES", "HOT_ENERGY_RADIATED", "BH_MASS"]
    # the satellite's stripped gas joins the host's outflowed reservoir (mass
    # and metals; hot_halo/standard/_class.F90:532-597)
    stripped = [("HOT_STRIPPED_MASS", "HOT_OUTFLOWED_MASS"), ("HOT_STRIPPED_ABUNDANCES", "HOT_OUTFLOWED_ABUNDANCES")]
    for a, b in to_sph + [(k, k) for k in same] + stripped:
        y[V[b], host] += y[V[a], sat]
        y[V[a], sat] = 0.0
    if major:
        for a, b in to_sph:
            y[V[b], host] += y[V[a], host]
            y[V[a], host] = 0.0
