"""Seeded synthetic node batches.

``isolated_halos`` draws the batch of independent isolated-halo node ODE
systems used for the throughput measurement: halo mass log-uniform in
[1e9, 1e13] Msun, initial hot-gas fraction uniform in [0.05, 0.17], hot-gas
metallicity log-uniform in [1e-4, 0.02], interval length uniform in
[0.1, 2] Gyr, cosmology Omega_m=0.3, Omega_L=0.7, h=0.7.  Every other initial
component value is drawn from the same generator so a seed fixes the batch;
disks start bar-stable.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import params as gp

NV = gp.NVAR

# State indices (enum gal_var of include/gal_ode.h).
V = dict(
    HOT_MASS=0, HOT_ABUNDANCES=1, HOT_ANGULAR_MOMENTUM=2, HOT_OUTFLOWED_MASS=3,
    HOT_OUTFLOWED_ANGULAR_MOMENTUM=4, HOT_OUTFLOWED_ABUNDANCES=5, HOT_UNACCRETED_MASS=6,
    HOT_UNACCRETED_ABUNDANCES=7, HOT_OUTER_RADIUS=8, HOT_STRIPPED_MASS=9, HOT_STRIPPED_ABUNDANCES=10,
    HOT_ENERGY_RADIATED=11, DISK_MASS_STELLAR=12, DISK_MASS_STELLAR_FORMED=13,
    DISK_FRACTION_MASS_RETAINED=14, DISK_ABUNDANCES_STELLAR=15, DISK_MASS_GAS=16, DISK_ABUNDANCES_GAS=17,
    DISK_ANGULAR_MOMENTUM=18, SPHEROID_MASS_STELLAR=19, SPHEROID_MASS_STELLAR_FORMED=20,
    SPHEROID_ABUNDANCES_STELLAR=21, SPHEROID_MASS_GAS=22, SPHEROID_ABUNDANCES_GAS=23,
    SPHEROID_ANGULAR_MOMENTUM=24, BH_MASS=25, BH_SPIN=26, BH_RADIAL_POSITION=27,
    SPIN_ANGULAR_MOMENTUM=28, SATELLITE_BOUND_MASS=29,
)
VAR_NAMES = [k for k, _ in sorted(V.items(), key=lambda kv: kv[1])]
NP = dict(TIME_REFERENCE=0, MASS_REFERENCE=1, MASS_RATE=2, SPIN_RATE=3, IS_SATELLITE=4, RAM_PRESSURE_FORCE=5,
          HOST_VELOCITY_VIRIAL=6)

# Masses and metal masses (the quantities compared per component).
MASS_VARS = [V[k] for k in ("HOT_MASS", "HOT_OUTFLOWED_MASS", "HOT_UNACCRETED_MASS", "HOT_STRIPPED_MASS",
                            "DISK_MASS_STELLAR", "DISK_MASS_GAS", "SPHEROID_MASS_STELLAR",
                            "SPHEROID_MASS_GAS", "BH_MASS")]
METAL_VARS = [V[k] for k in ("HOT_ABUNDANCES", "HOT_OUTFLOWED_ABUNDANCES", "HOT_UNACCRETED_ABUNDANCES",
                             "HOT_STRIPPED_ABUNDANCES", "DISK_ABUNDANCES_STELLAR", "DISK_ABUNDANCES_GAS",
                             "SPHEROID_ABUNDANCES_STELLAR", "SPHEROID_ABUNDANCES_GAS")]

# Constants matching the engine (see galacticus_amd/csrc/constants.hpp).
_PARSEC = 3.0856775814913673e16
_MPC = 1.0e6 * _PARSEC
_GYR = 1.0e9 * 3.15581497635456e7
_MSUN = 1.98847e30
_G_SI = 6.67430e-11
G_INTERNAL = _G_SI * _MSUN / 1.0e6 / _MPC
MPC_PER_KMS_GYR = _MPC / 1.0e3 / _GYR


def cosmic_time_to_scale_factor(t, om=0.3, ol=0.7, h0=70.0):
    h0g = h0 / MPC_PER_KMS_GYR
    return (om / ol) ** (1.0 / 3.0) * np.sinh(1.5 * math.sqrt(ol) * h0g * np.asarray(t)) ** (2.0 / 3.0)


def scale_factor_to_cosmic_time(a, om=0.3, ol=0.7, h0=70.0):
    h0g = h0 / MPC_PER_KMS_GYR
    a = np.asarray(a, dtype=np.float64)
    return np.arcsinh((a / (om / ol) ** (1.0 / 3.0)) ** 1.5) / (1.5 * math.sqrt(ol) * h0g)


def virial(mass, t, om=0.3, ol=0.7, h0=70.0):
    """(r_vir [Mpc], V_vir [km/s]) with the engine's Bryan-Norman contrast."""
    a = cosmic_time_to_scale_factor(t, om, ol, h0)
    a3 = a ** 3
    rho_crit0 = 3.0 * h0 * h0 / (8.0 * math.pi * G_INTERNAL)
    rho_mean = om * rho_crit0 / a3
    om_a = (om / a3) / (om / a3 + ol)
    x = om_a - 1.0
    delta_m = (18.0 * math.pi ** 2 + 82.0 * x - 39.0 * x * x) / om_a
    r = np.cbrt(3.0 * mass / (4.0 * math.pi * delta_m * rho_mean))
    v = np.sqrt(G_INTERNAL * mass / r)
    return r, v


@dataclasses.dataclass
class NodeBatch:
    y: np.ndarray          # (30, n)
    node_par: np.ndarray   # (7, n): GAL_NPARAM_NODE rows
    t_start: np.ndarray
    t_end: np.ndarray
    h_step: np.ndarray

    @property
    def n(self):
        return self.y.shape[1]

    @property
    def node_gyr(self):
        return float(np.sum(np.abs(self.t_end - self.t_start)))


def isolated_halos(n: int, seed: int = 0, params=None) -> NodeBatch:
    p = gp.physics_params() if params is None else params
    om, ol, h0 = p[gp.P["OmegaMatter"]], p[gp.P["OmegaDarkEnergy"]], p[gp.P["HubbleConstant"]]
    fb = p[gp.P["OmegaBaryon"]] / om
    rng = np.random.Generator(np.random.PCG64(seed))
    mass = 10.0 ** rng.uniform(9.0, 13.0, n)
    f_hot = rng.uniform(0.05, 0.17, n)
    metal = 10.0 ** rng.uniform(-4.0, math.log10(0.02), n)
    dt = rng.uniform(0.1, 2.0, n)
    t0 = rng.uniform(2.0, 11.5, n)
    spin = np.clip(np.exp(rng.normal(math.log(0.035), 0.5, n)), 0.01, 0.1)
    growth = rng.uniform(0.0, 0.1, n)
    f_out = rng.uniform(0.0, 0.02, n)
    f_dg = rng.uniform(0.01, 0.1, n)
    f_ds = rng.uniform(0.0, 0.1, n)
    f_sg = rng.uniform(0.0, 0.02, n)
    f_ss = rng.uniform(0.0, 0.1, n)
    bh_spin = rng.uniform(0.0, 0.9, n)
    r_vir, v_vir = virial(mass, t0, om, ol, h0)
    j_spin = spin * mass * v_vir * r_vir
    r_disk = p[gp.P["galacticStructureSolver.factor"]] * r_vir * spin
    y = np.zeros((NV, n))
    mb = fb * mass
    # Disks start below the bar-instability threshold (efstathiou1982: eps =
    # 1.18 V / sqrt(G M_disk / r_disk) at or above the threshold interpolated
    # in gas fraction) as a disk grown by cooling would: a random disk drawn
    # above it is driven to the threshold, and the spheroid that it feeds
    # (star formation on 1e-3 Gyr and outflows (100 km/s / V)^3.5 times faster
    # in a dwarf) becomes a 1e-5 Gyr stiff reservoir.
    m_disk = (f_dg + f_ds) * mb
    f_gas = f_dg / (f_dg + f_ds)
    thr = (p[gp.P["barInstability.stabilityThresholdStellar"]] * (1.0 - f_gas) +
           p[gp.P["barInstability.stabilityThresholdGaseous"]] * f_gas)
    m_stable = 0.9 * r_disk * v_vir ** 2 * (1.1800237580 / thr) ** 2 / G_INTERNAL
    shrink = np.minimum(1.0, m_stable / m_disk)
    f_dg = f_dg * shrink
    f_ds = f_ds * shrink
    y[V["HOT_MASS"]] = f_hot * mass
    y[V["HOT_ABUNDANCES"]] = metal * y[V["HOT_MASS"]]
    y[V["HOT_ANGULAR_MOMENTUM"]] = y[V["HOT_MASS"]] * j_spin / mass
    y[V["HOT_OUTFLOWED_MASS"]] = f_out * mb
    y[V["HOT_OUTFLOWED_ABUNDANCES"]] = metal * y[V["HOT_OUTFLOWED_MASS"]]
    y[V["HOT_OUTFLOWED_ANGULAR_MOMENTUM"]] = y[V["HOT_OUTFLOWED_MASS"]] * j_spin / mass
    y[V["HOT_OUTER_RADIUS"]] = r_vir
    y[V["DISK_MASS_GAS"]] = f_dg * mb
    y[V["DISK_MASS_STELLAR"]] = f_ds * mb
    y[V["DISK_MASS_STELLAR_FORMED"]] = y[V["DISK_MASS_STELLAR"]]
    y[V["DISK_FRACTION_MASS_RETAINED"]] = 1.0
    y[V["DISK_ABUNDANCES_GAS"]] = metal * y[V["DISK_MASS_GAS"]]
    y[V["DISK_ABUNDANCES_STELLAR"]] = metal * y[V["DISK_MASS_STELLAR"]]
    y[V["DISK_ANGULAR_MOMENTUM"]] = (y[V["DISK_MASS_GAS"]] + y[V["DISK_MASS_STELLAR"]]) * 2.0 * r_disk * v_vir
    y[V["SPHEROID_MASS_GAS"]] = f_sg * mb
    y[V["SPHEROID_MASS_STELLAR"]] = f_ss * mb
    y[V["SPHEROID_MASS_STELLAR_FORMED"]] = y[V["SPHEROID_MASS_STELLAR"]]
    y[V["SPHEROID_ABUNDANCES_GAS"]] = metal * y[V["SPHEROID_MASS_GAS"]]
    y[V["SPHEROID_ABUNDANCES_STELLAR"]] = metal * y[V["SPHEROID_MASS_STELLAR"]]
    y[V["SPHEROID_ANGULAR_MOMENTUM"]] = (y[V["SPHEROID_MASS_GAS"]] + y[V["SPHEROID_MASS_STELLAR"]]) * 0.1 * r_disk * v_vir
    y[V["BH_MASS"]] = 100.0 + 1.0e-3 * y[V["SPHEROID_MASS_STELLAR"]]
    y[V["BH_SPIN"]] = bh_spin
    y[V["SPIN_ANGULAR_MOMENTUM"]] = j_spin
    y[V["SATELLITE_BOUND_MASS"]] = mass
    npar = np.zeros((gp.NPARAM_NODE, n))
    npar[NP["TIME_REFERENCE"]] = t0
    npar[NP["MASS_REFERENCE"]] = mass
    npar[NP["MASS_RATE"]] = growth * mass
    npar[NP["SPIN_RATE"]] = growth * j_spin
    npar[NP["IS_SATELLITE"]] = 0.0
    return NodeBatch(y, npar, t0, t0 + dt, -np.ones(n))


def satellite_halos(n: int, seed: int = 0, params=None) -> NodeBatch:
    """The isolated-halo batch turned into satellites: each halo stopped
    growing (became a satellite) up to 2 Gyr before its interval starts, its
    host has a virial velocity 1.5-6 times its own, and the host's hot gas
    presses on it with a ram pressure log-uniform in [1e-3, 0.3] times the
    satellite's own binding scale formFactor V_vir^2 rho_0 of its hot-gas
    profile, so that some halos keep their gas and some are stripped to a
    radius inside their outer radius."""
    p = gp.physics_params() if params is None else params
    b = isolated_halos(n, seed, p)
    rng = np.random.Generator(np.random.PCG64(seed + 7919))
    om, ol, h0 = p[gp.P["OmegaMatter"]], p[gp.P["OmegaDarkEnergy"]], p[gp.P["HubbleConstant"]]
    t_iso = b.t_start - rng.uniform(0.0, 2.0, n)
    t_iso = np.maximum(t_iso, 0.5 * b.t_start)
    mass = b.node_par[NP["MASS_REFERENCE"]]
    r_vir, v_vir = virial(mass, t_iso, om, ol, h0)
    y = b.y.copy()
    y[V["HOT_OUTER_RADIUS"]] = r_vir * rng.uniform(0.3, 1.0, n)
    rc = p[gp.P["coreRadiusOverVirialRadius"]] * r_vir
    xo = y[V["HOT_OUTER_RADIUS"]] / rc
    rho0 = y[V["HOT_MASS"]] / (4.0 * math.pi * rc ** 3 * (xo - np.arctan(xo)))
    bind = p[gp.P["hotHaloRamPressureStripping.formFactor"]] * v_vir ** 2 * rho0
    npar = b.node_par.copy()
    npar[NP["TIME_REFERENCE"]] = t_iso
    npar[NP["MASS_RATE"]] = 0.0
    npar[NP["SPIN_RATE"]] = 0.0
    npar[NP["IS_SATELLITE"]] = 1.0
    npar[NP["RAM_PRESSURE_FORCE"]] = bind * 10.0 ** rng.uniform(-3.0, math.log10(0.3), n)
    npar[NP["HOST_VELOCITY_VIRIAL"]] = v_vir * rng.uniform(1.5, 6.0, n)
    return NodeBatch(y, npar, b.t_start, b.t_end, b.h_step)


def closed_box(timescale=0.25, outflow=0.0, recycled=0.4, yield_=0.025):
    """The reference's closed/leaky-box single-node known answers
    (testSuite/test-reproducibility.py:47-68, testSuite/parameters/reproducibility/{closed,leaky}Box*.xml):
    one disk with 1e11 Msun of metal-free gas at t=12.47 Gyr evolved to 13.47 Gyr."""
    p = gp.physics_params(physicsMode=gp.PHYSICS_BOX, OmegaMatter=0.3, OmegaDarkEnergy=0.7, OmegaBaryon=0.05,
                          HubbleConstant=70.0, recycledFraction=recycled, metalYield=yield_,
                          **{"starFormationTimescale.fixed.timescale": timescale,
                             "stellarFeedbackOutflows.fixed.fraction": outflow})
    y = np.zeros((NV, 1))
    y[V["DISK_MASS_GAS"], 0] = 1.0e11
    y[V["DISK_FRACTION_MASS_RETAINED"], 0] = 1.0
    y[V["SPIN_ANGULAR_MOMENTUM"], 0] = 1.635e12
    y[V["SATELLITE_BOUND_MASS"], 0] = 1.0e12
    npar = np.zeros((gp.NPARAM_NODE, 1))
    npar[NP["TIME_REFERENCE"], 0] = 12.47
    npar[NP["MASS_REFERENCE"], 0] = 1.0e12
    return p, NodeBatch(y, npar, np.array([12.47]), np.array([13.47]), -np.ones(1))


def rapid_depletion():
    """The reference's non-negativity model (testSuite/parameters/enforceNonNegativity/
    rapidDepletion.xml with tree.xml): a 1e12 Msun halo at t=12.47 Gyr whose disk
    (1e11 Msun of stars, 1e10 of gas) and spheroid (1e9 Msun of gas) form stars on
    a fixed 0.25 Gyr timescale with fixed outflow fractions 1 (disk) and 10
    (spheroid), no hot halo, and no post-step zeroing of negative masses,
    evolved to 13.47 Gyr."""
    p = gp.physics_params(physicsMode=gp.PHYSICS_BOX, OmegaMatter=0.3, OmegaDarkEnergy=0.7, OmegaBaryon=0.05,
                          HubbleConstant=70.0, recycledFraction=0.4, metalYield=0.025,
                          postStepZeroNegativeMasses=0.0,
                          **{"starFormationTimescale.fixed.timescale": 0.25,
                             "stellarFeedbackOutflows.fixed.fraction": 1.0,
                             "stellarFeedbackOutflows.fixed.fractionSpheroid": 10.0})
    y = np.zeros((NV, 1))
    y[V["DISK_MASS_STELLAR"], 0] = 1.0e11
    y[V["DISK_MASS_GAS"], 0] = 1.0e10
    y[V["DISK_ANGULAR_MOMENTUM"], 0] = 7.7e10
    y[V["DISK_FRACTION_MASS_RETAINED"], 0] = 1.0
    y[V["SPHEROID_MASS_GAS"], 0] = 1.0e9
    y[V["SPHEROID_ANGULAR_MOMENTUM"], 0] = 7.7e8
    y[V["SPIN_ANGULAR_MOMENTUM"], 0] = 1.635e12
    y[V["SATELLITE_BOUND_MASS"], 0] = 1.0e12
    npar = np.zeros((gp.NPARAM_NODE, 1))
    npar[NP["TIME_REFERENCE"], 0] = 12.47
    npar[NP["MASS_REFERENCE"], 0] = 1.0e12
    return p, NodeBatch(y, npar, np.array([12.47]), np.array([13.47]), -np.ones(1))


def hubble_time(t, om=0.3, ol=0.7, h0=70.0):
    """1/H(t) in Gyr for flat LCDM."""
    a = cosmic_time_to_scale_factor(t, om, ol, h0)
    return MPC_PER_KMS_GYR / (h0 * np.sqrt(om / a ** 3 + ol))


def simple_timestep(t, om=0.3, ol=0.7, h0=70.0, relative=0.1, absolute=1.0):
    """mergerTreeEvolveTimestepSimple (merger_trees/evolve/timesteps/simple.F90:124-155):
    min(timeStepRelative / H(t), timeStepAbsolute)."""
    return np.minimum(relative * hubble_time(t, om, ol, h0), absolute)


@dataclasses.dataclass
class HaloHistories:
    """Initial galaxies on growing halos, evolved interval by interval to z=0."""
    batch: NodeBatch       # state at each history's start; t_end = time at z=0
    t_final: float
    seed: int

    def intervals(self, t, rng):
        """Next interval end for histories now at time t: the simple timestep,
        shortened by a uniform factor in [0.25, 1] that stands in for the
        clipping of intervals at the times of tree events, then clipped at z=0."""
        p = gp.physics_params()
        om, ol, h0 = p[gp.P["OmegaMatter"]], p[gp.P["OmegaDarkEnergy"]], p[gp.P["HubbleConstant"]]
        dt = simple_timestep(t, om, ol, h0) * rng.uniform(0.25, 1.0, np.shape(t))
        return np.minimum(t + dt, self.t_final)


def halo_histories(n: int, seed: int = 0, params=None, start=None) -> HaloHistories:
    """Node histories for the surrogate: each node's halo grows linearly from
    M_s at its start time t_s to M_0 at z=0, with M_0 log-uniform in
    [1e10, 1e13] Msun and log10(M_0/M_s) uniform in [0.3, 1.5] scaled by the
    fraction of the z=5..0 time span still ahead; t_s is uniform between z=5
    and 1 Gyr before z=0 (or fixed at ``start``).  The galaxy at t_s is drawn as
    in ``isolated_halos`` with the halo at t_s."""
    p = gp.physics_params() if params is None else params
    om, ol, h0 = p[gp.P["OmegaMatter"]], p[gp.P["OmegaDarkEnergy"]], p[gp.P["HubbleConstant"]]
    fb = p[gp.P["OmegaBaryon"]] / om
    t_final = float(scale_factor_to_cosmic_time(1.0, om, ol, h0))
    t_z5 = float(scale_factor_to_cosmic_time(1.0 / 6.0, om, ol, h0))
    rng = np.random.Generator(np.random.PCG64(seed))
    ts = rng.uniform(t_z5, t_final - 1.0, n) if start is None else np.full(n, float(start))
    m0 = 10.0 ** rng.uniform(10.0, 13.0, n)
    ahead = (t_final - ts) / (t_final - t_z5)
    ms = m0 * 10.0 ** (-rng.uniform(0.3, 1.5, n) * ahead)
    m_rate = (m0 - ms) / (t_final - ts)
    f_hot = rng.uniform(0.05, 0.17, n)
    metal = 10.0 ** rng.uniform(-4.0, math.log10(0.02), n)
    spin = np.clip(np.exp(rng.normal(math.log(0.035), 0.5, n)), 0.01, 0.1)
    f_out = rng.uniform(0.0, 0.02, n)
    f_dg = rng.uniform(0.0, 0.1, n)
    f_ds = rng.uniform(0.0, 0.1, n)
    f_sg = rng.uniform(0.0, 0.02, n) * (rng.uniform(size=n) < 0.5)
    f_ss = rng.uniform(0.0, 0.1, n) * (rng.uniform(size=n) < 0.5)
    bh_spin = rng.uniform(0.0, 0.9, n)
    r_vir, v_vir = virial(ms, ts, om, ol, h0)
    j_spin = spin * ms * v_vir * r_vir
    r_disk = p[gp.P["galacticStructureSolver.factor"]] * r_vir * spin
    y = np.zeros((NV, n))
    mb = fb * ms
    y[V["HOT_MASS"]] = f_hot * ms
    y[V["HOT_ABUNDANCES"]] = metal * y[V["HOT_MASS"]]
    y[V["HOT_ANGULAR_MOMENTUM"]] = y[V["HOT_MASS"]] * j_spin / ms
    y[V["HOT_OUTFLOWED_MASS"]] = f_out * mb
    y[V["HOT_OUTFLOWED_ABUNDANCES"]] = metal * y[V["HOT_OUTFLOWED_MASS"]]
    y[V["HOT_OUTFLOWED_ANGULAR_MOMENTUM"]] = y[V["HOT_OUTFLOWED_MASS"]] * j_spin / ms
    y[V["HOT_OUTER_RADIUS"]] = r_vir
    y[V["DISK_MASS_GAS"]] = f_dg * mb
    y[V["DISK_MASS_STELLAR"]] = f_ds * mb
    y[V["DISK_MASS_STELLAR_FORMED"]] = y[V["DISK_MASS_STELLAR"]]
    y[V["DISK_FRACTION_MASS_RETAINED"]] = 1.0
    y[V["DISK_ABUNDANCES_GAS"]] = metal * y[V["DISK_MASS_GAS"]]
    y[V["DISK_ABUNDANCES_STELLAR"]] = metal * y[V["DISK_MASS_STELLAR"]]
    y[V["DISK_ANGULAR_MOMENTUM"]] = (y[V["DISK_MASS_GAS"]] + y[V["DISK_MASS_STELLAR"]]) * 2.0 * r_disk * v_vir
    y[V["SPHEROID_MASS_GAS"]] = f_sg * mb
    y[V["SPHEROID_MASS_STELLAR"]] = f_ss * mb
    y[V["SPHEROID_MASS_STELLAR_FORMED"]] = y[V["SPHEROID_MASS_STELLAR"]]
    y[V["SPHEROID_ABUNDANCES_GAS"]] = metal * y[V["SPHEROID_MASS_GAS"]]
    y[V["SPHEROID_ABUNDANCES_STELLAR"]] = metal * y[V["SPHEROID_MASS_STELLAR"]]
    y[V["SPHEROID_ANGULAR_MOMENTUM"]] = (y[V["SPHEROID_MASS_GAS"]] + y[V["SPHEROID_MASS_STELLAR"]]) * 0.1 * r_disk * v_vir
    y[V["BH_MASS"]] = 100.0 + 1.0e-3 * y[V["SPHEROID_MASS_STELLAR"]]
    y[V["BH_SPIN"]] = bh_spin
    y[V["SPIN_ANGULAR_MOMENTUM"]] = j_spin
    y[V["SATELLITE_BOUND_MASS"]] = ms
    npar = np.zeros((gp.NPARAM_NODE, n))
    npar[NP["TIME_REFERENCE"]] = ts
    npar[NP["MASS_REFERENCE"]] = ms
    npar[NP["MASS_RATE"]] = m_rate
    npar[NP["SPIN_RATE"]] = j_spin * m_rate / ms
    npar[NP["IS_SATELLITE"]] = 0.0
    return HaloHistories(NodeBatch(y, npar, ts, np.full(n, t_final), -np.ones(n)), t_final, seed)
