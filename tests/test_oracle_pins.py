"""Closed-form pins of the oracle's RHS terms, one process at a time.

Each test sets up one node in which the process under test is the only one
that moves a chosen variable (other reservoirs empty, or the process isolated
as the difference of two RHS evaluations with the node operator switched on
and off), and compares the oracle's rate with a value the test derives from
the physics: a special case that fixes the answer (V = 200 km/s in the
dynamical-time star-formation law, V = V_0 in the power-law outflow, the
rate-limit cap, a Schwarzschild black hole), a textbook formula evaluated here
with its own constants (Bondi-Hoyle, Eddington, Meier's jet power, the
radiated energy of a uniform sphere), or a derivative taken numerically
(virial-radius growth, cooling-radius growth).  A dropped factor, a wrong
exponent or a transposed operand in the oracle fails one of them.

References: star_formation/timescales/dynamical_time.F90:164-188,
stellar_feedback/outflows/power_law/_class.F90:118-164, rate_limit.F90:111-170,
hot_halo/outflow_reincorporation/halo_dynamical_time.F90:113-128,
cooling/cooling_rate/White-Frenk.F90:132-185, cooling_radius/simple.F90:229-427,
black_holes/accretion_rates/standard.F90:236-440, Bondi_Hoyle_Lyttleton.F90:34-78,
accretion_disks/Shakura_Sunyaev.F90:70-179, cooling_energy_radiated.F90:200-243,
bar_instability.F90:145-248 + galactic_dynamics/bar_instability/Efstathiou1982.F90:153-258,
black_holes/winds.F90:102-133 + winds/Ciotti2009.F90:113-193 +
spheroid/standard/_class.F90:673-725, satellite/mass_loss/_class.F90:176-200,
circumgalactic_medium/outer_radius/ram_pressure_stripping.F90:111-242,
hot_halo/ram_pressure_stripping/Font2008.F90:115-267,
ram_pressure_stripping/timescale/ram_pressure_acceleration.F90:79-134,
hot_halo/standard/_class.F90:598-720, circumgalactic_medium/cooling_heating.F90:216-240.
"""
import math

import numpy as np
import pytest

import oracle
from galacticus_amd import params as gp
from galacticus_amd import workloads as wl
from galacticus_amd.cooling import synthetic_cie_table

V = wl.V
NP = wl.NP

# Constants, SI unless stated (CODATA 2018 / IAU 2015), restated here.
PARSEC = 3.0856775814913673e16
MPC = 1.0e6 * PARSEC
GYR = 1.0e9 * 3.15581497635456e7
MSUN = 1.98847e30
G_SI = 6.67430e-11
G = G_SI * MSUN / 1.0e6 / MPC                  # Mpc (km/s)^2 / Msun
KMS_MPC_GYR = MPC / 1.0e3 / GYR                # Gyr per (Mpc / (km/s))
AMU = 1.66053906660e-27
M_H = 1.0078250322 * AMU
K_B = 1.380649e-23
C_LIGHT = 2.99792458e8
SIGMA_T = 6.6524587321e-29
MU = 1.0 / (2.0 * 0.7514 / 1.0078250322 + 3.0 * 0.2486 / 4.002602)


@pytest.fixture(scope="module")
def table():
    return synthetic_cie_table()


def _node(mass, t, spin=0.035, satellite=False, m_rate=0.0):
    """An empty node (every reservoir zero) of halo mass `mass` at time t."""
    r_vir, v_vir = wl.virial(mass, t)
    y = np.zeros((30, 1))
    y[V["SPIN_ANGULAR_MOMENTUM"], 0] = spin * mass * v_vir * r_vir
    y[V["HOT_OUTER_RADIUS"], 0] = r_vir
    npar = np.zeros((gp.NPARAM_NODE, 1))
    npar[NP["TIME_REFERENCE"]] = t
    npar[NP["MASS_REFERENCE"]] = mass
    npar[NP["MASS_RATE"]] = m_rate
    npar[NP["IS_SATELLITE"]] = 1.0 if satellite else 0.0
    return y, npar, float(r_vir), float(v_vir)


def _mass_for_velocity(v_target, t):
    """Halo mass whose virial velocity is v_target at time t (V ~ M^(1/3))."""
    m = 1e12
    for _ in range(60):
        _, v = wl.virial(m, t)
        m *= (v_target / float(v)) ** 3
    return m


def _rhs(y, npar, t, table, **over):
    return oracle.rhs(np.array([t]), y, npar, gp.physics_params(**over), table)[:, 0]


T0 = 8.0


# ---- star formation: dynamical-time timescale --------------------------------
def test_star_formation_timescale_at_200_kms(table):
    """At V = 200 km/s the (V/200)^alpha factor is 1 whatever alpha: SFR = M_gas eps V / r."""
    m = _mass_for_velocity(200.0, T0)
    y, npar, rv, vv = _node(m, T0)
    assert abs(vv - 200.0) < 1e-9
    y[V["DISK_MASS_GAS"], 0] = 1e9
    y[V["DISK_ANGULAR_MOMENTUM"], 0] = 1e9 * 0.1 * vv
    r_disk = math.sqrt(0.5) * rv * 0.035
    for alpha in (-1.5, 0.7, 2.0):
        f = _rhs(y, npar, T0, table, nodeOperators=0, **{"starFormationDisks.exponentVelocity": alpha})
        expected = 1e9 * 0.01 * vv / (r_disk * KMS_MPC_GYR)
        assert f[V["DISK_MASS_STELLAR_FORMED"]] == pytest.approx(expected, rel=1e-12)


def test_star_formation_velocity_exponent(table):
    """SFR t_dyn / (eps M_gas) = (V/200)^(-alpha): the exponent's sign and value."""
    for v_target, alpha in ((100.0, -1.5), (400.0, -1.5), (60.0, 2.0)):
        m = _mass_for_velocity(v_target, T0)
        y, npar, rv, vv = _node(m, T0)
        y[V["DISK_MASS_GAS"], 0] = 1e8
        y[V["DISK_ANGULAR_MOMENTUM"], 0] = 1e8 * 0.1 * vv
        f = _rhs(y, npar, T0, table, nodeOperators=0, **{"starFormationDisks.exponentVelocity": alpha,
                                                          "starFormationDisks.timescaleMinimum": 0.0})
        t_dyn = math.sqrt(0.5) * rv * 0.035 / vv * KMS_MPC_GYR
        got = f[V["DISK_MASS_STELLAR_FORMED"]] * t_dyn / (0.01 * 1e8)
        assert got == pytest.approx((vv / 200.0) ** (-alpha), rel=1e-12)


def test_star_formation_timescale_minimum(table):
    y, npar, rv, vv = _node(_mass_for_velocity(200.0, T0), T0)
    y[V["DISK_MASS_GAS"], 0] = 1e8
    y[V["DISK_ANGULAR_MOMENTUM"], 0] = 1e8 * 0.1 * vv
    f = _rhs(y, npar, T0, table, nodeOperators=0, **{"starFormationDisks.timescaleMinimum": 5.0})
    assert f[V["DISK_MASS_STELLAR_FORMED"]] == pytest.approx(1e8 / 5.0, rel=1e-14)


# ---- feedback: power law and rate limit --------------------------------------
def test_outflow_equals_sfr_at_characteristic_velocity(table):
    """(V0/V)^alpha = 1 at V = V0 (disk 250 km/s): outflow rate = SFR."""
    y, npar, rv, vv = _node(_mass_for_velocity(250.0, T0), T0)
    y[V["DISK_MASS_GAS"], 0] = 1e9
    y[V["DISK_ANGULAR_MOMENTUM"], 0] = 1e9 * 0.1 * vv
    f = _rhs(y, npar, T0, table, nodeOperators=0)
    sfr = f[V["DISK_MASS_STELLAR_FORMED"]]
    assert sfr > 0.0
    assert f[V["HOT_OUTFLOWED_MASS"]] == pytest.approx(sfr, rel=1e-12)
    # and the gas budget: -(1-R) SFR - outflow
    assert f[V["DISK_MASS_GAS"]] == pytest.approx(-(1.0 - 0.46) * sfr - sfr, rel=1e-12)


def test_outflow_power_law_exponent(table):
    """At V = V0/2 the outflow is 2^alpha times the SFR (alpha = 3.5)."""
    y, npar, rv, vv = _node(_mass_for_velocity(125.0, T0), T0)
    y[V["DISK_MASS_GAS"], 0] = 1e9
    y[V["DISK_ANGULAR_MOMENTUM"], 0] = 1e9 * 0.1 * vv
    f = _rhs(y, npar, T0, table, nodeOperators=0)
    assert f[V["HOT_OUTFLOWED_MASS"]] == pytest.approx(2.0 ** 3.5 * f[V["DISK_MASS_STELLAR_FORMED"]], rel=1e-12)


def test_outflow_rate_limit_cap(table):
    """A large fractional minimum time makes the cap M_gas / (t_dyn f) bind."""
    y, npar, rv, vv = _node(_mass_for_velocity(50.0, T0), T0)
    y[V["DISK_MASS_GAS"], 0] = 1e9
    y[V["DISK_ANGULAR_MOMENTUM"], 0] = 1e9 * 0.1 * vv
    frac = 1e4
    f = _rhs(y, npar, T0, table, nodeOperators=0,
             **{"stellarFeedbackDisks.timescaleOutflowFractionalMinimum": frac})
    t_dyn = math.sqrt(0.5) * rv * 0.035 / vv * KMS_MPC_GYR
    assert f[V["HOT_OUTFLOWED_MASS"]] == pytest.approx(1e9 / (t_dyn * frac), rel=1e-12)


def test_outflow_angular_momentum_boost(table):
    """Outflowing angular momentum enters the outflowed reservoir divided by
    (1 - fractionLossAngularMomentum), here 1 - 0.3 (hot_halo/standard/_class.F90:690-693)."""
    y, npar, rv, vv = _node(_mass_for_velocity(250.0, T0), T0)
    y[V["DISK_MASS_GAS"], 0] = 1e9
    y[V["DISK_MASS_STELLAR"], 0] = 3e9
    y[V["DISK_ANGULAR_MOMENTUM"], 0] = 4e9 * 0.1 * vv
    f = _rhs(y, npar, T0, table, nodeOperators=0)
    out = f[V["HOT_OUTFLOWED_MASS"]]
    j_lost = out * y[V["DISK_ANGULAR_MOMENTUM"], 0] / 4e9
    assert f[V["HOT_OUTFLOWED_ANGULAR_MOMENTUM"]] == pytest.approx(j_lost / 0.7, rel=1e-12)


# ---- reincorporation ----------------------------------------------------------
def test_reincorporation_rate(table):
    y, npar, rv, vv = _node(1e12, T0)
    y[V["HOT_OUTFLOWED_MASS"], 0] = 3e9
    y[V["HOT_OUTFLOWED_ABUNDANCES"], 0] = 3e7
    f = _rhs(y, npar, T0, table, nodeOperators=0)
    t_dyn = rv / vv * KMS_MPC_GYR
    assert f[V["HOT_OUTFLOWED_MASS"]] == pytest.approx(-3e9 * 5.0 / t_dyn, rel=1e-12)
    assert f[V["HOT_MASS"]] == pytest.approx(3e9 * 5.0 / t_dyn, rel=1e-12)
    assert f[V["HOT_ABUNDANCES"]] == pytest.approx(3e7 * 5.0 / t_dyn, rel=1e-12)


# ---- cooling -------------------------------------------------------------------
def _hot_node(mass, t, f_hot, metallicity, table):
    y, npar, rv, vv = _node(mass, t)
    y[V["HOT_MASS"], 0] = f_hot * mass
    y[V["HOT_ABUNDANCES"], 0] = metallicity * f_hot * mass
    y[V["HOT_ANGULAR_MOMENTUM"], 0] = f_hot * mass * 0.035 * rv * vv
    return y, npar, rv, vv


def test_white_frenk_rapid_cooling_limit(table):
    """r_cool >= r_outer (short cooling time): the rate is M_hot / t_dyn."""
    y, npar, rv, vv = _hot_node(3e10, 2.0, 0.15, 0.02, table)
    props = oracle.halo_properties(gp.physics_params(), table, npar[:, 0], 2.0, y[:, 0])
    assert props["r_cool"] == pytest.approx(rv)
    f = _rhs(y, npar, 2.0, table, nodeOperators=0)
    t_dyn = rv / vv * KMS_MPC_GYR
    assert f[V["DISK_MASS_GAS"]] == pytest.approx(y[V["HOT_MASS"], 0] / t_dyn, rel=1e-12)


def test_white_frenk_interior_rate_and_radius_growth(table):
    """Inside the halo: rate = 4 pi r_c^2 rho(r_c) dr_c/dt, with dr_c/dt the
    derivative of the cooling radius with respect to the time available (which
    grows at rate 1), taken here numerically from the cooling time t_cool =
    K / rho of the beta profile."""
    p = gp.physics_params()
    y, npar, rv, vv = _hot_node(1e12, T0, 0.1, 0.02, table)
    props = oracle.halo_properties(p, table, npar[:, 0], T0, y[:, 0])
    rc_core = 0.3 * rv
    xo = rv / rc_core
    rho0 = y[V["HOT_MASS"], 0] / (4.0 * math.pi * rc_core ** 3 * (xo - math.atan(xo)))
    rho = lambda r: rho0 / (1.0 + (r / rc_core) ** 2)
    t_vir = 0.5 * AMU * MU * (1e3 * vv) ** 2 / K_B
    zf = 0.02
    k_const = oracle.cooling_time(p, table, t_vir, 1.0, zf)       # t_cool = K / rho
    r_of = lambda tau: rc_core * math.sqrt(rho0 * tau / k_const - 1.0)
    t_avail = rv / vv * KMS_MPC_GYR
    r_c = r_of(t_avail)
    assert 0.0 < r_c < rv
    assert props["r_cool"] == pytest.approx(r_c, rel=1e-9)
    d = 1e-6 * t_avail
    rdot = (r_of(t_avail + d) - r_of(t_avail - d)) / (2.0 * d)
    assert props["r_cool_rate"] == pytest.approx(rdot, rel=1e-7)
    f = _rhs(y, npar, T0, table, nodeOperators=0, **{"blackHoleCGMHeating.efficiencyRadioMode": 0.0})
    assert f[V["DISK_MASS_GAS"]] == pytest.approx(4.0 * math.pi * r_c ** 2 * rho(r_c) * rdot, rel=1e-7)


def test_energy_radiated_uniform_sphere(table):
    """Lambda(T, Z) n_H n_tot V for the hot gas spread uniformly inside r_outer."""
    p = gp.physics_params()
    y, npar, rv, vv = _hot_node(1e12, T0, 0.1, 0.004, table)
    f = _rhs(y, npar, T0, table, nodeOperators=0)
    t_vir = 0.5 * AMU * MU * (1e3 * vv) ** 2 / K_B
    m = y[V["HOT_MASS"], 0]
    r_cm = rv * MPC * 100.0
    vol = 4.0 / 3.0 * math.pi * r_cm ** 3
    rho = m * MSUN / (vol * 1e-6)                       # kg / m^3
    x_h = (1.0 - 0.004) * 0.7514 / (0.7514 + 0.2486)
    x_he = (1.0 - 0.004) * 0.2486 / (0.7514 + 0.2486)
    n_h = rho * x_h / M_H * 1e-6                         # cm^-3
    n_he = n_h * (x_he / 4.002602) / (x_h / 1.0078250322)
    n_tot = n_h + n_he + (n_h + 2.0 * n_he)             # ions + electrons, fully ionised
    lam = oracle.cooling_function(table, t_vir, 0.004 / 0.0188)
    assert f[V["HOT_ENERGY_RADIATED"]] == pytest.approx(lam * n_h * n_tot * vol, rel=1e-10)


# ---- black holes ---------------------------------------------------------------
def _sound_speed(temperature):
    return math.sqrt(5.0 / 3.0 * K_B * temperature / (MU * AMU)) / 1e3     # km/s


def _bh_node(bh_mass, gas, table, mass=1e12, t=T0):
    y, npar, rv, vv = _node(mass, t)
    y[V["BH_MASS"], 0] = bh_mass
    y[V["SPHEROID_MASS_GAS"], 0] = gas
    y[V["SPHEROID_ANGULAR_MOMENTUM"], 0] = gas * 0.01 * rv * vv
    return y, npar, rv, vv


def _mdot_from_spin_rate(f, bh_mass):
    # Schwarzschild: d(spin)/dt = L_isco Mdot / M with L_isco = 2 sqrt(3)
    return f[V["BH_SPIN"]] * bh_mass / (2.0 * math.sqrt(3.0))


def _eddington(bh_mass):
    return 4.0 * math.pi * G_SI * bh_mass * M_H * GYR / (SIGMA_T * C_LIGHT)       # Msun / Gyr


def test_bondi_hoyle_spheroid_rate(table):
    """Mdot = 5 x 4 pi G^2 M^2 rho / c_s^3 at the Bondi radius of 100 K gas, with
    rho the Hernquist density of the spheroid gas there."""
    mbh, gas = 1e5, 1e8
    y, npar, rv, vv = _bh_node(mbh, gas, table)
    f = _rhs(y, npar, T0, table, nodeOperators=0)
    cs = _sound_speed(100.0)
    r_b = G * mbh / cs ** 2
    a = math.sqrt(0.5) * rv * 0.035
    rho_b = gas * a / (2.0 * math.pi * r_b * (r_b + a) ** 3)
    l_j = min(cs / math.sqrt(G * rho_b), a)
    rho = gas * a / (2.0 * math.pi * l_j * (l_j + a) ** 3) if l_j > r_b else rho_b
    mdot = 5.0 * 4.0 * math.pi * (G * mbh) ** 2 * rho / cs ** 3 / KMS_MPC_GYR
    eff = 1.0 - math.sqrt(8.0 / 9.0)
    mdot = min(mdot, _eddington(mbh) / eff)
    assert _mdot_from_spin_rate(f, mbh) == pytest.approx(mdot, rel=1e-9)


def test_eddington_limit_and_meier_jet(table):
    """A small hole in dense gas accretes at L_Edd / (eff c^2); Meier's (2001)
    Schwarzschild jet power is 10^41.7 erg/s (M/1e9)^0.9 (Mdot/Mdot_Edd)^1.2 and
    the hole keeps Mdot (1 - eff - P_jet / (Mdot c^2))."""
    mbh, gas = 1e3, 1e10
    y, npar, rv, vv = _bh_node(mbh, gas, table, mass=1e11)
    f = _rhs(y, npar, T0, table, nodeOperators=0)
    eff = 1.0 - math.sqrt(8.0 / 9.0)
    mdot = _eddington(mbh) / eff
    assert _mdot_from_spin_rate(f, mbh) == pytest.approx(mdot, rel=1e-9)
    p_erg_s = 10.0 ** 41.7 * (mbh / 1e9) ** 0.9 * (mdot / _eddington(mbh)) ** 1.2
    p_jet = p_erg_s * 1e-7 * GYR / MSUN / 1e6            # Msun (km/s)^2 / Gyr
    c_kms = C_LIGHT / 1e3
    assert f[V["BH_MASS"]] == pytest.approx(mdot * (1.0 - eff) - p_jet / c_kms ** 2, rel=1e-9)


# ---- bar instability -------------------------------------------------------------
def _disk(y, rv, vv, m_gas, m_star):
    y[V["DISK_MASS_GAS"], 0] = m_gas
    y[V["DISK_MASS_STELLAR"], 0] = m_star
    y[V["DISK_ABUNDANCES_GAS"], 0] = 0.01 * m_gas
    y[V["DISK_ABUNDANCES_STELLAR"], 0] = 0.02 * m_star
    y[V["DISK_ANGULAR_MOMENTUM"], 0] = (m_gas + m_star) * 0.05 * rv * vv


def _bar_difference(y, npar, table):
    f_on = _rhs(y, npar, T0, table, nodeOperators=gp.OPERATOR_BITS["barInstability"])
    f_off = _rhs(y, npar, T0, table, nodeOperators=0)
    return f_on - f_off


def test_bar_instability_stable_disk_moves_nothing(table):
    y, npar, rv, vv = _node(1e12, T0)
    r_d = math.sqrt(0.5) * rv * 0.035
    m_stable = 0.5 * r_d * vv ** 2 * (1.1800237580 / 1.1) ** 2 / G
    _disk(y, rv, vv, 0.0, m_stable)
    assert np.all(_bar_difference(y, npar, table) == 0.0)


@pytest.mark.parametrize("gas_fraction", [0.0, 0.3, 1.0])
def test_bar_instability_timescale(table, gas_fraction):
    """eps = 1.18 V / sqrt(G M_d / r_d) below eps_c = 1.1 (1 - f_g) + 0.7 f_g:
    everything but the disk's angular momentum moves to the spheroid on
    t_disk ((eps_c - 0.6221297315) / (eps_c - eps))^2."""
    y, npar, rv, vv = _node(1e12, T0)
    r_d = math.sqrt(0.5) * rv * 0.035
    eps_c = 1.1 * (1.0 - gas_fraction) + 0.7 * gas_fraction
    eps = 0.8 * eps_c
    m_d = r_d * vv ** 2 * (1.1800237580 / eps) ** 2 / G
    _disk(y, rv, vv, gas_fraction * m_d, (1.0 - gas_fraction) * m_d)
    d = _bar_difference(y, npar, table)
    eps = max(eps, 0.6221297315)          # an isolated disk's estimator is the floor
    tau = r_d / vv * KMS_MPC_GYR * ((eps_c - 0.6221297315) / (eps_c - eps)) ** 2
    for disk, sph in (("DISK_MASS_GAS", "SPHEROID_MASS_GAS"), ("DISK_MASS_STELLAR", "SPHEROID_MASS_STELLAR"),
                      ("DISK_ABUNDANCES_GAS", "SPHEROID_ABUNDANCES_GAS"),
                      ("DISK_ABUNDANCES_STELLAR", "SPHEROID_ABUNDANCES_STELLAR")):
        assert d[V[disk]] == pytest.approx(-y[V[disk], 0] / tau, rel=1e-10, abs=1e-300)
        assert d[V[sph]] == pytest.approx(y[V[disk], 0] / tau, rel=1e-10, abs=1e-300)
    # retained fractions 1: the disk keeps its angular momentum, the spheroid gains J_d / tau
    assert d[V["DISK_ANGULAR_MOMENTUM"]] == 0.0
    assert d[V["SPHEROID_ANGULAR_MOMENTUM"]] == pytest.approx(y[V["DISK_ANGULAR_MOMENTUM"], 0] / tau, rel=1e-10)
    assert d[V["DISK_MASS_STELLAR_FORMED"]] == 0.0


def test_bar_instability_dimensionless_cap(table):
    """Just below the threshold the squared ratio is capped at 1e10."""
    y, npar, rv, vv = _node(1e12, T0)
    r_d = math.sqrt(0.5) * rv * 0.035
    eps = 1.1 * (1.0 - 1e-12)
    m_d = r_d * vv ** 2 * (1.1800237580 / eps) ** 2 / G
    _disk(y, rv, vv, 0.0, m_d)
    d = _bar_difference(y, npar, table)
    tau = r_d / vv * KMS_MPC_GYR * 1e10
    assert d[V["DISK_MASS_STELLAR"]] == pytest.approx(-m_d / tau, rel=1e-6)


# ---- black-hole winds -----------------------------------------------------------
def _winds_difference(y, npar, table, **over):
    f_on = _rhs(y, npar, T0, table, nodeOperators=gp.OPERATOR_BITS["blackHolesWinds"], **over)
    f_off = _rhs(y, npar, T0, table, nodeOperators=0, **over)
    return f_on - f_off


# (black-hole mass, spheroid gas, efficiencyWind): wind pressure above,
# comparable to and below the ISM pressure.  The Bondi rate grows with the gas
# mass, so the pressure ratio is set by the wind efficiency until the
# Eddington cap binds.
WIND_CASES = ((1e6, 1e5, 0.1), (1e6, 1e5, 0.0024), (1e6, 1e11, 0.0024))


def _wind_pressure_ratio(y, npar, rv, mbh, gas, efficiency, table):
    """(x = P_ISM / P_wind - 1/2, total accretion rate, wind efficiency)."""
    mdot = _mdot_from_spin_rate(_rhs(y, npar, T0, table, nodeOperators=0), mbh)
    eff = efficiency * (1.0 - math.sqrt(8.0 / 9.0))
    r = math.sqrt(0.5) * rv * 0.035 * MPC
    p_wind = eff * mdot * MSUN / GYR * C_LIGHT ** 2 / (4.0 * math.pi) / 1.0e7 / r ** 2
    p_ism = 3.0 / (4.0 * math.pi) * gas * MSUN / r ** 3 / M_H * 1.5 * K_B * 1.0e4
    return p_ism / p_wind - 0.5, mdot, eff


@pytest.mark.parametrize("mbh,gas,efficiency", WIND_CASES)
def test_black_hole_wind_outflow(table, mbh, gas, efficiency):
    """P = efficiencyWind eff_rad Mdot c^2 f(x), f the smoothstep of
    x = P_ISM / P_wind - 1/2; the spheroid loses 0.01 P / V^2 of gas to the
    outflowed reservoir."""
    y, npar, rv, vv = _bh_node(mbh, gas, table)
    y[V["SPHEROID_MASS_STELLAR"], 0] = 2.0 * gas
    y[V["SPHEROID_ABUNDANCES_GAS"], 0] = 0.01 * gas
    x, mdot, eff = _wind_pressure_ratio(y, npar, rv, mbh, gas, efficiency, table)
    assert mdot > 0.0
    coupled = 0.0 if x <= 0.0 else (1.0 if x >= 1.0 else 3 * x * x - 2 * x ** 3)
    power = eff * coupled * mdot * (C_LIGHT / 1e3) ** 2
    mo = 0.01 * power / vv ** 2
    d = _winds_difference(y, npar, table, **{"blackHoleWind.efficiencyWind": efficiency})
    assert d[V["SPHEROID_MASS_GAS"]] == pytest.approx(-mo, rel=1e-9, abs=1e-300)
    assert d[V["HOT_OUTFLOWED_MASS"]] == pytest.approx(mo, rel=1e-9, abs=1e-300)
    assert d[V["SPHEROID_ANGULAR_MOMENTUM"]] == pytest.approx(-mo / (3.0 * gas) * y[V["SPHEROID_ANGULAR_MOMENTUM"], 0],
                                                              rel=1e-9, abs=1e-300)
    assert d[V["HOT_OUTFLOWED_ANGULAR_MOMENTUM"]] == pytest.approx(-d[V["SPHEROID_ANGULAR_MOMENTUM"]] / 0.7,
                                                                   rel=1e-9, abs=1e-300)


def test_black_hole_wind_coupling_cases_are_covered(table):
    """The three cases above span x <= 0, 0 < x < 1 and x >= 1."""
    xs = []
    for mbh, gas, efficiency in WIND_CASES:
        y, npar, rv, vv = _bh_node(mbh, gas, table)
        xs.append(_wind_pressure_ratio(y, npar, rv, mbh, gas, efficiency, table)[0])
    assert xs[0] <= 0.0 and 0.0 < xs[1] < 1.0 and xs[2] >= 1.0, xs


# ---- satellite mass loss ---------------------------------------------------------
def test_bound_mass_follows_basic_mass(table):
    y, npar, rv, vv = _node(1e12, T0, m_rate=3e10)
    f = _rhs(y, npar, T0, table)
    assert f[V["SATELLITE_BOUND_MASS"]] == 3e10
    npar[NP["IS_SATELLITE"]] = 1.0
    assert _rhs(y, npar, T0, table)[V["SATELLITE_BOUND_MASS"]] == 0.0
    npar[NP["IS_SATELLITE"]] = 0.0
    assert _rhs(y, npar, T0, table, nodeOperators=0)[V["SATELLITE_BOUND_MASS"]] == 0.0


# ---- CGM outer radius -----------------------------------------------------------
def test_central_outer_radius_follows_virial_radius(table):
    """dr_outer/dt = dr_vir/dt, the analytic growth of the virial radius of a
    halo growing at Mdot, checked against a numerical derivative of r_vir(t)."""
    p = gp.physics_params()
    mass, mdot = 1e12, 5e10
    y, npar, rv, vv = _node(mass, T0, m_rate=mdot)
    f = _rhs(y, npar, T0, table, nodeOperators=gp.OPERATOR_BITS["CGMOuterRadiusRamPressureStripping"])
    d = 1e-5
    r = lambda t: oracle.halo_properties(p, table, npar[:, 0], t, y[:, 0])["r_vir"]
    assert f[V["HOT_OUTER_RADIUS"]] == pytest.approx((r(T0 + d) - r(T0 - d)) / (2.0 * d), rel=1e-7)
    # a halo that stops growing still has r_vir shrinking... or growing with the falling density:
    npar[NP["MASS_RATE"]] = 0.0
    f0 = _rhs(y, npar, T0, table, nodeOperators=gp.OPERATOR_BITS["CGMOuterRadiusRamPressureStripping"])
    assert f0[V["HOT_OUTER_RADIUS"]] == pytest.approx((r(T0 + d) - r(T0 - d)) / (2.0 * d), rel=1e-6)
    assert f0[V["HOT_OUTER_RADIUS"]] > 0.0


def test_outer_radius_return_of_outflows(table):
    """Inside the virial radius, returning gas pushes the outer radius out at
    Mdot_return / (4 pi r^2 max(rho(r), rho_min))."""
    p = gp.physics_params()
    mass = 1e12
    y, npar, rv, vv = _node(mass, T0, satellite=True)
    ro = 0.6 * rv
    y[V["HOT_OUTER_RADIUS"], 0] = ro
    y[V["HOT_MASS"], 0] = 5e10
    y[V["HOT_OUTFLOWED_MASS"], 0] = 2e9
    bit = gp.OPERATOR_BITS["CGMOuterRadiusRamPressureStripping"]
    f = _rhs(y, npar, T0, table, nodeOperators=bit)
    rc = 0.3 * rv
    xo = ro / rc
    rho = 5e10 / (4.0 * math.pi * rc ** 3 * (xo - math.atan(xo))) / (1.0 + xo * xo)
    m_ret = 2e9 * 5.0 / (rv / vv * KMS_MPC_GYR)
    rho_min = 0.045 / 0.3 * mass / rv ** 3 / (4.0 * math.pi)
    assert rho > rho_min
    assert f[V["HOT_OUTER_RADIUS"]] == pytest.approx(m_ret / (4.0 * math.pi * ro ** 2 * rho), rel=1e-10)
    # with no hot gas the radius grows as r_vir Mdot_return / M
    y[V["HOT_MASS"], 0] = 0.0
    f = _rhs(y, npar, T0, table, nodeOperators=bit)
    assert f[V["HOT_OUTER_RADIUS"]] == pytest.approx(m_ret / mass * rv, rel=1e-12)


@pytest.mark.parametrize("target", [0.5, 0.9])
def test_satellite_ram_pressure_stripping(table, target):
    """The binding force formFactor V^2 rho(r) equals the ram pressure F at
    r = target r_outer; the radius shrinks at (r - r_o)/tau with
    tau = max(sqrt(2 r_o Sigma / F), r_o / (10 V_host)), Sigma = r_o rho(r_o),
    and the gas beyond goes to the stripped reservoir at 4 pi rho(r_o) r_o^2 dr/dt."""
    mass = 1e11
    y, npar, rv, vv = _node(mass, T0, satellite=True)
    ro = 0.8 * rv
    y[V["HOT_OUTER_RADIUS"], 0] = ro
    y[V["HOT_MASS"], 0] = 1e10
    y[V["HOT_ABUNDANCES"], 0] = 1e8
    y[V["HOT_ANGULAR_MOMENTUM"], 0] = 1e10 * 0.03 * rv * vv
    rc = 0.3 * rv
    xo = ro / rc
    rho0 = 1e10 / (4.0 * math.pi * rc ** 3 * (xo - math.atan(xo)))
    rho = lambda r: rho0 / (1.0 + (r / rc) ** 2)
    r_star = target * ro
    force = 2.0 * vv ** 2 * rho(r_star)
    npar[NP["RAM_PRESSURE_FORCE"]] = force
    npar[NP["HOST_VELOCITY_VIRIAL"]] = 3.0 * vv
    bit = gp.OPERATOR_BITS["CGMOuterRadiusRamPressureStripping"]
    d = _rhs(y, npar, T0, table, nodeOperators=bit) - _rhs(y, npar, T0, table, nodeOperators=0)
    tau = max(math.sqrt(2.0 * ro * ro * rho(ro) / force), ro / 10.0 / (3.0 * vv)) * KMS_MPC_GYR
    dr = (r_star - ro) / tau
    assert d[V["HOT_OUTER_RADIUS"]] == pytest.approx(dr, rel=1e-9)
    loss = 4.0 * math.pi * rho(ro) * ro ** 2 * dr
    assert d[V["HOT_MASS"]] == pytest.approx(loss, rel=1e-9)
    assert d[V["HOT_STRIPPED_MASS"]] == pytest.approx(-loss, rel=1e-9)
    assert d[V["HOT_STRIPPED_ABUNDANCES"]] == pytest.approx(-0.01 * loss, rel=1e-9)
    assert d[V["HOT_ANGULAR_MOMENTUM"]] == pytest.approx(loss / 1e10 * y[V["HOT_ANGULAR_MOMENTUM"], 0], rel=1e-9)


def test_satellite_ram_pressure_limits(table):
    """No force: nothing moves.  A force above the binding force at 1e-6 r_vir
    strips to zero radius."""
    mass = 1e11
    y, npar, rv, vv = _node(mass, T0, satellite=True)
    y[V["HOT_OUTER_RADIUS"], 0] = 0.8 * rv
    y[V["HOT_MASS"], 0] = 1e10
    y[V["HOT_ANGULAR_MOMENTUM"], 0] = 1e10 * 0.03 * rv * vv
    npar[NP["HOST_VELOCITY_VIRIAL"]] = 3.0 * vv
    bit = gp.OPERATOR_BITS["CGMOuterRadiusRamPressureStripping"]
    d = _rhs(y, npar, T0, table, nodeOperators=bit) - _rhs(y, npar, T0, table, nodeOperators=0)
    assert np.all(d == 0.0)
    rc = 0.3 * rv
    xo = 0.8 * rv / rc
    rho0 = 1e10 / (4.0 * math.pi * rc ** 3 * (xo - math.atan(xo)))
    npar[NP["RAM_PRESSURE_FORCE"]] = 2.0 * vv ** 2 * rho0 * 1.01
    d = _rhs(y, npar, T0, table, nodeOperators=bit) - _rhs(y, npar, T0, table, nodeOperators=0)
    ro = 0.8 * rv
    rho_o = rho0 / (1.0 + xo * xo)
    tau = max(math.sqrt(2.0 * ro * ro * rho_o / npar[NP["RAM_PRESSURE_FORCE"], 0]), ro / 10.0 / (3.0 * vv)) * KMS_MPC_GYR
    assert d[V["HOT_OUTER_RADIUS"]] == pytest.approx(-ro / tau, rel=1e-9)


# ---- excess heating -------------------------------------------------------------
def test_excess_heating_expels_gas_from_the_node(table):
    """Heating beyond cooling removes hot gas at min(heat - cool, M/t_dyn): for a
    central it leaves the node (no reservoir gains it), for a satellite it joins
    the stripped reservoir."""
    mbh, gas = 1e9, 1e10
    for satellite in (False, True):
        y, npar, rv, vv = _bh_node(mbh, gas, table, mass=1e13)
        y[V["HOT_MASS"], 0] = 1e11
        y[V["HOT_ABUNDANCES"], 0] = 1e9
        y[V["HOT_ANGULAR_MOMENTUM"], 0] = 1e11 * 0.03 * rv * vv
        npar[NP["IS_SATELLITE"]] = 1.0 if satellite else 0.0
        on = _rhs(y, npar, T0, table, nodeOperators=0)
        off = _rhs(y, npar, T0, table, nodeOperators=0, **{"CGMCoolingHeating.excessHeatDrivesOutflow": 0.0})
        expelled = off[V["HOT_MASS"]] - on[V["HOT_MASS"]]
        assert expelled > 0.0
        assert expelled <= 1e11 / (rv / vv * KMS_MPC_GYR) * (1.0 + 1e-12)
        assert on[V["HOT_OUTFLOWED_MASS"]] == off[V["HOT_OUTFLOWED_MASS"]]
        stripped = on[V["HOT_STRIPPED_MASS"]] - off[V["HOT_STRIPPED_MASS"]]
        assert stripped == (pytest.approx(expelled, rel=1e-12) if satellite else 0.0)
