"""Parameter handling: the reference's parameter-XML keys -> the engine's flat arrays.

The engine's physics parameters are one float64 array indexed by the
``GAL_P_*`` enumeration of ``include/gal_ode.h``.  Names are the reference's
parameter-XML element names; defaults are those of ``parameters/quickTest.xml``
(reference) where that file sets them and the class defaults otherwise:

* cosmology: ``cosmologyParameters`` (quickTest.xml:37-43) overridden to the
  Omega_m=0.3, Omega_L=0.7, h=0.7 cosmology of the node-evolution workloads;
* stellar populations: ``stellarPopulation`` recycledFraction=0.46,
  metalYield=0.035 (quickTest.xml:168-171);
* star formation: ``starFormationTimescale dynamicalTime`` with the class
  defaults efficiency=0.01, exponentVelocity=-1.5, timescaleMinimum=1e-3 for
  disks (star_formation/timescales/dynamical_time.F90:69-87) and 0.04/2.0/1e-3
  for spheroids (quickTest.xml:154-160);
* feedback: ``stellarFeedbackOutflows rateLimit(powerLaw)`` with
  (250 km/s, 3.5) for disks and (100 km/s, 3.5) for spheroids, fractional
  minimum timescale 1e-3 (quickTest.xml:240-257);
* CGM: haloDynamicalTime multiplier 5 (quickTest.xml:116-118), White-Frenk
  cooling with velocityCutOff 1e4 (quickTest.xml:127-129), core radius 0.3
  r_vir (quickTest.xml:109-111), 3 degrees of freedom (quickTest.xml:130-132),
  infall torque fraction 0.3 (quickTest.xml:122-124), rateMaximumExpulsion 1
  (quickTest.xml:282-287), jet-power heating efficiency 1 (quickTest.xml:178-180);
* black holes: Bondi-Hoyle enhancements 5 (spheroid) and 6 (hot halo), spheroid
  gas temperature 100 K (quickTest.xml:198-203);
* accretion: reionization z=10.5, V=35 km/s (quickTest.xml:101-104);
* structure: fixed solver factor sqrt(1/2) (galactic_structure/radius_solver/fixed.F90:97-106);
* the optional operators of the nodeOperator list (quickTest.xml:233-288 lists
  all four): satelliteMassLoss, barInstability with efstathiou1982 thresholds
  1.1 (stellar) and 0.7 (gaseous) (quickTest.xml:259-264) and the class
  defaults timescaleMinimum 1e-9 Gyr and retained angular momentum fractions
  1 (galactic_dynamics/bar_instability/Efstathiou1982.F90:89-128),
  blackHolesWinds with ciotti2009 efficiencyWind 0.0024 scaled by the
  radiative efficiency (quickTest.xml:174-177) and the spheroid's
  efficiencyEnergeticOutflow 0.01 (spheroid/standard/_class.F90:225-232),
  CGMOuterRadiusRamPressureStripping with the font2008 formFactor 2
  (hot_halo/ram_pressure_stripping/Font2008.F90:90-97);
* CGMCoolingHeating excessHeatDrivesOutflow true (quickTest.xml:285).
"""
from __future__ import annotations

import math
import xml.etree.ElementTree as ET

import numpy as np

NVAR = 30
NPARAM_NODE = 7
# Rows of the per-node parameters (enum gal_node_param in include/gal_ode.h).
NP_TIME_REFERENCE, NP_MASS_REFERENCE, NP_MASS_RATE, NP_SPIN_RATE, NP_IS_SATELLITE, NP_RAM_PRESSURE_FORCE, \
    NP_HOST_VELOCITY_VIRIAL = range(NPARAM_NODE)

# Indices mirror enum gal_param in include/gal_ode.h.
PARAM_NAMES = [
    "physicsMode",
    "OmegaMatter",
    "OmegaDarkEnergy",
    "OmegaBaryon",
    "HubbleConstant",
    "recycledFraction",
    "metalYield",
    "starFormationDisks.efficiency",
    "starFormationDisks.exponentVelocity",
    "starFormationDisks.timescaleMinimum",
    "starFormationSpheroids.efficiency",
    "starFormationSpheroids.exponentVelocity",
    "starFormationSpheroids.timescaleMinimum",
    "stellarFeedbackDisks.velocityCharacteristic",
    "stellarFeedbackDisks.exponent",
    "stellarFeedbackDisks.timescaleOutflowFractionalMinimum",
    "stellarFeedbackSpheroids.velocityCharacteristic",
    "stellarFeedbackSpheroids.exponent",
    "stellarFeedbackSpheroids.timescaleOutflowFractionalMinimum",
    "hotHaloOutflowReincorporation.multiplier",
    "coolingRate.velocityCutOff",
    "coreRadiusOverVirialRadius",
    "coolingTime.degreesOfFreedom",
    "coolingInfallTorque.fractionLossAngularMomentum",
    "CGMCoolingHeating.rateMaximumExpulsion",
    "blackHoleCGMHeating.efficiencyRadioMode",
    "bondiHoyleAccretionEnhancementSpheroid",
    "bondiHoyleAccretionEnhancementHotHalo",
    "bondiHoyleAccretionTemperatureSpheroid",
    "redshiftReionization",
    "velocitySuppressionReionization",
    "galacticStructureSolver.factor",
    "toleranceRelativeMetallicity",
    "starFormationTimescale.fixed.timescale",
    "stellarFeedbackOutflows.fixed.fraction",
    "nodeOperators",
    "CGMCoolingHeating.excessHeatDrivesOutflow",
    "barInstability.stabilityThresholdStellar",
    "barInstability.stabilityThresholdGaseous",
    "barInstability.timescaleMinimum",
    "barInstability.fractionAngularMomentumRetainedDisk",
    "barInstability.fractionAngularMomentumRetainedSpheroid",
    "blackHoleWind.efficiencyWind",
    "blackHoleWind.efficiencyWindScalesWithEfficiencyRadiative",
    "componentSpheroid.efficiencyEnergeticOutflow",
    "hotHaloRamPressureStripping.formFactor",
    "stellarFeedbackOutflows.fixed.fractionSpheroid",
    "postStepZeroNegativeMasses",
]
P = {name: i for i, name in enumerate(PARAM_NAMES)}
NPARAM = len(PARAM_NAMES)

PHYSICS_STANDARD = 0
PHYSICS_BOX = 1
PHYSICS_TEST = 2

DEFAULTS = {
    "physicsMode": PHYSICS_STANDARD,
    "OmegaMatter": 0.3,
    "OmegaDarkEnergy": 0.7,
    "OmegaBaryon": 0.045,
    "HubbleConstant": 70.0,
    "recycledFraction": 0.46,
    "metalYield": 0.035,
    "starFormationDisks.efficiency": 0.01,
    "starFormationDisks.exponentVelocity": -1.5,
    "starFormationDisks.timescaleMinimum": 1.0e-3,
    "starFormationSpheroids.efficiency": 0.04,
    "starFormationSpheroids.exponentVelocity": 2.0,
    "starFormationSpheroids.timescaleMinimum": 1.0e-3,
    "stellarFeedbackDisks.velocityCharacteristic": 250.0,
    "stellarFeedbackDisks.exponent": 3.5,
    "stellarFeedbackDisks.timescaleOutflowFractionalMinimum": 1.0e-3,
    "stellarFeedbackSpheroids.velocityCharacteristic": 100.0,
    "stellarFeedbackSpheroids.exponent": 3.5,
    "stellarFeedbackSpheroids.timescaleOutflowFractionalMinimum": 1.0e-3,
    "hotHaloOutflowReincorporation.multiplier": 5.0,
    "coolingRate.velocityCutOff": 1.0e4,
    "coreRadiusOverVirialRadius": 0.3,
    "coolingTime.degreesOfFreedom": 3.0,
    "coolingInfallTorque.fractionLossAngularMomentum": 0.3,
    "CGMCoolingHeating.rateMaximumExpulsion": 1.0,
    "blackHoleCGMHeating.efficiencyRadioMode": 1.0,
    "bondiHoyleAccretionEnhancementSpheroid": 5.0,
    "bondiHoyleAccretionEnhancementHotHalo": 6.0,
    "bondiHoyleAccretionTemperatureSpheroid": 100.0,
    "redshiftReionization": 10.5,
    "velocitySuppressionReionization": 35.0,
    "galacticStructureSolver.factor": math.sqrt(0.5),
    "toleranceRelativeMetallicity": 1.0e-4,
    "starFormationTimescale.fixed.timescale": 1.0,
    "stellarFeedbackOutflows.fixed.fraction": 0.0,
    "nodeOperators": 0xF,
    "CGMCoolingHeating.excessHeatDrivesOutflow": 1.0,
    "barInstability.stabilityThresholdStellar": 1.1,
    "barInstability.stabilityThresholdGaseous": 0.7,
    "barInstability.timescaleMinimum": 1.0e-9,
    "barInstability.fractionAngularMomentumRetainedDisk": 1.0,
    "barInstability.fractionAngularMomentumRetainedSpheroid": 1.0,
    "blackHoleWind.efficiencyWind": 0.0024,
    "blackHoleWind.efficiencyWindScalesWithEfficiencyRadiative": 1.0,
    "componentSpheroid.efficiencyEnergeticOutflow": 1.0e-2,
    "hotHaloRamPressureStripping.formFactor": 2.0,
    "stellarFeedbackOutflows.fixed.fractionSpheroid": 0.0,
    "postStepZeroNegativeMasses": 1.0,
}

# Optional node operators (GAL_OPERATOR_* bits of include/gal_ode.h), by the
# nodeOperator value that lists them in a parameter file.
OPERATOR_BITS = {
    "satelliteMassLoss": 0x1,
    "barInstability": 0x2,
    "blackHolesWinds": 0x4,
    "CGMOuterRadiusRamPressureStripping": 0x8,
}

# Node-evolver keys and defaults (merger_trees/node_evolver/standard.F90:165-256).
EVOLVER_DEFAULTS = {
    "odeToleranceAbsolute": 0.01,
    "odeToleranceRelative": 0.01,
    "odeAlgorithm": "rungeKuttaCashKarp",
    "reuseODEStepSize": True,
    "enforceNonNegativity": False,
    "profileOdeEvolver": False,
}

FLAG_ENFORCE_NON_NEGATIVITY = 0x1
FLAG_REUSE_STEP_SIZE = 0x2
FLAG_SCALES_PROVIDED = 0x4
FLAG_COST_ORDER = 0x8
# per-node statuses (include/gal_ode.h; error/_module.F90 errorStatus*)
STATUS_SUCCESS = 0
STATUS_INTERRUPT = 9
STATUS_UNDERFLOW = 15
STATUS_XCPU = 1025


def physics_params(**overrides) -> np.ndarray:
    """Flat GAL_NPARAM array from defaults plus overrides (keys as PARAM_NAMES)."""
    vals = dict(DEFAULTS)
    for k, v in overrides.items():
        if k not in vals:
            raise KeyError(f"unknown physics parameter {k!r}")
        vals[k] = v
    return np.array([float(vals[n]) for n in PARAM_NAMES], dtype=np.float64)


# Parameter-XML paths (element names joined by '/') -> engine keys.  The
# reference's tree of <class value="impl"><param value="..."/></class>
# elements is walked and matched on these suffixes.
_XML_MAP = {
    "cosmologyParameters/OmegaMatter": "OmegaMatter",
    "cosmologyParameters/OmegaDarkEnergy": "OmegaDarkEnergy",
    "cosmologyParameters/OmegaBaryon": "OmegaBaryon",
    "cosmologyParameters/HubbleConstant": "HubbleConstant",
    "stellarPopulation/recycledFraction": "recycledFraction",
    "stellarPopulation/metalYield": "metalYield",
    "starFormationRateSpheroids/starFormationTimescale/efficiency": "starFormationSpheroids.efficiency",
    "starFormationRateSpheroids/starFormationTimescale/exponentVelocity": "starFormationSpheroids.exponentVelocity",
    "starFormationRateSpheroids/starFormationTimescale/timescaleMinimum": "starFormationSpheroids.timescaleMinimum",
    "starFormationRateDisks/starFormationTimescale/efficiency": "starFormationDisks.efficiency",
    "starFormationRateDisks/starFormationTimescale/exponentVelocity": "starFormationDisks.exponentVelocity",
    "starFormationRateDisks/starFormationTimescale/timescaleMinimum": "starFormationDisks.timescaleMinimum",
    "starFormationRateDisks/starFormationTimescale/timescale": "starFormationTimescale.fixed.timescale",
    "hotHaloOutflowReincorporation/multiplier": "hotHaloOutflowReincorporation.multiplier",
    "coolingRate/velocityCutOff": "coolingRate.velocityCutOff",
    "hotHaloMassDistributionCoreRadius/coreRadiusOverVirialRadius": "coreRadiusOverVirialRadius",
    "coolingTime/degreesOfFreedom": "coolingTime.degreesOfFreedom",
    "coolingInfallTorque/fractionLossAngularMomentum": "coolingInfallTorque.fractionLossAngularMomentum",
    "blackHoleCGMHeating/efficiencyRadioMode": "blackHoleCGMHeating.efficiencyRadioMode",
    "blackHoleAccretionRate/bondiHoyleAccretionEnhancementSpheroid": "bondiHoyleAccretionEnhancementSpheroid",
    "blackHoleAccretionRate/bondiHoyleAccretionEnhancementHotHalo": "bondiHoyleAccretionEnhancementHotHalo",
    "blackHoleAccretionRate/bondiHoyleAccretionTemperatureSpheroid": "bondiHoyleAccretionTemperatureSpheroid",
    "accretionHalo/redshiftReionization": "redshiftReionization",
    "accretionHalo/velocitySuppressionReionization": "velocitySuppressionReionization",
    "galacticStructureSolver/factor": "galacticStructureSolver.factor",
    "componentDisk/toleranceRelativeMetallicity": "toleranceRelativeMetallicity",
    "galacticDynamicsBarInstability/stabilityThresholdStellar": "barInstability.stabilityThresholdStellar",
    "galacticDynamicsBarInstability/stabilityThresholdGaseous": "barInstability.stabilityThresholdGaseous",
    "galacticDynamicsBarInstability/timescaleMinimum": "barInstability.timescaleMinimum",
    "galacticDynamicsBarInstability/fractionAngularMomentumRetainedDisk":
        "barInstability.fractionAngularMomentumRetainedDisk",
    "galacticDynamicsBarInstability/fractionAngularMomentumRetainedSpheroid":
        "barInstability.fractionAngularMomentumRetainedSpheroid",
    "blackHoleWind/efficiencyWind": "blackHoleWind.efficiencyWind",
    "componentSpheroid/efficiencyEnergeticOutflow": "componentSpheroid.efficiencyEnergeticOutflow",
    "hotHaloRamPressureStripping/formFactor": "hotHaloRamPressureStripping.formFactor",
}

_BOOLEAN_XML = {
    "blackHoleWind/efficiencyWindScalesWithEfficiencyRadiative":
        "blackHoleWind.efficiencyWindScalesWithEfficiencyRadiative",
    "componentDisk/postStepZeroNegativeMasses": "postStepZeroNegativeMasses",
    "componentSpheroid/postStepZeroNegativeMasses": "postStepZeroNegativeMasses",
}


def _truth(value: str) -> bool:
    return value.strip().lower() in ("true", ".true.", "1", "yes")


def _walk(elem, path, out):
    for child in elem:
        if not isinstance(child.tag, str):
            continue
        p = path + [child.tag]
        if "value" in child.attrib:
            out.append(("/".join(p), child.attrib["value"], child))
        _walk(child, p, out)


def parse_parameter_xml(path_or_text: str):
    """Read a reference-format parameter file.

    Returns (physics_params array, evolver dict).  Recognised keys are those
    of ``_XML_MAP`` plus the ``mergerTreeNodeEvolver`` block; stellar feedback
    blocks are matched by their enclosing ``nodeOperator`` value
    (stellarFeedbackDisks / stellarFeedbackSpheroids) as in quickTest.xml.
    """
    text = path_or_text
    if "<" not in path_or_text:
        with open(path_or_text, "r", encoding="utf-8") as f:
            text = f.read()
    root = ET.fromstring(text)
    items = []
    _walk(root, [], items)
    overrides = {}
    evolver = dict(EVOLVER_DEFAULTS)
    for path, value, _ in items:
        for suffix, key in _XML_MAP.items():
            if path.endswith(suffix):
                overrides[key] = float(value)
        for suffix, key in _BOOLEAN_XML.items():
            if path.endswith(suffix):
                overrides[key] = 1.0 if _truth(value) else 0.0
        if path.startswith("mergerTreeNodeEvolver/"):
            k = path.split("/", 1)[1]
            if k in ("odeToleranceAbsolute", "odeToleranceRelative"):
                evolver[k] = float(value)
            elif k in ("reuseODEStepSize", "enforceNonNegativity", "profileOdeEvolver"):
                evolver[k] = _truth(value)
            elif k == "odeAlgorithm":
                evolver[k] = value
    # Feedback and star-formation blocks nested under nodeOperator entries;
    # the optional operators are on exactly when listed.
    listed = {op.attrib.get("value") for op in root.iter("nodeOperator")}
    if "multi" in listed:
        overrides["nodeOperators"] = float(sum(bit for name, bit in OPERATOR_BITS.items() if name in listed))
    for op in root.iter("nodeOperator"):
        kind = op.attrib.get("value")
        if kind in ("stellarFeedbackDisks", "stellarFeedbackSpheroids"):
            tag = "stellarFeedbackDisks" if kind == "stellarFeedbackDisks" else "stellarFeedbackSpheroids"
            for e in op.iter("stellarFeedbackOutflows"):
                for c in e:
                    if c.tag == "velocityCharacteristic":
                        overrides[tag + ".velocityCharacteristic"] = float(c.attrib["value"])
                    elif c.tag == "exponent":
                        overrides[tag + ".exponent"] = float(c.attrib["value"])
                    elif c.tag == "timescaleOutflowFractionalMinimum":
                        overrides[tag + ".timescaleOutflowFractionalMinimum"] = float(c.attrib["value"])
                    elif c.tag == "fraction":
                        key = ("stellarFeedbackOutflows.fixed.fraction" if tag == "stellarFeedbackDisks"
                               else "stellarFeedbackOutflows.fixed.fractionSpheroid")
                        overrides[key] = float(c.attrib["value"])
        if kind == "CGMCoolingHeating":
            for c in op:
                if c.tag == "rateMaximumExpulsion":
                    overrides["CGMCoolingHeating.rateMaximumExpulsion"] = float(c.attrib["value"])
                elif c.tag == "excessHeatDrivesOutflow":
                    overrides["CGMCoolingHeating.excessHeatDrivesOutflow"] = 1.0 if _truth(c.attrib["value"]) else 0.0
    if evolver["odeAlgorithm"] != "rungeKuttaCashKarp":
        raise ValueError("only odeAlgorithm=rungeKuttaCashKarp is implemented by the batched evolver")
    return physics_params(**overrides), evolver


def solver_flags(evolver: dict, scales_provided: bool = False, cost_order: bool = True) -> int:
    f = FLAG_COST_ORDER if cost_order else 0
    if evolver.get("enforceNonNegativity"):
        f |= FLAG_ENFORCE_NON_NEGATIVITY
    if evolver.get("reuseODEStepSize", True):
        f |= FLAG_REUSE_STEP_SIZE
    if scales_provided:
        f |= FLAG_SCALES_PROVIDED
    return f
