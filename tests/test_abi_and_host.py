"""CPU tests of the boundary and the host logic (no GPU calls).

* the HIP library loads and exports every function declared in include/*.h;
* the C enums (state layout, parameters, node parameters, flags, statuses)
  agree with the host-side tables in galacticus_amd;
* the parameter-XML reader maps the reference's keys (parameters/quickTest.xml
  layout) onto the engine's parameter array;
* Values_Agree semantics (numerical/comparison.F90:129-155);
* the static FLOP count of the RHS.
"""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = []
    for hdr in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(hdr).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names += re.findall(r"^\s*(?:int64_t|int|void|size_t|double)\s+(gal_\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


@pytest.fixture(scope="module")
def lib():
    from galacticus_amd import _lib, build
    if not os.path.exists(_lib.LIB_PATH):
        build.build()
    return ctypes.CDLL(_lib.LIB_PATH)


def test_library_exports_every_declared_symbol(lib):
    names = _declared_functions()
    assert len(names) >= 12
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_symbol_without_gpu(lib):
    lib.gal_ode_version.restype = ctypes.c_int
    assert lib.gal_ode_version() >= 10000


def _enum_values(name):
    src = open(os.path.join(ROOT, "include", "gal_ode.h")).read()
    body = re.search(r"enum %s \{(.*?)\};" % name, src, flags=re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    out, nxt = [], 0
    for item in [x.strip() for x in body.split(",") if x.strip()]:
        if "=" in item:
            k, v = [z.strip() for z in item.split("=")]
            nxt = int(v)
        else:
            k = item
        out.append((k, nxt))
        nxt += 1
    return out


def test_state_layout_matches_host_tables():
    from galacticus_amd import workloads as wl
    vals = _enum_values("gal_var")
    assert vals[-1] == ("GAL_NVAR", 30)
    for k, v in vals[:-1]:
        assert wl.V[k[len("GAL_V_"):]] == v, k


def test_parameter_layout_matches_host_tables():
    from galacticus_amd import params as gp
    vals = _enum_values("gal_param")
    assert vals[-1][0] == "GAL_NPARAM" and vals[-1][1] == gp.NPARAM
    assert len(vals) - 1 == len(gp.PARAM_NAMES)
    node = _enum_values("gal_node_param")
    assert node[-1] == ("GAL_NPARAM_NODE", gp.NPARAM_NODE)
    for name, value in node[:-1]:
        assert getattr(gp, name[len("GAL_"):]) == value


def test_flags_and_statuses_match():
    from galacticus_amd import params as gp, evolver
    src = open(os.path.join(ROOT, "include", "gal_ode.h")).read()
    d = dict(re.findall(r"#define (GAL_\w+)\s+(-?0x[0-9a-fA-F]+|-?\d+)", src))
    assert int(d["GAL_FLAG_ENFORCE_NON_NEGATIVITY"], 0) == gp.FLAG_ENFORCE_NON_NEGATIVITY
    assert int(d["GAL_FLAG_REUSE_STEP_SIZE"], 0) == gp.FLAG_REUSE_STEP_SIZE
    assert int(d["GAL_FLAG_SCALES_PROVIDED"], 0) == gp.FLAG_SCALES_PROVIDED
    assert int(d["GAL_FLAG_COST_ORDER"], 0) == gp.FLAG_COST_ORDER
    assert int(d["GAL_STATUS_UNDERFLOW"]) == evolver.STATUS_UNDERFLOW == 15
    assert int(d["GAL_STATUS_INTERRUPT"]) == evolver.STATUS_INTERRUPT == 9
    assert int(d["GAL_STATUS_XCPU"]) == evolver.STATUS_XCPU == 1025


PARAMETER_XML = """<?xml version='1.0' encoding='UTF-8'?>
<parameters>
  <formatVersion>2</formatVersion>
  <cosmologyParameters value="simple">
    <HubbleConstant value="70.2"/>
    <OmegaMatter value="0.2725"/>
    <OmegaDarkEnergy value="0.7275"/>
    <OmegaBaryon value="0.0455"/>
  </cosmologyParameters>
  <coolingRate value="whiteFrenk1991">
    <velocityCutOff value="10000"/>
  </coolingRate>
  <starFormationRateSpheroids value="timescale">
    <starFormationTimescale value="dynamicalTime">
      <efficiency value="0.04"/>
      <exponentVelocity value="2.0"/>
      <timescaleMinimum value="0.001"/>
    </starFormationTimescale>
  </starFormationRateSpheroids>
  <stellarPopulation value="standard">
    <recycledFraction value="0.46"/>
    <metalYield value="0.035"/>
  </stellarPopulation>
  <nodeOperator value="multi">
    <nodeOperator value="stellarFeedbackDisks">
      <stellarFeedbackOutflows value="rateLimit">
        <timescaleOutflowFractionalMinimum value="0.002"/>
        <stellarFeedbackOutflows value="powerLaw">
          <velocityCharacteristic value="240.0"/>
          <exponent value="3.0"/>
        </stellarFeedbackOutflows>
      </stellarFeedbackOutflows>
    </nodeOperator>
    <nodeOperator value="CGMCoolingHeating">
      <rateMaximumExpulsion value="0.5"/>
    </nodeOperator>
  </nodeOperator>
  <mergerTreeNodeEvolver value="standard">
    <odeToleranceAbsolute value="1.0e-6"/>
    <odeToleranceRelative value="0.001"/>
    <enforceNonNegativity value="true"/>
  </mergerTreeNodeEvolver>
</parameters>
"""


def test_parameter_xml_reader():
    from galacticus_amd import params as gp
    p, ev = gp.parse_parameter_xml(PARAMETER_XML)
    P = gp.P
    assert p[P["HubbleConstant"]] == 70.2 and p[P["OmegaMatter"]] == 0.2725
    assert p[P["starFormationSpheroids.efficiency"]] == 0.04
    assert p[P["stellarFeedbackDisks.velocityCharacteristic"]] == 240.0
    assert p[P["stellarFeedbackDisks.exponent"]] == 3.0
    assert p[P["stellarFeedbackDisks.timescaleOutflowFractionalMinimum"]] == 0.002
    assert p[P["stellarFeedbackSpheroids.velocityCharacteristic"]] == 100.0   # default kept
    assert p[P["CGMCoolingHeating.rateMaximumExpulsion"]] == 0.5
    assert ev["odeToleranceAbsolute"] == 1e-6 and ev["odeToleranceRelative"] == 1e-3
    assert ev["enforceNonNegativity"] is True and ev["reuseODEStepSize"] is True
    flags = gp.solver_flags(ev)
    assert flags & gp.FLAG_ENFORCE_NON_NEGATIVITY and flags & gp.FLAG_REUSE_STEP_SIZE


def test_parameter_xml_rejects_other_algorithms():
    from galacticus_amd import params as gp
    bad = PARAMETER_XML.replace('<enforceNonNegativity value="true"/>', '<odeAlgorithm value="bdf"/>')
    with pytest.raises(ValueError):
        gp.parse_parameter_xml(bad)


def test_values_agree():
    from galacticus_amd.evolver import values_agree
    assert values_agree(1.0, 1.0, 0.0, 0.0)
    assert values_agree(1.0, 1.005, None, 1e-2)            # |d| <= 0.5|a+b| rel
    assert not values_agree(1.0, 1.05, None, 1e-2)
    assert values_agree(0.0, 1e-3, 1e-2, None)
    assert not values_agree(0.0, 1e-3, 1e-4, 1e-4)
    assert np.array_equal(values_agree(np.array([1.0, 2.0]), np.array([1.0, 3.0]), 0.1, 0.1), [True, False])


def test_static_flop_count():
    from galacticus_amd import flops
    f = flops.rhs_flops()
    assert 400 < f < 2000
    assert flops.attempt_flops() == 6 * f + 52 * 30
