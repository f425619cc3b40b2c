"""The Fortran node-evolver type's parameter reader, without a GPU.

The Fortran binding (galacticus_amd/fortran/galacticus_amd_ode.F90,
readParameters) takes odeToleranceAbsolute, odeToleranceRelative,
reuseODEStepSize and enforceNonNegativity from the mergerTreeNodeEvolver
element of a Galacticus parameter file, with the reference's defaults
(node_evolver/standard.F90:165-256) for absent keys.  Its values must equal
those of the Python reader (galacticus_amd.params.parse_parameter_xml, built
on ElementTree) on files shaped like the reference's: quickTest.xml's block,
the enforceNonNegativity test's block, commented-out keys, an element whose
name begins with a key's name, blank and tab variants, and no block at all.
"""
import os
import subprocess

import pytest

from galacticus_amd import params as gp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "galacticus_amd", "fortran", "build", "node_evolver_xml")

OTHER = """  <cosmologyParameters value="simple">
    <HubbleConstant value="70.2"/>
    <OmegaMatter value="0.2725"/>
  </cosmologyParameters>
"""

CASES = {
    "quickTest": """<parameters>
%s  <!-- Numerical tolerances -->
  <mergerTreeNodeEvolver value="standard">
    <odeToleranceAbsolute value="0.01"/>
    <odeToleranceRelative value="0.01"/>
  </mergerTreeNodeEvolver>
</parameters>
""" % OTHER,
    "rapidDepletion": """<parameters>
  <mergerTreeNodeEvolver value="standard">
    <odeToleranceAbsolute value="0.1"/>
    <odeToleranceRelative value="0.1"/>
    <reuseODEStepSize value="true"/>
    <enforceNonNegativity value="false"/>
  </mergerTreeNodeEvolver>
%s</parameters>
""" % OTHER,
    "enforce_tight_commented": """<parameters>
  <mergerTreeNodeEvolver value="standard">
    <!-- <odeToleranceAbsolute value="5.0"/> -->
    <odeToleranceAbsoluteScale value="7.0"/>
    <odeToleranceAbsolute	value="1.0e-6" />
    <odeToleranceRelative value="3.5e-4"/>
    <reuseODEStepSize value="false"/>
    <enforceNonNegativity value="true"/>
  </mergerTreeNodeEvolver>
</parameters>
""",
    "defaults": """<parameters>
%s  <mergerTreeEvolver value="standard">
    <timestepHostAbsolute value="1.0"/>
  </mergerTreeEvolver>
</parameters>
""" % OTHER,
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_fortran_reader_equals_python_reader(tmp_path, case):
    assert os.path.exists(EXE), "build the Fortran binding first (__graft_entry__.build())"
    path = tmp_path / "parameters.xml"
    path.write_text(CASES[case])
    r = subprocess.run([EXE, str(path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    got = {}
    for line in r.stdout.splitlines():
        k, v = line.split()
        got[k] = float(v) if k.startswith("odeTolerance") else (v == "T")
    _, expect = gp.parse_parameter_xml(CASES[case])
    for k in ("odeToleranceAbsolute", "odeToleranceRelative", "reuseODEStepSize", "enforceNonNegativity"):
        assert got[k] == expect[k], (case, k, got[k], expect[k])
