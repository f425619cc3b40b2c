"""The Fortran host binding (galacticus_amd/fortran, ISO_C_BINDING over
include/gal_ode.h) drives the GPU engine: a batch evolved by the Fortran
program equals, bit for bit, the same batch evolved through the Python
binding, and agrees with the CPU oracle within the ODE tolerance."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from galacticus_amd import params as gp
from galacticus_amd import workloads as wl

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "galacticus_amd", "fortran", "build", "evolve_batch_file")


def _write_input(path, b, physics, table, tol_abs, tol_rel):
    with open(path, "wb") as f:
        np.array([b.n, table.n_t, table.n_z, 1, 0], dtype=np.int32).tofile(f)
        np.array([tol_abs, tol_rel, table.ln_t_min, table.d_ln_t, table.ln_z_min, table.d_ln_z]).tofile(f)
        np.ascontiguousarray(physics, dtype=np.float64).tofile(f)
        np.ascontiguousarray(table.ln_lambda, dtype=np.float64).ravel().tofile(f)
        for a in (b.y, b.node_par, b.t_start, b.t_end, b.h_step):
            np.ascontiguousarray(a, dtype=np.float64).ravel().tofile(f)


def test_fortran_binding_evolves_a_batch(tmp_path, cooling_table):
    assert os.path.exists(EXE), "build the Fortran binding first (__graft_entry__.build())"
    from galacticus_amd.evolver import NodeEvolverBatched
    physics = gp.physics_params()
    b = wl.isolated_halos(500, seed=12, params=physics)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_input(fin, b, physics, cooling_table, 1e-6, 1e-3)
    print("running", EXE, flush=True)
    r = subprocess.run([EXE, fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = open(fout, "rb").read()
    n = b.n
    yf = np.frombuffer(raw, dtype=np.float64, count=30 * n).reshape(30, n)
    hf = np.frombuffer(raw, dtype=np.float64, count=n, offset=30 * n * 8)
    stf = np.frombuffer(raw, dtype=np.int32, count=n, offset=61 * n * 8)
    rc = np.frombuffer(raw, dtype=np.int32, count=1, offset=61 * n * 8 + 4 * n)[0]
    assert rc == 0
    ev = NodeEvolverBatched(n, physics, cooling_table, odeToleranceAbsolute=1e-6, odeToleranceRelative=1e-3)
    yp, hp, _, stp, _ = ev.evolve(b.y, b.node_par, b.t_start, b.t_end, b.h_step)
    assert np.array_equal(yf, yp) and np.array_equal(hf, hp) and np.array_equal(stf, stp)
    yo, _, sco, sto, _ = oracle.evolve(b.y, b.node_par, b.t_start, b.t_end, b.h_step, physics, cooling_table, 1e-6,
                                       1e-3, gp.FLAG_REUSE_STEP_SIZE, root_tol_rel=1e-15)
    assert (sto == stf).all()
    assert (np.abs(yf - yo) <= 10.0 * (1e-3 * np.abs(yo) + 1e-6 * sco)).all()


FOREST_EXE = os.path.join(ROOT, "galacticus_amd", "fortran", "build", "evolve_forest_file")


def test_fortran_forest_driver_equals_python_driver(tmp_path, cooling_table):
    """A Fortran host program builds a PCH forest, records the walk, keeps the
    forest state in device memory and runs the schedule on the GPU through
    the ISO_C_BINDING modules; the same forest driven from Python ends in the
    same state, bit for bit, with the same statuses."""
    import torch
    from galacticus_amd import trees as tr
    from galacticus_amd.evolver import NodeEvolverBatched
    assert os.path.exists(FOREST_EXE), "build the Fortran binding first (__graft_entry__.build())"
    physics = gp.physics_params()
    t_z0 = float(wl.scale_factor_to_cosmic_time(1.0))
    t_z5 = float(wl.scale_factor_to_cosmic_time(1.0 / 6.0))
    cosmo, pch, walk = tr.cosmology(physics), tr.pch_params(2e10, t_z5), tr.walk_params(t_z0, merging_multiplier=0.2)
    roots, seed = np.geomspace(1e11, 1e12, 6), 3
    t = cooling_table
    fin, fout = str(tmp_path / "forest.in"), str(tmp_path / "forest.out")
    with open(fin, "wb") as f:
        np.array([len(roots), t.n_t, t.n_z], dtype=np.int32).tofile(f)
        np.array([seed], dtype=np.int64).tofile(f)
        np.array([1e-6, 1e-3, t.ln_t_min, t.d_ln_t, t.ln_z_min, t.d_ln_z]).tofile(f)
        np.ascontiguousarray(physics, dtype=np.float64).tofile(f)
        np.ascontiguousarray(t.ln_lambda, dtype=np.float64).ravel().tofile(f)
        roots.astype(np.float64).tofile(f)
        np.array([t_z0]).tofile(f)
        f.write(bytes(cosmo))
        f.write(bytes(pch))
        f.write(bytes(walk))
        np.array([1e-4, 100.0, 0.25]).tofile(f)
    print("running", FOREST_EXE, flush=True)
    r = subprocess.run([FOREST_EXE, fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = open(fout, "rb").read()
    n = int(np.frombuffer(raw, dtype=np.int64, count=1)[0])
    offsets = np.frombuffer(raw, dtype=np.int64, count=len(roots) + 1, offset=8)
    o = 8 + 8 * (len(roots) + 1)
    yf = np.frombuffer(raw, dtype=np.float64, count=30 * n, offset=o).reshape(30, n)
    stf = np.frombuffer(raw, dtype=np.int32, count=n, offset=o + 240 * n)
    rc = np.frombuffer(raw, dtype=np.int32, count=1, offset=o + 244 * n)[0]
    assert rc == 0
    forest = tr.Forest.build_pch(len(roots), roots, t_z0, cosmo, pch, seed=seed)
    assert forest.n_nodes == n and np.array_equal(forest.tree_offset, offsets)
    sched = tr.Schedule(forest, cosmo, walk)
    ev = NodeEvolverBatched(n, physics, cooling_table, odeToleranceAbsolute=1e-6, odeToleranceRelative=1e-3)
    y = torch.from_numpy(forest.initial_state(cosmo, walk)).cuda()
    st = tr.evolve_forest(ev, sched, y).cpu().numpy()
    print("nodes", n, "operations", sched.n_ops, flush=True)
    assert np.array_equal(yf, y.cpu().numpy()) and np.array_equal(stf, st)


XML_EXE = os.path.join(ROOT, "galacticus_amd", "fortran", "build", "node_evolver_xml")
XML_PARAMETERS = """<parameters>
  <mergerTreeNodeEvolver value="standard">
    <odeToleranceAbsolute value="1.0e-6"/>
    <odeToleranceRelative value="1.0e-3"/>
    <reuseODEStepSize value="true"/>
    <enforceNonNegativity value="false"/>
  </mergerTreeNodeEvolver>
</parameters>
"""


def _write_xml_batch(path, b, physics, table, wall_seconds, events, ratio_major):
    with open(path, "wb") as f:
        np.array([b.n, table.n_t, table.n_z], dtype=np.int32).tofile(f)
        np.array([table.ln_t_min, table.d_ln_t, table.ln_z_min, table.d_ln_z, wall_seconds]).tofile(f)
        np.ascontiguousarray(physics, dtype=np.float64).tofile(f)
        np.ascontiguousarray(table.ln_lambda, dtype=np.float64).ravel().tofile(f)
        for a in (b.y, b.node_par, b.t_start, b.t_end, b.h_step):
            np.ascontiguousarray(a, dtype=np.float64).ravel().tofile(f)
        np.array([len(events)], dtype=np.int32).tofile(f)
        np.array([[k, a + 1, t + 1] for k, a, t in events], dtype=np.int32).ravel().tofile(f)
        np.array([ratio_major]).tofile(f)


def _read_xml_output(path, n):
    raw = open(path, "rb").read()
    y = np.frombuffer(raw, dtype=np.float64, count=30 * n).reshape(30, n)
    h = np.frombuffer(raw, dtype=np.float64, count=n, offset=30 * n * 8)
    st = np.frombuffer(raw, dtype=np.int32, count=n, offset=31 * n * 8)
    rc = np.frombuffer(raw, dtype=np.int32, count=1, offset=31 * n * 8 + 4 * n)[0]
    return y, h, st, rc


def test_fortran_node_evolver_method_set(tmp_path, cooling_table):
    """The Fortran type's method set (mergerTreeNodeEvolver, node_evolver/_class.F90:41-79):
    tolerances read from a parameter file, evolve, then promote and merge --
    bit for bit what the Python binding gives (evolve + gal_node_events_h);
    and evolve past an expired systemClockMaximum returns errorStatusXCPU for
    every node with its state as it came in."""
    assert os.path.exists(XML_EXE), "build the Fortran binding first (__graft_entry__.build())"
    from galacticus_amd.evolver import NodeEvolverBatched
    from galacticus_amd import trees as tr
    physics = gp.physics_params()
    b = wl.isolated_halos(300, seed=14, params=physics)
    xml = tmp_path / "parameters.xml"
    xml.write_text(XML_PARAMETERS)
    events = [(tr.OP_PROMOTE, 3, 10), (tr.OP_MERGE_GALAXY, 5, 6), (tr.OP_MERGE_GALAXY, 7, 8)]
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_xml_batch(fin, b, physics, cooling_table, 0.0, events, 0.25)
    print("running", XML_EXE, flush=True)
    r = subprocess.run([XML_EXE, str(xml), fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    yf, hf, stf, rc = _read_xml_output(fout, b.n)
    assert rc == 0 and (stf == 0).all()
    ev = NodeEvolverBatched(b.n, physics, cooling_table, odeToleranceAbsolute=1e-6, odeToleranceRelative=1e-3)
    yp, hp, _, stp, _ = ev.evolve(b.y, b.node_par, b.t_start, b.t_end, b.h_step)
    yp, hp = np.ascontiguousarray(yp), np.ascontiguousarray(hp)
    evolved = yp.copy()
    for k, a, t in events:      # one at a time, as the Fortran program applies them
        tr.node_events([k], [a], [t], yp, hp, 0.25)
    assert np.array_equal(yf, yp) and np.array_equal(hf, hp) and np.array_equal(stf, stp)
    # the promotion handed node 3's evolved galaxy to node 10
    assert np.array_equal(yf[:, 10], evolved[:, 3])

    print("running with an expired clock limit", flush=True)
    _write_xml_batch(fin, b, physics, cooling_table, 1.0e-9, [], 0.25)
    r = subprocess.run([XML_EXE, str(xml), fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    yx, hx, stx, rc = _read_xml_output(fout, b.n)
    assert rc == 0 and (stx == gp.STATUS_XCPU).all()
    assert np.array_equal(yx, b.y) and np.array_equal(hx, b.h_step)
