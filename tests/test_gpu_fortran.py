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
