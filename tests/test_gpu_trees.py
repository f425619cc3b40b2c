"""GPU tests of the tree path (include/gal_tree.h) through the C ABI.

* The reference's closed-box tree (testSuite/parameters/reproducibility/
  closedBoxTree.xml) run through the forest reader, the walk and the wave
  executor reproduces the closed-box known answers
  (testSuite/test-reproducibility.py:47-57) at the root.
* The operation-graph executor equals the wave executor bit for bit at
  every launch shape.
* A Parkinson-Cole-Helly forest run on the GPU equals the same
  schedule applied one operation at a time on the CPU with the oracle
  integrator and the oracle's merger mass movements (oracle/trees.py): median
  relative difference below 1e-9 (both sides take the same steps almost
  everywhere), every mass within 50x the per-interval ODE tolerance where
  they do not, identical statuses.
"""
import math
import os

import numpy as np
import pytest

import oracle
from oracle import trees as ot
from galacticus_amd import params as gp
from galacticus_amd import trees as tr
from galacticus_amd import workloads as wl

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
T0 = float(wl.scale_factor_to_cosmic_time(1.0))
T_Z5 = float(wl.scale_factor_to_cosmic_time(1.0 / 6.0))
V = wl.V


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def test_closed_box_tree(torch_cuda, cooling_table):
    torch = torch_cuda
    from galacticus_amd.evolver import NodeEvolverBatched
    physics, _ = wl.closed_box()
    text = open(os.path.join(ROOT, "tests", "golden", "closedBoxTree.xml")).read()
    forest, comps = tr.Forest.from_xml(text)
    y = np.zeros((gp.NVAR, forest.n_nodes))
    leaf = 1
    y[V["DISK_MASS_GAS"], leaf] = comps[leaf]["disk/massGas"]
    y[V["DISK_FRACTION_MASS_RETAINED"], leaf] = 1.0
    y[V["SPIN_ANGULAR_MOMENTUM"], leaf] = comps[leaf]["spin/angularMomentum"]
    y[V["SATELLITE_BOUND_MASS"], leaf] = forest.mass[leaf]
    cosmo = tr.cosmology(physics)
    sched = tr.Schedule(forest, cosmo, tr.walk_params(forest.time[0]))
    assert [int(k) for k in sched.ops["kind"]] == [tr.OP_EVOLVE, tr.OP_PROMOTE]
    ev = NodeEvolverBatched(forest.n_nodes, physics, cooling_table, odeToleranceAbsolute=1e-6,
                            odeToleranceRelative=1e-6)
    yd = torch.from_numpy(y).cuda()
    st = tr.evolve_forest(ev, sched, yd)
    root = yd[:, 0].cpu().numpy()
    assert int(st[1]) == 0
    # closedBox: M_gas = 1e11 exp(-(1-R) dt/tau), R=0.4, tau=0.25, dt=1 (test-reproducibility.py:47-57)
    expect = {"DISK_MASS_GAS": 9.0717953e9, "DISK_MASS_STELLAR": 9.0928205e10, "DISK_ABUNDANCES_GAS": 9.0717953e8,
              "DISK_ABUNDANCES_STELLAR": 2.8814957e9}
    for k, v in expect.items():
        assert abs(root[V[k]] / v - 1.0) < 1e-2, (k, root[V[k]], v)
    # the same known answers read back from the Galacticus-layout output, by
    # the property names and output index of test-reproducibility.py
    import tempfile
    from galacticus_amd import output as go
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "closedBox.npz")
        go.write_output(path, [go.output_group(forest, sched, yd.cpu().numpy(), forest.time[0])])
        f = go.read_output(path)
    for prop, v in (("diskMassGas", 9.0717953e9), ("diskMassStellar", 9.0928205e10),
                    ("diskAbundancesGasMetals", 9.0717953e8), ("diskAbundancesStellarMetals", 2.8814957e9)):
        got = f[f"Outputs/Output1/nodeData/{prop}"]
        assert got.shape == (1,) and abs(got[0] / v - 1.0) < 1e-2, (prop, got, v)


def _oracle_apply(sched, y0, physics, table, tol_abs, tol_rel):
    flags = gp.FLAG_REUSE_STEP_SIZE

    def evolve_one(col, npar, t0, t1, h):
        yo, ho, _, st, _ = oracle.evolve(col[:, None], np.asarray(npar)[:, None], [t0], [t1], [h], physics, table,
                                         tol_abs, tol_rel, flags, root_tol_rel=1e-15)
        return yo[:, 0], ho[0], st[0]

    log = [(int(o["kind"]), int(o["a"]), int(o["b"]), float(o["t0"]), float(o["t1"]), list(o["node_par"]))
           for o in sched.ops]
    return ot.run_sequential(log, y0, -np.ones(y0.shape[1]), evolve_one,
                             lambda y, a, b: ot.merge_galaxy_simple(y, a, b, 0.25))


def test_forest_waves_match_sequential_oracle(torch_cuda, cooling_table):
    torch = torch_cuda
    from galacticus_amd.evolver import NodeEvolverBatched
    physics = gp.physics_params()
    cosmo = tr.cosmology(physics)
    forest = tr.Forest.build_pch(8, np.geomspace(1e11, 1e12, 8), T0, cosmo, tr.pch_params(2e10, T_Z5), seed=2)
    wp = tr.walk_params(T0, merging_multiplier=0.2)
    sched = tr.Schedule(forest, cosmo, wp)
    kinds = np.bincount(sched.ops["kind"], minlength=4)
    print("nodes", forest.n_nodes, "ops", sched.n_ops, "waves", sched.n_waves, "kinds", kinds, flush=True)
    assert kinds[tr.OP_MERGE_GALAXY] > 0 and kinds[tr.OP_BECOME_SATELLITE] > 0
    y0 = forest.initial_state(cosmo, wp)
    ev = NodeEvolverBatched(forest.n_nodes, physics, cooling_table, odeToleranceAbsolute=1e-6,
                            odeToleranceRelative=1e-3)
    yd = torch.from_numpy(y0.copy()).cuda()
    sd = torch.zeros_like(yd)
    st = tr.evolve_forest(ev, sched, yd, scale=sd).cpu().numpy()
    yg, sc = yd.cpu().numpy(), sd.cpu().numpy()
    yo, _, sto = _oracle_apply(sched, y0, physics, cooling_table, 1e-6, 1e-3)
    evolved = np.unique(sched.ops["a"][sched.ops["kind"] == tr.OP_EVOLVE])
    assert (st[evolved] == 0).all() and (sto[evolved] == 0).all()
    mass = [V[k] for k in ("HOT_MASS", "DISK_MASS_GAS", "DISK_MASS_STELLAR", "SPHEROID_MASS_GAS",
                           "SPHEROID_MASS_STELLAR", "BH_MASS")]
    ref = np.abs(yo[mass])
    big = ref > 1e-6 * np.abs(yo[V["HOT_MASS"]]).max()
    rel = np.abs(yg[mass] - yo[mass])[big] / ref[big]
    print("median rel", np.median(rel), "max rel", rel.max(), flush=True)
    assert np.median(rel) < 1e-9                      # the same steps on both sides, almost everywhere
    # where the two integrations took different steps they agree to the ODE
    # tolerance per interval, accumulated over each node's history
    bad = ~(np.isfinite(yg) & np.isfinite(yo) & np.isfinite(sc))
    if bad.any():
        v, i = np.nonzero(bad)
        for vv, ii in list(zip(v, i))[:10]:
            print("non-finite", wl.VAR_NAMES[vv], "node", ii, "gpu", yg[vv, ii], "oracle", yo[vv, ii], "scale",
                  sc[vv, ii], "ops", [tuple(o)[:6] for o in sched.ops if o["a"] == ii or o["b"] == ii][-4:], flush=True)
    assert not bad.any()
    tol = 1e-3 * np.abs(yo[mass]) + 1e-6 * np.abs(sc[mass])
    worst = np.abs(yg[mass] - yo[mass]) / np.maximum(tol, 1e-300)
    print("max |dy| / tolerance", worst.max(), flush=True)
    assert worst.max() < 50.0
    roots = forest.roots()
    mstar_g = yg[V["DISK_MASS_STELLAR"], roots] + yg[V["SPHEROID_MASS_STELLAR"], roots]
    mstar_o = yo[V["DISK_MASS_STELLAR"], roots] + yo[V["SPHEROID_MASS_STELLAR"], roots]
    assert np.allclose(mstar_g, mstar_o, rtol=1e-3)


def test_concurrent_forests_equal_one_forest(torch_cuda, cooling_table):
    """A forest split into independent forests of consecutive trees, evolved
    concurrently on their own streams and host threads, ends in exactly the
    state of the whole forest evolved at once (trees never interact)."""
    torch = torch_cuda
    from galacticus_amd.evolver import NodeEvolverBatched
    physics = gp.physics_params()
    cosmo = tr.cosmology(physics)
    forest = tr.Forest.build_pch(12, np.geomspace(1e11, 1e12, 12), T0, cosmo, tr.pch_params(2e10, T_Z5), seed=5)
    wp = tr.walk_params(T0, merging_multiplier=0.2)
    y0 = forest.initial_state(cosmo, wp)
    ev = NodeEvolverBatched(forest.n_nodes, physics, cooling_table)
    yd = torch.from_numpy(y0.copy()).cuda()
    st = tr.evolve_forest(ev, tr.Schedule(forest, cosmo, wp), yd).cpu().numpy()
    parts = forest.split(3)
    print("nodes", forest.n_nodes, "parts", [f.n_nodes for f, _ in parts], flush=True)
    scheds = [tr.Schedule(f, cosmo, wp) for f, _ in parts]
    evs = [NodeEvolverBatched(f.n_nodes, physics, cooling_table) for f, _ in parts]
    ys = [torch.from_numpy(np.ascontiguousarray(y0[:, n0:n0 + f.n_nodes])).cuda() for f, n0 in parts]
    sts = tr.evolve_forests(evs, scheds, ys, [torch.cuda.Stream() for _ in parts])
    torch.cuda.synchronize()
    y_split = np.concatenate([y.cpu().numpy() for y in ys], axis=1)
    st_split = np.concatenate([s.cpu().numpy() for s in sts])
    assert np.array_equal(y_split, yd.cpu().numpy())
    assert np.array_equal(st_split, st)
    # the same three forests with two of them sharing one stream: their
    # launches interleave in that stream, the states do not change
    shared = torch.cuda.Stream()
    ys = [torch.from_numpy(np.ascontiguousarray(y0[:, n0:n0 + f.n_nodes])).cuda() for f, n0 in parts]
    sts = tr.evolve_forests(evs, scheds, ys, [shared, torch.cuda.Stream(), shared])
    torch.cuda.synchronize()
    assert np.array_equal(np.concatenate([y.cpu().numpy() for y in ys], axis=1), yd.cpu().numpy())
    assert np.array_equal(np.concatenate([s.cpu().numpy() for s in sts]), st)


@pytest.mark.parametrize("launch", [(64, 1, 1), (64, 2, 1), (64, 4, 1), (64, 1, 2), (64, 2, 2)])
def test_graph_executor_equals_wave_executor(torch_cuda, cooling_table, launch):
    """The operation-graph executor (one persistent launch, each operation
    waiting only for the earlier operations on its own nodes) ends in the
    state, step sizes, scales and statuses of the wave executor bit for bit,
    at every launch shape: every node sees the same operations in the same
    order with the same inputs."""
    torch = torch_cuda
    from galacticus_amd.evolver import NodeEvolverBatched
    physics = gp.physics_params()
    cosmo = tr.cosmology(physics)
    forest = tr.Forest.build_pch(24, np.geomspace(1e11, 1e13, 24), T0, cosmo, tr.pch_params(2e10, T_Z5), seed=7)
    wp = tr.walk_params(T0, merging_multiplier=0.2)
    sched = tr.Schedule(forest, cosmo, wp)
    kinds = np.bincount(sched.ops["kind"], minlength=4)
    print("launch", launch, "nodes", forest.n_nodes, "ops", sched.n_ops, "waves", sched.n_waves, "kinds", kinds,
          flush=True)
    assert kinds[tr.OP_PROMOTE] > 0 and kinds[tr.OP_MERGE_GALAXY] > 0 and kinds[tr.OP_BECOME_SATELLITE] > 0
    y0 = forest.initial_state(cosmo, wp)
    ev = NodeEvolverBatched(forest.n_nodes, physics, cooling_table, odeToleranceAbsolute=1e-6,
                            odeToleranceRelative=1e-3)
    ev.set_launch(launch[0], launch[1], rhs_waves=launch[2])
    out = {}
    for executor in ("waves", "graph"):
        yd = torch.from_numpy(y0.copy()).cuda()
        sd = torch.zeros_like(yd)
        st = tr.evolve_forest(ev, sched, yd, scale=sd, executor=executor)
        torch.cuda.synchronize()
        out[executor] = (yd.cpu().numpy(), sd.cpu().numpy(), st.cpu().numpy())
    evolved = np.unique(sched.ops["a"][sched.ops["kind"] == tr.OP_EVOLVE])
    assert (out["graph"][2][evolved] == 0).all()
    for a, b in zip(out["waves"], out["graph"]):
        assert np.array_equal(a, b)


def test_graph_executor_stall_limit_fails_loudly_and_recovers(torch_cuda, cooling_table):
    """The graph executor's safety net, forced from the host: with a stall
    limit far below one operation, lanes waiting on the ready list give up
    (every kernel still exits: a lane that gives up leaves the loop, the
    others finish their operations and then give up in turn), and the call
    raises instead of returning a half-evolved forest.  With the default
    limit the same handle and schedule then evolve the forest exactly as the
    wave executor does."""
    torch = torch_cuda
    from galacticus_amd import _lib
    from galacticus_amd.evolver import NodeEvolverBatched
    physics = gp.physics_params()
    cosmo = tr.cosmology(physics)
    forest = tr.Forest.build_pch(8, np.geomspace(1e11, 1e12, 8), T0, cosmo, tr.pch_params(2e10, T_Z5), seed=6)
    wp = tr.walk_params(T0, merging_multiplier=0.2)
    sched = tr.Schedule(forest, cosmo, wp)
    y0 = forest.initial_state(cosmo, wp)
    ev = NodeEvolverBatched(forest.n_nodes, physics, cooling_table)
    ev.set_graph_stall_limit(1e-9)
    with pytest.raises(_lib.GalError):
        tr.evolve_forest(ev, sched, torch.from_numpy(y0.copy()).cuda())
    torch.cuda.synchronize()
    ev.set_graph_stall_limit(0.0)
    ya = torch.from_numpy(y0.copy()).cuda()
    sa = tr.evolve_forest(ev, sched, ya)
    yb = torch.from_numpy(y0.copy()).cuda()
    sb = tr.evolve_forest(ev, sched, yb, executor="waves")
    assert torch.equal(ya, yb) and torch.equal(sa, sb)


def test_graph_executor_rejects_unknown_executor(torch_cuda, cooling_table):
    torch = torch_cuda
    from galacticus_amd.evolver import NodeEvolverBatched
    physics = gp.physics_params()
    cosmo = tr.cosmology(physics)
    forest = tr.Forest.build_pch(2, np.array([1e11, 2e11]), T0, cosmo, tr.pch_params(2e10, T_Z5), seed=1)
    wp = tr.walk_params(T0)
    ev = NodeEvolverBatched(forest.n_nodes, physics, cooling_table)
    yd = torch.from_numpy(forest.initial_state(cosmo, wp)).cuda()
    with pytest.raises(ValueError):
        tr.evolve_forest(ev, tr.Schedule(forest, cosmo, wp), yd, executor="levels")


def test_surrogate_route_rejecting_everything_equals_plain_forest(torch_cuda, cooling_table):
    """With a gate that rejects every step, the surrogate route of the tree
    executor (gather, network, gate, integrator on the compact copies,
    scatter back; satellites straight to the integrator) ends in exactly the
    plain executor's state, step sizes and statuses."""
    torch = torch_cuda
    from galacticus_amd.evolver import NodeEvolverBatched
    from galacticus_amd.surrogate import SurrogateMLP, SurrogateEvolver, GateConfig
    physics = gp.physics_params()
    cosmo = tr.cosmology(physics)
    forest = tr.Forest.build_pch(6, np.geomspace(1e11, 1e12, 6), T0, cosmo, tr.pch_params(2e10, T_Z5), seed=3)
    wp = tr.walk_params(T0, merging_multiplier=0.2)
    sched = tr.Schedule(forest, cosmo, wp)
    y0 = forest.initial_state(cosmo, wp)
    ev = NodeEvolverBatched(forest.n_nodes, physics, cooling_table)
    ya = torch.from_numpy(y0.copy()).cuda()
    sa = torch.zeros_like(ya)
    st_a = tr.evolve_forest(ev, sched, ya, scale=sa).cpu().numpy()
    norm = torch.zeros((4, 32), dtype=torch.float32, device="cuda")
    norm[1] = 1.0
    norm[3] = 1.0
    se = SurrogateEvolver(ev, SurrogateMLP(1024, seed=1), norm, GateConfig(error_max=1e-30))
    yb = torch.from_numpy(y0.copy()).cuda()
    sb = torch.zeros_like(yb)
    st_b = tr.evolve_forest(ev, sched, yb, scale=sb, surrogate=se).cpu().numpy()
    n_sat = int((sched.ops["node_par"][:, 4] != 0).sum())
    print("nodes", forest.n_nodes, "evolves", int((sched.ops["kind"] == 0).sum()), "satellite evolves", n_sat,
          "accepted", se.tree_accepted, flush=True)
    assert se.tree_accepted == 0
    assert np.array_equal(yb.cpu().numpy(), ya.cpu().numpy())
    assert np.array_equal(sb.cpu().numpy(), sa.cpu().numpy())
    assert np.array_equal(st_b, st_a)


def test_recorded_intervals_are_the_forest_trajectory(torch_cuda, cooling_table):
    """The recording executor leaves the forest exactly as the plain one, and
    every recorded interval is a piece of the trajectory: re-evolving its
    recorded start state over its interval lands on its recorded end state
    within the ODE tolerance (bit-exact is not expected: the recorded run
    reused step sizes from earlier intervals)."""
    torch = torch_cuda
    from galacticus_amd.evolver import NodeEvolverBatched
    physics = gp.physics_params()
    cosmo = tr.cosmology(physics)
    forest = tr.Forest.build_pch(6, np.geomspace(1e11, 1e12, 6), T0, cosmo, tr.pch_params(2e10, T_Z5), seed=3)
    wp = tr.walk_params(T0, merging_multiplier=0.2)
    sched = tr.Schedule(forest, cosmo, wp)
    y0 = forest.initial_state(cosmo, wp)
    ev = NodeEvolverBatched(forest.n_nodes, physics, cooling_table, odeToleranceAbsolute=1e-6,
                            odeToleranceRelative=1e-3)
    ya = torch.from_numpy(y0.copy()).cuda()
    tr.evolve_forest(ev, sched, ya)
    yb = torch.from_numpy(y0.copy()).cuda()
    r0, r1, npar, t0, t1, st = tr.record_forest(ev, sched, yb)
    n = int(t0.shape[0])
    print("recorded", n, "of", int((sched.ops["kind"] == 0).sum()), flush=True)
    assert n == int((sched.ops["kind"] == 0).sum())
    assert torch.equal(ya, yb) and bool((st == 0).all())
    ev2 = NodeEvolverBatched(n, physics, cooling_table, odeToleranceAbsolute=1e-6, odeToleranceRelative=1e-3)
    y = r0.clone()
    sc = torch.zeros_like(y)
    st2 = ev2.evolve_device(y, npar.contiguous(), t0, t1, -torch.ones(n, dtype=torch.float64, device="cuda"), scale=sc)
    assert bool((st2 == 0).all())
    mass = [V[k] for k in ("HOT_MASS", "DISK_MASS_GAS", "DISK_MASS_STELLAR", "SPHEROID_MASS_GAS",
                           "SPHEROID_MASS_STELLAR", "BH_MASS")]
    got, ref, s = y[mass].cpu().numpy(), r1[mass].cpu().numpy(), sc[mass].cpu().numpy()
    tol = 1e-3 * np.abs(ref) + 1e-6 * np.abs(s)
    worst = (np.abs(got - ref) / np.maximum(tol, 1e-300)).max()
    print("max |dy| / tolerance", worst, flush=True)
    assert worst < 50.0


def test_64_tree_workload_against_the_cpu_tree_runner(torch_cuda, cooling_table):
    """The 64-tree workload at full size (PCH, seed 0, roots log-uniform
    1e10-1e12 Msun, resolution 1e9 Msun, z=5 -> 0): the GPU wave executor
    against the CPU oracle's tree runner on the same schedule.  Both take the
    same steps almost everywhere (median relative difference of the masses
    below 1e-9); where they do not they agree to the ODE tolerance; the z=0
    stellar masses of the roots agree to 1e-4, a tenth of the relative ODE
    tolerance (a root integrates its progenitors' merged galaxies over
    hundreds of steps, and a last-place difference that changes one step
    size of one progenitor moves its z=0 mass by a few 1e-6)."""
    torch = torch_cuda
    from galacticus_amd.evolver import NodeEvolverBatched
    physics = gp.physics_params()
    cosmo = tr.cosmology(physics)
    rng = np.random.Generator(np.random.PCG64(0))
    forest = tr.Forest.build_pch(64, 10.0 ** rng.uniform(10.0, 12.0, 64), T0, cosmo, tr.pch_params(1e9, T_Z5), seed=0)
    wp = tr.walk_params(T0)
    sched = tr.Schedule(forest, cosmo, wp)
    y0 = forest.initial_state(cosmo, wp)
    print("nodes", forest.n_nodes, "operations", sched.n_ops, "waves", sched.n_waves, flush=True)
    ev = NodeEvolverBatched(forest.n_nodes, physics, cooling_table, odeToleranceAbsolute=1e-6,
                            odeToleranceRelative=1e-3)
    yd = torch.from_numpy(y0.copy()).cuda()
    sd = torch.zeros_like(yd)
    st = tr.evolve_forest(ev, sched, yd, scale=sd).cpu().numpy()
    yg, sc = yd.cpu().numpy(), sd.cpu().numpy()
    yc, _, stc, bad = oracle.tree_run(sched.ops, y0, -np.ones(forest.n_nodes), forest.tree_of_node(),
                                      forest.n_trees, physics, cooling_table, 1e-6, 1e-3, gp.FLAG_REUSE_STEP_SIZE,
                                      0.25, root_tol_rel=1e-15, nthreads=8)
    assert bad == 0 and (st == 0).all() and (stc == 0).all()
    mass = [V[k] for k in ("HOT_MASS", "DISK_MASS_GAS", "DISK_MASS_STELLAR", "SPHEROID_MASS_GAS",
                           "SPHEROID_MASS_STELLAR", "BH_MASS")]
    ref = np.abs(yc[mass])
    big = ref > 1e-6 * np.abs(yc[V["HOT_MASS"]]).max()
    rel = np.abs(yg[mass] - yc[mass])[big] / ref[big]
    tol = 1e-3 * np.abs(yc[mass]) + 1e-6 * np.abs(sc[mass])
    worst = (np.abs(yg[mass] - yc[mass]) / np.maximum(tol, 1e-300)).max()
    print("median rel", np.median(rel), "max |dy|/tol", worst, flush=True)
    assert np.median(rel) < 1e-9 and worst < 50.0
    r = forest.roots()
    mg = yg[V["DISK_MASS_STELLAR"], r] + yg[V["SPHEROID_MASS_STELLAR"], r]
    mc = yc[V["DISK_MASS_STELLAR"], r] + yc[V["SPHEROID_MASS_STELLAR"], r]
    print("root stellar mass max rel", float(np.max(np.abs(mg - mc) / mc)), flush=True)
    assert np.allclose(mg, mc, rtol=1e-4, atol=0)


BARYONS = ["HOT_MASS", "HOT_OUTFLOWED_MASS", "HOT_UNACCRETED_MASS", "HOT_STRIPPED_MASS", "DISK_MASS_GAS",
           "DISK_MASS_STELLAR", "SPHEROID_MASS_GAS", "SPHEROID_MASS_STELLAR", "BH_MASS"]
METALS = ["HOT_ABUNDANCES", "HOT_OUTFLOWED_ABUNDANCES", "HOT_UNACCRETED_ABUNDANCES", "HOT_STRIPPED_ABUNDANCES",
          "DISK_ABUNDANCES_GAS", "DISK_ABUNDANCES_STELLAR", "SPHEROID_ABUNDANCES_GAS", "SPHEROID_ABUNDANCES_STELLAR"]


def test_large_forest_sample_budgets_and_cpu_agreement(torch_cuda, cooling_table):
    """The first 8 trees of the 4096-tree large-forest workload (PCH, seed 0,
    roots log-uniform 1e11-1e14 Msun, resolution 1e9 Msun, z=8 -> 0; the
    builder seeds each tree by its index, so these have that workload's first
    seeds and roots, with branchings that differ in detail because the
    builder's rate table spans the call's largest root) as one operation
    graph on the GPU:
      * every evolve succeeds;
      * per tree, baryons and metals close over the nodes still alive at z=0
        (never promoted away or merged): the sources are the leaves' initial
        galaxies, accretion f_b dM/dt dt onto halos that are not satellites
        and the yield of new stars; the sinks are the black holes' radiated
        and jet energy, at most eff/(1 - eff) <= 0.75 of their growth, for
        baryons and metals alike (the expulsion by excess heating and the
        clipping of negative reservoirs are off so that every step conserves
        exactly); the residual is held to 1e-7 of the tree's budget;
      * the trees with the smallest schedules equal the CPU oracle's tree
        runner on the same operations to the ODE tolerance, with the z=0
        stellar masses of their roots within 1e-4."""
    torch = torch_cuda
    from galacticus_amd.evolver import NodeEvolverBatched
    physics = gp.physics_params(**{"CGMCoolingHeating.excessHeatDrivesOutflow": 0.0,
                                   "postStepZeroNegativeMasses": 0.0})
    cosmo = tr.cosmology(physics)
    t_z8 = float(wl.scale_factor_to_cosmic_time(1.0 / 9.0))
    n_trees = 8
    roots = 10.0 ** np.random.Generator(np.random.PCG64(0)).uniform(11.0, 14.0, 4096)[:n_trees]
    forest = tr.Forest.build_pch(n_trees, roots, T0, cosmo, tr.pch_params(1e9, t_z8), seed=0)
    wp = tr.walk_params(T0)
    sched = tr.Schedule(forest, cosmo, wp)
    ops = sched.ops
    tree = forest.tree_of_node()
    print("roots", roots, "nodes", forest.n_nodes, "operations", sched.n_ops, "waves", sched.n_waves, flush=True)
    y0 = forest.initial_state(cosmo, wp)
    ev = NodeEvolverBatched(forest.n_nodes, physics, cooling_table, odeToleranceAbsolute=1e-6,
                            odeToleranceRelative=1e-3)
    yd = torch.from_numpy(y0.copy()).cuda()
    sd = torch.zeros_like(yd)
    st = tr.evolve_forest(ev, sched, yd, scale=sd).cpu().numpy()
    yg, sc = yd.cpu().numpy(), sd.cpu().numpy()
    evolved = np.unique(ops["a"][ops["kind"] == tr.OP_EVOLVE])
    assert (st[evolved] == 0).all()
    # budgets per tree over the nodes alive at the end
    gone = np.zeros(forest.n_nodes, bool)
    gone[ops["a"][(ops["kind"] == tr.OP_PROMOTE) | (ops["kind"] == tr.OP_MERGE_GALAXY)]] = True
    alive = ~gone
    ev_ops = ops[ops["kind"] == tr.OP_EVOLVE]
    fb = physics[gp.P["OmegaBaryon"]] / physics[gp.P["OmegaMatter"]]
    yld = physics[gp.P["metalYield"]]
    acc_op = np.where(ev_ops["node_par"][:, gp.NP_IS_SATELLITE] != 0, 0.0,
                      fb * ev_ops["node_par"][:, gp.NP_MASS_RATE] * (ev_ops["t1"] - ev_ops["t0"]))
    acc = np.bincount(tree[ev_ops["a"]], weights=acc_op, minlength=n_trees)

    def per_tree(y, names, mask):
        return np.bincount(tree[mask], weights=sum(y[V[k]] for k in names)[mask], minlength=n_trees)
    every = np.ones(forest.n_nodes, bool)
    b0, b1 = per_tree(y0, BARYONS, every), per_tree(yg, BARYONS, alive)
    bh0, bh1 = per_tree(y0, ["BH_MASS"], every), per_tree(yg, ["BH_MASS"], alive)
    dbh = np.maximum(bh1 - bh0, 0.0)
    loss = acc - (b1 - b0)
    scale = b0 + acc
    print("baryon loss / budget", loss / scale, "bh growth / budget", dbh / scale, flush=True)
    assert (loss >= -1e-7 * scale).all()
    assert (loss <= 0.75 * dbh + 1e-7 * scale).all()
    formed = per_tree(yg, ["DISK_MASS_STELLAR_FORMED", "SPHEROID_MASS_STELLAR_FORMED"], alive) - \
        per_tree(y0, ["DISK_MASS_STELLAR_FORMED", "SPHEROID_MASS_STELLAR_FORMED"], every)
    z0, z1 = per_tree(y0, METALS, every), per_tree(yg, METALS, alive)
    zres = (z1 - z0) - yld * formed
    zscale = z0 + yld * np.abs(formed)
    print("metal residual / budget", zres / zscale, flush=True)
    assert (zres <= 1e-7 * zscale).all()
    assert (-zres <= 2.5 * dbh + 1e-7 * zscale).all()
    # the smallest trees against the CPU tree runner
    n_ops_tree = np.bincount(tree[ops["a"]], minlength=n_trees)
    small = np.nonzero(n_ops_tree <= 200000)[0]
    print("operations per tree", n_ops_tree, "compared", small, flush=True)
    assert len(small) >= 3
    sel = np.isin(tree[ops["a"]], small)
    yc, _, stc, bad = oracle.tree_run(ops[sel], y0, -np.ones(forest.n_nodes), tree, n_trees, physics, cooling_table,
                                      1e-6, 1e-3, gp.FLAG_REUSE_STEP_SIZE, 0.25, root_tol_rel=1e-15, nthreads=8)
    assert bad == 0
    nodes = np.isin(tree, small)
    mass = [V[k] for k in ("HOT_MASS", "DISK_MASS_GAS", "DISK_MASS_STELLAR", "SPHEROID_MASS_GAS",
                           "SPHEROID_MASS_STELLAR", "BH_MASS")]
    g, c, s = yg[mass][:, nodes], yc[mass][:, nodes], sc[mass][:, nodes]
    ref = np.abs(c)
    big = ref > 1e-6 * np.abs(yc[V["HOT_MASS"]][nodes]).max()
    rel = np.abs(g - c)[big] / ref[big]
    worst = (np.abs(g - c) / np.maximum(1e-3 * ref + 1e-6 * np.abs(s), 1e-300)).max()
    print("median rel", np.median(rel), "max |dy|/tol", worst, flush=True)
    assert np.median(rel) < 1e-9 and worst < 50.0
    r = forest.roots()[small]
    mg = yg[V["DISK_MASS_STELLAR"], r] + yg[V["SPHEROID_MASS_STELLAR"], r]
    mc = yc[V["DISK_MASS_STELLAR"], r] + yc[V["SPHEROID_MASS_STELLAR"], r]
    assert np.allclose(mg, mc, rtol=1e-4, atol=0)
