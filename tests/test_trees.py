"""CPU tests of the tree path's host side (include/gal_tree.h).

* cosmology: sigma(M) normalised to sigma_8; delta_c(t0) equals the spherical
  collapse value with the Navarro, Frenk & White (1997) Omega dependence;
  delta_c a -> constant in the matter era (D proportional to a);
* the Parkinson-Cole-Helly forest: mass bookkeeping of every split and
  accretion step, parents later than children, resolution respected, and the
  main-progenitor mass at z=1 in the range of published PCH trees;
* the tree walk: the engine's schedule equals, operation for operation and
  bit for bit, an independent restatement of the reference's walk
  (oracle/trees.py), which is itself pinned on small trees whose operation
  sequence follows by hand from the reference's rules;
* waves: every node sees its operations in schedule order, one per wave;
* the fully-specified tree reader on the reference's fixture trees.
"""
import math
import os

import numpy as np
import pytest

from galacticus_amd import trees as tr
from galacticus_amd import workloads as wl
from oracle import trees as ot

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
T0 = float(wl.scale_factor_to_cosmic_time(1.0))
T_Z5 = float(wl.scale_factor_to_cosmic_time(1.0 / 6.0))


@pytest.fixture(scope="module")
def cosmo():
    return tr.cosmology()


def _oracle_walk(forest, cosmo, wp):
    w = {k: getattr(wp, k) for k, _ in tr.WalkParams._fields_}
    cos = ot.Cosmo(cosmo.omega_matter, cosmo.omega_dark_energy, cosmo.hubble_constant)
    trees = [(int(forest.tree_offset[i]), int(forest.tree_offset[i + 1])) for i in range(forest.n_trees)]
    return ot.walk(forest.parent, forest.first_child, forest.sibling, forest.mass, forest.time, cos, w, trees)


def _same(ops, log):
    if len(ops) != len(log):
        return False
    for op, lg in zip(ops, log):
        if (int(op["kind"]), int(op["a"]), int(op["b"]), float(op["t0"]), float(op["t1"])) != tuple(lg[:5]):
            return False
        if lg[5] is not None and list(op["node_par"]) != list(lg[5]):
            return False
    return True


def test_sigma_normalisation_and_shape(cosmo):
    rho = cosmo.omega_matter * 3.0 * cosmo.hubble_constant ** 2 / (8.0 * math.pi * wl.G_INTERNAL)
    m8 = 4.0 / 3.0 * math.pi * (8.0 / (cosmo.hubble_constant / 100.0)) ** 3 * rho
    s = tr.sigma(cosmo, [m8])[0]
    assert abs(s - cosmo.sigma_8) < 1e-3 * cosmo.sigma_8
    m = np.geomspace(1e8, 1e15, 30)
    sig = tr.sigma(cosmo, m)
    assert np.all(np.diff(sig) < 0)                       # decreasing with mass
    slope = np.diff(np.log(sig)) / np.diff(np.log(m))
    assert np.all(slope < 0) and np.all(slope > -0.5)     # CDM: shallow at low mass, steeper at high
    assert slope[0] > slope[-1]


def test_delta_critical(cosmo):
    d0 = tr.delta_critical(cosmo, [T0])[0]
    expect = 0.15 * (12.0 * math.pi) ** (2.0 / 3.0) * (1.0 + 0.0123 * math.log10(cosmo.omega_matter))
    assert abs(d0 - expect) < 2e-4 * expect
    # matter era: linear growth D proportional to a, so delta_c * a is constant
    ts = wl.scale_factor_to_cosmic_time(np.array([0.01, 0.02]))
    a = np.array([0.01, 0.02])
    dc = tr.delta_critical(cosmo, ts)
    om_a = cosmo.omega_matter / a ** 3 / (cosmo.omega_matter / a ** 3 + cosmo.omega_dark_energy)
    base = 0.15 * (12.0 * math.pi) ** (2.0 / 3.0) * (1.0 + 0.0123 * np.log10(om_a))
    growth = base / dc
    assert abs(growth[1] / growth[0] - 2.0) < 2e-3


def test_pch_forest_bookkeeping(cosmo):
    mres = 1e9
    f = tr.Forest.build_pch(24, np.geomspace(1e10, 1e12, 24), T0, cosmo, tr.pch_params(mres, T_Z5), seed=1)
    assert f.n_trees == 24 and f.tree_offset[-1] == f.n_nodes
    for t in range(f.n_trees):
        lo, hi = f.tree_offset[t], f.tree_offset[t + 1]
        assert f.parent[lo] == -1 and f.time[lo] == T0
        for i in range(lo, hi):
            kids = []
            k = f.first_child[i]
            while k >= 0:
                kids.append(k)
                assert f.parent[k] == i and f.time[k] < f.time[i]
                k = f.sibling[k]
            if kids:
                tot = f.mass[kids].sum()
                # children carry the node's mass less the sub-resolution accretion
                # (<= the accretion limit) and less a split-off branch at or
                # below the resolution, which is not kept (Cole2000/_class.F90:827-850)
                assert tot <= f.mass[i] * (1 + 1e-12)
                assert tot >= f.mass[i] * (1 - 0.1) * (1 - 1e-9) - mres
                assert f.mass[kids[0]] >= f.mass[kids[-1]]      # the primary progenitor is the largest
            assert f.mass[i] > 0.85 * mres or i == lo
            assert f.time[i] >= T_Z5 * (1 - 1e-9)


def test_pch_main_progenitor_at_z1(cosmo):
    # Mean main-progenitor mass fraction of 1e12 Msun halos at z=1: published
    # PCH and Millennium trees give about one half.
    f = tr.Forest.build_pch(200, 1e12, T0, cosmo, tr.pch_params(1e9, T_Z5), seed=7)
    t1 = float(wl.scale_factor_to_cosmic_time(0.5))
    fr = []
    for t in range(f.n_trees):
        i = f.tree_offset[t]
        while f.first_child[i] >= 0 and f.time[f.first_child[i]] >= t1:
            i = f.first_child[i]
        c = f.first_child[i]
        m = f.mass[i] if c < 0 else f.mass[c] + (f.mass[i] - f.mass[c]) * (t1 - f.time[c]) / (f.time[i] - f.time[c])
        fr.append(m / 1e12)
    assert 0.35 < np.mean(fr) < 0.7, np.mean(fr)


def test_walk_matches_oracle_restatement(cosmo):
    for seed in range(4):
        f = tr.Forest.build_pch(12, np.geomspace(1e11, 1e12, 12), T0, cosmo, tr.pch_params(1e10, T_Z5), seed=seed)
        wp = tr.walk_params(T0)
        s = tr.Schedule(f, cosmo, wp)
        log = _oracle_walk(f, cosmo, wp)
        kinds = np.bincount(s.ops["kind"], minlength=4)
        assert kinds[tr.OP_PROMOTE] > 0 and kinds[tr.OP_BECOME_SATELLITE] > 0 and kinds[tr.OP_MERGE_GALAXY] > 0
        assert _same(s.ops, log), seed


def _steps(t, t_end, cos):
    """Hand restatement of consecutive simple timesteps (simple.F90:124-155)."""
    out = []
    while t < t_end:
        t1 = min(t + min(0.1 * cos.hubble_time(t), 1.0), t_end)
        out.append((t, t1))
        t = t1
    return out


def test_schedule_node_gyr_and_lazy_operations(cosmo):
    """gal_schedule_node_gyr sums the EVOLVE intervals without copying the
    operations; the operations are copied out on first use only, and a
    closed schedule refuses both."""
    f = tr.Forest.build_pch(6, np.geomspace(1e11, 1e12, 6), T0, cosmo, tr.pch_params(1e10, T_Z5), seed=3)
    s = tr.Schedule(f, cosmo, tr.walk_params(T0))
    assert s._ops is None
    g = s.node_gyr()
    ev = s.ops[s.ops["kind"] == tr.OP_EVOLVE]
    assert len(ev) > 0 and s.ops is s.ops
    assert g == pytest.approx(float(np.sum(ev["t1"] - ev["t0"])), rel=1e-13, abs=0.0)
    t = tr.Schedule(f, cosmo, tr.walk_params(T0))
    t.close()
    with pytest.raises(ValueError):
        t.ops
    with pytest.raises(Exception):
        t.node_gyr()


def test_walk_two_node_tree(cosmo):
    # closedBoxTree: node 1 (t=12.47) is the only progenitor of node 2 (t=13.47):
    # one interval to the parent's time, then the promotion.
    f = tr.Forest.from_arrays([0, 2], [-1, 0], [1, -1], [-1, -1], [1e12, 1e12], [13.47, 12.47])
    log = _oracle_walk(f, cosmo, tr.walk_params(13.47))
    assert [(k, a, b) for k, a, b, *_ in log] == [(ot.EVOLVE, 1, -1), (ot.PROMOTE, 1, 0)]
    assert log[0][3] == 12.47 and log[0][4] == 13.47
    s = tr.Schedule(f, cosmo, tr.walk_params(13.47))
    assert _same(s.ops, log)


def test_walk_three_node_tree(cosmo):
    # Root R (t=10) with progenitors A (primary, 8e11) and B (1e11) at t=8.
    # The walk visits A first, which may not pass its sibling B; B then
    # evolves in simple timesteps to t=10 and becomes a satellite of R; the
    # walker moves on to A, which evolves to t=10 and is promoted into R.
    cos = ot.Cosmo(cosmo.omega_matter, cosmo.omega_dark_energy, cosmo.hubble_constant)
    f = tr.Forest.from_arrays([0, 3], [-1, 0, 0], [1, -1, -1], [-1, 2, -1], [1e12, 8e11, 1e11], [10.0, 8.0, 8.0])
    log = _oracle_walk(f, cosmo, tr.walk_params(10.0))
    steps = _steps(8.0, 10.0, cos)
    expect = [(ot.EVOLVE, 2, -1, a, b) for a, b in steps] + [(ot.SATELLITE, 2, 0)] + \
             [(ot.EVOLVE, 1, -1, a, b) for a, b in steps] + [(ot.PROMOTE, 1, 0)]
    got = [lg[:5] if lg[0] == ot.EVOLVE else lg[:3] for lg in log]
    assert got == expect
    s = tr.Schedule(f, cosmo, tr.walk_params(10.0))
    assert _same(s.ops, log)


def test_walk_satellite_merger(cosmo):
    # Root R (t=T0) <- M (t=9) <- {A, B} (t=6, equal masses): B becomes a
    # satellite of M at t=9, A is promoted into M, and M then evolves towards
    # T0 with B in tow; with a short merging time B merges into M's galaxy
    # at its merging time, and M is finally promoted into R.
    f = tr.Forest.from_arrays([0, 4], [-1, 0, 1, 1], [1, 2, -1, -1], [-1, -1, 3, -1], [3e12, 2e12, 1e12, 1e12],
                              [T0, 9.0, 6.0, 6.0])
    wp = tr.walk_params(T0, merging_multiplier=0.05)
    log = _oracle_walk(f, cosmo, wp)
    kinds = [lg[0] for lg in log]
    assert kinds.count(ot.SATELLITE) == 1 and kinds.count(ot.MERGE) == 1 and kinds.count(ot.PROMOTE) == 2
    sat = next(lg for lg in log if lg[0] == ot.SATELLITE)
    merge = next(lg for lg in log if lg[0] == ot.MERGE)
    assert sat[1] == 3 and sat[2] == 1 and sat[3] == 9.0
    assert merge[1] == 3 and merge[2] == 1                 # satellite B merges into M's galaxy
    assert 9.0 < sat[4] < T0
    assert abs(merge[3] - sat[4]) <= 1e-2                  # at its merging time (satellite offsets)
    assert kinds.index(ot.MERGE) < len(kinds) - 1 and log[-1][0] == ot.PROMOTE and log[-1][1:3] == (1, 0)
    assert _same(tr.Schedule(f, cosmo, wp).ops, log)


def test_waves_keep_per_node_order(cosmo):
    f = tr.Forest.build_pch(16, np.geomspace(1e11, 3e12, 16), T0, cosmo, tr.pch_params(1e10, T_Z5), seed=11)
    s = tr.Schedule(f, cosmo, tr.walk_params(T0))
    last = {}
    per_wave = {}
    for op in s.ops:
        nodes = [int(op["a"])] + ([int(op["b"])] if op["b"] >= 0 and op["kind"] != tr.OP_BECOME_SATELLITE else [])
        for n in nodes:
            assert op["wave"] > last.get(n, -1)
            last[n] = int(op["wave"])
            key = (int(op["wave"]), n)
            assert key not in per_wave          # one operation per node per wave
            per_wave[key] = True
    assert s.n_waves == 1 + max(int(w) for w in s.ops["wave"])


def test_fully_specified_reader():
    text = open(os.path.join(GOLDEN, "closedBoxTree.xml")).read()
    f, comps = tr.Forest.from_xml(text)
    assert f.n_nodes == 2 and f.n_trees == 1
    assert f.parent[0] == -1 and f.first_child[0] == 1 and f.parent[1] == 0
    assert f.time[0] == 13.47 and f.time[1] == 12.47 and f.mass[1] == 1e12
    assert comps[1]["disk/massGas"] == 1e11 and comps[1]["spin/angularMomentum"] == 1.635e12
    text2 = open(os.path.join(GOLDEN, "enforceNonNegativityTree.xml")).read()
    f2, comps2 = tr.Forest.from_xml([text, text2])
    assert f2.n_trees == 2 and f2.n_nodes == 4
    assert comps2[3]["spheroid/massGas"] == 1e9 and comps2[3]["disk/angularMomentum"] == 7.7e10


def test_forest_from_arrays_rejects_bad_links():
    with pytest.raises(Exception):
        tr.Forest.from_arrays([0, 2], [-1, 5], [1, -1], [-1, -1], [1e12, 1e12], [13.47, 12.47])


def test_c_tree_runner_equals_sequential_restatement(cosmo, cooling_table):
    """oracle.tree_run (C, trees in parallel) applies a schedule exactly as the
    Python restatement does one operation at a time: identical states."""
    import oracle
    from galacticus_amd import params as gp
    physics = gp.physics_params()
    t0 = float(wl.scale_factor_to_cosmic_time(1.0))
    t5 = float(wl.scale_factor_to_cosmic_time(1.0 / 6.0))
    forest = tr.Forest.build_pch(3, np.geomspace(3e10, 1e11, 3), t0, cosmo, tr.pch_params(5e9, t5), seed=4)
    wp = tr.walk_params(t0, merging_multiplier=0.2)
    sched = tr.Schedule(forest, cosmo, wp)
    y0 = forest.initial_state(cosmo, wp)
    h0 = -np.ones(forest.n_nodes)
    flags = gp.FLAG_REUSE_STEP_SIZE

    def evolve_one(col, npar, a, b, h):
        yo, ho, _, st, _ = oracle.evolve(col[:, None], np.asarray(npar)[:, None], [a], [b], [h], physics,
                                         cooling_table, 1e-6, 1e-3, flags)
        return yo[:, 0], ho[0], st[0]

    log = [(int(o["kind"]), int(o["a"]), int(o["b"]), float(o["t0"]), float(o["t1"]), list(o["node_par"]))
           for o in sched.ops]
    yp, hp, stp = ot.run_sequential(log, y0, h0, evolve_one, lambda y, a, b: ot.merge_galaxy_simple(y, a, b, 0.25))
    yc, hc, stc, bad = oracle.tree_run(sched.ops, y0, h0, forest.tree_of_node(), forest.n_trees, physics,
                                       cooling_table, 1e-6, 1e-3, flags, 0.25, nthreads=3)
    assert bad == 0
    assert np.array_equal(yc, yp) and np.array_equal(hc, hp) and np.array_equal(stc, stp)


def test_forest_split_keeps_trees(cosmo):
    """Splitting a forest into parts of consecutive trees keeps every tree:
    the parts' schedules are the whole forest's operations, re-indexed."""
    t0 = float(wl.scale_factor_to_cosmic_time(1.0))
    t5 = float(wl.scale_factor_to_cosmic_time(1.0 / 6.0))
    forest = tr.Forest.build_pch(7, np.geomspace(3e10, 3e11, 7), t0, cosmo, tr.pch_params(5e9, t5), seed=8)
    wp = tr.walk_params(t0)
    whole = tr.Schedule(forest, cosmo, wp).ops
    parts = forest.split(3)
    assert sum(f.n_nodes for f, _ in parts) == forest.n_nodes and sum(f.n_trees for f, _ in parts) == 7
    ops = []
    for f, n0 in parts:
        o = tr.Schedule(f, cosmo, wp).ops.copy()
        o["a"] += n0
        o["b"] = np.where(o["b"] >= 0, o["b"] + n0, o["b"])
        ops.append(o)
    ops = np.concatenate(ops)
    key = lambda o: sorted(zip(o["kind"].tolist(), o["a"].tolist(), o["b"].tolist(), o["t0"].tolist(),
                               o["t1"].tolist()))
    assert key(ops) == key(whole)


def test_output_layout_and_tree_contiguity(cosmo, tmp_path):
    """The Galacticus-layout output: one entry per surviving node, trees in
    contiguous blocks located by mergerTreeStartIndex/Count (the check of
    testSuite/test-output-tree-contiguousity.py restated), properties equal
    to the state, promoted children and merged satellites absent."""
    from galacticus_amd import output as go
    t0 = float(wl.scale_factor_to_cosmic_time(1.0))
    t5 = float(wl.scale_factor_to_cosmic_time(1.0 / 6.0))
    forest = tr.Forest.build_pch(5, np.geomspace(3e10, 3e11, 5), t0, cosmo, tr.pch_params(5e9, t5), seed=6)
    wp = tr.walk_params(t0, merging_multiplier=0.2)
    sched = tr.Schedule(forest, cosmo, wp)
    y = np.random.default_rng(0).uniform(1.0, 2.0, (30, forest.n_nodes))
    path = str(tmp_path / "out.npz")
    go.write_output(path, [go.output_group(forest, sched, y, t0, 1.0)])
    f = go.read_output(path)
    out = f["Outputs"]["Output1"]
    nd = out["nodeData"]
    tree_by_node = nd["mergerTreeIndex"]
    for idx, start, count in zip(out["mergerTreeIndex"], out["mergerTreeStartIndex"], out["mergerTreeCount"]):
        assert (tree_by_node[start:start + count] == idx).all()
    assert out["mergerTreeCount"].sum() == len(nd["nodeIndex"])
    nodes = nd["nodeIndex"]
    assert set(forest.roots().tolist()) <= set(nodes.tolist())
    k = sched.ops["kind"]
    removed = set(sched.ops["a"][(k == tr.OP_PROMOTE) | (k == tr.OP_MERGE_GALAXY)].tolist())
    assert not (removed & set(nodes.tolist()))
    assert np.array_equal(f["Outputs/Output1/nodeData/diskMassStellar"], y[wl.V["DISK_MASS_STELLAR"], nodes])
    assert (nd["nodeIsIsolated"][np.isin(nodes, forest.roots())] == 1).all()


@pytest.mark.parametrize("major", [False, True])
def test_merge_galaxy_simple_hand_cases(major):
    """mergerMassMovementsSimple with spheroid destinations
    (satellites/merging/mass_movements/simple.F90:208-290), by hand: the
    satellite's cold gas and stars (and their metals, formed-star tallies and
    angular momentum) land in the host spheroid, its circumgalactic
    reservoirs and black hole join the host's, its stripped gas the host's
    outflowed reservoir; in a major merger (smaller
    galaxy at least 0.25 of the larger) the host disk moves to the spheroid
    too; the satellite is left empty."""
    V = wl.V
    y = np.zeros((30, 2))
    sat, host = 0, 1
    y[V["DISK_MASS_GAS"], sat], y[V["DISK_MASS_STELLAR"], sat], y[V["SPHEROID_MASS_STELLAR"], sat] = 1.0, 2.0, 0.5
    y[V["DISK_ABUNDANCES_GAS"], sat], y[V["HOT_MASS"], sat], y[V["BH_MASS"], sat] = 0.01, 3.0, 1e-3
    y[V["HOT_STRIPPED_MASS"], sat], y[V["HOT_STRIPPED_ABUNDANCES"], sat] = 0.7, 0.007
    y[V["HOT_OUTFLOWED_MASS"], host] = 1.5
    y[V["DISK_MASS_GAS"], host], y[V["DISK_MASS_STELLAR"], host] = 4.0, (6.0 if major else 40.0)
    y[V["SPHEROID_MASS_GAS"], host], y[V["HOT_MASS"], host], y[V["BH_MASS"], host] = 0.2, 10.0, 2e-3
    before = y.copy()
    ot.merge_galaxy_simple(y, sat, host, 0.25)
    m_sat = 1.0 + 2.0 + 0.5
    m_host = 4.0 + (6.0 if major else 40.0) + 0.2
    assert (min(m_sat, m_host) >= 0.25 * max(m_sat, m_host)) == major
    assert np.all(y[:, sat][[V[k] for k in ("DISK_MASS_GAS", "DISK_MASS_STELLAR", "SPHEROID_MASS_STELLAR",
                                             "DISK_ABUNDANCES_GAS", "HOT_MASS", "BH_MASS")]] == 0.0)
    assert y[V["HOT_MASS"], host] == 13.0 and y[V["BH_MASS"], host] == 3e-3
    # the satellite's stripped gas joins the host's outflowed reservoir
    assert y[V["HOT_OUTFLOWED_MASS"], host] == 1.5 + 0.7 and y[V["HOT_OUTFLOWED_ABUNDANCES"], host] == 0.007
    assert y[V["HOT_STRIPPED_MASS"], sat] == 0.0 and y[V["HOT_STRIPPED_ABUNDANCES"], sat] == 0.0
    if major:
        assert y[V["SPHEROID_MASS_GAS"], host] == 0.2 + 1.0 + 4.0
        assert y[V["SPHEROID_MASS_STELLAR"], host] == 0.5 + 2.0 + 6.0
        assert y[V["DISK_MASS_GAS"], host] == 0.0 and y[V["DISK_MASS_STELLAR"], host] == 0.0
    else:
        assert y[V["SPHEROID_MASS_GAS"], host] == 0.2 + 1.0
        assert y[V["SPHEROID_MASS_STELLAR"], host] == 0.5 + 2.0
        assert y[V["DISK_MASS_GAS"], host] == 4.0 and y[V["DISK_MASS_STELLAR"], host] == 40.0
    assert y[V["SPHEROID_ABUNDANCES_GAS"], host] == 0.01
    # total baryons conserved
    mass = [V[k] for k in ("HOT_MASS", "HOT_OUTFLOWED_MASS", "HOT_STRIPPED_MASS", "DISK_MASS_GAS",
                           "DISK_MASS_STELLAR", "SPHEROID_MASS_GAS", "SPHEROID_MASS_STELLAR", "BH_MASS")]
    assert np.isclose(y[mass].sum(), before[mass].sum(), rtol=1e-15)
