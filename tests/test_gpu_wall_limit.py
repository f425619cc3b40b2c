"""The node evolver's wall-time limit on the GPU (errorStatusXCPU, 1025).

The reference checks System_Clock against systemClockMaximum in the trial
loop and in every RHS call (node_evolver/standard.F90:693-705, 860-866) and
returns errorStatusXCPU with the node's state restored.  Here the limit is
checked on the device before each Cash-Karp attempt, in ticks of the GPU's
realtime counter counted from the start of the launch:
  * a limit far below one attempt: every node returns 1025 with its state
    and step size untouched;
  * a limit that the stiffest node of the 1e5-node benchmark batch cannot
    meet (its chain of 3701 attempts takes over 100 ms) but most nodes can:
    every status is 0 or 1025, the stiff node's is 1025, nodes that succeed
    are bit-identical to a run without the limit, nodes that do not keep
    their input state;
  * removing the limit restores full success;
  * a forest under a limit below one attempt: the operation graph drains
    with every evolve at 1025 and equals the per-wave executor.
The limit is forced from the host: every kernel exits normally.
"""
import numpy as np
import pytest

from galacticus_amd import params as gp
from galacticus_amd import workloads as wl

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def _run(ev, torch, b):
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    y, h = d(b.y), d(b.h_step)
    st = ev.evolve_device(y, d(b.node_par), d(b.t_start), d(b.t_end), h)
    torch.cuda.synchronize()
    return y.cpu().numpy(), h.cpu().numpy(), st.cpu().numpy()


def test_wall_time_limit_returns_xcpu(torch_cuda, cooling_table):
    torch = torch_cuda
    from galacticus_amd.evolver import NodeEvolverBatched
    p = gp.physics_params()
    b = wl.isolated_halos(100000, seed=0, params=p)
    stiff = 19753                  # 3701 attempts at rel 1e-3
    ev = NodeEvolverBatched(b.n, p, cooling_table, odeToleranceAbsolute=1e-6, odeToleranceRelative=1e-3)
    print("no limit", flush=True)
    y_ref, h_ref, st_ref = _run(ev, torch, b)
    assert (st_ref == 0).all()

    print("limit 1e-7 s", flush=True)
    ev.set_wall_limit(1e-7)
    y, h, st = _run(ev, torch, b)
    assert (st == gp.STATUS_XCPU).all()
    assert np.array_equal(y, b.y) and np.array_equal(h, b.h_step)

    print("limit 0.03 s", flush=True)
    ev.set_wall_limit(0.03)
    y, h, st = _run(ev, torch, b)
    ok, xcpu = st == 0, st == gp.STATUS_XCPU
    print("succeeded", int(ok.sum()), "xcpu", int(xcpu.sum()), flush=True)
    assert (ok | xcpu).all()
    assert xcpu[stiff] and ok.sum() > b.n // 2
    assert np.array_equal(y[:, ok], y_ref[:, ok]) and np.array_equal(h[ok], h_ref[ok])
    assert np.array_equal(y[:, xcpu], b.y[:, xcpu]) and np.array_equal(h[xcpu], b.h_step[xcpu])

    print("limit removed", flush=True)
    ev.set_wall_limit(0.0)
    y, h, st = _run(ev, torch, b)
    assert (st == 0).all() and np.array_equal(y, y_ref)


def test_wall_time_limit_on_a_forest(torch_cuda, cooling_table):
    """A forest under a limit far below one attempt: in the operation graph
    (one launch, the limit covers the whole forest) every evolve returns 1025
    with its input state and still releases the operations after it, so the
    launch drains, events move the untouched galaxies, and the result equals
    the per-wave executor's (each wave's launch returns 1025 the same way)."""
    torch = torch_cuda
    from galacticus_amd import trees as tr
    from galacticus_amd.evolver import NodeEvolverBatched
    physics = gp.physics_params()
    cosmo = tr.cosmology(physics)
    t0 = float(wl.scale_factor_to_cosmic_time(1.0))
    t5 = float(wl.scale_factor_to_cosmic_time(1.0 / 6.0))
    forest = tr.Forest.build_pch(8, np.geomspace(1e11, 1e12, 8), t0, cosmo, tr.pch_params(2e10, t5), seed=4)
    wp = tr.walk_params(t0, merging_multiplier=0.2)
    sched = tr.Schedule(forest, cosmo, wp)
    y0 = forest.initial_state(cosmo, wp)
    ev = NodeEvolverBatched(forest.n_nodes, physics, cooling_table)
    ev.set_wall_limit(1e-7)
    out = {}
    for executor in ("graph", "waves"):
        y = torch.from_numpy(y0.copy()).cuda()
        st = tr.evolve_forest(ev, sched, y, executor=executor)
        torch.cuda.synchronize()
        out[executor] = (y.cpu().numpy(), st.cpu().numpy())
    evolved = np.unique(sched.ops["a"][(sched.ops["kind"] == tr.OP_EVOLVE) & (sched.ops["t1"] > sched.ops["t0"])])
    print("evolved nodes", len(evolved), "operations", sched.n_ops, flush=True)
    assert (out["graph"][1][evolved] == gp.STATUS_XCPU).all()
    assert np.array_equal(out["graph"][0], out["waves"][0])
    assert np.array_equal(out["graph"][1], out["waves"][1])
    assert np.isfinite(out["graph"][0]).all()
    # without the limit the forest evolves normally
    ev.set_wall_limit(0.0)
    y = torch.from_numpy(y0.copy()).cuda()
    st = tr.evolve_forest(ev, sched, y).cpu().numpy()
    assert (st[evolved] == 0).all()
