"""GPU parity: the HIP engine against the CPU oracle and the reference's known answers.

All calls go through the C ABI (include/gal_ode.h) via galacticus_amd.evolver.
Bars:
  * controller (cscal2.c:93-169): decision identical; h within 4 ulp (the
    device pow and glibc's pow differ in the last bits of rmax^(-1/ord)).
  * RHS: |f_gpu - f_oracle| <= 1e-9 * (|f_oracle| + rate scale) elementwise.
  * evolution: every node within the ODE tolerance of the oracle
    (|dy| <= 10 * (eps_rel*|y| + eps_abs*scale)), median relative
    difference of the masses below 1e-9, identical statuses.
  * known answers: the reference's closed/leaky box numbers at its 1%
    tolerance (testSuite/test-reproducibility.py:47-68) and the ODE solver
    closed forms at absTol 1e-6 / relTol 5e-6 (source/tests/ODE_solver.F90:73-104).
"""
import math

import numpy as np
import pytest

import oracle
from galacticus_amd import params as gp
from galacticus_amd import workloads as wl

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def _evolver(n, physics=None, **kw):
    from galacticus_amd.evolver import NodeEvolverBatched
    return NodeEvolverBatched(n, physics, **kw)


def test_library_loads_and_device(torch_cuda):
    from galacticus_amd import _lib
    import ctypes
    L = _lib.load()
    nd, ncu = ctypes.c_int(), ctypes.c_int()
    assert L.gal_ode_device_check(ctypes.byref(nd), ctypes.byref(ncu)) == 0
    print("devices", nd.value, "CUs", ncu.value, flush=True)
    assert ncu.value >= 64


def test_controller_matches_oracle(torch_cuda):
    torch = torch_cuda
    from galacticus_amd.evolver import hadjust_device
    rng = np.random.default_rng(1)
    dim, n = 30, 4096
    y = rng.normal(0, 1, (dim, n)) * 10.0 ** rng.uniform(-3, 3, (dim, n))
    scale = 10.0 ** rng.uniform(-2, 2, (dim, n))
    # error magnitudes spanning every branch: rmax >> 1.1, in (0.5,1.1), < 0.5, tiny
    lev = 10.0 ** rng.uniform(-8, 1, n)
    yerr = rng.normal(0, 1, (dim, n)) * (1e-3 * np.abs(y) + 1e-6 * scale) * lev[None, :] * 3.0
    yp = rng.normal(0, 1, (dim, n))
    nonneg = (rng.random((dim, n)) < 0.3).astype(np.int32)
    h = 10.0 ** rng.uniform(-4, 0, n)
    # a few zero-scale / zero-error entries (NaN ratio ignored by the max)
    scale[0, :16] = 0.0
    yerr[0, :16] = 0.0
    y[0, :16] = 0.0
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    hd = dev(h.copy())
    res = hadjust_device(dev(y), dev(yerr), dev(yp), dev(scale), dev(nonneg), hd, 4, 1e-6, 1e-3, 1.0, 0.0)
    torch.cuda.synchronize()
    res = res.cpu().numpy()
    hg = hd.cpu().numpy()
    dec = np.zeros(n, dtype=int)
    ho = np.zeros(n)
    for i in range(n):
        dec[i], ho[i] = oracle.hadjust(4, 1e-6, 1e-3, 1.0, 0.0, y[:, i], yerr[:, i], yp[:, i], scale[:, i],
                                       nonneg[:, i], h[i])
    assert np.array_equal(res, dec)
    assert set(np.unique(dec)) == {-1, 0, 1}
    ulp = np.abs(hg - ho) / np.spacing(np.abs(ho))
    assert ulp.max() <= 4.0, ulp.max()


def test_rhs_matches_oracle(torch_cuda, cooling_table):
    torch = torch_cuda
    b = wl.isolated_halos(512, seed=3)
    p = gp.physics_params()
    ev = _evolver(512, p, table=cooling_table)
    rng = np.random.default_rng(4)
    t = b.t_start + rng.uniform(0, 1, b.n) * (b.t_end - b.t_start)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    fg = ev.rhs_device(dev(t), dev(b.y), dev(b.node_par)).cpu().numpy()
    fo = oracle.rhs(t, b.y, b.node_par, p, cooling_table, root_tol_rel=1e-15)
    sc = oracle.scales(b.t_start, b.y, b.node_par, p)
    tol = 1e-9 * (np.abs(fo) + sc)
    bad = np.abs(fg - fo) > tol
    assert not bad.any(), [(v, int(np.sum(bad[v]))) for v in range(30) if bad[v].any()]


@pytest.mark.parametrize("case", ["closedBox", "leakyBox"])
def test_box_known_answers(torch_cuda, cooling_table, case):
    """testSuite/test-reproducibility.py:47-68 (closed box tau=0.25; leaky box tau=0.5, beta=1)."""
    if case == "closedBox":
        p, b = wl.closed_box(0.25, 0.0)
        expect = dict(DISK_MASS_GAS=9.0717953e9, DISK_MASS_STELLAR=9.0928205e10, DISK_ABUNDANCES_GAS=9.0717953e8,
                      DISK_ABUNDANCES_STELLAR=2.8814957e9)
        tols = dict(DISK_MASS_GAS=1e-2)
    else:
        p, b = wl.closed_box(0.5, 1.0)
        expect = dict(DISK_MASS_GAS=4.0762204e9, DISK_MASS_STELLAR=3.5971417e10, DISK_ABUNDANCES_GAS=2.03811e8,
                      DISK_ABUNDANCES_STELLAR=4.85624e8)
        tols = dict(DISK_MASS_GAS=1.1e-2)
    ev = _evolver(1, p, table=cooling_table, odeToleranceAbsolute=1e-6, odeToleranceRelative=1e-3)
    y, h, sc, st, cnt = ev.evolve(b.y, b.node_par, b.t_start, b.t_end, b.h_step)
    assert st[0] == 0
    for k, v in expect.items():
        got = y[wl.V[k], 0]
        assert abs(got / v - 1.0) <= tols.get(k, 1e-2), (k, got, v)
    # and the same answer as the oracle's restatement, step for step
    yo, ho, _, sto, cnto = oracle.evolve(b.y, b.node_par, b.t_start, b.t_end, b.h_step, p, cooling_table, 1e-6,
                                         1e-3, gp.FLAG_REUSE_STEP_SIZE)
    assert np.array_equal(cnt[0], cnto[0]) and np.array_equal(cnt[1], cnto[1])
    np.testing.assert_allclose(y, yo, rtol=1e-12, atol=1e-3)


def test_ode_solver_closed_forms(torch_cuda, cooling_table):
    """source/tests/ODE_solver.F90:73-104: y'=sin(x) reversed from x to 0, and y''=-y from 0 to x,
    RKCK, tolerances 1e-9, scale 1."""
    p = gp.physics_params(physicsMode=gp.PHYSICS_TEST)
    xs = np.arange(1.0, 11.0)
    n = 2 * len(xs)
    y = np.zeros((30, n))
    t0 = np.zeros(n)
    t1 = np.zeros(n)
    for i, x in enumerate(xs):
        y[0, i] = 1.0 - math.cos(x)      # reversed sine
        t0[i], t1[i] = x, 0.0
        j = len(xs) + i
        y[1, j], y[2, j] = 1.0, 0.0      # oscillator
        t0[j], t1[j] = 0.0, x
    npar = np.zeros((gp.NPARAM_NODE, n))
    sc = np.ones((30, n))
    ev = _evolver(n, p, table=cooling_table, odeToleranceAbsolute=1e-9, odeToleranceRelative=1e-9,
                  reuseODEStepSize=False, scales_provided=True)
    yg, h, _, st, cnt = ev.evolve(y, npar, t0, t1, -np.ones(n), scale=sc)
    assert (st == 0).all()
    for i, x in enumerate(xs):
        assert abs(yg[0, i] - 0.0) <= 1e-6
        j = len(xs) + i
        assert abs(yg[1, j] - math.cos(x)) <= 1e-6 * abs(math.cos(x)) + 1e-9
        assert abs(yg[2, j] + math.sin(x)) <= 1e-6 * abs(math.sin(x)) + 1e-9
        # identical step counts to the oracle's restatement
        st1, yo, ho, co = oracle.test_ode_solve(1, x, 0.0, [1.0 - math.cos(x)], 1e-9, 1e-9)
        assert cnt[0, i] == co[0] and cnt[1, i] == co[1]
        st2, yo2, ho2, co2 = oracle.test_ode_solve(2, 0.0, x, [1.0, 0.0], 1e-9, 1e-9)
        assert cnt[0, j] == co2[0] and cnt[1, j] == co2[1]
        assert abs(yg[0, i] - yo[0]) <= 1e-12
        assert np.allclose(yg[1:3, j], yo2, rtol=1e-11, atol=1e-14)


def _compare_evolution(yg, yo, sc, eps_abs, eps_rel):
    tol = 10.0 * (eps_rel * np.abs(yo) + eps_abs * sc)
    bad = np.abs(yg - yo) > tol
    masses = wl.MASS_VARS
    rel = np.abs(yg[masses] - yo[masses]) / np.maximum(np.abs(yo[masses]), 1.0)
    return bad, float(np.median(rel)), float(rel.max())


@pytest.mark.parametrize("reuse", [True, False])
def test_evolve_matches_oracle(torch_cuda, cooling_table, reuse):
    n = 3000
    b = wl.isolated_halos(n, seed=11)
    p = gp.physics_params()
    flags = gp.FLAG_REUSE_STEP_SIZE if reuse else 0
    print(f"evolve parity n={n} reuse={reuse}", flush=True)
    ev = _evolver(n, p, table=cooling_table, odeToleranceAbsolute=1e-6, odeToleranceRelative=1e-3,
                  reuseODEStepSize=reuse)
    hin = b.h_step.copy()
    if reuse:
        hin = (b.t_end - b.t_start) * 0.05
    yg, hg, scg, stg, cntg = ev.evolve(b.y, b.node_par, b.t_start, b.t_end, hin)
    yo, ho, sco, sto, cnto = oracle.evolve(b.y, b.node_par, b.t_start, b.t_end, hin, p, cooling_table, 1e-6, 1e-3,
                                           flags, root_tol_rel=1e-15, nthreads=8)
    assert np.array_equal(stg, sto)
    assert (stg == 0).all()
    np.testing.assert_allclose(scg, sco, rtol=1e-13)
    bad, med, mx = _compare_evolution(yg, yo, sco, 1e-6, 1e-3)
    same_steps = np.mean(cntg[0] == cnto[0])
    print(f"median rel {med:.3e} max rel {mx:.3e} same-step-count fraction {same_steps:.4f}", flush=True)
    for v, i in zip(*np.nonzero(bad)):
        print(f"  beyond 10x tolerance: node {i} {wl.VAR_NAMES[v]} gpu {yg[v, i]:.9e} oracle {yo[v, i]:.9e} "
              f"attempts {cntg[0][i]}/{cnto[0][i]}", flush=True)
    assert not bad.any(), int(bad.sum())
    assert med < 1e-9
    assert same_steps > 0.95
    np.testing.assert_allclose(hg[cntg[0] == cnto[0]], ho[cntg[0] == cnto[0]], rtol=1e-6)


def test_launch_variants_bit_identical(torch_cuda, cooling_table):
    """One workgroup per CU (every lane vector in LDS), two (state in the HBM
    workspace) and four (stage derivatives in the HBM workspace), each with
    one wave per group of nodes, and one or two two-wave groups per CU (the
    RHS split over two SIMDs) run the same arithmetic: identical results,
    step sizes, statuses and counters, bit for bit."""
    torch = torch_cuda
    physics = gp.physics_params()
    b = wl.isolated_halos(4000, seed=21, params=physics)
    ev = _evolver(b.n, physics)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    out = []
    shapes = ((64, 1, 1), (64, 2, 1), (64, 4, 1), (64, 1, 2), (64, 2, 2))
    for shape in shapes:
        print("launch", shape, flush=True)
        ev.set_launch(*shape)
        assert ev.launch() == shape
        y, h = d(b.y), d(b.h_step)
        cnt = torch.zeros(3 * b.n, dtype=torch.int64, device="cuda")
        st = ev.evolve_device(y, d(b.node_par), d(b.t_start), d(b.t_end), h, counters=cnt)
        torch.cuda.synchronize()
        out.append((y.cpu(), h.cpu(), st.cpu(), cnt.cpu()))
    assert (out[0][2] == 0).all()
    for o in out[1:]:
        assert all(torch.equal(r, g) for r, g in zip(out[0], o))
    shape = ev.autotune(d(b.y), d(b.node_par), d(b.t_start), d(b.t_end), d(b.h_step))
    assert shape in shapes


def test_evolve_captured_in_a_hip_graph(torch_cuda, cooling_table):
    """The batch evolve (claim order by node index) records into a hipGraph on
    the caller's stream -- two kernels, no host synchronisation, no memset --
    and every replay, with eager work and synchronisations in between,
    reproduces the eager result bit for bit."""
    torch = torch_cuda
    from galacticus_amd.evolver import NodeEvolverBatched
    physics = gp.physics_params()
    b = wl.isolated_halos(3000, seed=31, params=physics)
    ev = NodeEvolverBatched(b.n, physics, cooling_table, odeToleranceAbsolute=1e-6, odeToleranceRelative=1e-3,
                            cost_order=False)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    npar, t0, t1 = d(b.node_par), d(b.t_start), d(b.t_end)
    y_ref, h_ref = d(b.y), d(b.h_step)
    st_ref = ev.evolve_device(y_ref, npar, t0, t1, h_ref)
    torch.cuda.synchronize()
    y_s, h_s = d(b.y), d(b.h_step)
    sc_s = torch.zeros_like(y_s)
    st_s = torch.zeros(b.n, dtype=torch.int32, device="cuda")
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):        # warm-up on the capture stream
        ev.evolve_device(y_s, npar, t0, t1, h_s, scale=sc_s, status=st_s, stream=side)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        ev.evolve_device(y_s, npar, t0, t1, h_s, scale=sc_s, status=st_s, stream=torch.cuda.current_stream())
    y_in, h_in = d(b.y), d(b.h_step)
    for rep in range(3):
        y_s.copy_(y_in)
        h_s.copy_(h_in)
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(y_s, y_ref) and torch.equal(h_s, h_ref) and torch.equal(st_s, st_ref), rep
        # eager work between replays
        y_e, h_e = d(b.y), d(b.h_step)
        ev.evolve_device(y_e, npar, t0, t1, h_e)
        torch.cuda.synchronize()
        assert torch.equal(y_e, y_ref)


def test_planned_evolve_equals_evolve(torch_cuda, cooling_table):
    """gal_ode_batch_plan_d then gal_ode_batch_evolve_planned_d (the claim
    order computed ahead of the launch) gives the same states, step sizes,
    statuses and counters as gal_ode_batch_evolve_d, bit for bit, also with
    several handles planned on their streams before any evolve; a plan serves
    one evolve, needs the same y, n and ld, and another evolve discards it."""
    torch = torch_cuda
    from galacticus_amd import _lib
    physics = gp.physics_params()
    b = wl.isolated_halos(3000, seed=33, params=physics)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    npar, t0, t1 = d(b.node_par), d(b.t_start), d(b.t_end)
    ref = _evolver(b.n, physics)
    y_r, h_r = d(b.y), d(b.h_step)
    c_r = torch.zeros(3 * b.n, dtype=torch.int64, device="cuda")
    st_r = ref.evolve_device(y_r, npar, t0, t1, h_r, counters=c_r)
    torch.cuda.synchronize()
    assert (st_r.cpu() == 0).all()
    evs = [_evolver(b.n, physics) for _ in range(3)]
    streams = [torch.cuda.Stream() for _ in evs]
    ys = [d(b.y) for _ in evs]
    hs = [d(b.h_step) for _ in evs]
    cs = [torch.zeros(3 * b.n, dtype=torch.int64, device="cuda") for _ in evs]
    sts = [torch.full((b.n,), -7, dtype=torch.int32, device="cuda") for _ in evs]
    torch.cuda.synchronize()
    for ev, y, s in zip(evs, ys, streams):
        print("plan", flush=True)
        ev.plan_device(y, npar, t0, t1, stream=s)
    for ev, y, h, c, st, s in zip(evs, ys, hs, cs, sts, streams):
        print("planned evolve", flush=True)
        ev.evolve_device(y, npar, t0, t1, h, status=st, counters=c, stream=s, planned=True)
    torch.cuda.synchronize()
    for y, h, c, st in zip(ys, hs, cs, sts):
        assert torch.equal(y, y_r) and torch.equal(h, h_r) and torch.equal(c, c_r) and torch.equal(st, st_r)
    ev = evs[0]
    y = d(b.y)
    # no plan left: the evolve above consumed it
    with pytest.raises(_lib.GalError):
        ev.evolve_device(y, npar, t0, t1, d(b.h_step), planned=True)
    # a plan for other inputs does not serve this y
    ev.plan_device(d(b.y), npar, t0, t1)
    with pytest.raises(_lib.GalError):
        ev.evolve_device(y, npar, t0, t1, d(b.h_step), planned=True)
    # an evolve of its own discards the plan
    ev.plan_device(y, npar, t0, t1)
    ev.evolve_device(d(b.y), npar, t0, t1, d(b.h_step))
    with pytest.raises(_lib.GalError):
        ev.evolve_device(y, npar, t0, t1, d(b.h_step), planned=True)
    torch.cuda.synchronize()
