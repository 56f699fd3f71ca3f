"""Adaptive meshes with hanging nodes on the GPU against the oracle: the
condensed Jacobian and residual (deal.II distribute_local_to_global
semantics, constraints derived independently on the test side), Newton to
the oracle's fixed point, FP32 mixed precision, and one adaptive cycle
(Kelly, refinement, transfer) that keeps the solution continuous."""
import sys

import numpy as np
import pytest

import oracle_lib as O
from product_util import adaptive_forest, glhip, oracle_adaptive

pytestmark = pytest.mark.gpu
G = glhip()
LO, HI = (-20.0, -20.0, -6.0), (20.0, 20.0, 6.0)
SIDES = [1, 1, 1, 1, 0, 0]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = G.Context(0)
    yield c
    c.close()


def _case(ctx, steps, seed, p=22.0, t=0.8, amp=0.05, sides=SIDES, n=(3, 3, 2), lo=LO, hi=HI, degree=1, bt=1.0e10):
    f = adaptive_forest(G, n, lo, hi, sides, steps, seed, degree=degree)
    m = f.mesh()
    say("adaptive case: steps %d seed %d cells %d nodes %d hanging %d" % (steps, seed, m.n_cells, m.n_nodes,
                                                                         m.constraints()["n_hanging"]))
    ctx.set_material_pt(p, t, bt=bt)
    ctx.upload_mesh(m)
    om = oracle_adaptive(m, lo, hi, sides)
    U = O.initial_guess(m.n_nodes, O.matep(p, t)["gap_B"], amp=amp, seed=seed)
    Ud = O.distribute(om, U)
    ctx.set_solution(U)
    return f, m, om, Ud, O.coef(p, t, bt=bt)


def _dense(row_ptr, col, val, n_nodes):
    A = np.zeros((n_nodes * 18, n_nodes * 18))
    v = val.reshape(-1, 18, 18)
    for r in range(n_nodes):
        for b in range(row_ptr[r], row_ptr[r + 1]):
            s = col[b]
            A[18 * r:18 * r + 18, 18 * s:18 * s + 18] = v[b]
    return A


@pytest.mark.parametrize("steps,seed,sides", [(1, 0, SIDES), (2, 1, SIDES), (2, 2, [1, 1, 1, 1, 4, 4])])
def test_condensed_jacobian_and_rhs_match_oracle(ctx, steps, seed, sides):
    f, m, om, Ud, co = _case(ctx, steps, seed, sides=sides)
    assert np.abs(ctx.solution() - Ud).max() <= 1e-15 * np.abs(Ud).max()  # distribute
    ctx.assemble()
    Jo, ro = O.assemble(om, co, Ud)
    rg = ctx.rhs()
    assert np.abs(rg - ro).max() <= 1e-12 * np.abs(ro).max()
    a = m.arrays()
    Ag = _dense(a["row_ptr"], a["col"], ctx.jacobian(), m.n_nodes)
    Ao = _dense(om.row_ptr, om.col, Jo, om.n_nodes)
    err = np.abs(Ag - Ao).max() / np.abs(Ao).max()
    say("J rel err", err)
    assert err <= 1e-12


def test_newton_on_adaptive_mesh_matches_oracle(ctx):
    f, m, om, Ud, co = _case(ctx, 2, 3)
    st = ctx.newton_solve(rtol=1e-10, max_newton=60, lin_tol=1e-3, restart=50)
    say("gpu newton", st)
    assert st["converged"] == 1.0
    Ug = ctx.solution()
    Uo, rc, ost = O.newton(om, co, Ud, rtol=1e-10, max_newton=60, lin_tol=1e-3, restart=50)
    say("oracle newton", ost)
    assert rc == 0
    Fo = O.free_energy(om, co, Uo)
    assert st["F"] == pytest.approx(Fo, rel=1e-10)
    assert np.abs(Ug - Uo).max() <= 1e-8 * np.abs(Uo).max()
    # the converged field is continuous: hanging values follow their lines
    assert np.abs(O.distribute(om, Ug) - Ug).max() <= 1e-15 * np.abs(Ug).max()


def test_fp32_mixed_precision_on_adaptive_mesh(ctx):
    f, m, om, Ud, co = _case(ctx, 2, 4)
    st64 = ctx.newton_solve(rtol=1e-10, max_newton=60, lin_tol=1e-2, restart=50)
    F64 = st64["F"]
    ctx.set_solution(Ud)
    st32 = ctx.newton_solve(rtol=1e-10, max_newton=60, lin_tol=1e-2, restart=50, fp32=True)
    say(st64, st32)
    assert st32["converged"] == 1.0
    assert st32["F"] == pytest.approx(F64, rel=1e-10)


def test_adaptive_cycle_refines_and_transfers(ctx):
    # one cycle of the reference's run(): converge, Kelly, refine 30%,
    # transfer, converge again, on a mesh fine enough (h <= 3 xi0) for the
    # transferred state to be a good guess; the oracle repeats it on the CPU
    lo, hi = (-5.0, -5.0, -6.0), (5.0, 5.0, 6.0)
    f, m, om, Ud, co = _case(ctx, 1, 5, n=(4, 4, 4), lo=lo, hi=hi)
    st = ctx.newton_solve(rtol=1e-8, max_newton=60, lin_tol=1e-2, restart=50)
    assert st["converged"] == 1.0
    U = ctx.solution()
    eta = f.kelly(m, U)
    fl = G.flag_fixed_number(eta, 0.3)
    f2 = f.refine(fl)
    m2 = f2.mesh()
    U2 = f.transfer(m, U, f2, m2)
    ctx.upload_mesh(m2)
    ctx.set_solution(U2)
    # the transferred field is already continuous on the new mesh
    assert np.abs(ctx.solution() - U2).max() <= 1e-13 * np.abs(U2).max()
    st2 = ctx.newton_solve(rtol=1e-8, max_newton=60, lin_tol=1e-2, restart=50)
    say("cycle 1", st2, "cells", m2.n_cells, "hanging", m2.constraints()["n_hanging"])
    assert st2["converged"] == 1.0
    om2 = oracle_adaptive(m2, lo, hi, SIDES)
    Uo2, rc, ost = O.newton(om2, co, O.distribute(om2, U2), rtol=1e-8, max_newton=60, lin_tol=1e-2, restart=50)
    assert rc == 0
    # same stationary state, reached from the same transferred guess
    assert st2["F"] == pytest.approx(O.free_energy(om2, co, Uo2), rel=1e-10)


def test_run_adaptive_driver_fp32_against_fp64(ctx):
    # the cycle loop of FemGL::run in the library: BnA guess (seed 1),
    # two refinement cycles, FP32 mixed precision checked against FP64 per
    # cycle.  The A/B wall is not stationary (B is the lower phase at 0.8 Tc,
    # 22 bar): the damped Newton of the reference stalls on it, pseudo-
    # transient continuation reaches the minimum.
    lo, hi = (-5.0, -5.0, -6.0), (5.0, 5.0, 6.0)
    p, t = 22.0, 0.8
    f = G.Forest((4, 4, 4), lo, hi, SIDES)
    m = f.mesh()
    mp = O.matep(p, t)
    U0 = G.initial_guess_bna(m.arrays()["node_xyz"], 0.0, mp["gap_A"], mp["gap_B"], amp=0.05, seed=1)
    ctx.set_material_pt(p, t)
    ctx.set_continuation(1.0)
    try:
        rows, f2, m2 = ctx.run_adaptive(f, U0, n_refine=2, refine_frac=0.3, rtol=1e-9, lin_tol=1e-2, restart=50,
                                        fp32=True, check_fp64=True)
    finally:
        ctx.set_continuation(0.0)
    for r in rows:
        say(r)
        assert r["converged"] == 1.0 and r["converged64"] == 1.0
        assert r["F_rel_err_vs_fp64"] <= 1e-10
    assert rows[1]["hanging"] > 0 and rows[2]["cells"] > rows[1]["cells"] > rows[0]["cells"] == 64
    # the last field is a fixed point of the oracle's condensed system
    om = oracle_adaptive(m2, lo, hi, SIDES)
    co = O.coef(p, t)
    U = ctx.solution()
    _, ro = O.assemble(om, co, U, J=False)
    assert np.linalg.norm(ro) <= 1.01e-9 * rows[2]["r0"]
    assert rows[2]["F"] == pytest.approx(O.free_energy(om, co, U), rel=1e-12)


def test_continuation_reaches_the_minimum_where_damped_newton_stalls(ctx):
    lo, hi = (-5.0, -5.0, -6.0), (5.0, 5.0, 6.0)
    p, t = 22.0, 0.8
    m = G.Forest((4, 4, 4), lo, hi, SIDES).mesh()
    mp = O.matep(p, t)
    U0 = G.initial_guess_bna(m.arrays()["node_xyz"], 0.0, mp["gap_A"], mp["gap_B"], amp=0.05, seed=1)
    ctx.set_material_pt(p, t)
    ctx.upload_mesh(m)
    ctx.set_solution(U0)
    st0 = ctx.newton_solve(rtol=1e-10, max_newton=40, lin_tol=1e-2, restart=50)
    ctx.set_solution(U0)
    ctx.set_continuation(1.0)
    try:
        st1 = ctx.newton_solve(rtol=1e-10, max_newton=40, lin_tol=1e-2, restart=50)
    finally:
        ctx.set_continuation(0.0)
    say("damped newton", st0)
    say("continuation", st1)
    assert st0["converged"] == 0.0 and st1["converged"] == 1.0
    assert st1["F"] < st0["F"]
    # the continuation fixed point is the oracle's: the oracle's residual at it
    # is the one the GPU reports, and plain Newton from it keeps the energy
    om = oracle_adaptive(m, lo, hi, SIDES)
    co = O.coef(p, t)
    U = ctx.solution()
    _, R = O.assemble(om, co, U, J=False, R=True)
    assert np.linalg.norm(R) == pytest.approx(st1["r_final"], rel=1e-3)
    Uo, rc, ost = O.newton(om, co, U, rtol=1e-6, max_newton=5, lin_tol=1e-3, restart=50)
    assert O.free_energy(om, co, Uo) == pytest.approx(st1["F"], rel=1e-10)
    # this state sits where damped Newton stalls: the Jacobian is nearly
    # singular, and the oracle's polish moves U by several hundred times the
    # final residual norm (about 5e-8 here), more than the 1e-8 of a
    # well-conditioned solve
    assert np.abs(Uo - U).max() <= 1e-7 * np.abs(U).max()


def test_run_prm_follows_the_reference_cycle_loop(ctx):
    # FemGL::run from a .prm file: cycle 0 leaves through the stuck test
    # (|R - R_last| < 1, the default threshold) and refines by Kelly; the last
    # cycle runs to the absolute converge accuracy
    prm = G.Prm("""
subsection physical parameters
  set pressure in bar = 22.0
  set t_reduced = 0.8
end
subsection control parameters
  set half x length of retangle = 5
  set half y length of retangle = 5
  set half z length of retangle = 6
  set Number of initial global refinments = 2
  set Number of refinements = 1
  set Number of interations = 60
  set converge accuracy = 1e-7
  set Cycle 0 linear solver tol = 1e-2
  set Cycle 1 linear solver tol = 1e-2
end
subsection glhip
  set boundary ids = 1 1 1 1 0 0
  set random seed = 1
  set GMRES restart = 50
end
""")
    rows = ctx.run_prm(prm)
    for r in rows:
        say(r)
    assert len(rows) == 2 and rows[0]["cells"] == 64 and rows[1]["cells"] > 64
    assert rows[0]["residual"] > 1e-7 >= rows[1]["residual"]
    # the context holds the last cycle's state: recompute its residual and energy
    assert ctx.residual_norm() == pytest.approx(rows[1]["residual"], rel=1e-6)
    assert ctx.free_energy() == pytest.approx(rows[1]["F"], rel=1e-12)


@pytest.mark.parametrize("deg,sides", [(1, [5, 5, 5, 5, 0, 0]), (2, [5, 5, 1, 1, 4, 4]), (1, [5, 5, 5, 5, 5, 5])])
def test_periodic_box_parity_with_oracle(ctx, deg, sides):
    from product_util import side_masks
    n = (3, 3, 2) if deg == 1 else (2, 2, 2)
    lo, hi = (-6.0, -6.0, -4.0), (6.0, 6.0, 4.0)
    axes = tuple(a for a in range(3) if sides[2 * a] == 5)
    m = G.Mesh(deg, n, lo, hi, sides)
    say("periodic case", deg, sides, "hanging", m.constraints()["n_hanging"])
    ctx.set_material_pt(22.0, 0.8)
    ctx.upload_mesh(m)
    om = O.periodic_oracle_mesh(deg, n, lo, hi, side_masks(sides), axes)
    U = O.initial_guess(m.n_nodes, O.matep(22.0, 0.8)["gap_B"], amp=0.05, seed=3)
    Ud = O.distribute(om, U)
    ctx.set_solution(U)
    assert np.abs(ctx.solution() - Ud).max() <= 1e-15 * np.abs(Ud).max()
    ctx.assemble()
    co = O.coef(22.0, 0.8)
    Jo, ro = O.assemble(om, co, Ud)
    assert np.abs(ctx.rhs() - ro).max() <= 1e-12 * np.abs(ro).max()
    a = m.arrays()
    Ag = _dense(a["row_ptr"], a["col"], ctx.jacobian(), m.n_nodes)
    Ao = _dense(om.row_ptr, om.col, Jo, om.n_nodes)
    assert np.abs(Ag - Ao).max() <= 1e-12 * np.abs(Ao).max()


def test_periodic_slab_converges_to_a_laterally_uniform_state(ctx):
    # x and y periodic, A = 0 on the z faces: the minimum depends on z only
    lo, hi = (-5.0, -5.0, -6.0), (5.0, 5.0, 6.0)
    m = G.Mesh(1, (4, 4, 6), lo, hi, [5, 5, 5, 5, 0, 0])
    ctx.set_material_pt(22.0, 0.8)
    ctx.upload_mesh(m)
    ctx.set_solution(O.initial_guess(m.n_nodes, O.matep(22.0, 0.8)["gap_B"], amp=0.05, seed=4))
    st = ctx.newton_solve(rtol=1e-10, max_newton=60, lin_tol=1e-3, restart=50)
    say("periodic slab", st)
    assert st["converged"] == 1.0
    U = ctx.solution().reshape(-1, 18)
    xyz = m.arrays()["node_xyz"].reshape(-1, 3)
    # the bulk order parameter norm is the same on every node of one z plane
    # (the state may be rotated, so compare invariants: |A|^2 per node)
    a2 = (U ** 2).sum(1)
    for z in np.unique(np.round(xyz[:, 2], 9)):
        sel = np.abs(xyz[:, 2] - z) < 1e-9
        assert np.ptp(a2[sel]) <= 1e-7 * max(a2.max(), 1.0)


@pytest.mark.parametrize("guess", ["BinA", "PDW"])
def test_run_prm_with_analytic_guesses(ctx, guess):
    # the setup variants' initial configurations (B-in-A ball on a cube with
    # A = 0 on every face; weak-coupling PDW on a z-walled slab), one cycle
    # each; the context's final state is a fixed point at converge accuracy
    geo = "cube" if guess == "BinA" else "retangle"
    ids = "0 0 0 0 0 0" if guess == "BinA" else "1 1 1 1 0 0"
    prm = G.Prm("""
subsection physical parameters
  set pressure in bar = 20.0
  set t_reduced = 0.5
  set trun on Strong Coupling Correction = %s
end
subsection control parameters
  set cube half side length = 6
  set half x length of retangle = 8
  set half y length of retangle = 8
  set half z length of retangle = 5
  set Number of initial global refinments = 2
  set Number of refinements = 0
  set Number of interations = 80
  set converge accuracy = 1e-8
  set Cycle 0 linear solver tol = 1e-3
end
subsection glhip
  set geometry = %s
  set boundary ids = %s
  set initial guess = %s
  set noise amplitude = 0
  set GMRES restart = 50
end
""" % ("true" if guess == "BinA" else "false", geo, ids, guess))
    rows = ctx.run_prm(prm)
    for r in rows:
        say(guess, r)
    assert len(rows) == 1 and rows[0]["residual"] <= 1e-8
    assert ctx.residual_norm() == pytest.approx(rows[0]["residual"], rel=1e-6)


def test_device_transfer_equals_host_transfer(ctx):
    # two refinements: the second starts from a mesh with hanging nodes; the
    # device path must leave the state set_solution leaves with the host field
    lo, hi = (-5.0, -5.0, -6.0), (5.0, 5.0, 6.0)
    f = G.Forest((3, 3, 2), lo, hi, SIDES)
    m = f.mesh()
    rng = np.random.default_rng(7)
    xyz = m.arrays()["node_xyz"].reshape(-1, 3)
    A = rng.standard_normal((3, 18))
    U = (np.sin(xyz @ A) + 0.1 * rng.standard_normal((m.n_nodes, 18))).reshape(-1)
    ctx.set_material_pt(22.0, 0.8)
    ctx.upload_mesh(m)
    ctx.set_solution(U)
    for step in range(2):
        Uc = ctx.solution()
        f2 = f.refine((rng.random(f.n_cells) < 0.35).astype(np.uint8))
        m2 = f2.mesh()
        say("transfer step", step, "cells", f.n_cells, "->", f2.n_cells, "nodes", m2.n_nodes)
        Uh = f.transfer(m, Uc, f2, m2)
        ctx.transfer_solution(f, m, f2, m2)
        Ud = ctx.solution()
        ctx.set_solution(Uh)
        assert np.array_equal(Ud, ctx.solution())
        if step == 0:
            assert m2.constraints()["n_hanging"] > 0
        f, m = f2, m2


def _blocks(row_ptr, col, val, n_nodes):
    v = val.reshape(-1, 18, 18)
    return {(r, int(col[b])): v[b] for r in range(n_nodes) for b in range(row_ptr[r], row_ptr[r + 1])}


@pytest.mark.parametrize("steps,seed,n", [(1, 0, (2, 2, 2)), (2, 1, (2, 2, 1))])
def test_q2_condensed_jacobian_and_rhs_match_oracle(ctx, steps, seed, n):
    f, m, om, Ud, co = _case(ctx, steps, seed, n=n, degree=2)
    assert m.deg == 2 and m.constraints()["n_hanging"] > 0
    assert np.abs(ctx.solution() - Ud).max() <= 1e-15 * np.abs(Ud).max()  # distribute
    ctx.assemble()
    Jo, ro = O.assemble(om, co, Ud)
    rg = ctx.rhs()
    assert np.abs(rg - ro).max() <= 1e-12 * np.abs(ro).max()
    a = m.arrays()
    bg = _blocks(a["row_ptr"], a["col"], ctx.jacobian(), m.n_nodes)
    bo = _blocks(om.row_ptr, om.col, Jo, om.n_nodes)
    scale = max(np.abs(b).max() for b in bo.values())
    err = 0.0
    for k in set(bg) | set(bo):
        z = np.zeros((18, 18))
        err = max(err, np.abs(bg.get(k, z) - bo.get(k, z)).max() / scale)
    say("Q2 J rel err", err)
    assert err <= 1e-12


def test_q2_newton_on_adaptive_mesh_matches_oracle(ctx):
    f, m, om, Ud, co = _case(ctx, 1, 3, n=(2, 2, 2), degree=2)
    st = ctx.newton_solve(rtol=1e-10, max_newton=60, lin_tol=1e-3, restart=50)
    say("gpu newton Q2", st)
    assert st["converged"] == 1.0
    Ug = ctx.solution()
    Uo, rc, ost = O.newton(om, co, Ud, rtol=1e-10, max_newton=60, lin_tol=1e-3, restart=50)
    say("oracle newton Q2", ost)
    assert rc == 0
    assert st["F"] == pytest.approx(O.free_energy(om, co, Uo), rel=1e-10)
    assert np.abs(Ug - Uo).max() <= 1e-8 * np.abs(Uo).max()
    assert np.abs(O.distribute(om, Ug) - Ug).max() <= 1e-15 * np.abs(Ug).max()


def test_q2_device_transfer_equals_host_transfer(ctx):
    lo, hi = (-5.0, -5.0, -6.0), (5.0, 5.0, 6.0)
    f = G.Forest((2, 2, 2), lo, hi, SIDES, degree=2)
    m = f.mesh()
    rng = np.random.default_rng(11)
    xyz = m.arrays()["node_xyz"].reshape(-1, 3)
    U = (np.sin(xyz @ rng.standard_normal((3, 18))) + 0.1 * rng.standard_normal((m.n_nodes, 18))).reshape(-1)
    ctx.set_material_pt(22.0, 0.8)
    ctx.upload_mesh(m)
    ctx.set_solution(U)
    for step in range(2):
        Uc = ctx.solution()
        f2 = f.refine((rng.random(f.n_cells) < 0.35).astype(np.uint8))
        m2 = f2.mesh()
        say("Q2 transfer step", step, "cells", f.n_cells, "->", f2.n_cells, "nodes", m2.n_nodes)
        Uh = f.transfer(m, Uc, f2, m2)
        ctx.transfer_solution(f, m, f2, m2)
        Ud = ctx.solution()
        ctx.set_solution(Uh)
        assert np.array_equal(Ud, ctx.solution())
        f, m = f2, m2


def test_q2_adaptive_cycles_fp32_against_fp64(ctx):
    lo, hi = (-5.0, -5.0, -6.0), (5.0, 5.0, 6.0)
    p, t = 22.0, 0.8
    f = G.Forest((2, 2, 2), lo, hi, SIDES, degree=2)
    m = f.mesh()
    ctx.set_material_pt(p, t)
    U0 = O.initial_guess(m.n_nodes, O.matep(p, t)["gap_B"], amp=0.05, seed=2)
    rows, f2, m2 = ctx.run_adaptive(f, U0, n_refine=2, refine_frac=0.3, rtol=1e-9, lin_tol=1e-2, restart=50,
                                    fp32=True, check_fp64=True)
    for r in rows:
        say(r)
        assert r["converged"] == 1.0 and r["converged64"] == 1.0
        assert r["F_rel_err_vs_fp64"] <= 1e-10
    assert m2.deg == 2 and any(r["hanging"] > 0 for r in rows)
    assert rows[2]["cells"] > rows[1]["cells"] > rows[0]["cells"]
    om = oracle_adaptive(m2, lo, hi, SIDES)
    co = O.coef(p, t)
    U = ctx.solution()
    _, ro = O.assemble(om, co, U, J=False)
    assert np.linalg.norm(ro) <= 1.01e-9 * rows[2]["r0"]


PERIODIC = [5, 5, 5, 5, 0, 0]


@pytest.mark.parametrize("deg,steps,seed", [(1, 2, 1), (2, 1, 2)])
def test_periodic_forest_condensed_parity_with_oracle(ctx, deg, steps, seed):
    lo, hi = (-6.0, -6.0, -4.0), (6.0, 6.0, 4.0)
    f, m, om, Ud, co = _case(ctx, steps, seed, sides=PERIODIC, n=(3, 3, 2) if deg == 1 else (2, 2, 2), lo=lo, hi=hi,
                             degree=deg)
    assert m.constraints()["n_hanging"] > 0
    assert np.abs(ctx.solution() - Ud).max() <= 1e-15 * np.abs(Ud).max()
    ctx.assemble()
    Jo, ro = O.assemble(om, co, Ud)
    assert np.abs(ctx.rhs() - ro).max() <= 1e-12 * np.abs(ro).max()
    a = m.arrays()
    bg = _blocks(a["row_ptr"], a["col"], ctx.jacobian(), m.n_nodes)
    bo = _blocks(om.row_ptr, om.col, Jo, om.n_nodes)
    scale = max(np.abs(b).max() for b in bo.values())
    z = np.zeros((18, 18))
    err = max(np.abs(bg.get(k, z) - bo.get(k, z)).max() / scale for k in set(bg) | set(bo))
    say("periodic forest J rel err", err)
    assert err <= 1e-12


def test_periodic_forest_newton_matches_oracle(ctx):
    lo, hi = (-6.0, -6.0, -4.0), (6.0, 6.0, 4.0)
    f, m, om, Ud, co = _case(ctx, 2, 5, sides=PERIODIC, lo=lo, hi=hi)
    st = ctx.newton_solve(rtol=1e-10, max_newton=60, lin_tol=1e-3, restart=50)
    say("gpu newton periodic forest", st)
    assert st["converged"] == 1.0
    Uo, rc, ost = O.newton(om, co, Ud, rtol=1e-10, max_newton=60, lin_tol=1e-3, restart=50)
    assert rc == 0
    assert st["F"] == pytest.approx(O.free_energy(om, co, Uo), rel=1e-10)
    assert np.abs(ctx.solution() - Uo).max() <= 1e-8 * np.abs(Uo).max()


def test_run_prm_periodic_adaptive_like_the_reference_default_build(ctx):
    # makegrid_retangle-z-AdGR_xy-periodic.cc: x and y periodic, AdGR walls
    # (id 4: the z-normal column vanishes) on the z faces; two adaptive cycles
    prm = G.Prm("""
subsection physical parameters
  set pressure in bar = 22.0
  set t_reduced = 0.8
end
subsection control parameters
  set half x length of retangle = 6
  set half y length of retangle = 6
  set half z length of retangle = 5
  set Number of initial global refinments = 2
  set Number of refinements = 2
  set Number of interations = 60
  set converge accuracy = 1e-8
  set Cycle 0 refinement threshold = 1e-3
  set Cycle 1 refinement threshold = 1e-3
  set Cycle 0 linear solver tol = 1e-2
  set Cycle 1 linear solver tol = 1e-2
  set Cycle 2 linear solver tol = 1e-2
end
subsection glhip
  set boundary ids = 5 5 5 5 4 4
  set random seed = 3
  set GMRES restart = 50
end
""")
    rows = ctx.run_prm(prm)
    for r in rows:
        say(r)
    assert len(rows) >= 2 and rows[-1]["cells"] > rows[0]["cells"]
    assert rows[-1]["residual"] <= 1e-8
    assert ctx.residual_norm() == pytest.approx(rows[-1]["residual"], rel=1e-6)


def test_solver_executable_runs_a_parameter_file_and_writes_output(tmp_path):
    # the reference's driver: read configuration.prm, run, write
    # refine-cycle_<cycle>/solution_<iteration>.pvtu per Newton iteration
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "verkko-hem-repo_amd",
                       "glhip_run")
    cfg = tmp_path / "configuration.prm"
    cfg.write_text("""
subsection physical parameters
  set pressure in bar = 22.0
  set t_reduced = 0.8
end
subsection control parameters
  set half x length of retangle = 5
  set half y length of retangle = 5
  set half z length of retangle = 6
  set Number of initial global refinments = 2
  set Number of refinements = 1
  set Number of interations = 60
  set converge accuracy = 1e-7
  set Cycle 0 linear solver tol = 1e-2
  set Cycle 1 linear solver tol = 1e-2
end
subsection glhip
  set boundary ids = 1 1 1 1 0 0
  set random seed = 1
  set GMRES restart = 50
  set output directory = %s
  set verbose = true
end
""" % (tmp_path / "out"))
    say("running", exe)
    r = subprocess.run([exe, str(cfg)], capture_output=True, text=True, timeout=240)
    say(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    lines = [l for l in r.stdout.splitlines() if l.startswith("cycle ")]
    assert len(lines) == 2 and "Refinement cycle : 1, iteration_loop: 0" in r.stdout
    for cyc in (0, 1):
        d = tmp_path / "out" / ("refine-cycle_%d" % cyc)
        assert (d / "solution_00.pvtu").exists() and (d / "solution_00.0.vtu").exists()


@pytest.mark.parametrize("deg", [1, 2])
def test_robin_walls_on_adaptive_mesh_match_oracle(ctx, deg):
    # AdGR diffuse walls (bt < 1e10, assemble.cc:286-348) on the z faces of
    # a refined forest: the face terms of the coarse and the fine boundary cells
    sides = [1, 1, 1, 1, 4, 4]
    f, m, om, Ud, co = _case(ctx, 1, 7, sides=sides, n=(2, 2, 2), degree=deg, bt=3.0)
    ctx.assemble()
    Jo, ro = O.assemble(om, co, Ud)
    assert np.abs(ctx.rhs() - ro).max() <= 1e-12 * np.abs(ro).max()
    a = m.arrays()
    bg = _blocks(a["row_ptr"], a["col"], ctx.jacobian(), m.n_nodes)
    bo = _blocks(om.row_ptr, om.col, Jo, om.n_nodes)
    scale = max(np.abs(b).max() for b in bo.values())
    z = np.zeros((18, 18))
    err = max(np.abs(bg.get(k, z) - bo.get(k, z)).max() / scale for k in set(bg) | set(bo))
    say("Robin adaptive J rel err", err)
    assert err <= 1e-12
    # the wall term is present: without it the matrices differ
    Jn, rn = O.assemble(om, O.coef(22.0, 0.8), Ud)
    assert np.abs(rn - ro).max() > 1e-6 * np.abs(ro).max()


def test_run_prm_coarsens_with_the_coarsen_ratio(ctx):
    # refine_and_coarsen_fixed_number with a nonzero "adaptive coarsen ratio"
    # (refine.cc:140-157): the cells of lowest Kelly indicator are coarsened
    # where whole families are flagged and balance allows (on 8^3 cells every
    # such family touches a refined one), so the refined 16^3 mesh has fewer
    # cells than without coarsening, and the last cycle still converges
    def prm(cratio):
        return G.Prm(f"""
subsection physical parameters
  set pressure in bar = 22.0
  set t_reduced = 0.8
end
subsection control parameters
  set half x length of retangle = 5
  set half y length of retangle = 5
  set half z length of retangle = 6
  set Number of initial global refinments = 4
  set Number of refinements = 1
  set Number of interations = 60
  set converge accuracy = 1e-7
  set adaptive refinment ratio = 0.2
  set adaptive coarsen ratio = {cratio}
  set Cycle 0 linear solver tol = 1e-2
  set Cycle 1 linear solver tol = 1e-2
end
subsection glhip
  set boundary ids = 1 1 1 1 0 0
  set random seed = 1
  set GMRES restart = 50
end
""")
    plain = ctx.run_prm(prm(0.0))
    rows = ctx.run_prm(prm(0.5))
    for r in plain + rows:
        say(r)
    assert plain[0]["cells"] == rows[0]["cells"] == 4096
    assert rows[1]["cells"] < plain[1]["cells"]
    assert rows[1]["residual"] <= 1e-7
    assert ctx.residual_norm() == pytest.approx(rows[1]["residual"], rel=1e-6)


CALLER_CASES = [
    # degree, refinement steps, seed, boundary ids, Robin length bt
    (1, 2, 1, PERIODIC, 1.0e10),
    (2, 1, 2, PERIODIC, 1.0e10),
    (1, 2, 3, [2, 2, 3, 3, 4, 4], 2.0),
]


@pytest.mark.parametrize("deg,steps,seed,sides,bt", CALLER_CASES)
def test_caller_built_mesh_matches_oracle(ctx, deg, steps, seed, sides, bt):
    # The per-cycle upload from the arrays a deal.II caller holds
    # (glh_mesh_from_arrays): its own node numbering (randomly permuted),
    # FE_Q's hierarchic local order, periodic images as constraint lines and
    # the hanging-node lines; the library builds colours and patterns.  The
    # oracle of that constraint set (hanging lines derived from the geometry,
    # images from the coordinates, closed as deal.II closes them) gives R and
    # the condensed J to 1e-12, constrained rows included.
    from product_util import caller_arrays, oracle_caller
    lo, hi = (-6.0, -6.0, -4.0), (6.0, 6.0, 4.0)
    f = adaptive_forest(G, (3, 3, 2) if deg == 1 else (2, 2, 2), lo, hi, sides, steps, seed, degree=deg)
    m = f.mesh()
    kw, perm, images = caller_arrays(m, lo, hi, seed=11 + seed, hierarchic=True)
    m2 = G.Mesh.from_arrays(**kw)
    say("caller mesh: deg %d cells %d nodes %d (library %d) lines %d" % (deg, m2.n_cells, m2.n_nodes, m.n_nodes,
                                                                         len(kw["lines"])))
    p, t = 22.0, 0.8
    om2 = oracle_caller(oracle_adaptive(m, lo, hi, sides), perm, images, m2)
    co = O.coef(p, t, bt=bt)
    U = O.initial_guess(m2.n_nodes, O.matep(p, t)["gap_B"], amp=0.05, seed=seed)
    Ud = O.distribute(om2, U)
    ctx.set_material_pt(p, t, bt=bt)
    ctx.upload_mesh(m2)
    ctx.set_solution(U)
    assert np.abs(ctx.solution() - Ud).max() <= 1e-15 * np.abs(Ud).max()  # distribute
    ctx.assemble()
    Jo, ro = O.assemble(om2, co, Ud)
    er = np.abs(ctx.rhs() - ro).max() / np.abs(ro).max()
    ej = np.abs(ctx.jacobian() - Jo).max() / np.abs(Jo).max()
    say("caller mesh rhs rel err", er, "J rel err", ej, "images", len(images))
    assert er <= 1e-12 and ej <= 1e-12
