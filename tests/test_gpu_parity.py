"""GPU parity of the HIP path against the FP64 CPU oracle, through the C ABI.

Tolerances: assembled Jacobian and residual 1e-12
relative (normwise, max|diff| / max|ref|); converged free energy 1e-10
relative; converged order parameter 1e-8 relative.  FP32 Jacobian: 2e-6.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
from product_util import cfg_mesh, glhip, initial_state, oracle_mesh

pytestmark = pytest.mark.gpu
G = glhip()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = G.Context(0)
    yield c
    c.close()


def _setup(ctx, mesh, p, t, amp, seed, bt=1.0e10, U=None):
    ctx.set_material_pt(p, t, bt=bt)
    ctx.upload_mesh(mesh)
    if U is None:
        U = initial_state(mesh.n_nodes, p, t, amp, seed)
    ctx.set_solution(U)
    om = oracle_mesh(mesh)
    Ud = O.distribute(om, U)
    return om, Ud, O.coef(p, t, bt=bt)


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("case", [
    ("cube_q1_ref3", 20.0, 0.5, 1.0e10),   # config 1: Q1 8^3, T/Tc 0.5, p 20 bar
    ("cube_q2_small", 26.0, 0.7, 1.0e10),  # config 2 physics on 4^3 Q2 cells
    ("cube_q1_small", 22.0, 0.8, 1.0e10),
])
def test_jacobian_and_rhs_parity_fp64(ctx, case):
    name, p, t, bt = case
    say("case", name, p, t, bt)
    mesh = cfg_mesh(G, name)
    om, U, co = _setup(ctx, mesh, p, t, 0.05, 0, bt)
    ctx.assemble(J=True, R=True)
    Jg, rg = ctx.jacobian(), ctx.rhs()
    Jo, ro = O.assemble(om, co, U)
    assert _rel(Jg, Jo) <= 1e-12, _rel(Jg, Jo)
    assert _rel(rg, ro) <= 1e-12, _rel(rg, ro)
    # constrained rows carry zero rhs exactly and zero off-diagonal entries
    cons = om.con_of >= 0
    assert np.all(rg[cons] == 0)


@pytest.mark.parametrize("deg", [1, 2])
def test_robin_wall_parity(ctx, deg):
    # AdGR diffuse walls (boundary ids 2 and 4) with bt = 2.0: Robin face term on
    mesh = G.Mesh(deg, (3, 2, 2), (-3, -2, -2), (3, 2, 2), [2, 2, 1, 1, 4, 4])
    say("robin deg", deg)
    om, U, co = _setup(ctx, mesh, 20.0, 0.5, 0.3, 5, bt=2.0)
    ctx.assemble()
    Jo, ro = O.assemble(om, co, U)
    assert _rel(ctx.jacobian(), Jo) <= 1e-12
    assert _rel(ctx.rhs(), ro) <= 1e-12
    F = ctx.free_energy()
    assert F == pytest.approx(O.free_energy(om, co, U), rel=1e-12)


def test_fp32_jacobian_close_to_fp64(ctx):
    mesh = cfg_mesh(G, "cube_q1_ref3")
    om, U, co = _setup(ctx, mesh, 20.0, 0.5, 0.05, 0)
    ctx.assemble(J=True, R=True, fp32=True)
    Jo, ro = O.assemble(om, co, U)
    assert _rel(ctx.jacobian(), Jo) <= 2e-6
    assert _rel(ctx.rhs(), ro) <= 1e-12  # residual stays FP64


def test_assembly_is_deterministic(ctx):
    mesh = cfg_mesh(G, "cube_q2_small")
    _setup(ctx, mesh, 26.0, 0.7, 0.05, 1)
    ctx.assemble()
    J1, r1 = ctx.jacobian(), ctx.rhs()
    ctx.assemble()
    J2, r2 = ctx.jacobian(), ctx.rhs()
    assert np.array_equal(J1, J2) and np.array_equal(r1, r2)


def test_free_energy_and_residual_norm(ctx):
    mesh = cfg_mesh(G, "cube_q1_ref3")
    om, U, co = _setup(ctx, mesh, 20.0, 0.5, 0.05, 0)
    assert ctx.free_energy() == pytest.approx(O.free_energy(om, co, U), rel=1e-12)
    _, ro = O.assemble(om, co, U, J=False)
    assert ctx.residual_norm() == pytest.approx(np.linalg.norm(ro), rel=1e-12)


@pytest.mark.parametrize("fp32", [False, True])
def test_gmres_solves_the_assembled_system(ctx, fp32):
    mesh = cfg_mesh(G, "cube_q1_ref3")
    om, U, co = _setup(ctx, mesh, 20.0, 0.5, 0.05, 0)
    ctx.assemble(fp32=fp32)
    its, rel = ctx.solve(rtol=1e-8, maxit=5000, restart=30)
    say("gmres its", its, rel, "fp32", fp32)
    dU = ctx.update()
    Jo, ro = O.assemble(om, co, U)
    y = np.zeros_like(dU)
    O.lib().orc_bsr_spmv(om.n_nodes, om.row_ptr, om.col, O.f64(Jo.reshape(-1)), O.f64(dU), y)
    assert np.linalg.norm(y - ro) <= 1e-8 * np.linalg.norm(ro) * (1.05 if not fp32 else 50)
    assert its > 0 and rel <= 1e-8


def test_gmres_no_convergence_is_reported(ctx):
    mesh = cfg_mesh(G, "cube_q1_ref3")
    _setup(ctx, mesh, 20.0, 0.5, 0.05, 0)
    ctx.assemble()
    with pytest.raises(G.GLHError, match="no convergence"):
        ctx.solve(rtol=1e-14, maxit=2, restart=30)


@pytest.mark.parametrize("bad", [dict(lin_tol=0.0), dict(lin_maxit=0), dict(max_newton=-1), dict(rtol=-1.0),
                                 dict(ls_step=1.0), dict(ls_step=0.0)])
def test_newton_solve_rejects_bad_controls_before_any_work(ctx, bad):
    mesh = cfg_mesh(G, "cube_q1_small")
    _setup(ctx, mesh, 22.0, 0.8, 0.05, 2)
    U = ctx.solution()
    with pytest.raises(G.GLHError, match="bad Newton controls"):
        ctx.newton_solve(**bad)
    assert np.array_equal(ctx.solution(), U)


def test_newton_step_line_search_matches_oracle_semantics(ctx):
    mesh = cfg_mesh(G, "cube_q1_small")
    om, U, co = _setup(ctx, mesh, 22.0, 0.8, 0.05, 2)
    ctx.assemble()
    ctx.solve(rtol=1e-10, maxit=5000)
    dU = ctx.update()
    out = ctx.newton_step(0.83, True)
    prev, cur, a, trials = out
    Jo, ro = O.assemble(om, co, U)
    assert prev == pytest.approx(np.linalg.norm(ro), rel=1e-12)
    X = O.distribute(om, U + a * dU)
    _, rx = O.assemble(om, co, X, J=False)
    assert cur == pytest.approx(np.linalg.norm(rx), rel=1e-10)
    assert cur < prev
    assert np.allclose(ctx.solution(), X, rtol=0, atol=1e-14 * np.abs(X).max())


def test_newton_converges_like_the_oracle_config1(ctx):
    # config 1: Q1 [-10,10]^3 ref 3, T/Tc 0.5, p 20 bar, A = 0 on z faces,
    # B phase + U[-0.05,0.05] Delta_B noise (seed 0), ||R||/||R0|| <= 1e-10
    mesh = cfg_mesh(G, "cube_q1_ref3")
    om, U0, co = _setup(ctx, mesh, 20.0, 0.5, 0.05, 0)
    st = ctx.newton_solve(rtol=1e-10, max_newton=60, lin_tol=1e-3, restart=50)
    say("gpu newton", st)
    assert st["r_final"] <= 1e-10 * st["r0"]
    Ug = ctx.solution()
    Uo, rc, ost = O.newton(om, co, U0, rtol=1e-10, max_newton=60, lin_tol=1e-3, restart=50)
    say("oracle newton", ost)
    assert rc == 0 and ost[3] <= 1e-10 * ost[2]
    Fo = O.free_energy(om, co, Uo)
    assert st["F"] == pytest.approx(Fo, rel=1e-10)
    assert np.abs(Ug - Uo).max() <= 1e-8 * np.abs(Uo).max()
    # converged state has the oracle's residual
    _, ro = O.assemble(om, co, Ug, J=False)
    assert np.linalg.norm(ro) <= 1.0001e-10 * st["r0"] * 1.01


GOLD_CASES = [
    ("q1_cube", 1, (3, 3, 3), (-3.0, -3.0, -3.0), (3.0, 3.0, 3.0), [1, 1, 1, 1, 0, 0], 20.0, 0.5, 1.0e10),
    ("q2_slice", 2, (2, 2, 2), (-20.0, -20.0, -20.0), (20.0, 20.0, 20.0), [1, 1, 1, 1, 0, 0], 26.0, 0.7, 1.0e10),
    ("q1_robin", 1, (2, 2, 1), (-2.0, -2.0, -1.0), (2.0, 2.0, 1.0), [2, 2, 3, 3, 4, 4], 22.0, 0.8, 2.0),
]


@pytest.mark.parametrize("case", GOLD_CASES, ids=[c[0] for c in GOLD_CASES])
def test_gpu_reproduces_committed_oracle_fixtures(ctx, case):
    import os
    gold = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_vectors.npz"))
    name, deg, n, lo, hi, sides, p, t, bt = case
    mesh = G.Mesh(deg, n, lo, hi, sides)
    ctx.set_material_pt(p, t, bt=bt)
    ctx.upload_mesh(mesh)
    ctx.set_solution(gold[name + "_U"])
    ctx.assemble()
    rg = ctx.rhs()
    ro = gold[name + "_rhs"]
    assert np.abs(rg - ro).max() <= 1e-12 * np.abs(ro).max()
    assert ctx.free_energy() == pytest.approx(gold[name + "_F"][0], rel=1e-12)
    a = mesh.arrays()
    Jv = ctx.jacobian()
    for k in range(3):
        v = gold[name + "_V"][k].reshape(-1, 18)
        y = np.zeros_like(v)
        for r in range(mesh.n_nodes):
            for b in range(a["row_ptr"][r], a["row_ptr"][r + 1]):
                y[r] += Jv[b] @ v[a["col"][b]]
        ref = gold[name + "_JV"][k]
        assert np.abs(y.reshape(-1) - ref).max() <= 1e-12 * np.abs(ref).max()
