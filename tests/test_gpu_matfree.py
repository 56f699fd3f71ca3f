"""The matrix-free Jacobian (glh_set_matrix_free, GLH_ASM_MATRIX_FREE) against
the assembled BSR Jacobian and the oracle's: J x to 1e-12 relative (normwise,
max|diff| / max|ref|) on Q1 and Q2 box meshes with Dirichlet (id 0), AdGR
(ids 2-4) and Robin walls, x random over every DoF so that the Dirichlet
columns and rows are exercised, and the FP32 forms to 2e-6; a Newton solve,
FP64 or with the FP32 matrix-free inner solve, reaches the assembled FP64
fixed point (free energy 1e-10, order parameter 1e-8); meshes with
constraint lines solve matrix-free too (their parity:
test_gpu_matfree_constrained.py)."""
import sys

import numpy as np
import pytest

import oracle_lib as O
from product_util import adaptive_forest, glhip, initial_state, oracle_mesh

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


def _bsr_mv(om, val, x):
    x = x.reshape(-1, 18)
    y = np.zeros_like(x)
    for r in range(len(om.row_ptr) - 1):
        for b in range(om.row_ptr[r], om.row_ptr[r + 1]):
            y[r] += val[b] @ x[om.col[b]]
    return y.reshape(-1)


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


CASES = [
    # name, degree, cells, lo, hi, boundary ids, p, T/Tc, bt
    ("q1_dirichlet_z", 1, (3, 3, 3), (-3.0, -3.0, -3.0), (3.0, 3.0, 3.0), [1, 1, 1, 1, 0, 0], 22.0, 0.8, 1.0e10),
    ("q2_dirichlet_z", 2, (3, 2, 2), (-6.0, -4.0, -4.0), (6.0, 4.0, 4.0), [1, 1, 1, 1, 0, 0], 26.0, 0.7, 1.0e10),
    ("q1_adgr_robin", 1, (2, 2, 1), (-2.0, -2.0, -1.0), (2.0, 2.0, 1.0), [2, 2, 3, 3, 4, 4], 22.0, 0.8, 2.0),
    ("q2_adgr_robin", 2, (2, 2, 2), (-3.0, -3.0, -3.0), (3.0, 3.0, 3.0), [2, 2, 3, 0, 4, 1], 20.0, 0.6, 1.5),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_matrix_free_product_equals_the_assembled_and_the_oracle(ctx, case):
    name, deg, n, lo, hi, sides, p, t, bt = case
    say("case", name)
    mesh = G.Mesh(deg, n, lo, hi, sides)
    ctx.set_material_pt(p, t, bt=bt)
    ctx.upload_mesh(mesh)
    U = initial_state(mesh.n_nodes, p, t, 0.05, 3)
    ctx.set_solution(U)
    om = oracle_mesh(mesh)
    co = O.coef(p, t, bt=bt)
    Jo, _ = O.assemble(om, co, O.distribute(om, U), R=False)
    rng = np.random.default_rng(deg)
    for k in range(2):
        x = rng.standard_normal(mesh.n_nodes * 18)
        yo = _bsr_mv(om, Jo, x)
        ctx.assemble(J=True, R=False)
        ya = ctx.apply_jacobian(x)
        ctx.assemble(J=True, R=False, matrix_free=True)
        ym = ctx.apply_jacobian(x)
        say(" ", k, "assembled", _rel(ya, yo), "matrix-free", _rel(ym, yo), "mf vs assembled", _rel(ym, ya))
        assert _rel(ya, yo) <= 1e-12
        assert _rel(ym, yo) <= 1e-12
    # the matrix-free form holds no BSR values
    with pytest.raises(G.GLHError, match="matrix-free"):
        ctx.jacobian()
    # FP32 forms (FP32 values and state): the two agree to FP32 rounding
    ctx.assemble(J=True, R=False, fp32=True)
    ya32 = ctx.apply_jacobian(x)
    ctx.assemble(J=True, R=False, fp32=True, matrix_free=True)
    ym32 = ctx.apply_jacobian(x)
    say("  fp32: assembled", _rel(ya32, yo), "matrix-free", _rel(ym32, yo))
    assert _rel(ya32, yo) <= 2e-6 and _rel(ym32, yo) <= 2e-6


def test_matrix_free_newton_reaches_the_assembled_fixed_point(ctx):
    mesh = G.Mesh(2, (4, 4, 4), (-20.0, -20.0, -20.0), (20.0, 20.0, 20.0), [1, 1, 1, 1, 0, 0])
    p, t = 26.0, 0.7
    ctx.set_material_pt(p, t)
    ctx.upload_mesh(mesh)
    U0 = initial_state(mesh.n_nodes, p, t, 0.05, 0)
    kw = dict(rtol=1e-10, max_newton=50, lin_tol=0.1, lin_maxit=20000, restart=50)
    ctx.set_solution(U0)
    st_a = ctx.newton_solve(**kw)
    Ua = ctx.solution()
    ctx.set_solution(U0)
    ctx.set_matrix_free(True)
    try:
        st_m = ctx.newton_solve(**kw)
    finally:
        ctx.set_matrix_free(False)
    Um = ctx.solution()
    say("assembled", st_a)
    say("matrix-free", st_m)
    assert st_a["converged"] == 1.0 and st_m["converged"] == 1.0
    assert abs(st_m["newton_steps"] - st_a["newton_steps"]) <= 1
    assert st_m["F"] == pytest.approx(st_a["F"], rel=1e-10)
    assert np.abs(Um - Ua).max() <= 1e-8 * np.abs(Ua).max()
    om = oracle_mesh(mesh)
    assert O.free_energy(om, O.coef(p, t), Um) == pytest.approx(st_m["F"], rel=1e-10)
    # the last Jacobian was matrix-free
    with pytest.raises(G.GLHError, match="matrix-free"):
        ctx.jacobian()


def test_mixed_precision_matrix_free_newton_reaches_the_fp64_fixed_point(ctx):
    # FP32 matrix-free inner solve (FP32 iterate state, Krylov basis and block
    # Jacobi) inside FP64 iterative refinement with an FP64 residual
    mesh = G.Mesh(2, (4, 4, 4), (-20.0, -20.0, -20.0), (20.0, 20.0, 20.0), [1, 1, 1, 1, 0, 0])
    p, t = 26.0, 0.7
    ctx.set_material_pt(p, t)
    ctx.upload_mesh(mesh)
    U0 = initial_state(mesh.n_nodes, p, t, 0.05, 0)
    kw = dict(rtol=1e-10, max_newton=50, lin_tol=0.1, lin_maxit=20000, restart=50)
    ctx.set_solution(U0)
    st64 = ctx.newton_solve(**kw)
    U64 = ctx.solution()
    ctx.set_solution(U0)
    ctx.set_matrix_free(True)
    try:
        st32 = ctx.newton_solve(fp32=True, **kw)
    finally:
        ctx.set_matrix_free(False)
    U32 = ctx.solution()
    say("fp64", st64)
    say("fp32 matrix-free", st32)
    assert st32["converged"] == 1.0
    assert st32["F"] == pytest.approx(st64["F"], rel=1e-10)
    assert np.abs(U32 - U64).max() <= 1e-8 * np.abs(U64).max()


def test_meshes_with_constraints_solve_matrix_free(ctx):
    lo, hi, sides = (-5.0, -5.0, -6.0), (5.0, 5.0, 6.0), [1, 1, 1, 1, 0, 0]
    f = adaptive_forest(G, (2, 2, 2), lo, hi, sides, 1, 0, frac=0.3)
    m = f.mesh()
    assert m.constraints()["n_hanging"] > 0
    p, t = 22.0, 0.8
    ctx.set_material_pt(p, t)
    ctx.upload_mesh(m)
    ctx.set_solution(initial_state(m.n_nodes, p, t, 0.05, 1))
    ctx.set_matrix_free(True)
    try:
        st = ctx.newton_solve(rtol=1e-8, max_newton=30, lin_tol=0.1, restart=50)
    finally:
        ctx.set_matrix_free(False)
    assert st["converged"] == 1.0
    # the Newton steps kept the matrix-free form (condensed under the hanging
    # lines; its parity is in test_gpu_matfree_constrained): no BSR values
    with pytest.raises(G.GLHError):
        ctx.jacobian()
    assert ctx.diagonal_blocks().shape == (m.n_nodes, 18, 18)


def test_material_change_on_one_mesh_matches_a_fresh_context(ctx):
    # The GMRES cycle is replayed from captured HIP graphs whose matrix-free
    # product carries the material by value (K1, K2 + K3 and the Robin
    # coefficient K1 / bt).  Changing bt on the same uploaded mesh must give
    # the solve of a fresh context at the new bt: same Newton and GMRES
    # iterations, free energy to 1e-12.
    # a slab with AdGR walls on the z faces, thick enough for the B phase
    mesh = G.Mesh(1, (3, 3, 3), (-6.0, -6.0, -6.0), (6.0, 6.0, 6.0), [1, 1, 1, 1, 4, 4])
    p, t = 22.0, 0.5
    U0 = initial_state(mesh.n_nodes, p, t, 0.05, 4)
    kw = dict(rtol=1e-10, max_newton=40, lin_tol=0.1, lin_maxit=20000, restart=30)
    ctx.set_matrix_free(True)
    try:
        ctx.set_material_pt(p, t, bt=2.0)
        ctx.upload_mesh(mesh)
        ctx.set_solution(U0)
        st_a = ctx.newton_solve(**kw)
        ctx.set_material_pt(p, t, bt=1.5)
        ctx.set_solution(U0)
        st_b = ctx.newton_solve(**kw)
    finally:
        ctx.set_matrix_free(False)
    fresh = G.Context(0)
    try:
        fresh.set_matrix_free(True)
        fresh.set_material_pt(p, t, bt=1.5)
        fresh.upload_mesh(mesh)
        fresh.set_solution(U0)
        st_f = fresh.newton_solve(**kw)
    finally:
        fresh.close()
    say("bt 2.0", st_a)
    say("bt 1.5 on the same mesh", st_b)
    say("bt 1.5 fresh context", st_f)
    assert st_b["converged"] == 1.0 and st_f["converged"] == 1.0
    assert st_a["F"] < 0 and st_b["F"] < 0  # an ordered state
    assert abs(st_b["F"] - st_a["F"]) > 1e-6 * abs(st_a["F"])  # the wall term changed it
    assert st_b["newton_steps"] == st_f["newton_steps"]
    assert st_b["gmres_its"] == st_f["gmres_its"]
    assert st_b["F"] == pytest.approx(st_f["F"], rel=1e-12)
