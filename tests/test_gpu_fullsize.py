"""The Newton step at the full size of the headline workload (32^3 Q2 cells,
4.94 M DoFs, T/Tc 0.7, p 26 bar, B phase + 0.05 noise from seed 0).

Against the oracle, cell by cell without a global J (orc_cells_apply, its
colours in parallel on the host cores): the residual at the seeded state to
1e-12; J x of the assembled and of the matrix-free Jacobian for random x to
1e-12 (FP64) and 2e-6 (the FP32 forms); at the state the headline solve
converges to, the oracle's free energy within 1e-10 of the GPU's and the
oracle's ||R|| at most 1e-10 of its ||R0|| (FP64 and the FP32 mixed solve).

Size-independent properties besides: the assembled J is the derivative of
the assembled residual (the linearised residual along the GMRES update
vanishes to second order), the residual is half the gradient of the free
energy, and assembly repeats bit for bit."""
import sys

import numpy as np
import pytest

import oracle_lib as O
from product_util import cfg_mesh, glhip, initial_state

pytestmark = pytest.mark.gpu
G = glhip()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def full():
    say("building the 32^3 Q2 mesh")
    mesh = cfg_mesh(G, "cube_q2_ref5")
    ctx = G.Context(0)
    ctx.set_material_pt(26.0, 0.7)
    ctx.upload_mesh(mesh)
    U = initial_state(mesh.n_nodes, 26.0, 0.7, 0.05, 0)
    ctx.set_solution(U)
    U = ctx.solution()  # with the Dirichlet components distributed
    yield ctx, mesh, U
    ctx.close()
    mesh.close()


def _rhs_at(ctx, U):
    ctx.set_solution(U)
    ctx.assemble(J=False, R=True)
    return ctx.rhs()


def test_jacobian_is_the_derivative_of_the_residual(full):
    ctx, mesh, U = full
    ctx.set_solution(U)
    ctx.assemble()
    rhs0 = ctx.rhs()  # -R(U)
    its, rel = ctx.solve(rtol=1e-8, maxit=20000, restart=100)
    say("gmres", its, rel)
    assert rel <= 1e-8
    d = ctx.update()  # J d = -R
    # R(U + t d) = R + t J d + O(t^2) = (1 - t) R + O(t^2): the first-order
    # remainder divided by t must shrink in proportion to t
    norm0 = np.linalg.norm(rhs0)
    n = {}
    for t in (4e-3, 2e-3, 1e-3):
        r = _rhs_at(ctx, U + t * d)
        n[t] = np.linalg.norm(r - (1 - t) * rhs0) / (t * norm0)
        say("t", t, "normalised remainder", n[t])
    assert n[1e-3] <= 1e-3
    # above the floor left by the linear solve (rtol 1e-8) the remainder is
    # the second-order term, which halves with t; a wrong Jacobian term
    # would leave a part that does not shrink
    if n[1e-3] > 1e-6:
        assert n[1e-3] <= 0.6 * n[2e-3] and n[2e-3] <= 0.6 * n[4e-3]


def test_residual_is_half_the_energy_gradient(full):
    ctx, mesh, U = full
    rhs0 = _rhs_at(ctx, U)  # -R(U)
    rng = np.random.default_rng(3)
    v = rng.standard_normal(U.size) * np.abs(U).max()
    ctx.set_solution(v)
    v = ctx.solution()  # a direction that respects the Dirichlet constraints
    t = 1e-4
    ctx.set_solution(U + t * v)
    Fp = ctx.free_energy()
    ctx.set_solution(U - t * v)
    Fm = ctx.free_energy()
    dF = (Fp - Fm) / (2 * t)
    want = 2.0 * float(np.dot(-rhs0, v))
    say("dF/dt", dF, "2 R.v", want)
    assert dF == pytest.approx(want, rel=1e-6)


def test_full_size_assembly_repeats_bit_for_bit(full):
    ctx, mesh, U = full
    ctx.set_solution(U)
    ctx.assemble()
    r1 = ctx.rhs()
    its1, rel1 = ctx.solve(rtol=1e-3, maxit=5000, restart=100)
    d1 = ctx.update()
    ctx.assemble()
    r2 = ctx.rhs()
    its2, rel2 = ctx.solve(rtol=1e-3, maxit=5000, restart=100)
    d2 = ctx.update()
    assert np.array_equal(r1, r2)
    assert its1 == its2 and rel1 == rel2 and np.array_equal(d1, d2)


P, T = 26.0, 0.7


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def test_full_size_residual_and_products_match_the_oracle(full):
    ctx, mesh, U = full
    co = O.coef(P, T)
    a = mesh.arrays()
    rng = np.random.default_rng(12)
    X = rng.standard_normal((2, U.size))
    say("oracle: residual and two products on %d cells" % mesh.n_cells)
    ro, Yo = O.cells_apply(mesh.deg, a, mesh.n_nodes, co, U, X)
    ctx.set_solution(U)
    ctx.assemble()  # FP64 BSR J and -R
    er = _rel(ctx.rhs(), ro)
    say("rhs rel err", er)
    assert er <= 1e-12
    for x, yo in zip(X, Yo):
        ya = ctx.apply_jacobian(x)
        say("assembled J x rel err", _rel(ya, yo))
        assert _rel(ya, yo) <= 1e-12
    ctx.assemble(J=True, R=False, matrix_free=True)
    for x, yo in zip(X, Yo):
        ym = ctx.apply_jacobian(x)
        say("matrix-free J x rel err", _rel(ym, yo))
        assert _rel(ym, yo) <= 1e-12
    ctx.assemble(J=True, R=False, fp32=True, matrix_free=True)
    for x, yo in zip(X, Yo):
        say("matrix-free FP32 J x rel err", _rel(ctx.apply_jacobian(x), yo))
        assert _rel(ctx.apply_jacobian(x), yo) <= 2e-6
    ctx.assemble(J=True, R=False, fp32=True)
    ya32 = ctx.apply_jacobian(X[0])
    say("assembled FP32 J x rel err", _rel(ya32, Yo[0]))
    assert _rel(ya32, Yo[0]) <= 2e-6


@pytest.mark.parametrize("fp32", [False, True], ids=["fp64", "fp32_mixed"])
def test_headline_solve_reaches_the_oracle_fixed_point(full, fp32):
    # the headline solve (matrix-free J, GMRES(100), eta 0.1, damped Newton
    # to ||R|| <= 1e-10 ||R0||); the oracle judges the state it reaches
    ctx, mesh, U = full
    co = O.coef(P, T)
    a = mesh.arrays()
    ctx.set_solution(U)
    ctx.set_matrix_free(True)
    try:
        st = ctx.newton_solve(rtol=1e-10, max_newton=50, lin_tol=0.1, lin_maxit=20000, restart=100, fp32=fp32)
    finally:
        ctx.set_matrix_free(False)
    say("headline solve", "fp32" if fp32 else "fp64", st)
    assert st["converged"] == 1.0
    Us = ctx.solution()
    r0, _ = O.cells_apply(mesh.deg, a, mesh.n_nodes, co, U)
    rs, _ = O.cells_apply(mesh.deg, a, mesh.n_nodes, co, Us)
    om_F = O.lib().orc_free_energy(mesh.deg, mesh.n_cells, O.i32(a["cell_nodes"]), O.f64(a["cell_h"]),
                                   len(a["faces"]) // 3, O.i32(a["faces"]), O.f64(co), O.f64(Us))
    say("oracle ||R0|| %.6e ||R*|| %.6e ratio %.3e F %.15e GPU F %.15e" % (np.linalg.norm(r0), np.linalg.norm(rs),
                                                                         np.linalg.norm(rs) / np.linalg.norm(r0),
                                                                         om_F, st["F"]))
    assert np.linalg.norm(rs) <= 1e-10 * np.linalg.norm(r0)
    assert st["F"] == pytest.approx(om_F, rel=1e-10)
