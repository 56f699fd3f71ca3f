"""The matrix-free Jacobian on meshes with constraint lines (hanging nodes of
adaptive forests; periodic images and hanging lines of caller-built meshes)
against the assembled condensed Jacobian and the oracle: J x to 1e-12,
hanging and Dirichlet rows included; the condensed diagonal blocks that the
point-block Jacobi inverts to 1e-12; the FP32 form to 2e-6; Newton solves
(FP64, FP32 mixed, with pseudo-transient continuation) that take the
assembled path's Newton steps and reach its free energy."""
import sys

import numpy as np
import pytest

import oracle_lib as O
from product_util import adaptive_forest, caller_arrays, glhip, oracle_adaptive, oracle_caller

pytestmark = pytest.mark.gpu
G = glhip()
P, T = 22.0, 0.8
LO, HI = (-6.0, -6.0, -4.0), (6.0, 6.0, 4.0)
PERIODIC = [5, 5, 5, 5, 0, 0]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = G.Context(0)
    yield c
    c.close()


def bsr_mv(row_ptr, col, val, x):
    """y = J x for BSR values val[nnzb, 18, 18]."""
    X = x.reshape(-1, 18)
    Y = np.zeros_like(X)
    for r in range(len(row_ptr) - 1):
        for b in range(row_ptr[r], row_ptr[r + 1]):
            Y[r] += val[b] @ X[col[b]]
    return Y.reshape(-1)


def build(ctx, deg, steps, seed, sides, bt, caller):
    """Upload an adaptive forest mesh (or the caller-built arrays of one),
    set a seeded state; returns (mesh, oracle mesh, distributed state,
    coefficients)."""
    f = adaptive_forest(G, (3, 3, 2) if deg == 1 else (2, 2, 2), LO, HI, sides, steps, seed, degree=deg)
    m = f.mesh()
    om = oracle_adaptive(m, LO, HI, sides)
    if caller:
        kw, perm, images = caller_arrays(m, LO, HI, seed=21 + seed, hierarchic=True)
        m = G.Mesh.from_arrays(**kw)
        om = oracle_caller(om, perm, images, m)
    say("case deg %d steps %d seed %d sides %s bt %g caller %d: cells %d nodes %d" % (
        deg, steps, seed, sides, bt, caller, m.n_cells, m.n_nodes))
    ctx.set_material_pt(P, T, bt=bt)
    ctx.upload_mesh(m)
    U = O.initial_guess(m.n_nodes, O.matep(P, T)["gap_B"], amp=0.05, seed=seed)
    Ud = O.distribute(om, U)
    ctx.set_solution(U)
    return m, om, Ud, O.coef(P, T, bt=bt)


CASES = [
    # degree, refinement steps, seed, boundary ids, Robin length bt, caller-built
    (1, 2, 1, [1, 1, 1, 1, 0, 0], 1.0e10, False),
    (2, 1, 2, [1, 1, 1, 1, 4, 4], 2.0, False),
    (1, 2, 3, PERIODIC, 1.0e10, True),
    (2, 1, 2, PERIODIC, 1.0e10, True),
    (1, 2, 4, [2, 2, 3, 3, 4, 4], 2.0, True),
]


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("deg,steps,seed,sides,bt,caller", CASES)
def test_products_and_diagonal_blocks_match_the_assembled_condensed_jacobian(ctx, deg, steps, seed, sides, bt,
                                                                             caller):
    m, om, Ud, co = build(ctx, deg, steps, seed, sides, bt, caller)
    assert m.constraints()["n_hanging"] > 0
    rng = np.random.default_rng(100 + seed)
    X = rng.standard_normal((2, m.n_nodes * 18))
    Jo, _ = O.assemble(om, co, Ud, R=False)
    ctx.assemble(J=True, R=False)
    ya = [ctx.apply_jacobian(x) for x in X]
    Da = ctx.diagonal_blocks()
    ctx.assemble(J=True, R=False, matrix_free=True)
    ym = [ctx.apply_jacobian(x) for x in X]
    Dm = ctx.diagonal_blocks()
    for x, a, b in zip(X, ya, ym):
        yo = bsr_mv(om.row_ptr, om.col, Jo, x)
        say("J x rel err: matrix-free vs assembled %.3e, vs oracle %.3e, assembled vs oracle %.3e" % (
            _rel(b, a), _rel(b, yo), _rel(a, yo)))
        assert _rel(b, a) <= 1e-12 and _rel(b, yo) <= 1e-12
    say("diagonal blocks rel err %.3e" % _rel(Dm, Da))
    assert _rel(Dm, Da) <= 1e-12
    # the FP32 form: FP32 state and arithmetic, FP64 diagonal blocks
    ctx.assemble(J=True, R=False, matrix_free=True, fp32=True)
    y32 = ctx.apply_jacobian(X[0])
    say("FP32 matrix-free J x rel err %.3e" % _rel(y32, ya[0]))
    assert _rel(y32, ya[0]) <= 2e-6
    assert _rel(ctx.diagonal_blocks(), Da) <= 1e-12


def _solve(ctx, U, mf, fp32=False, **kw):
    ctx.set_solution(U)
    ctx.set_matrix_free(mf)
    try:
        return ctx.newton_solve(fp32=fp32, **kw), ctx.solution()
    finally:
        ctx.set_matrix_free(False)


@pytest.mark.parametrize("deg,steps,seed,sides,bt,caller", [CASES[0], CASES[3]])
def test_newton_with_the_matrix_free_jacobian_matches_the_assembled_solve(ctx, deg, steps, seed, sides, bt, caller):
    m, om, Ud, co = build(ctx, deg, steps, seed, sides, bt, caller)
    kw = dict(rtol=1e-10, max_newton=60, lin_tol=1e-3, restart=50)
    sa, Ua = _solve(ctx, Ud, False, **kw)
    sm, Um = _solve(ctx, Ud, True, **kw)
    s32, U32 = _solve(ctx, Ud, True, fp32=True, **kw)
    say("assembled", sa)
    say("matrix-free", sm)
    say("matrix-free FP32", s32)
    assert sa["converged"] == sm["converged"] == s32["converged"] == 1.0
    assert sm["newton_steps"] == sa["newton_steps"]
    assert abs(sm["gmres_its"] - sa["gmres_its"]) <= 2  # the same operator and preconditioner, to rounding
    assert sm["F"] == pytest.approx(sa["F"], rel=1e-12)
    assert np.abs(Um - Ua).max() <= 1e-8 * np.abs(Ua).max()
    assert s32["F"] == pytest.approx(sa["F"], rel=1e-10)
    # the fixed point is the oracle's: its condensed residual there
    _, r0 = O.assemble(om, co, Ud, J=False)
    _, rs = O.assemble(om, co, Um, J=False)
    assert np.linalg.norm(rs) <= 1.01e-10 * np.linalg.norm(r0)


@pytest.mark.parametrize("adaptive", [False, True])
def test_continuation_with_the_matrix_free_jacobian(ctx, adaptive):
    # J + sigma_k |diag J| applied matrix-free takes the steps of the
    # assembled shifted matrix and reaches the same state: on the uniform
    # mesh from the B-in-A start where damped Newton stalls (see the adaptive
    # tests), on an adaptive mesh from a noisy B phase
    lo, hi = (-5.0, -5.0, -6.0), (5.0, 5.0, 6.0)
    sides = [1, 1, 1, 1, 0, 0]
    f = adaptive_forest(G, (4, 4, 4), lo, hi, sides, 1 if adaptive else 0, seed=6)
    m = f.mesh()
    mp = O.matep(P, T)
    if adaptive:
        U0 = O.initial_guess(m.n_nodes, mp["gap_B"], amp=0.05, seed=6)
    else:
        U0 = G.initial_guess_bna(m.arrays()["node_xyz"], 0.0, mp["gap_A"], mp["gap_B"], amp=0.05, seed=1)
    ctx.set_material_pt(P, T)
    ctx.upload_mesh(m)
    ctx.set_solution(U0)
    U0 = ctx.solution()
    kw = dict(rtol=1e-10, max_newton=40, lin_tol=1e-2, restart=50)
    ctx.set_continuation(1.0)
    try:
        sa, Ua = _solve(ctx, U0, False, **kw)
        sm, Um = _solve(ctx, U0, True, **kw)
    finally:
        ctx.set_continuation(0.0)
    say("hanging", m.constraints()["n_hanging"], "assembled", sa)
    say("matrix-free", sm)
    assert (m.constraints()["n_hanging"] > 0) == adaptive
    assert sa["converged"] == sm["converged"] == 1.0
    assert sm["newton_steps"] == sa["newton_steps"]
    assert abs(sm["gmres_its"] - sa["gmres_its"]) <= 2
    assert sm["F"] == pytest.approx(sa["F"], rel=1e-11)
