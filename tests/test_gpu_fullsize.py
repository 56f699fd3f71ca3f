"""Size-independent properties of the Newton step at the full size of the
headline workload (32^3 Q2 cells, 4.94 M DoFs), where the oracle would take
hours: the assembled J is the derivative of the assembled residual (the
linearised residual along the GMRES update vanishes to second order), the
residual is half the gradient of the free energy, and assembly repeats bit
for bit."""
import sys

import numpy as np
import pytest

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
