"""The point-block Jacobi preconditioner against the inverses of its own
diagonal blocks: D^-1 x to 1e-12 (FP64, assembled and matrix-free, with
Dirichlet rows, Robin walls and hanging nodes) and to FP32 rounding for an
FP32 Jacobian, on Q1 and Q2 meshes; a block count that is not a multiple of
the kernel's nodes per workgroup is included."""
import numpy as np
import pytest

import oracle_lib as O
from product_util import adaptive_forest, glhip

pytestmark = pytest.mark.gpu
G = glhip()


@pytest.fixture(scope="module")
def ctx():
    c = G.Context(0)
    yield c
    c.close()


def _check(ctx, mesh, seed, fp32=False, mf=False, tol=1e-12):
    ctx.upload_mesh(mesh)
    U = O.initial_guess(mesh.n_nodes, O.matep(22.0, 0.8)["gap_B"], amp=0.3, seed=seed)
    ctx.set_solution(U)
    ctx.assemble(J=True, R=False, fp32=fp32, matrix_free=mf)
    D = ctx.diagonal_blocks()
    x = np.random.default_rng(seed).standard_normal(mesh.n_nodes * 18)
    y = ctx.apply_block_jacobi(x)
    ref = np.linalg.solve(D, x.reshape(-1, 18, 1)).reshape(-1)
    err = np.abs(y - ref).max() / np.abs(ref).max()
    print("nodes %d fp32 %d matrix-free %d: rel err %.3e" % (mesh.n_nodes, fp32, mf, err), flush=True)
    assert err <= tol


@pytest.mark.parametrize("deg,cells", [(1, (5, 4, 2)), (2, (3, 3, 2))])
def test_block_jacobi_is_the_inverse_of_the_diagonal_blocks(ctx, deg, cells):
    ctx.set_material_pt(22.0, 0.8, bt=2.0)
    m = G.Mesh(deg, cells, (-5.0, -4.0, -3.0), (5.0, 4.0, 3.0), [1, 1, 2, 2, 4, 4])
    assert m.n_nodes % 12 != 0
    _check(ctx, m, 1)
    _check(ctx, m, 2, mf=True)
    _check(ctx, m, 3, fp32=True, tol=2e-6)
    _check(ctx, m, 4, fp32=True, mf=True, tol=2e-6)


def test_block_jacobi_on_a_mesh_with_hanging_nodes(ctx):
    ctx.set_material_pt(22.0, 0.8)
    lo, hi = (-5.0, -5.0, -6.0), (5.0, 5.0, 6.0)
    m = adaptive_forest(G, (3, 3, 3), lo, hi, [1, 1, 1, 1, 0, 0], 1, 0, frac=0.3).mesh()
    assert m.constraints()["n_hanging"] > 0
    _check(ctx, m, 5)
    _check(ctx, m, 6, mf=True)
