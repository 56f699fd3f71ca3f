"""The adaptive slab at its full size (40 x 40 x 12 xi0, 16 x 16 x 4 Q1 root
cells, A = 0 on the slab faces, four Kelly cycles refining 30%, T/Tc 0.8,
p 22 bar, B-in-A guess seed 1, pseudo-transient continuation sigma0 = 1):
the library's cycle loop with the matrix-free Jacobian and the FP32 inner
solve, each cycle checked against an FP64 solve from the same start.  At the
last cycle (about 2.2 M dofs, 43k hanging nodes) the oracle, under the
library's constraint lines, finds the GPU's state a fixed point of the
condensed system and the same free energy."""
import sys

import numpy as np
import pytest

import oracle_lib as O
from product_util import glhip, oracle_from_library

pytestmark = pytest.mark.gpu
G = glhip()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_adaptive_slab_four_cycles_at_full_size():
    lo, hi = (-20.0, -20.0, -6.0), (20.0, 20.0, 6.0)
    p, t = 22.0, 0.8
    f = G.Forest((16, 16, 4), lo, hi, [1, 1, 1, 1, 0, 0])
    m = f.mesh()
    mp = O.matep(p, t)
    U0 = G.initial_guess_bna(m.arrays()["node_xyz"], 0.0, mp["gap_A"], mp["gap_B"], amp=0.05, seed=1)
    ctx = G.Context(0)
    try:
        ctx.set_material_pt(p, t)
        ctx.set_continuation(1.0)
        ctx.set_matrix_free(True)
        say("running 4 cycles from %d cells" % m.n_cells)
        rows, f2, m2 = ctx.run_adaptive(f, U0, n_refine=4, refine_frac=0.3, rtol=1e-10, lin_tol=0.1, restart=100,
                                        fp32=True, check_fp64=True)
        for r in rows:
            say(r)
            assert r["converged"] == 1.0 and r["converged64"] == 1.0
            assert r["F_rel_err_vs_fp64"] <= 1e-10
        assert len(rows) == 5 and rows[0]["cells"] == 1024 and rows[-1]["hanging"] > 0
        assert rows[-1]["dofs"] > 1.0e6
        co = O.coef(p, t)
        U = ctx.solution()
        om = oracle_from_library(m2)
        _, R = O.assemble(om, co, U, J=False)
        Fo = O.free_energy(om, co, U)
        say("oracle ||R|| %.6e (cycle r0 %.6e), F %.15e GPU F %.15e" % (np.linalg.norm(R), rows[-1]["r0"], Fo,
                                                                        rows[-1]["F"]))
        assert np.linalg.norm(R) <= 1.01e-10 * rows[-1]["r0"]
        assert rows[-1]["F"] == pytest.approx(Fo, rel=1e-12)
    finally:
        ctx.set_continuation(0.0)
        ctx.close()
