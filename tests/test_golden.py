"""The committed oracle fixtures (tests/golden/oracle_vectors.npz, written by
tests/golden/make_oracle_vectors.py): the oracle still gives them, and the
library's matep gives the fixture's coefficients bit for bit."""
import os

import numpy as np
import pytest

import oracle_lib as O
from product_util import glhip, oracle_mesh

G = glhip()
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_vectors.npz")
CASES = [
    ("q1_cube", 1, (3, 3, 3), (-3.0, -3.0, -3.0), (3.0, 3.0, 3.0), [1, 1, 1, 1, 0, 0], 20.0, 0.5, 1.0e10),
    ("q2_slice", 2, (2, 2, 2), (-20.0, -20.0, -20.0), (20.0, 20.0, 20.0), [1, 1, 1, 1, 0, 0], 26.0, 0.7, 1.0e10),
    ("q1_robin", 1, (2, 2, 1), (-2.0, -2.0, -1.0), (2.0, 2.0, 1.0), [2, 2, 3, 3, 4, 4], 22.0, 0.8, 2.0),
]


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)  # allow_pickle is False by default


def test_library_matep_equals_fixture_bitwise(gold):
    for p, t, scc, *coef in gold["matep"]:
        rc, m = G.matep(p, t, bool(scc))
        assert rc == 0
        assert np.array_equal(m[:6], np.array(coef)), (p, t, scc)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_oracle_reproduces_fixture(gold, case):
    name, deg, n, lo, hi, sides, p, t, bt = case
    mesh = G.Mesh(deg, n, lo, hi, sides)
    om = oracle_mesh(mesh)
    co = O.coef(p, t, bt=bt)
    U = gold[name + "_U"]
    Jo, ro = O.assemble(om, co, U)
    np.testing.assert_allclose(ro, gold[name + "_rhs"], rtol=0, atol=1e-13 * np.abs(ro).max())
    assert O.free_energy(om, co, U) == pytest.approx(gold[name + "_F"][0], rel=1e-13)
    for k in range(3):
        y = np.zeros(mesh.n_nodes * 18)
        O.lib().orc_bsr_spmv(om.n_nodes, om.row_ptr, om.col, O.f64(Jo.reshape(-1)), O.f64(gold[name + "_V"][k]), y)
        np.testing.assert_allclose(y, gold[name + "_JV"][k], rtol=0, atol=1e-13 * np.abs(y).max())
