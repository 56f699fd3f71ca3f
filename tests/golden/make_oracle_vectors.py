"""Writes tests/golden/oracle_vectors.npz: fixed inputs and the oracle's
outputs for them (matep coefficients on a (p, T, SCC) grid; residual, free
energy and Jacobian-vector products on three small seeded meshes).  The
fixtures freeze the oracle's answers, so that a change of the oracle or of
the GPU path against them shows up as a failed comparison.

    python3 tests/golden/make_oracle_vectors.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402
from product_util import glhip, initial_state, oracle_mesh  # noqa: E402

# name, degree, cells, lo, hi, side ids, p, T, bt
CASES = [
    ("q1_cube", 1, (3, 3, 3), (-3.0, -3.0, -3.0), (3.0, 3.0, 3.0), [1, 1, 1, 1, 0, 0], 20.0, 0.5, 1.0e10),
    ("q2_slice", 2, (2, 2, 2), (-20.0, -20.0, -20.0), (20.0, 20.0, 20.0), [1, 1, 1, 1, 0, 0], 26.0, 0.7, 1.0e10),
    ("q1_robin", 1, (2, 2, 1), (-2.0, -2.0, -1.0), (2.0, 2.0, 1.0), [2, 2, 3, 3, 4, 4], 22.0, 0.8, 2.0),
]
MATEP_P = [0.0, 2.5, 20.0, 22.0, 26.0, 33.9]
MATEP_T = [0.3, 0.5, 0.7, 0.8, 0.95]


def case_vectors(G, name, deg, n, lo, hi, sides, p, t, bt):
    mesh = G.Mesh(deg, n, lo, hi, sides)
    om = oracle_mesh(mesh)
    co = O.coef(p, t, bt=bt)
    U = O.distribute(om, initial_state(mesh.n_nodes, p, t, 0.05, 0))
    Jo, ro = O.assemble(om, co, U)
    rng = np.random.default_rng(17)
    V = rng.standard_normal((3, mesh.n_nodes * 18))
    JV = np.zeros_like(V)
    for k in range(3):
        y = np.zeros(mesh.n_nodes * 18)
        O.lib().orc_bsr_spmv(om.n_nodes, om.row_ptr, om.col, O.f64(Jo.reshape(-1)), O.f64(V[k]), y)
        JV[k] = y
    return {name + "_U": U, name + "_rhs": ro, name + "_F": np.array([O.free_energy(om, co, U)]),
            name + "_V": V, name + "_JV": JV}


def main():
    G = glhip()
    out = {}
    rows = []
    for scc in (1, 0):
        for p in MATEP_P:
            for t in MATEP_T:
                m = O.matep(p, t, bool(scc))
                rows.append([p, t, scc] + [m[k] for k in ("alpha", "beta1", "beta2", "beta3", "beta4", "beta5")])
    out["matep"] = np.array(rows)
    for c in CASES:
        out.update(case_vectors(G, *c))
    np.savez_compressed(os.path.join(HERE, "oracle_vectors.npz"), **out)
    print("wrote", sorted(out))


if __name__ == "__main__":
    main()
