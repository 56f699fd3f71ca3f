"""glh_residual_into, the variant of glh_residual that allocates nothing
(caller's device buffers in and out, the norm's chunk partials in the
library's fixed device scratch array), called through the C ABI and checked
against the existing entry points on the same context: R equal to the rhs of
glh_assemble(GLH_ASM_R) bit for bit, and the per-instance ||R||^2 equal to
the squares of glh_instance_residuals.  Two calls in a row, the second on a
smaller input (a batch of two small instances after an adaptive mesh with
hanging nodes), so the scratch contents of the first call cannot leak into
the second."""
import sys

import numpy as np
import pytest

import oracle_lib as O
from product_util import adaptive_forest, glhip, initial_state

pytestmark = pytest.mark.gpu
G = glhip()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class _Dev:
    """Device buffers for the test through the HIP runtime the library itself
    links (/opt/rocm; a second runtime in the process, such as torch's, does
    not see the GPU once this one holds it)."""

    def __init__(self):
        import ctypes as C
        self.C = C
        self.hip = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.ptrs = []

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float64)
        p = self.C.c_void_p()
        assert self.hip.hipMalloc(self.C.byref(p), max(8, a.nbytes)) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(self.C.c_void_p), a.nbytes, 1) == 0  # host to device
        return p.value

    def download(self, ptr, n):
        out = np.empty(n, np.float64)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(self.C.c_void_p), self.C.c_void_p(ptr), out.nbytes, 2) == 0
        return out

    def close(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _check(ctx, n_dofs, n_inst, U):
    ctx.set_solution(U)
    ctx.assemble(J=False, R=True)
    rhs = ctx.rhs()
    r_inst = ctx.instance_residuals()
    dev = _Dev()
    try:
        Ud = dev.upload(ctx.solution())
        Rd = dev.upload(np.full(n_dofs, np.nan))
        N2 = dev.upload(np.full(n_inst, np.nan))
        assert dev.hip.hipDeviceSynchronize() == 0
        ctx.residual_into(Ud, Rd, N2)
        R = dev.download(Rd, n_dofs)
        n2 = dev.download(N2, n_inst)
    finally:
        dev.close()
    say("no-alloc residual: dofs %d instances %d max|R - rhs| %.3e norms %s vs %s" %
        (n_dofs, n_inst, np.abs(R - rhs).max(), np.sqrt(n2), r_inst))
    assert np.array_equal(R, rhs)
    assert np.allclose(np.sqrt(n2), r_inst, rtol=1e-15, atol=0.0)


def test_no_alloc_residual_matches_the_allocating_entry_points():
    ctx = G.Context(0)
    try:
        p, t = 22.0, 0.8
        # call 1: an adaptive Q1 mesh with hanging nodes (condensed residual)
        lo, hi, sides = (-5.0, -5.0, -6.0), (5.0, 5.0, 6.0), [1, 1, 1, 1, 0, 0]
        m1 = adaptive_forest(G, (3, 3, 3), lo, hi, sides, 1, 0, frac=0.3).mesh()
        assert m1.constraints()["n_hanging"] > 0
        ctx.set_material_pt(p, t)
        ctx.upload_mesh(m1)
        _check(ctx, m1.n_nodes * 18, 1, initial_state(m1.n_nodes, p, t, 0.05, 1))
        # call 2: a smaller batch of two box instances with their own materials
        ma = G.Mesh(1, (2, 2, 2), (-2.0, -2.0, -2.0), (2.0, 2.0, 2.0), [1, 1, 1, 1, 0, 0])
        mb = G.Mesh(1, (3, 2, 1), (-3.0, -2.0, -1.0), (3.0, 2.0, 1.0), [2, 2, 3, 3, 4, 4])
        mbat = G.Mesh.concat([ma, mb])
        assert mbat.n_nodes < m1.n_nodes
        ctx.upload_mesh(mbat)
        ctx.set_materials(np.array([O.coef(p, t), O.coef(26.0, 0.7, bt=2.0)]))
        U2 = np.concatenate([initial_state(ma.n_nodes, p, t, 0.05, 2), initial_state(mb.n_nodes, 26.0, 0.7, 0.05, 3)])
        _check(ctx, mbat.n_nodes * 18, 2, U2)
    finally:
        ctx.close()
