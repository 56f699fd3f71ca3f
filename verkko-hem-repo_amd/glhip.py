"""ctypes binding of libglhip.so (include/glhip.h) for tests and bench.

This is plumbing only: every computation runs in the C++/HIP library.  The
binding refuses to work without the built library -- there is no CPU
fallback.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libglhip.so")

GLH_ASM_J, GLH_ASM_R, GLH_ASM_FP32, GLH_ASM_MATRIX_FREE = 1, 2, 4, 8
ERRORS = {-1: "bad argument", -2: "HIP error", -3: "bad state", -4: "no convergence", -5: "singular block",
          -6: "pressure out of range"}

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")

EXPORTS = ["glh_create", "glh_destroy", "glh_last_error", "glh_set_material", "glh_set_materials", "glh_matep",
           "glh_mesh_box", "glh_mesh_concat", "glh_mesh_instances", "glh_instance_energy", "glh_instance_stats",
           "glh_mesh_destroy", "glh_mesh_sizes", "glh_mesh_get", "glh_upload_mesh", "glh_initial_guess_host", "glh_set_solution", "glh_store_initial", "glh_restore_initial",
           "glh_get_solution", "glh_get_update", "glh_get_rhs", "glh_get_jacobian", "glh_assemble", "glh_rhs_norm",
           "glh_residual", "glh_solve", "glh_newton_step", "glh_free_energy", "glh_newton_solve",
           "glh_last_assembly_ms", "glh_set_verbose", "glh_set_timing", "glh_timing", "glh_bench_spmv", "glh_bench_copy",
           "glh_mesh_constraint_sizes", "glh_mesh_constraints", "glh_forest_create", "glh_forest_destroy",
           "glh_forest_n_cells", "glh_forest_refine", "glh_forest_mesh", "glh_forest_transfer", "glh_kelly",
           "glh_flag_fixed_number", "glh_initial_guess_bna", "glh_run_adaptive",
           "glh_set_continuation", "glh_set_matrix_free", "glh_apply_jacobian", "glh_diagonal_blocks", "glh_apply_block_jacobi", "glh_mlp_create", "glh_mlp_destroy", "glh_mlp_last_error", "glh_mlp_n_params",
           "glh_mlp_get_params", "glh_mlp_set_params", "glh_mlp_set_data", "glh_mlp_train", "glh_mlp_loss_grad",
           "glh_mlp_predict", "glh_instance_residuals", "glh_set_forcing", "glh_set_residual_reference",
           "glh_set_schedule", "glh_set_reorth", "glh_autotune", "glh_prm_parse", "glh_prm_destroy", "glh_prm_get",
           "glh_prm_get_double", "glh_run_prm", "glh_write_vtu", "glh_initial_guess_field",
           "glh_transfer_solution", "glh_forest_set_degree", "glh_bench_read",
           "glh_flag_refine_coarsen", "glh_mesh_from_arrays", "glh_mesh_last_error", "glh_residual_into"]

GLH_ORDER_LEXICOGRAPHIC, GLH_ORDER_HIERARCHIC = 0, 1


class MeshDesc(C.Structure):
    """glh_mesh_desc of include/glhip.h."""
    _fields_ = [("degree", C.c_int), ("n_cells", C.c_int), ("n_nodes", C.c_int), ("local_order", C.c_int),
                ("cell_nodes", C.c_void_p), ("cell_lo", C.c_void_p), ("cell_h", C.c_void_p),
                ("node_xyz", C.c_void_p), ("n_faces", C.c_int), ("faces", C.c_void_p), ("node_mask", C.c_void_p),
                ("n_lines", C.c_int), ("line_node", C.c_void_p), ("line_ptr", C.c_void_p),
                ("line_master", C.c_void_p), ("line_w", C.c_void_p)]

_lib = None


class GLHError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GLHError("libglhip.so is not built: run __graft_entry__.build() (no CPU fallback exists)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.glh_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.glh_destroy.argtypes = [vp]
    L.glh_last_error.argtypes = [vp]
    L.glh_last_error.restype = C.c_char_p
    L.glh_set_material.argtypes = [vp, C.c_double, _dp, _dp, C.c_double]
    L.glh_matep.argtypes = [C.c_double, C.c_double, C.c_int, _dp]
    L.glh_mesh_box.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _ip, C.POINTER(vp)]
    L.glh_mesh_destroy.argtypes = [vp]
    L.glh_set_materials.argtypes = [vp, C.c_int, _dp]
    L.glh_mesh_concat.argtypes = [C.POINTER(vp), C.c_int, C.POINTER(vp)]
    L.glh_mesh_instances.argtypes = [vp, _ip, _ip]
    L.glh_instance_energy.argtypes = [vp, _dp]
    L.glh_instance_residuals.argtypes = [vp, _dp]
    L.glh_instance_stats.argtypes = [vp, _dp]
    L.glh_mesh_sizes.argtypes = [vp, _i64p]
    L.glh_mesh_get.argtypes = [vp] + [vp] * 9
    L.glh_upload_mesh.argtypes = [vp, vp]
    for f in ("glh_set_solution", "glh_get_solution", "glh_get_update", "glh_get_rhs", "glh_get_jacobian"):
        getattr(L, f).argtypes = [vp, _dp]
    L.glh_initial_guess_host.argtypes = [C.c_int, C.c_double, C.c_int, C.c_double, C.c_uint64, _dp]
    L.glh_store_initial.argtypes = [vp]
    L.glh_restore_initial.argtypes = [vp]
    L.glh_assemble.argtypes = [vp, C.c_int]
    L.glh_rhs_norm.argtypes = [vp, C.POINTER(C.c_double)]
    L.glh_residual.argtypes = [vp, C.POINTER(C.c_double)]
    L.glh_solve.argtypes = [vp, C.c_double, C.c_int, C.c_int, _dp]
    L.glh_newton_step.argtypes = [vp, C.c_double, C.c_int, _dp]
    L.glh_free_energy.argtypes = [vp, C.POINTER(C.c_double)]
    L.glh_newton_solve.argtypes = [vp, C.c_double, C.c_int, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int,
                                   C.c_int, _dp]
    L.glh_last_assembly_ms.argtypes = [vp, _dp]
    L.glh_set_timing.argtypes = [vp, C.c_int]
    L.glh_set_verbose.argtypes = [vp, C.c_int]
    L.glh_set_continuation.argtypes = [vp, C.c_double]
    L.glh_set_matrix_free.argtypes = [vp, C.c_int]
    L.glh_apply_jacobian.argtypes = [vp, _dp, _dp]
    L.glh_diagonal_blocks.argtypes = [vp, _dp]
    L.glh_apply_block_jacobi.argtypes = [vp, _dp, _dp]
    L.glh_set_forcing.argtypes = [vp, C.c_int]
    L.glh_write_vtu.argtypes = [vp, _dp, vp, C.c_char_p, C.c_int]
    L.glh_prm_parse.argtypes = [C.c_char_p, C.POINTER(vp), C.c_char_p, C.c_int]
    L.glh_prm_destroy.argtypes = [vp]
    L.glh_prm_get.argtypes = [vp, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int]
    L.glh_prm_get_double.argtypes = [vp, C.c_char_p, C.c_char_p, C.POINTER(C.c_double)]
    L.glh_run_prm.argtypes = [vp, vp, C.c_int, _dp, C.POINTER(C.c_int)]
    L.glh_set_schedule.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    L.glh_set_reorth.argtypes = [vp, C.c_double]
    L.glh_autotune.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_int, _dp]
    L.glh_set_residual_reference.argtypes = [vp, C.c_int, vp]
    L.glh_mlp_create.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.c_uint64, C.POINTER(vp)]
    L.glh_mlp_destroy.argtypes = [vp]
    L.glh_mlp_last_error.argtypes = [vp]
    L.glh_mlp_last_error.restype = C.c_char_p
    L.glh_mlp_n_params.argtypes = [vp]
    L.glh_mlp_n_params.restype = C.c_int64
    L.glh_mlp_get_params.argtypes = [vp, vp, vp]
    L.glh_mlp_set_params.argtypes = [vp, _f32p]
    L.glh_mlp_set_data.argtypes = [vp, C.c_int64, _f32p, _f32p]
    L.glh_mlp_train.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_uint64, C.c_int, vp]
    L.glh_mlp_loss_grad.argtypes = [vp, C.c_int, _f32p, _f32p, C.POINTER(C.c_float), vp]
    L.glh_mlp_predict.argtypes = [vp, C.c_int64, _f32p, _f32p]
    L.glh_timing.argtypes = [vp, _dp]
    L.glh_bench_spmv.argtypes = [vp, C.c_int, C.POINTER(C.c_double)]
    L.glh_bench_copy.argtypes = [vp, C.c_size_t, C.c_int, C.POINTER(C.c_double)]
    L.glh_mesh_constraint_sizes.argtypes = [vp, _i64p]
    L.glh_mesh_constraints.argtypes = [vp] + [vp] * 5
    L.glh_forest_create.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _ip, C.POINTER(vp)]
    L.glh_forest_destroy.argtypes = [vp]
    L.glh_forest_n_cells.argtypes = [vp]
    L.glh_forest_n_cells.restype = C.c_int64
    L.glh_forest_refine.argtypes = [vp, _u8p, C.POINTER(vp)]
    L.glh_forest_mesh.argtypes = [vp, C.POINTER(vp)]
    L.glh_forest_transfer.argtypes = [vp, vp, _dp, vp, vp, _dp]
    L.glh_kelly.argtypes = [vp, vp, _dp, _f32p]
    L.glh_flag_fixed_number.argtypes = [C.c_int64, _f32p, C.c_double, _u8p]
    L.glh_run_adaptive.argtypes = [vp, vp, _dp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_double, C.c_int,
                                   C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, _dp, C.POINTER(vp), C.POINTER(vp)]
    L.glh_initial_guess_bna.argtypes = [C.c_int, _dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                        _dp]
    L.glh_transfer_solution.argtypes = [vp, vp, vp, vp, vp]
    L.glh_forest_set_degree.argtypes = [vp, C.c_int]
    L.glh_bench_read.argtypes = [vp, C.c_size_t, C.c_int, C.POINTER(C.c_double)]
    L.glh_flag_refine_coarsen.argtypes = [C.c_int64, _f32p, C.c_double, C.c_double, _u8p]
    L.glh_initial_guess_field.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_double, C.c_uint64, _dp]
    L.glh_mesh_from_arrays.argtypes = [C.POINTER(MeshDesc), C.POINTER(vp)]
    L.glh_mesh_last_error.argtypes = []
    L.glh_mesh_last_error.restype = C.c_char_p
    L.glh_residual_into.argtypes = [vp, vp, vp, vp]
    _lib = L
    return L


def matep(p, t, scc=True):
    out = np.zeros(12)
    rc = lib().glh_matep(float(p), float(t), int(scc), out)
    return rc, out


K_CONST = (0.42072, 0.42072, 0.42072)  # femgl/inc/femgl.h:320-322


def initial_guess(n_nodes, gap, phase=0, amp=0.05, seed=0):
    U = np.zeros(n_nodes * 18)
    rc = lib().glh_initial_guess_host(int(n_nodes), float(gap), int(phase), float(amp), int(seed), U)
    if rc:
        raise GLHError("glh_initial_guess_host failed (%d)" % rc)
    return U


def initial_guess_bna(node_xyz, z_a, gap_A, gap_B, amp=0.05, seed=1):
    xyz = np.ascontiguousarray(node_xyz, np.float64).reshape(-1)
    n = xyz.size // 3
    U = np.zeros(n * 18)
    rc = lib().glh_initial_guess_bna(n, xyz, float(z_a), float(gap_A), float(gap_B), float(amp), int(seed), U)
    if rc:
        raise GLHError("glh_initial_guess_bna failed (%d)" % rc)
    return U


GUESS_KINDS = {"BinA": 1, "BnB": 2, "PDW": 3}


def initial_guess_field(kind, node_xyz, par, amp=0.0, seed=0):
    """Analytic guess at the nodes: kind "BinA" par (radius, gap_A, gap_B),
    "BnB" par (radius,), "PDW" par (u, half_D, gap_B) (glh_initial_guess_field)."""
    xyz = np.ascontiguousarray(node_xyz, np.float64).reshape(-1)
    n = xyz.size // 3
    U = np.zeros(n * 18)
    pv = np.zeros(4)
    pv[:len(par)] = par
    rc = lib().glh_initial_guess_field(GUESS_KINDS[kind], n, xyz, pv, float(amp), int(seed), U)
    if rc:
        raise GLHError("glh_initial_guess_field failed (%d)" % rc)
    return U


class Mesh:
    """Host mesh layer object (glh_mesh) with numpy views of its arrays."""

    def __init__(self, degree=None, n=None, lo=None, hi=None, side_id=None, handle=None):
        self.h = C.c_void_p()
        if handle is not None:
            self.h = handle
        else:
            rc = lib().glh_mesh_box(int(degree), int(n[0]), int(n[1]), int(n[2]), np.asarray(lo, np.float64),
                                    np.asarray(hi, np.float64), np.asarray(side_id, np.int32), C.byref(self.h))
            if rc:
                raise GLHError("glh_mesh_box failed (%d)" % rc)
        s = np.zeros(8, np.int64)
        lib().glh_mesh_sizes(self.h, s)
        self.deg, self.n_cells, self.n_nodes, self.n_faces, self.nnzb, self.n_colors = [int(v) for v in s[:6]]
        self.n_inst = int(s[6])
        self.ns = (self.deg + 1) ** 3

    @staticmethod
    def from_arrays(degree, cell_nodes, cell_lo, cell_h, node_xyz=None, faces=None, node_mask=None, lines=None,
                    order=GLH_ORDER_LEXICOGRAPHIC):
        """glh_mesh_from_arrays: a mesh from caller-built arrays.  lines is a
        list of (node, masters, weights) constraint lines or None."""
        cn = np.ascontiguousarray(cell_nodes, np.int32).reshape(-1)
        ns = (int(degree) + 1) ** 3
        n_cells = cn.size // ns
        n_nodes = int(cn.max()) + 1 if node_xyz is None else len(node_xyz) // 3 if np.ndim(node_xyz) == 1 else \
            len(node_xyz)
        keep = []

        def ptr(a, dt):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dt).reshape(-1)
            keep.append(a)
            return a.ctypes.data_as(C.c_void_p)

        d = MeshDesc()
        d.degree, d.n_cells, d.n_nodes, d.local_order = int(degree), n_cells, n_nodes, int(order)
        d.cell_nodes = ptr(cn, np.int32)
        d.cell_lo, d.cell_h = ptr(cell_lo, np.float64), ptr(cell_h, np.float64)
        d.node_xyz = ptr(node_xyz, np.float64)
        fa = None if faces is None else np.asarray(faces, np.int32).reshape(-1, 3)
        d.n_faces = 0 if fa is None else len(fa)
        d.faces = ptr(fa, np.int32)
        d.node_mask = ptr(node_mask, np.uint32)
        if lines:
            d.n_lines = len(lines)
            d.line_node = ptr([ln[0] for ln in lines], np.int32)
            d.line_ptr = ptr(np.concatenate([[0], np.cumsum([len(ln[1]) for ln in lines])]), np.int32)
            d.line_master = ptr(np.concatenate([np.asarray(ln[1], np.int32) for ln in lines]), np.int32)
            d.line_w = ptr(np.concatenate([np.asarray(ln[2], np.float64) for ln in lines]), np.float64)
        h = C.c_void_p()
        rc = lib().glh_mesh_from_arrays(C.byref(d), C.byref(h))
        if rc:
            raise GLHError("glh_mesh_from_arrays: " + lib().glh_mesh_last_error().decode())
        return Mesh(handle=h)

    def cell_lo(self):
        """Lower corner of every cell (the coordinates of its local node 0)."""
        a = self.arrays()
        return a["node_xyz"].reshape(-1, 3)[a["cell_nodes"].reshape(-1, self.ns)[:, 0]]

    @staticmethod
    def concat(meshes):
        arr = (C.c_void_p * len(meshes))(*[m.h for m in meshes])
        h = C.c_void_p()
        rc = lib().glh_mesh_concat(arr, len(meshes), C.byref(h))
        if rc:
            raise GLHError("glh_mesh_concat failed (%d)" % rc)
        return Mesh(handle=h)

    def instances(self):
        no = np.zeros(self.n_inst + 1, np.int32)
        co = np.zeros(self.n_inst + 1, np.int32)
        lib().glh_mesh_instances(self.h, no, co)
        return no, co

    def arrays(self):
        a = dict(cell_nodes=np.zeros(self.n_cells * self.ns, np.int32), cell_h=np.zeros(self.n_cells * 3),
                 node_xyz=np.zeros(self.n_nodes * 3), faces=np.zeros(max(1, self.n_faces * 3), np.int32),
                 node_mask=np.zeros(self.n_nodes, np.uint32), color_ptr=np.zeros(self.n_colors + 1, np.int32),
                 color_cells=np.zeros(self.n_cells, np.int32), row_ptr=np.zeros(self.n_nodes + 1, np.int32),
                 col=np.zeros(self.nnzb, np.int32))
        lib().glh_mesh_get(self.h, *[v.ctypes.data_as(C.c_void_p) for v in a.values()])
        a["faces"] = a["faces"][:self.n_faces * 3]
        return a

    def write_vtu(self, U, directory, counter=0, dU=None):
        """VTU/PVTU output of U (and dU) on this mesh (glh_write_vtu)."""
        dptr = None if dU is None else np.ascontiguousarray(dU, np.float64).ctypes.data_as(C.c_void_p)
        rc = lib().glh_write_vtu(self.h, np.ascontiguousarray(U, np.float64), dptr, str(directory).encode(),
                                 int(counter))
        if rc:
            raise GLHError("glh_write_vtu failed (%d)" % rc)

    def constraints(self):
        """Hanging-node lines (hang_ptr, hang_m, hang_w) and the cell-loop
        pattern (arow_ptr, acol); empty lines for conforming meshes."""
        s = np.zeros(3, np.int64)
        lib().glh_mesh_constraint_sizes(self.h, s)
        a = dict(hang_ptr=np.zeros(self.n_nodes + 1, np.int32), hang_m=np.zeros(max(1, s[1]), np.int32),
                 hang_w=np.zeros(max(1, s[1])), arow_ptr=np.zeros(self.n_nodes + 1, np.int32),
                 acol=np.zeros(max(1, s[2]), np.int32))
        lib().glh_mesh_constraints(self.h, *[v.ctypes.data_as(C.c_void_p) for v in a.values()])
        a["hang_m"], a["hang_w"], a["acol"] = a["hang_m"][:s[1]], a["hang_w"][:s[1]], a["acol"][:s[2]]
        a["n_hanging"] = int(s[0])
        return a

    def close(self):
        if self.h:
            lib().glh_mesh_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Forest:
    """Adaptive Q1 or Q2 hex mesh (glh_forest): refinement with 2:1 balance,
    the mesh with hanging-node constraints, Kelly indicators and transfer."""

    def __init__(self, n=None, lo=None, hi=None, side_id=None, handle=None, degree=1):
        self.h = C.c_void_p()
        if handle is not None:
            self.h = handle
        else:
            rc = lib().glh_forest_create(int(n[0]), int(n[1]), int(n[2]), np.asarray(lo, np.float64),
                                         np.asarray(hi, np.float64), np.asarray(side_id, np.int32),
                                         C.byref(self.h))
            if rc:
                raise GLHError("glh_forest_create failed (%d)" % rc)
            if degree != 1:
                rc = lib().glh_forest_set_degree(self.h, int(degree))
                if rc:
                    lib().glh_forest_destroy(self.h)
                    raise GLHError("glh_forest_set_degree failed (%d)" % rc)

    @property
    def n_cells(self):
        return int(lib().glh_forest_n_cells(self.h))

    def refine(self, flags):
        f = np.ascontiguousarray(flags, np.uint8)
        if f.size != self.n_cells:
            raise GLHError("one flag per cell expected")
        h = C.c_void_p()
        rc = lib().glh_forest_refine(self.h, f, C.byref(h))
        if rc:
            raise GLHError("glh_forest_refine failed (%d)" % rc)
        return Forest(handle=h)

    def mesh(self):
        h = C.c_void_p()
        rc = lib().glh_forest_mesh(self.h, C.byref(h))
        if rc:
            raise GLHError("glh_forest_mesh failed (%d)" % rc)
        return Mesh(handle=h)

    def kelly(self, mesh, U):
        eta = np.zeros(self.n_cells, np.float32)
        rc = lib().glh_kelly(self.h, mesh.h, np.ascontiguousarray(U, np.float64), eta)
        if rc:
            raise GLHError("glh_kelly failed (%d)" % rc)
        return eta

    def transfer(self, mesh, U, new_forest, new_mesh):
        Un = np.zeros(new_mesh.n_nodes * 18)
        rc = lib().glh_forest_transfer(self.h, mesh.h, np.ascontiguousarray(U, np.float64), new_forest.h,
                                       new_mesh.h, Un)
        if rc:
            raise GLHError("glh_forest_transfer failed (%d)" % rc)
        return Un

    def close(self):
        if self.h:
            lib().glh_forest_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MLP:
    """Coordinate MLP (glh_mlp): Fourier features of (x, y, z, T, p, D), SiLU
    hidden layers, trained and evaluated by BF16 matrix-core kernels."""

    def __init__(self, n_freq=64, sigma=4.0, width=256, depth=5, n_out=18, seed=0, device=0):
        self.h = C.c_void_p()
        rc = lib().glh_mlp_create(int(device), int(n_freq), float(sigma), int(width), int(depth), int(n_out),
                                  int(seed), C.byref(self.h))
        if rc:
            raise GLHError("glh_mlp_create failed (%d)" % rc)
        self.n_freq, self.width, self.depth, self.n_out = n_freq, width, depth, n_out
        self.n_params = int(lib().glh_mlp_n_params(self.h))

    def _chk(self, rc, what):
        if rc:
            msg = lib().glh_mlp_last_error(self.h)
            raise GLHError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    def params(self):
        p = np.zeros(self.n_params, np.float32)
        b = np.zeros(6 * self.n_freq, np.float32)
        self._chk(lib().glh_mlp_get_params(self.h, p.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p)),
                  "glh_mlp_get_params")
        return p, b.reshape(6, self.n_freq)

    def set_params(self, p):
        self._chk(lib().glh_mlp_set_params(self.h, np.ascontiguousarray(p, np.float32)), "glh_mlp_set_params")

    def set_data(self, X, Y):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        self._chk(lib().glh_mlp_set_data(self.h, X.shape[0], X.reshape(-1), Y.reshape(-1)), "glh_mlp_set_data")

    def train(self, steps, batch=65536, lr=1e-3, seed=0, loss_every=100):
        n = steps // loss_every if loss_every else 0
        L = np.zeros(max(1, n), np.float32)
        self._chk(lib().glh_mlp_train(self.h, int(steps), int(batch), float(lr), int(seed), int(loss_every),
                                      L.ctypes.data_as(C.c_void_p)), "glh_mlp_train")
        return L[:n]

    def loss_grad(self, X, Y):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        g = np.zeros(self.n_params, np.float32)
        L = C.c_float()
        self._chk(lib().glh_mlp_loss_grad(self.h, X.shape[0], X.reshape(-1), Y.reshape(-1), C.byref(L),
                                          g.ctypes.data_as(C.c_void_p)), "glh_mlp_loss_grad")
        return float(L.value), g

    def predict(self, X):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros((X.shape[0], self.n_out), np.float32)
        self._chk(lib().glh_mlp_predict(self.h, X.shape[0], X.reshape(-1), Y.reshape(-1)), "glh_mlp_predict")
        return Y

    def close(self):
        if self.h:
            lib().glh_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Prm:
    """Parsed .prm configuration (glh_prm)."""

    def __init__(self, text):
        self.h = C.c_void_p()
        err = C.create_string_buffer(512)
        rc = lib().glh_prm_parse(text.encode(), C.byref(self.h), err, 512)
        if rc:
            raise GLHError("glh_prm_parse: " + err.value.decode())

    def get(self, subsection, key):
        buf = C.create_string_buffer(256)
        if lib().glh_prm_get(self.h, subsection.encode(), key.encode(), buf, 256):
            raise GLHError("no entry '%s' in '%s'" % (key, subsection))
        return buf.value.decode()

    def get_double(self, subsection, key):
        v = C.c_double()
        if lib().glh_prm_get_double(self.h, subsection.encode(), key.encode(), C.byref(v)):
            raise GLHError("no numeric entry '%s' in '%s'" % (key, subsection))
        return v.value

    def close(self):
        if self.h:
            lib().glh_prm_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def flag_fixed_number(eta, frac):
    eta = np.ascontiguousarray(eta, np.float32)
    fl = np.zeros(eta.size, np.uint8)
    rc = lib().glh_flag_fixed_number(eta.size, eta, float(frac), fl)
    if rc:
        raise GLHError("glh_flag_fixed_number failed (%d)" % rc)
    return fl


def flag_refine_coarsen(eta, refine, coarsen):
    """1 (refine) for the top fraction, 2 (coarsen) for the bottom fraction
    (glh_flag_refine_coarsen)."""
    eta = np.ascontiguousarray(eta, np.float32)
    fl = np.zeros(eta.size, np.uint8)
    rc = lib().glh_flag_refine_coarsen(eta.size, eta, float(refine), float(coarsen), fl)
    if rc:
        raise GLHError("glh_flag_refine_coarsen failed (%d)" % rc)
    return fl


class Context:
    """One device context (glh_ctx): the FemGL step methods on the GPU."""

    def __init__(self, device=0):
        self.h = C.c_void_p()
        rc = lib().glh_create(int(device), C.byref(self.h))
        if rc:
            raise GLHError("glh_create failed (%d): is a GPU visible?" % rc)
        self.n_dofs = 0
        self.nnzb = 0
        self.n_inst = 1

    def _chk(self, rc, what):
        if rc:
            msg = lib().glh_last_error(self.h)
            raise GLHError("%s failed (%d, %s): %s" % (what, rc, ERRORS.get(rc, "?"),
                                                      msg.decode() if msg else ""))

    def set_material(self, alpha, beta, K=K_CONST, bt=1.0e10):
        self._chk(lib().glh_set_material(self.h, float(alpha), np.asarray(beta, np.float64),
                                         np.asarray(K, np.float64), float(bt)), "glh_set_material")

    def set_material_pt(self, p, t, scc=True, bt=1.0e10, K=K_CONST):
        rc, m = matep(p, t, scc)
        if rc:
            raise GLHError("glh_matep: pressure %g outside the strong-coupling table" % p)
        self.set_material(m[0], m[1:6], K, bt)
        return m

    def set_materials(self, coefs):
        """coefs: (n_inst, 10) rows (alpha, beta1..5, K1, K2, K3, bt)."""
        c = np.ascontiguousarray(coefs, np.float64)
        self._chk(lib().glh_set_materials(self.h, c.shape[0], c.reshape(-1)), "glh_set_materials")

    def instance_energy(self):
        return self._get("glh_instance_energy", self.n_inst)

    def instance_residuals(self):
        return self._get("glh_instance_residuals", self.n_inst)

    def instance_stats(self):
        out = np.zeros(self.n_inst * 8)
        self._chk(lib().glh_instance_stats(self.h, out), "glh_instance_stats")
        return out.reshape(-1, 8)

    def upload_mesh(self, mesh):
        self._chk(lib().glh_upload_mesh(self.h, mesh.h), "glh_upload_mesh")
        self.n_dofs = mesh.n_nodes * 18
        self.nnzb = mesh.nnzb
        self.n_inst = mesh.n_inst

    def transfer_solution(self, fo, mo, fn, mn):
        """Refine on the device: upload mn and carry the solution of mo over
        (glh_transfer_solution)."""
        self._chk(lib().glh_transfer_solution(self.h, fo.h, mo.h, fn.h, mn.h), "glh_transfer_solution")
        self.n_dofs = mn.n_nodes * 18
        self.nnzb = mn.nnzb
        self.n_inst = mn.n_inst

    def set_solution(self, U):
        self._chk(lib().glh_set_solution(self.h, np.ascontiguousarray(U, np.float64)), "glh_set_solution")

    def store_initial(self):
        self._chk(lib().glh_store_initial(self.h), "glh_store_initial")

    def restore_initial(self):
        self._chk(lib().glh_restore_initial(self.h), "glh_restore_initial")

    def _get(self, fn, n):
        out = np.zeros(n)
        self._chk(getattr(lib(), fn)(self.h, out), fn)
        return out

    def solution(self):
        return self._get("glh_get_solution", self.n_dofs)

    def update(self):
        return self._get("glh_get_update", self.n_dofs)

    def rhs(self):
        return self._get("glh_get_rhs", self.n_dofs)

    def apply_jacobian(self, x):
        """J x with the current Jacobian, assembled or matrix-free."""
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros_like(x)
        self._chk(lib().glh_apply_jacobian(self.h, x, y), "glh_apply_jacobian")
        return y

    def diagonal_blocks(self):
        """(n_nodes, 18, 18) diagonal blocks of the current Jacobian, the ones
        the point-block Jacobi inverts; glh_diagonal_blocks."""
        out = np.zeros(self.n_dofs * 18)
        self._chk(lib().glh_diagonal_blocks(self.h, out), "glh_diagonal_blocks")
        return out.reshape(-1, 18, 18)

    def apply_block_jacobi(self, x):
        """D^-1 x with the point-block Jacobi preconditioner of the current
        Jacobian, as GMRES applies it; glh_apply_block_jacobi."""
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros_like(x)
        self._chk(lib().glh_apply_block_jacobi(self.h, x, y), "glh_apply_block_jacobi")
        return y

    def jacobian(self):
        return self._get("glh_get_jacobian", self.nnzb * 324).reshape(-1, 18, 18)

    def assemble(self, J=True, R=True, fp32=False, matrix_free=False):
        flags = ((GLH_ASM_J if J else 0) | (GLH_ASM_R if R else 0) | (GLH_ASM_FP32 if fp32 else 0) |
                 (GLH_ASM_MATRIX_FREE if matrix_free else 0))
        self._chk(lib().glh_assemble(self.h, flags), "glh_assemble")

    def rhs_norm(self):
        v = C.c_double()
        self._chk(lib().glh_rhs_norm(self.h, C.byref(v)), "glh_rhs_norm")
        return v.value

    def residual_norm(self):
        v = C.c_double()
        self._chk(lib().glh_residual(self.h, C.byref(v)), "glh_residual")
        return v.value

    def residual_into(self, U_ptr, R_ptr, norm2_ptr):
        """glh_residual_into on device pointers (ints): enqueues -R at U into
        R and the per-instance ||R||^2 into norm2; allocates nothing and does
        not synchronise."""
        self._chk(lib().glh_residual_into(self.h, C.c_void_p(U_ptr), C.c_void_p(R_ptr), C.c_void_p(norm2_ptr)),
                  "glh_residual_into")

    def solve(self, rtol=0.1, maxit=10000, restart=30):
        st = np.zeros(2)
        self._chk(lib().glh_solve(self.h, rtol, maxit, restart, st), "glh_solve")
        return int(st[0]), st[1]

    def newton_step(self, step=0.83, damped=True):
        out = np.zeros(4)
        self._chk(lib().glh_newton_step(self.h, step, int(damped), out), "glh_newton_step")
        return out

    def free_energy(self):
        v = C.c_double()
        self._chk(lib().glh_free_energy(self.h, C.byref(v)), "glh_free_energy")
        return v.value

    def run_prm(self, prm, max_rows=16):
        """FemGL::run from a parsed configuration (glh_run_prm): per-cycle rows
        (cycle, cells, dofs, Newton iterations, ||R||, F)."""
        rows = np.zeros(max_rows * 6)
        n = C.c_int()
        self._chk(lib().glh_run_prm(self.h, prm.h, int(max_rows), rows, C.byref(n)), "glh_run_prm")
        return [dict(zip(["cycle", "cells", "dofs", "iterations", "residual", "F"], r))
                for r in rows.reshape(-1, 6)[:n.value].tolist()]

    CYCLE_KEYS = ["cycle", "cells", "dofs", "hanging", "newton_steps", "gmres_its", "r0", "r_final", "F",
                  "converged", "t_solve", "t_cycle", "F64", "F_rel_err_vs_fp64", "U_rel_err_vs_fp64", "converged64"]

    def run_adaptive(self, forest, U0, n_refine=4, refine_frac=0.3, rtol=1e-10, max_newton=60, lin_tol=0.1,
                     lin_maxit=10000, restart=100, ls_step=0.83, damped=True, fp32=False, check_fp64=False):
        """Adaptive cycles (glh_run_adaptive): returns (per-cycle stats dicts,
        last Forest, last Mesh); the context keeps the last mesh and field."""
        st = np.zeros((n_refine + 1) * 16)
        fo, mo = C.c_void_p(), C.c_void_p()
        self._chk(lib().glh_run_adaptive(self.h, forest.h, np.ascontiguousarray(U0, np.float64), int(n_refine),
                                         float(refine_frac), float(rtol), int(max_newton), float(lin_tol),
                                         int(lin_maxit), int(restart), float(ls_step), int(damped), int(fp32),
                                         int(check_fp64), st, C.byref(fo), C.byref(mo)), "glh_run_adaptive")
        mesh = Mesh(handle=mo)
        self.n_dofs = mesh.n_nodes * 18
        self.nnzb = mesh.nnzb
        self.n_inst = 1
        rows = [dict(zip(self.CYCLE_KEYS, r.tolist())) for r in st.reshape(-1, 16)]
        return rows, Forest(handle=fo), mesh

    def newton_solve(self, rtol=1e-10, max_newton=50, lin_tol=0.1, lin_maxit=20000, restart=30, ls_step=0.83,
                     damped=True, fp32=False):
        st = np.zeros(10)
        self._chk(lib().glh_newton_solve(self.h, rtol, max_newton, lin_tol, lin_maxit, restart, ls_step,
                                         int(damped), int(fp32), st), "glh_newton_solve")
        keys = ["newton_steps", "gmres_its", "r0", "r_final", "t_assemble", "t_solve", "t_linesearch", "F",
                "ls_trials", "converged"]
        return dict(zip(keys, st.tolist()))

    def last_assembly_ms(self):
        v = np.zeros(2)
        self._chk(lib().glh_last_assembly_ms(self.h, v), "glh_last_assembly_ms")
        return v[0]

    def set_verbose(self, level=1):
        self._chk(lib().glh_set_verbose(self.h, int(level)), "glh_set_verbose")

    def set_reorth(self, ratio):
        """Second Gram-Schmidt pass when ||w'||^2 < ratio ||w||^2 (glh_set_reorth)."""
        self._chk(lib().glh_set_reorth(self.h, float(ratio)), "glh_set_reorth")

    def set_schedule(self, restart=100, q_block=0, row_cap=0, chunk_cap=0, gs_tile=0):
        self._chk(lib().glh_set_schedule(self.h, int(restart), int(q_block), int(row_cap), int(chunk_cap),
                                         int(gs_tile)), "glh_set_schedule")

    def autotune(self, fp32=True, trial_newton=2, lin_tol=0.1, lin_maxit=3000):
        """Time candidate schedules and kernel tiles on the current problem and
        keep the fastest (see glh_autotune); returns the choice as a dict."""
        out = np.zeros(7)
        self._chk(lib().glh_autotune(self.h, int(bool(fp32)), int(trial_newton), float(lin_tol), int(lin_maxit), out),
                  "glh_autotune")
        return dict(restart=int(out[0]), q_block=int(out[1]), row_cap=int(out[2]), chunk_cap=int(out[3]),
                    trial_s=float(out[4]), candidates=int(out[5]), gs_tile=int(out[6]))

    def set_forcing(self, eisenstat_walker):
        self._chk(lib().glh_set_forcing(self.h, int(bool(eisenstat_walker))), "glh_set_forcing")

    def set_residual_reference(self, r_ref=None):
        if r_ref is None:
            self._chk(lib().glh_set_residual_reference(self.h, 0, None), "glh_set_residual_reference")
        else:
            r = np.ascontiguousarray(r_ref, np.float64)
            self._chk(lib().glh_set_residual_reference(self.h, r.size, r.ctypes.data_as(C.c_void_p)),
                      "glh_set_residual_reference")

    def set_continuation(self, sigma0):
        """Pseudo-transient continuation (J + sigma0 ||R||/||R0|| |diag J|); 0: off."""
        self._chk(lib().glh_set_continuation(self.h, float(sigma0)), "glh_set_continuation")

    def set_matrix_free(self, on=True):
        """Matrix-free Jacobian in newton_solve (FP64 or FP32; batches on
        meshes without constraint lines); glh_set_matrix_free."""
        self._chk(lib().glh_set_matrix_free(self.h, int(bool(on))), "glh_set_matrix_free")

    def set_timing(self, on=True):
        self._chk(lib().glh_set_timing(self.h, int(on)), "glh_set_timing")

    def timing(self):
        v = np.zeros(6)
        self._chk(lib().glh_timing(self.h, v), "glh_timing")
        return dict(spmv_ms=v[0], spmv_n=int(v[1]), jasm_ms=v[2], jasm_n=int(v[3]), rasm_ms=v[4], rasm_n=int(v[5]))

    def bench_spmv(self, k=10):
        v = C.c_double()
        self._chk(lib().glh_bench_spmv(self.h, int(k), C.byref(v)), "glh_bench_spmv")
        return v.value

    def bench_read(self, nbytes, k=5):
        v = C.c_double()
        self._chk(lib().glh_bench_read(self.h, int(nbytes), int(k), C.byref(v)), "glh_bench_read")
        return v.value

    def bench_copy(self, nbytes, k=5):
        v = C.c_double()
        self._chk(lib().glh_bench_copy(self.h, int(nbytes), int(k), C.byref(v)), "glh_bench_copy")
        return v.value

    def close(self):
        if self.h:
            lib().glh_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
