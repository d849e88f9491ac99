"""Python wrappers over the C ABI (include/afepack_amd.h).

Host arrays are numpy; device vectors are DeviceVector buffers allocated
through the library (afe_dev_alloc), so a process loads exactly one HIP
runtime.  Every call goes through libafepack_amd.so.
"""
import ctypes as C

import numpy as np

from . import _sigs

__all__ = [
    "P1_TRI", "P2_TRI", "P1_TET", "OP_STIFF", "OP_MASS", "OP_VARSTIFF",
    "F_POISSON", "F_VARCOEF", "F_NODAL", "F_ONE", "PREC_FP64", "PREC_FP32",
    "set_tuning", "mesh_square", "mesh_cube", "mesh_displace", "field_random", "Field",
    "Context", "Space", "Matrix", "Amg", "HostAmg", "AmgParams", "DeviceVector", "Batch", "Server",
]

P1_TRI, P2_TRI, P1_TET = 1, 2, 3
OP_STIFF, OP_MASS, OP_VARSTIFF = 0, 1, 2
F_POISSON, F_VARCOEF, F_NODAL, F_ONE = 0, 1, 2, 3
PREC_FP64, PREC_FP32 = 0, 1


def _lib():
    from . import load
    return load()


def _chk(code):
    from . import check
    return check(code)


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, DeviceVector):
        return a.ptr
    return a.ctypes.data_as(C.c_void_p)


def set_tuning(xcd=0, spmv_blocks=2048, tail_rows=4096):
    """Launch tuning of the sliced-ELL kernels (process-wide)."""
    _chk(_lib().afe_set_tuning(xcd, spmv_blocks, tail_rows))


def mesh_square(n, jitter=0.0, seed=0):
    """Structured unit-square triangulation: (coords, elem_vert, vert_mark)."""
    nv, ne = (n + 1) ** 2, 2 * n * n
    coords = np.empty(2 * nv, dtype=np.float64)
    ev = np.empty(3 * ne, dtype=np.int32)
    vm = np.empty(nv, dtype=np.int32)
    _chk(_lib().afe_mesh_square(n, jitter, seed, _ptr(coords), _ptr(ev), _ptr(vm)))
    return coords, ev, vm


def mesh_cube(n, jitter=0.0, seed=0):
    """Unit cube, n^3 cubes x 6 tetrahedra: (coords, elem_vert, vert_mark)."""
    nv, ne = (n + 1) ** 3, 6 * n ** 3
    coords = np.empty(3 * nv, dtype=np.float64)
    ev = np.empty(4 * ne, dtype=np.int32)
    vm = np.empty(nv, dtype=np.int32)
    _chk(_lib().afe_mesh_cube(n, jitter, seed, _ptr(coords), _ptr(ev), _ptr(vm)))
    return coords, ev, vm


def mesh_displace(dim, coords, vert_mark, amp_h, k):
    out = np.array(coords, dtype=np.float64, copy=True)
    _chk(_lib().afe_mesh_displace(dim, len(vert_mark), _ptr(out), _ptr(vert_mark), amp_h, k))
    return out


class Field:
    """Seeded Fourier field g; a = exp(g).  Keeps the arrays alive."""

    def __init__(self, k, phi, w, scale):
        self.k = np.ascontiguousarray(k, dtype=np.float64)
        self.phi = np.ascontiguousarray(phi, dtype=np.float64)
        self.w = np.ascontiguousarray(w, dtype=np.float64)
        self.scale = float(scale)
        self.n_modes = len(self.phi)
        self.dim = self.k.size // self.n_modes

    def cstruct(self):
        pf = C.POINTER(C.c_double)
        return _sigs.Field(self.n_modes, self.dim, self.k.ctypes.data_as(pf),
                           self.phi.ctypes.data_as(pf), self.w.ctypes.data_as(pf), self.scale)


def field_random(n_modes, dim, amplitude, seed):
    k = np.empty(n_modes * dim)
    phi = np.empty(n_modes)
    w = np.empty(n_modes)
    sc = C.c_double()
    _chk(_lib().afe_field_random(n_modes, dim, amplitude, seed, _ptr(k), _ptr(phi), _ptr(w),
                                 C.byref(sc)))
    return Field(k, phi, w, sc.value)


def _field_ptr(f):
    if f is None:
        return None, None
    s = f.cstruct()
    return C.pointer(s), s


class Context:
    def __init__(self, device=0):
        h = C.c_void_p()
        _chk(_lib().afe_ctx_create(device, C.byref(h)))
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            _lib().afe_ctx_destroy(self.h)
            self.h = None

    def device_bytes(self):
        b = C.c_int64()
        _chk(_lib().afe_ctx_device_bytes(self.h, C.byref(b)))
        return b.value

    def copy_kernel_timing(self, nbytes, reps=20):
        """Average ms of a plain streaming copy of nbytes (read + write)."""
        ms = C.c_double()
        _chk(_lib().afe_copy_kernel_timing(self.h, int(nbytes), reps, C.byref(ms)))
        return ms.value

    def synchronize(self):
        _chk(_lib().afe_ctx_synchronize(self.h))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceVector:
    """A float64 (or other dtype) buffer in HBM owned by a Context."""

    def __init__(self, ctx, n, dtype=np.float64, zero=True):
        self.ctx = ctx
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        _chk(_lib().afe_dev_alloc(ctx.h, self.nbytes, C.byref(p)))
        self.ptr = p
        if zero:
            self.zero()

    @classmethod
    def from_numpy(cls, ctx, a):
        a = np.ascontiguousarray(a)
        v = cls(ctx, a.size, a.dtype, zero=False)
        v.upload(a)
        return v

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        _chk(_lib().afe_dev_upload(self.ctx.h, self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes))

    def numpy(self):
        out = np.empty(self.n, dtype=self.dtype)
        _chk(_lib().afe_dev_download(self.ctx.h, out.ctypes.data_as(C.c_void_p), self.ptr,
                                     self.nbytes))
        return out

    def zero(self):
        _chk(_lib().afe_dev_zero(self.ctx.h, self.ptr, self.nbytes))

    def copy_from(self, other):
        assert other.nbytes == self.nbytes
        _chk(_lib().afe_dev_copy(self.ctx.h, self.ptr, other.ptr, self.nbytes))

    def __len__(self):
        return self.n

    def free(self):
        if self.ptr and self.ctx is not None and self.ctx.h:
            _lib().afe_dev_free(self.ctx.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Space:
    """FE space (DOF numbering + pattern); ctx=None builds it on the host only."""

    def __init__(self, ctx, kind, coords, elem_vert, vert_mark, sides=None):
        """sides: optional (side_vert (n_side x 2), side_mark (n_side)) of the
        mesh; P2 edge DOFs then take their side's boundary mark"""
        self.ctx = ctx
        self.kind = kind
        self.dim = 3 if kind == P1_TET else 2
        self.coords = np.ascontiguousarray(coords, dtype=np.float64)
        self.elem_vert = np.ascontiguousarray(elem_vert, dtype=np.int32)
        self.vert_mark = np.ascontiguousarray(vert_mark, dtype=np.int32)
        nv = len(self.vert_mark)
        ne = len(self.elem_vert) // (4 if kind == P1_TET else 3)
        h = C.c_void_p()
        if sides is None:
            _chk(_lib().afe_space_create(ctx.h if ctx else None, kind, nv, _ptr(self.coords), ne,
                                         _ptr(self.elem_vert), _ptr(self.vert_mark), C.byref(h)))
        else:
            sv = np.ascontiguousarray(sides[0], dtype=np.int32).ravel()
            sm = np.ascontiguousarray(sides[1], dtype=np.int32)
            _chk(_lib().afe_space_create_sides(ctx.h if ctx else None, kind, nv, _ptr(self.coords), ne,
                                               _ptr(self.elem_vert), _ptr(self.vert_mark), len(sm), _ptr(sv),
                                               _ptr(sm), C.byref(h)))
        self.h = h
        self.n_vert, self.n_elem = nv, ne
        nd, nz, nl = C.c_int64(), C.c_int64(), C.c_int32()
        _chk(_lib().afe_space_info(h, C.byref(nd), C.byref(nz), C.byref(nl)))
        self.n_dof, self.nnz, self.n_loc = nd.value, nz.value, nl.value

    def dofs(self):
        ed = np.empty(self.n_elem * self.n_loc, dtype=np.int32)
        dm = np.empty(self.n_dof, dtype=np.int32)
        dp = np.empty(self.n_dof * self.dim, dtype=np.float64)
        _chk(_lib().afe_space_dofs(self.h, _ptr(ed), _ptr(dm), _ptr(dp)))
        return ed, dm, dp

    def pattern(self):
        rp = np.empty(self.n_dof + 1, dtype=np.int64)
        col = np.empty(self.nnz, dtype=np.int64)
        _chk(_lib().afe_space_pattern(self.h, _ptr(rp), _ptr(col)))
        return rp, col

    def update_coords(self, coords):
        self.coords = np.ascontiguousarray(coords, dtype=np.float64)
        _chk(_lib().afe_space_update_coords(self.h, _ptr(self.coords)))

    def clone(self, coords=None):
        """a space of the same connectivity (numbering, pattern and gather map
        copied) with new vertex coordinates"""
        c = self.coords if coords is None else np.ascontiguousarray(coords, dtype=np.float64)
        h = C.c_void_p()
        _chk(_lib().afe_space_clone(self.h, _ptr(c), C.byref(h)))
        out = Space.__new__(Space)
        out.__dict__.update(self.__dict__)
        out.coords = c
        out.h = h
        return out

    def rhs(self, acc, fkind, field=None, nodal=None, out=None):
        if out is None:
            out = DeviceVector(self.ctx, self.n_dof, zero=False)
        if nodal is not None and not isinstance(nodal, DeviceVector):
            nodal = DeviceVector.from_numpy(self.ctx, np.asarray(nodal, dtype=np.float64))
        fp, _keep = _field_ptr(field)
        _chk(_lib().afe_rhs(self.h, acc, fkind, fp, _ptr(nodal), _ptr(out)))
        return out

    def qpoints(self, acc):
        """Physical quadrature points (n_elem, n_qp, dim) and Jxw (n_elem, n_qp)."""
        nq = C.c_int64()
        _chk(_lib().afe_space_qpoints(self.h, acc, C.byref(nq), None, None))
        pts = np.empty(self.n_elem * nq.value * self.dim, dtype=np.float64)
        jxw = np.empty(self.n_elem * nq.value, dtype=np.float64)
        _chk(_lib().afe_space_qpoints(self.h, acc, C.byref(nq), _ptr(pts), _ptr(jxw)))
        return pts.reshape(self.n_elem, nq.value, self.dim), jxw.reshape(self.n_elem, nq.value)

    def rhs_qp(self, acc, fq, out=None):
        """L2Discretize of f given at the quadrature points (n_elem, n_qp)."""
        if out is None:
            out = DeviceVector(self.ctx, self.n_dof, zero=False)
        fq = np.ascontiguousarray(fq, dtype=np.float64)
        _chk(_lib().afe_rhs_qp(self.h, acc, _ptr(fq), _ptr(out)))
        return out

    def error_qp(self, uh, acc, exact, h1=False):
        """L2 error (h1=False, exact values) or H1 semi-error (exact gradients)."""
        e = C.c_double()
        ex = np.ascontiguousarray(exact, dtype=np.float64)
        _chk(_lib().afe_error_qp(self.h, _ptr(uh), acc, 1 if h1 else 0, _ptr(ex), C.byref(e)))
        return e.value

    def error_l2(self, uh, acc):
        e = C.c_double()
        _chk(_lib().afe_error_l2(self.h, _ptr(uh), acc, C.byref(e)))
        return e.value

    def error_h1(self, uh, acc):
        e = C.c_double()
        _chk(_lib().afe_error_h1(self.h, _ptr(uh), acc, C.byref(e)))
        return e.value

    def close(self):
        if self.h:
            _lib().afe_space_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Matrix:
    def __init__(self, space=None, ctx=None, csr=None, space1=None):
        h = C.c_void_p()
        if space is not None and space1 is not None:
            # BilinearOperator(sp0, sp1): two spaces on one mesh
            _chk(_lib().afe_matrix_create_pair(space.h, space1.h, C.byref(h)))
            m, n, nnz, mc = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int32()
            _chk(_lib().afe_matrix_info(h, C.byref(m), C.byref(n), C.byref(nnz), C.byref(mc)))
            self.m, self.n, self.nnz, self.max_coupling = m.value, n.value, nnz.value, mc.value
        elif space is not None:
            _chk(_lib().afe_matrix_create(space.h, C.byref(h)))
            self.n, self.nnz = space.n_dof, space.nnz
            self.m = self.n
        else:
            rp, col, val = csr
            rp = np.ascontiguousarray(rp, dtype=np.int64)
            col = np.ascontiguousarray(col, dtype=np.int64)
            val = np.ascontiguousarray(val, dtype=np.float64)
            _chk(_lib().afe_matrix_create_csr(ctx.h, len(rp) - 1, _ptr(rp), _ptr(col), _ptr(val),
                                              C.byref(h)))
            self.n, self.nnz = len(rp) - 1, int(rp[-1])
            self.m = self.n
        self.space = space
        self.space1 = space1
        self.h = h

    def pattern(self):
        """(row_ptr, col) of the compressed pattern"""
        rp = np.empty(self.m + 1, dtype=np.int64)
        col = np.empty(self.nnz, dtype=np.int64)
        _chk(_lib().afe_matrix_pattern(self.h, _ptr(rp), _ptr(col)))
        return rp, col

    def assemble(self, op, acc, field=None):
        fp, _keep = _field_ptr(field)
        _chk(_lib().afe_assemble(self.h, op, acc, fp))

    def assemble_elements(self, elem_mats):
        em = np.ascontiguousarray(elem_mats, dtype=np.float64)
        _chk(_lib().afe_assemble_elements(self.h, _ptr(em)))

    def values(self):
        v = np.empty(self.nnz, dtype=np.float64)
        _chk(_lib().afe_matrix_values(self.h, _ptr(v)))
        return v

    def set_values(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        _chk(_lib().afe_matrix_set_values(self.h, _ptr(v)))

    def apply_dirichlet(self, f, u=None):
        _chk(_lib().afe_apply_dirichlet(self.h, _ptr(u), _ptr(f)))

    def apply_dirichlet_values(self, marks, g, f, u=None, preserve_symmetry=True):
        """Dirichlet elimination on the DOFs whose mark is in marks, data g per
        DOF (BoundaryConditionAdmin::apply; without preserve_symmetry only the
        marked rows change)."""
        m = np.ascontiguousarray(marks, dtype=np.int32)
        gg = np.ascontiguousarray(g, dtype=np.float64)
        _chk(_lib().afe_apply_dirichlet_bc(self.h, _ptr(m), len(m), _ptr(gg), _ptr(u), _ptr(f),
                                           1 if preserve_symmetry else 0))

    def vmult(self, x, y):
        _chk(_lib().afe_matrix_vmult(self.h, _ptr(x), _ptr(y)))

    def tvmult(self, x, y):
        _chk(_lib().afe_matrix_tvmult(self.h, _ptr(x), _ptr(y)))

    def assemble_timing(self, op, acc, field=None, reps=10):
        """(element kernel ms, gather kernel ms, algorithmic bytes) of afe_assemble"""
        fp, _keep = _field_ptr(field)
        te, tg, by = C.c_double(), C.c_double(), C.c_double()
        _chk(_lib().afe_assemble_kernel_timing(self.h, op, acc, fp, reps, C.byref(te), C.byref(tg),
                                               C.byref(by)))
        return te.value, tg.value, by.value

    def assemble_pipelined_ms(self, op, acc, field=None, reps=10):
        """device ms of afe_assemble as it runs (element chunks overlapping the gathers)"""
        fp, _keep = _field_ptr(field)
        ms = C.c_double()
        _chk(_lib().afe_assemble_timing(self.h, op, acc, fp, reps, C.byref(ms)))
        return ms.value

    def close(self):
        if self.h:
            _lib().afe_matrix_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def AmgParams(alpha=0.25, nmpo=50, nmpsd=0.382, lazy=0, smooth=3, build_fp32=1):
    return _sigs.AmgParams(alpha, nmpo, nmpsd, lazy, smooth, build_fp32)


class _AmgBase:
    def info(self):
        L = C.c_int32()
        n = np.empty(128, dtype=np.int64)
        nz = np.empty(128, dtype=np.int64)
        full = C.c_int32()
        _chk(_lib().afe_amg_info(self.h, C.byref(L), _ptr(n), _ptr(nz), C.byref(full)))
        return {"levels": L.value, "n": n[:L.value].tolist(), "nnz": nz[:L.value].tolist(),
                "coarse_full": bool(full.value)}

    def level_shape(self, level, which=0):
        """(m, n, nnz) of A_l (which 0) or P_l (which 1) without copying it"""
        m, n, nnz = C.c_int64(), C.c_int64(), C.c_int64()
        _chk(_lib().afe_amg_level_csr(self.h, level, which, C.byref(m), C.byref(n), C.byref(nnz),
                                      None, None, None))
        return m.value, n.value, nnz.value

    def level_csr(self, level, which=0):
        m, n, nnz = C.c_int64(), C.c_int64(), C.c_int64()
        _chk(_lib().afe_amg_level_csr(self.h, level, which, C.byref(m), C.byref(n), C.byref(nnz),
                                      None, None, None))
        rp = np.empty(m.value + 1, dtype=np.int64)
        col = np.empty(nnz.value, dtype=np.int64)
        val = np.empty(nnz.value, dtype=np.float64)
        _chk(_lib().afe_amg_level_csr(self.h, level, which, C.byref(m), C.byref(n), C.byref(nnz),
                                      _ptr(rp), _ptr(col), _ptr(val)))
        return m.value, n.value, rp, col, val

    def level_colours(self, level):
        """Multicolour Gauss-Seidel colour of every row of a device level."""
        n = self.info()["n"][level]
        out = np.empty(n, dtype=np.int32)
        _chk(_lib().afe_amg_level_colours(self.h, level, _ptr(out)))
        return out

    def coarse_inverse(self):
        info = self.info()
        nc = info["n"][-1]
        out = np.empty(nc * nc, dtype=np.float64)
        _chk(_lib().afe_amg_coarse_inverse(self.h, _ptr(out)))
        return out.reshape(nc, nc)

    def close(self):
        if self.h:
            _lib().afe_amg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HostAmg(_AmgBase):
    """Host-only hierarchy of a CSR matrix (no device)."""

    def __init__(self, rp, col, val, params=None):
        rp = np.ascontiguousarray(rp, dtype=np.int64)
        col = np.ascontiguousarray(col, dtype=np.int64)
        val = np.ascontiguousarray(val, dtype=np.float64)
        h = C.c_void_p()
        prm = params if params is not None else AmgParams(build_fp32=0)
        _chk(_lib().afe_amg_create_host(len(rp) - 1, _ptr(rp), _ptr(col), _ptr(val), C.byref(prm),
                                        C.byref(h)))
        self.h = h

    def lazy_reinit(self, val):
        val = np.ascontiguousarray(val, dtype=np.float64)
        _chk(_lib().afe_amg_lazy_reinit_host(self.h, _ptr(val)))


class Amg(_AmgBase):
    def __init__(self, matrix, params=None):
        h = C.c_void_p()
        prm = params if params is not None else AmgParams()
        _chk(_lib().afe_amg_create(matrix.h, C.byref(prm), C.byref(h)))
        self.h = h
        self.matrix = matrix

    def lazy_reinit(self, matrix):
        _chk(_lib().afe_amg_lazy_reinit(self.h, matrix.h))
        self.matrix = matrix

    def clone(self, matrix):
        """the hierarchy's structure for another matrix of its pattern, with
        that matrix's Galerkin products (equal to a fresh lazy build)"""
        h = C.c_void_p()
        _chk(_lib().afe_amg_clone(self.h, matrix.h, C.byref(h)))
        out = Amg.__new__(Amg)
        out.h = h
        out.matrix = matrix
        return out

    def solve(self, x, b, tol=0.0, step=20, mode=1, rtol=0.0):
        st = _sigs.SolveStats()
        _chk(_lib().afe_amg_solve(self.h, _ptr(x), _ptr(b), tol, step, mode, rtol, C.byref(st)))
        return st

    def pcg(self, x, b, rtol=1e-10, max_it=500, prec=PREC_FP32, check_conv=True):
        st = _sigs.SolveStats()
        code = _lib().afe_amg_pcg(self.h, _ptr(x), _ptr(b), rtol, max_it, prec, C.byref(st))
        if check_conv or code not in (0, -5):
            _chk(code)
        return st

    def vcycle(self, z, r, prec=PREC_FP64):
        _chk(_lib().afe_amg_vcycle(self.h, _ptr(z), _ptr(r), prec))

    def precondition(self, x, b, iterate_step=5):
        _chk(_lib().afe_amg_precondition(self.h, _ptr(x), _ptr(b), iterate_step))

    def set_smooth(self, s):
        _chk(_lib().afe_amg_set_smooth(self.h, s))

    def set_coarse_exact(self, exact):
        """AMGSolver::isSolveMostProjectExactly(): False smooths the coarsest
        level with Gauss-Seidel sweeps instead of solving it exactly"""
        _chk(_lib().afe_amg_set_coarse_exact(self.h, 1 if exact else 0))

    def live_timing(self, enable):
        """Start (True) or stop (False) live timing of the CG's level-0 SpMV;
        returns (average ms so far, launches)."""
        ms, n = C.c_double(), C.c_int64()
        _chk(_lib().afe_amg_live_timing(self.h, 1 if enable else 0, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def part_timing(self, enable):
        """Start (True) or stop (False) graphs-on timing of the two parts of
        every CG iteration; returns (part 0 ms, part 1 ms, iterations timed)"""
        a, b, n = C.c_double(), C.c_double(), C.c_int64()
        _chk(_lib().afe_amg_part_timing(self.h, 1 if enable else 0, C.byref(a), C.byref(b), C.byref(n)))
        return a.value, b.value, n.value

    def set_tail_rows(self, rows):
        _chk(_lib().afe_amg_set_tail_rows(self.h, int(rows)))

    def kernel_timing(self, which, reps=20):
        ms, by = C.c_double(), C.c_double()
        _chk(_lib().afe_amg_kernel_timing(self.h, which, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def device_info(self):
        out = []
        for l in range(self.info()["levels"]):
            r, c, s, t, g = C.c_int64(), C.c_int32(), C.c_int64(), C.c_int32(), C.c_int32()
            _chk(_lib().afe_amg_device_info(self.h, l, C.byref(r), C.byref(c), C.byref(s), C.byref(t),
                                            C.byref(g)))
            out.append({"rows": r.value, "colours": c.value, "stored": s.value, "tail": t.value,
                        "row_lanes": g.value})
        return out


class Batch:
    """Continuous batching of FP64 CG solves preconditioned by one V-cycle
    (FP32 or FP64) per iteration; every step runs one iteration of every
    active member with shared launches."""

    def __init__(self, ctx, prec=PREC_FP32, smooth=1):
        self.ctx = ctx
        h = C.c_void_p()
        _chk(_lib().afe_batch_create(ctx.h, prec, smooth, C.byref(h)))
        self.h = h
        self._keep = {}

    def add(self, amg, x, b, rtol=1e-10, max_it=1000):
        slot = C.c_int32()
        _chk(_lib().afe_batch_add(self.h, amg.h, _ptr(x), _ptr(b), rtol, max_it, C.byref(slot)))
        self._keep[slot.value] = (amg, x, b)
        return slot.value

    def step(self):
        n = C.c_int32()
        _chk(_lib().afe_batch_step(self.h, C.byref(n)))
        return n.value

    def run(self):
        _chk(_lib().afe_batch_run(self.h))

    def result(self, slot):
        """(done, stats, status code)"""
        done = C.c_int32()
        st = _sigs.SolveStats()
        code = _lib().afe_batch_result(self.h, slot, C.byref(done), C.byref(st))
        if code not in (0, -5):
            _chk(code)
        return bool(done.value), st, code

    def remove(self, slot):
        _chk(_lib().afe_batch_remove(self.h, slot))
        self._keep.pop(slot, None)

    def set_tiles(self, tile_level0, tile_coarse):
        """slices per wavefront of the batched level launches (1, 2, 4, 8)"""
        _chk(_lib().afe_batch_set_tiles(self.h, tile_level0, tile_coarse))

    def close(self):
        if self.h:
            _lib().afe_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Solve server: set-up thread pool, HBM mesh/hierarchy cache, one
    continuously running batch (FP64 CG + FP32 V-cycle)."""

    def __init__(self, device=0, hbm_budget=200 << 30, max_slots=256, setup_threads=8, smooth=1):
        h = C.c_void_p()
        _chk(_lib().afe_server_create(device, int(hbm_budget), max_slots, setup_threads, smooth, C.byref(h)))
        self.h = h

    def submit(self, mesh_key, kind, n=0, jitter=0.0, seed=0, move_amp=0.0, move_k=1, acc=4,
               op=OP_STIFF, fkind=F_POISSON, field=None, rtol=1e-10, max_it=1000, keep_x=False,
               mesh=None):
        """a structured mesh (kind, n, jitter, seed), or mesh=(coords,
        elem_vert, vert_mark) of the caller's (copied by the call)"""
        r = _sigs.Request()
        r.mesh_key, r.kind, r.n, r.jitter, r.seed = mesh_key, kind, n, jitter, seed
        r.move_amp, r.move_k, r.acc, r.op, r.fkind = move_amp, move_k, acc, op, fkind
        r.rtol, r.max_it, r.keep_x = rtol, max_it, 1 if keep_x else 0
        if field is not None:
            r.coef = field.cstruct()  # the field's arrays are copied by the call
        if mesh is not None:
            c = np.ascontiguousarray(mesh[0], dtype=np.float64)
            ev = np.ascontiguousarray(mesh[1], dtype=np.int32)
            vm = np.ascontiguousarray(mesh[2], dtype=np.int32)
            r.n_vert, r.n_elem = len(vm), len(ev) // (4 if kind == P1_TET else 3)
            r.coords, r.elem_vert, r.vert_mark = _ptr(c), _ptr(ev), _ptr(vm)
        rid = C.c_int64()
        _chk(_lib().afe_server_submit(self.h, C.byref(r), C.byref(rid)))
        return rid.value

    def poll(self, rid):
        d = C.c_int32()
        _chk(_lib().afe_server_poll(self.h, rid, C.byref(d)))
        return bool(d.value)

    def wait(self, rid, n_dof=None):
        """(solution or None, stats, latency seconds, status)"""
        nd = C.c_int64()
        st = _sigs.SolveStats()
        lat = C.c_double()
        x = np.empty(n_dof) if n_dof else None
        code = _lib().afe_server_wait(self.h, rid, _ptr(x), C.byref(nd), C.byref(st), C.byref(lat))
        if code not in (0, -5):
            _chk(code)
        return x, st, lat.value, code

    def release(self, rid):
        _chk(_lib().afe_server_release(self.h, rid))

    def info(self):
        b, c, d, e = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        _chk(_lib().afe_server_info(self.h, C.byref(b), C.byref(c), C.byref(d), C.byref(e)))
        return {"device_bytes": b.value, "cached": c.value, "completed": d.value, "evictions": e.value}

    def timing(self):
        t = (C.c_double * 13)()
        _chk(_lib().afe_server_timing_n(self.h, t, 13))
        keys = ["setup_new_s", "setup_moving_s", "wait_setup_s", "wait_slot_s", "solve_s", "n_new", "n_moving",
                "setup_cloned_s", "n_cloned", "wait_room_s", "setup_cpu_user_s", "setup_cpu_sys_s",
                "setup_lock_wait_s"]
        return dict(zip(keys, list(t)))

    def close(self):
        if self.h:
            _lib().afe_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
