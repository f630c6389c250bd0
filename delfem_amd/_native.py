"""ctypes binding of libdfm_gpu.so (include/dfm_gpu.h). Thin: every call goes straight to the C-ABI.

The product runs only through the HIP library; there is no CPU fallback. Loading fails loudly when the
library has not been built (python -c "import __graft_entry__; __graft_entry__.build()")."""
import atexit
import ctypes as C
import os
import weakref

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libdfm_gpu.so")

POISSON, SOLID2D, SOLID3D, STVK2D, STVK3D, DIFFUSION2D, SOLID2D_NEWMARK, SOLID3D_NEWMARK, SOLID2D_BEULER, STVK3D_NEWMARK, STVK2D_NEWMARK, STVK2D_BEULER = range(12)
ORDER_NATURAL, ORDER_MULTICOLOR, ORDER_BLOCKCOLOR, ORDER_RCM, ORDER_RCM2, ORDER_AMD = range(6)
FP64, MIXED, FP32_ILU, MIXED_AUTO, FP64_TRUE = 0, 1, 2, 3, 4
ERR_SINGULAR = -5
NPAR = {POISSON: 2, SOLID2D: 5, SOLID3D: 6, STVK2D: 5, STVK3D: 6, DIFFUSION2D: 3, SOLID2D_NEWMARK: 5, SOLID3D_NEWMARK: 6, SOLID2D_BEULER: 5, STVK3D_NEWMARK: 6, STVK2D_NEWMARK: 5, STVK2D_BEULER: 5}

_lib = None
_live = weakref.WeakSet()  # systems and contexts still open, closed in order at interpreter exit


@atexit.register
def _close_all():
    objs = list(_live)
    for o in objs:
        if isinstance(o, (Batch, Server)):
            o.close()
    for o in objs:
        if isinstance(o, (System, MultiSegSystem)):
            o.close()
    for o in objs:
        if isinstance(o, Context):
            o.close()


class DfmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"dfm error {code}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
        _lib = C.CDLL(LIB_PATH)
        _lib.dfm_last_error.restype = C.c_char_p
    return _lib


def _check(rc):
    if rc < 0:
        raise DfmError(rc, lib().dfm_last_error().decode())
    return rc


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


def device_bytes():
    b = C.c_longlong(0)
    _check(lib().dfm_device_bytes(C.byref(b)))
    return b.value


def device_count():
    c = C.c_int(0)
    _check(lib().dfm_device_count(C.byref(c)))
    return c.value


# ---- seeded meshes -----------------------------------------------------------------------------------------
def rect_tri(nx, ny, lx, ly, seed, jitter=0.1):
    coords = np.empty((nx + 1) * (ny + 1) * 2)
    tri = np.empty(nx * ny * 6, dtype=np.int32)
    _check(lib().dfm_mesh_rect_tri(nx, ny, C.c_double(lx), C.c_double(ly), C.c_ulonglong(seed),
                                   C.c_double(jitter), _p(coords), _p(tri)))
    return coords, tri


def box_tet(nx, ny, nz, lx, ly, lz, seed, jitter=0.1):
    coords = np.empty((nx + 1) * (ny + 1) * (nz + 1) * 3)
    tet = np.empty(nx * ny * nz * 24, dtype=np.int32)
    _check(lib().dfm_mesh_box_tet(nx, ny, nz, C.c_double(lx), C.c_double(ly), C.c_double(lz), C.c_ulonglong(seed),
                                  C.c_double(jitter), _p(coords), _p(tet)))
    return coords, tet


def rng_uniform(seed, stream, n):
    out = np.empty(n)
    _check(lib().dfm_rng_uniform(C.c_ulonglong(seed), C.c_ulonglong(stream), C.c_size_t(n), _p(out)))
    return out


def pattern_ordering(kind, rowptr, colind):
    """COrdering_Blk's ordering (ORDER_RCM / ORDER_RCM2 / ORDER_AMD) of a block pattern, new -> old (host only)."""
    rp = np.ascontiguousarray(rowptr, dtype=np.uint32)
    ci = np.ascontiguousarray(colind, dtype=np.uint32)
    n = rp.size - 1
    out = np.zeros(max(n, 1), np.int32)
    _check(lib().dfm_pattern_ordering(kind, n, _p(rp), _p(ci), _p(out)))
    return out[:n]


def read_gid(path):
    """GiD .msh of P1 triangles / tetrahedra (dfm_mesh_read_gid): (ndim, coords[nnode*ndim], lnods[nelem*npoel])."""
    nd, nn, ne, npe = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    p = path.encode()
    _check(lib().dfm_mesh_read_gid(p, C.byref(nd), C.byref(nn), C.byref(ne), C.byref(npe), None, None))
    co = np.zeros(nn.value * nd.value)
    ln = np.zeros(ne.value * npe.value, np.int32)
    _check(lib().dfm_mesh_read_gid(p, C.byref(nd), C.byref(nn), C.byref(ne), C.byref(npe), _p(co), _p(ln)))
    return nd.value, co, ln


def rng_normal(seed, stream, n):
    out = np.empty(n)
    _check(lib().dfm_rng_normal(C.c_ulonglong(seed), C.c_ulonglong(stream), C.c_size_t(n), _p(out)))
    return out


# ---- context / system ----------------------------------------------------------------------------------------
class Context:
    def __init__(self, device=0):
        self.h = C.c_void_p()
        _check(lib().dfm_ctx_create(device, C.byref(self.h)))
        _live.add(self)

    def synchronize(self):
        _check(lib().dfm_ctx_synchronize(self.h))

    def close(self):
        if self.h:
            lib().dfm_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class System:
    """Device linear system: CLinearSystem_Field + CPreconditioner_ILU(0) on one GPU."""

    def __init__(self, ctx, ndim, coords, lnods, length):
        self.ctx = ctx
        self.ndim, self.len = ndim, length
        coords, lnods = _f64(coords), _i32(lnods)
        self.nnode = coords.size // ndim
        self.nelem = lnods.size // (ndim + 1)
        self.h = C.c_void_p()
        _check(lib().dfm_sys_create_mesh(ctx.h, ndim, self.nnode, _p(coords), self.nelem, _p(lnods), length,
                                         C.byref(self.h)))
        n, ln, nc = C.c_int(), C.c_int(), C.c_longlong()
        _check(lib().dfm_sys_info(self.h, C.byref(n), C.byref(ln), C.byref(nc)))
        self.ncrs = nc.value
        _live.add(self)

    @classmethod
    def from_pattern(cls, ctx, nblk, length, rowptr, colind):
        """A system on a host block pattern without a mesh (dfm_sys_create_pattern); values from set_matrix."""
        c = cls.__new__(cls)
        rowptr, colind = np.ascontiguousarray(rowptr, np.uint32), np.ascontiguousarray(colind, np.uint32)
        c.ctx, c.ndim, c.len, c.nnode, c.nelem = ctx, 0, length, nblk, 0
        c.h = C.c_void_p()
        _check(lib().dfm_sys_create_pattern(ctx.h, nblk, length, rowptr.size - 1, _p(rowptr), _p(colind),
                                            C.byref(c.h)))
        n, ln, nc = C.c_int(), C.c_int(), C.c_longlong()
        _check(lib().dfm_sys_info(c.h, C.byref(n), C.byref(ln), C.byref(nc)))
        c.ncrs = nc.value
        _live.add(c)
        return c

    def set_matrix(self, crs, dia):
        """Matrix values from the host in get_matrix's layout (dfm_sys_set_matrix)."""
        _check(lib().dfm_sys_set_matrix(self.h, _p(_f64(crs)), _p(_f64(dia))))

    def clone(self):
        """A new system on the same mesh, ordering and Dirichlet flags (dfm_sys_clone)."""
        c = System.__new__(System)
        c.ctx, c.ndim, c.len, c.nnode, c.nelem, c.ncrs = self.ctx, self.ndim, self.len, self.nnode, self.nelem, self.ncrs
        c.h = C.c_void_p()
        _check(lib().dfm_sys_clone(self.h, C.byref(c.h)))
        _live.add(c)
        return c

    def close(self):
        if self.h:
            lib().dfm_sys_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def ndof(self):
        return self.nnode * self.len

    def pattern(self):
        rp = np.empty(self.nnode + 1, np.uint32)
        ci = np.empty(max(self.ncrs, 1), np.uint32)
        _check(lib().dfm_sys_get_pattern(self.h, _p(rp), _p(ci)))
        return rp, ci[: self.ncrs]

    def make_ordering(self, kind, tile=0):
        order = np.empty(self.nnode, np.int32)
        nc = C.c_int(0)
        _check(lib().dfm_sys_make_ordering(self.h, kind, tile, _p(order), C.byref(nc)))
        return order, nc.value

    def set_ordering(self, order):
        _check(lib().dfm_sys_set_ordering(self.h, _p(_i32(order))))

    def schedule_info(self):
        """(forward phases, backward phases, most levels in a group, sweep launches per PCG iteration)"""
        a, b, c, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _check(lib().dfm_sys_schedule_info(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        return a.value, b.value, c.value, d.value

    def set_bc(self, flags):
        _check(lib().dfm_sys_set_bc(self.h, _p(_i32(flags))))

    def set_field(self, u=None, v=None):
        _check(lib().dfm_sys_set_field(self.h, _p(_f64(u)), _p(_f64(v))))

    def get_field(self):
        u = np.empty(self.ndof)
        _check(lib().dfm_sys_get_field(self.h, _p(u)))
        return u

    def set_acceleration(self, a=None):
        _check(lib().dfm_sys_set_acceleration(self.h, _p(_f64(a))))

    def get_fields(self):
        """(u, v, a): VALUE, VELOCITY and ACCELERATION in the original numbering."""
        out = [np.empty(self.nnode * self.len) for _ in range(3)]
        _check(lib().dfm_sys_get_fields(self.h, *[_p(o) for o in out]))
        return out

    def update_field_newmark(self, gamma, beta, dt, is_initial=True):
        """CLinearSystem_Field::UpdateValueOfField_NewmarkBeta with the update as the acceleration increment."""
        _check(lib().dfm_sys_update_field_newmark(self.h, C.c_double(gamma), C.c_double(beta), C.c_double(dt),
                                                  1 if is_initial else 0))

    def set_velocity_prev(self, v0=None):
        _check(lib().dfm_sys_set_velocity_prev(self.h, _p(_f64(v0))))

    def update_field_beuler(self, dt, is_initial=True):
        """CLinearSystem_Field::UpdateValueOfField_BackwardEular with the update as the velocity increment."""
        _check(lib().dfm_sys_update_field_beuler(self.h, C.c_double(dt), 1 if is_initial else 0))

    def update_field(self):
        _check(lib().dfm_sys_update_field(self.h))

    def assemble(self, kind, params, glob=None, faces=None, trac=None):
        params = _f64(params)
        if params.ndim == 1:
            npar, pstride = params.size, 0
        else:
            npar, pstride = params.shape[1], params.shape[1]
        glob = _f64(glob) if glob is not None else None
        faces, trac = _i32(faces), _f64(trac)
        nface = 0 if faces is None else faces.size // 3
        rn = C.c_double(0.0)
        _check(lib().dfm_sys_assemble(self.h, kind, _p(params), npar, pstride, _p(glob), nface, _p(faces), _p(trac),
                                      C.byref(rn)))
        return rn.value

    def set_params(self, params):
        params = _f64(params)
        npar, pstride = (params.size, 0) if params.ndim == 1 else (params.shape[1], params.shape[1])
        _check(lib().dfm_sys_set_params(self.h, _p(params), npar, pstride))

    def set_loads(self, faces=None, trac=None):
        faces, trac = _i32(faces), _f64(trac)
        nface = 0 if faces is None else faces.size // 3
        _check(lib().dfm_sys_set_loads(self.h, nface, _p(faces), _p(trac)))

    def assemble_resident(self, kind, glob=None, norm=True):
        glob = _f64(glob) if glob is not None else None
        rn = C.c_double(0.0)
        _check(lib().dfm_sys_assemble_resident(self.h, kind, _p(glob), C.byref(rn) if norm else None))
        return rn.value

    def bench_kernel(self, which, reps=20):
        ms, nbytes = C.c_double(), C.c_double()
        _check(lib().dfm_sys_bench_kernel(self.h, which, reps, C.byref(ms), C.byref(nbytes)))
        return ms.value, nbytes.value

    def get_matrix(self):
        bs = self.len * self.len
        crs = np.empty(max(self.ncrs, 1) * bs)
        dia = np.empty(self.nnode * bs)
        _check(lib().dfm_sys_get_matrix(self.h, _p(crs), _p(dia)))
        return crs[: self.ncrs * bs], dia

    # saved stiffness (CLinearSystem_Save)
    def assemble_save(self, kind, params):
        """InitializeMarge + *_SaveStiffMat + FinalizeMarge of CLinearSystem_Save: matrix, force, boundary
        coupling (include/delfem/femls/linearsystem_fieldsave.h:59-106). Set the Dirichlet flags first."""
        params = _f64(params)
        npar, pstride = (params.size, 0) if params.ndim == 1 else (params.shape[1], params.shape[1])
        _check(lib().dfm_sys_assemble_save(self.h, kind, _p(params), npar, pstride))

    def assemble_savedia(self, kind, params, glob):
        """InitializeMarge + element loop + FinalizeMarge of the lumped-mass saved systems: kind DIFFUSION2D
        (CLinearSystem_SaveDiaM_Newmark, params (rho, alpha, source), glob (dt, gamma)) or SOLID2D
        (CLinearSystem_SaveDiaM_NewmarkBeta, (lambda, myu, rho, gx, gy), glob (dt, gamma, beta)). Set the Dirichlet
        flags first; make_residual and update_field_save then follow these classes."""
        params = _f64(params)
        npar, pstride = (params.size, 0) if params.ndim == 1 else (params.shape[1], params.shape[1])
        g = np.zeros(3)
        g[: len(glob)] = glob
        _check(lib().dfm_sys_assemble_save_dia(self.h, kind, _p(params), npar, pstride, _p(g)))

    def make_residual(self):
        """MakeResidual of the saved systems: residual = f - K_b w, update = w (w = u_field for CLinearSystem_Save,
        the Newmark predictor for the lumped-mass ones); returns ||residual||."""
        rn = C.c_double(0.0)
        _check(lib().dfm_sys_make_residual(self.h, C.byref(rn)))
        return rn.value

    def update_field_save(self):
        """UpdateValueOfField of the saved systems: CLinearSystem_Save (VALUE): u_field = update, Dirichlet values
        kept; the lumped-mass systems: the Newmark update of u (and v), velocity / acceleration = update."""
        _check(lib().dfm_sys_update_field_save(self.h))

    def get_force(self):
        out = np.empty(self.nnode * self.len)
        _check(lib().dfm_sys_get_force(self.h, _p(out)))
        return out

    def boundary_terms(self):
        n = C.c_longlong(0)
        _check(lib().dfm_sys_boundary_terms(self.h, C.byref(n)))
        return n.value

    # handle ops (ILinearSystemPreconditioner_Sol)
    def resize_tmp(self, n):
        _check(lib().dfm_sys_resize_tmp(self.h, n))

    def dot(self, i, j):
        out = C.c_double()
        _check(lib().dfm_sys_dot(self.h, i, j, C.byref(out)))
        return out.value

    def copy(self, i, j):
        _check(lib().dfm_sys_copy(self.h, i, j))

    def scal(self, a, i):
        _check(lib().dfm_sys_scal(self.h, C.c_double(a), i))

    def axpy(self, a, i, j):
        _check(lib().dfm_sys_axpy(self.h, C.c_double(a), i, j))

    def matvec(self, a, i, b, j):
        _check(lib().dfm_sys_matvec(self.h, C.c_double(a), i, C.c_double(b), j))

    def get_vector(self, iv):
        out = np.empty(self.ndof)
        _check(lib().dfm_sys_get_vector(self.h, iv, _p(out)))
        return out

    def set_vector(self, iv, vals):
        _check(lib().dfm_sys_set_vector(self.h, iv, _p(_f64(vals))))

    # preconditioner
    def factor(self):
        _check(lib().dfm_prec_factor(self.h))

    def prec_apply(self, iv):
        _check(lib().dfm_prec_apply(self.h, iv))

    def set_fill_level(self, lev):
        """ILU(lev) (CPreconditioner_ILU::SetFillInLevel): 0 ILU(0), -1 no limit; factor again afterwards."""
        _check(lib().dfm_prec_set_fill_level(self.h, int(lev)))

    def factor_size(self):
        n = C.c_longlong()
        _check(lib().dfm_prec_factor_size(self.h, C.byref(n)))
        return n.value

    def get_factor(self):
        bs = self.len * self.len
        nf = self.factor_size()
        f = {"frowptr": np.empty(self.nnode + 1, np.uint32), "fcolind": np.empty(max(nf, 1), np.uint32),
             "diaind": np.empty(self.nnode, np.uint32), "fcrs": np.empty(max(nf, 1) * bs),
             "fdia": np.empty(self.nnode * bs)}
        _check(lib().dfm_prec_get_factor(self.h, _p(f["frowptr"]), _p(f["fcolind"]), _p(f["diaind"]), _p(f["fcrs"]),
                                         _p(f["fdia"])))
        f["fcolind"] = f["fcolind"][:nf]
        f["fcrs"] = f["fcrs"][: nf * bs]
        return f

    # solvers
    def solve_pcg(self, tol, max_iter, precision=FP64):
        conv = C.c_double(tol)
        it = C.c_uint(max_iter)
        ok = _check(lib().dfm_solve_pcg(self.h, precision, C.byref(conv), C.byref(it)))
        return bool(ok), conv.value, it.value

    def solve_pbicgstab(self, tol, max_iter, precision=FP64):
        conv = C.c_double(tol)
        it = C.c_uint(max_iter)
        ok = _check(lib().dfm_solve_pbicgstab(self.h, precision, C.byref(conv), C.byref(it)))
        return bool(ok), conv.value, it.value

    def _solve_fn(self, fn, tol, max_iter):
        conv = C.c_double(tol)
        it = C.c_uint(max_iter)
        ok = _check(fn(self.h, C.byref(conv), C.byref(it)))
        return bool(ok), conv.value, it.value

    def solve_cg(self, tol, max_iter):
        """LsSol::Solve_CG (no preconditioner), host-driven over the device handle operations."""
        return self._solve_fn(lib().dfm_solve_cg, tol, max_iter)

    def solve_bicgstab(self, tol, max_iter):
        """LsSol::Solve_BiCGSTAB (no preconditioner), host-driven."""
        return self._solve_fn(lib().dfm_solve_bicgstab, tol, max_iter)

    def solve_mat(self, method, tol, max_iter):
        """MatVec::Sol::Solve_CG / _PCG / _BiCGSTAB / _PBiCGSTAB (method "cg", "pcg", "bicgstab", "pbicgstab"), the
        matrix-level twins, host-driven over the device handle operations."""
        fn = {"cg": lib().dfm_solve_mat_cg, "pcg": lib().dfm_solve_mat_pcg, "bicgstab": lib().dfm_solve_mat_bicgstab,
              "pbicgstab": lib().dfm_solve_mat_pbicgstab}[method]
        return self._solve_fn(fn, tol, max_iter)

    def solve_pcg_host(self, tol, max_iter):
        """LsSol::Solve_PCG driven from the host over the seam (C++ layer)."""
        return self._solve_fn(lib().dfm_solve_pcg_host, tol, max_iter)

    def solve_pbicgstab_host(self, tol, max_iter):
        return self._solve_fn(lib().dfm_solve_pbicgstab_host, tol, max_iter)

    def true_residual(self):
        rel = C.c_double()
        _check(lib().dfm_sys_true_residual(self.h, C.byref(rel)))
        return rel.value

    def last_solve_info(self):
        """(reliable updates, fell back to FP64) of the last mixed-precision solve."""
        a, b = C.c_int(), C.c_int()
        _check(lib().dfm_sys_last_solve_info(self.h, C.byref(a), C.byref(b)))
        return a.value, bool(b.value)

    def last_solve_stats(self):
        a, b, c = C.c_double(), C.c_double(), C.c_uint()
        _check(lib().dfm_sys_last_solve_stats(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return {"ms": a.value, "ms_per_iter": b.value, "iters": c.value}


class MultiSegSystem:
    """Multi-segment linear system (dfm_msys_*): CLinearSystem with several segments + CPreconditioner_ILU(0).

    nblk[i], lens[i] per segment; matrices (i, j) from add_pattern / set_matrix; vectors concatenated."""

    def __init__(self, ctx, nblk, lens):
        self.ctx = ctx
        self.nblk, self.lens = [int(x) for x in nblk], [int(x) for x in lens]
        self.nseg = len(self.nblk)
        self.off = [0]
        for n, ln in zip(self.nblk, self.lens):
            self.off.append(self.off[-1] + n * ln)
        self.m = self.off[-1]
        nb, ln = _i32(self.nblk), _i32(self.lens)
        self.h = C.c_void_p()
        _check(lib().dfm_msys_create(ctx.h, self.nseg, _p(nb), _p(ln), C.byref(self.h)))
        _live.add(self)

    def close(self):
        if getattr(self, "h", None):
            _check(lib().dfm_msys_destroy(self.h))
            self.h = None

    def add_pattern(self, i, j, rowptr, colind):
        rowptr = np.ascontiguousarray(rowptr, np.uint32)
        colind = np.ascontiguousarray(colind, np.uint32)
        _check(lib().dfm_msys_add_pattern(self.h, i, j, _p(rowptr), _p(colind) if colind.size else None))

    def nnz(self, i, j):
        n = C.c_longlong()
        _check(lib().dfm_msys_info(self.h, i, j, C.byref(n)))
        return n.value

    def set_matrix(self, i, j, val, dia=None):
        _check(lib().dfm_msys_set_matrix(self.h, i, j, _p(_f64(val)), _p(_f64(dia))))

    def get_matrix(self, i, j):
        bs = self.lens[i] * self.lens[j]
        val = np.zeros(max(self.nnz(i, j), 1) * bs)
        dia = np.zeros(self.nblk[i] * bs) if i == j else None
        _check(lib().dfm_msys_get_matrix(self.h, i, j, _p(val), _p(dia)))
        return val[:self.nnz(i, j) * bs], dia

    def get_factor(self, i, j):
        bs = self.lens[i] * self.lens[j]
        val = np.zeros(max(self.nnz(i, j), 1) * bs)
        dia = np.zeros(self.nblk[i] * bs) if i == j else None
        _check(lib().dfm_msys_get_factor(self.h, i, j, _p(val), _p(dia)))
        return val[:self.nnz(i, j) * bs], dia

    def set_vector(self, iv, v):
        _check(lib().dfm_msys_set_vector(self.h, iv, _p(_f64(v))))

    def get_vector(self, iv):
        v = np.zeros(self.m)
        _check(lib().dfm_msys_get_vector(self.h, iv, _p(v)))
        return v

    def finalize(self, flags):
        nrm = C.c_double()
        _check(lib().dfm_msys_finalize(self.h, _p(_i32(flags)), C.byref(nrm)))
        return nrm.value

    def factor(self):
        _check(lib().dfm_msys_factor(self.h))

    def precond(self, iv):
        _check(lib().dfm_msys_precond(self.h, iv))

    def resize_tmp(self, n):
        _check(lib().dfm_msys_resize_tmp(self.h, n))

    def dot(self, a, b):
        d = C.c_double()
        _check(lib().dfm_msys_dot(self.h, a, b, C.byref(d)))
        return d.value

    def matvec(self, alpha, iv1, beta, iv2):
        _check(lib().dfm_msys_matvec(self.h, C.c_double(alpha), iv1, C.c_double(beta), iv2))

    def solve(self, method, tol, max_iter):
        """method "pcg" / "pbicgstab": (converged, conv_ratio, iteration)."""
        conv, it = C.c_double(tol), C.c_uint(max_iter)
        rc = _check(lib().dfm_msys_solve(self.h, 0 if method == "pcg" else 1, C.byref(conv), C.byref(it)))
        return rc == 1, conv.value, it.value

    @classmethod
    def from_arrays(cls, ctx, nblk, lens, mats):
        """mats: {(i, j): dict(rowptr, colind, val[, dia])} -> patterns and values installed."""
        ms = cls(ctx, nblk, lens)
        for (i, j), a in sorted(mats.items()):
            ms.add_pattern(i, j, a["rowptr"], a["colind"])
        for (i, j), a in sorted(mats.items()):
            ms.set_matrix(i, j, a["val"], a.get("dia"))
        return ms


class Batch:
    """Many systems solved together: one launch per kernel for all running systems (dfm_batch_*)."""

    def __init__(self, ctx, precision=MIXED, max_systems=1024):
        self.ctx = ctx
        self.h = C.c_void_p()
        _check(lib().dfm_batch_create(ctx.h, precision, max_systems, C.byref(self.h)))
        _live.add(self)

    def add(self, sys, tol, max_iter):
        _check(lib().dfm_batch_add(self.h, sys.h, C.c_double(tol), C.c_uint(max_iter)))

    def step(self, iterations):
        n = C.c_int()
        _check(lib().dfm_batch_step(self.h, iterations, C.byref(n)))
        return n.value

    def status(self, sys):
        """(state, conv_ratio, iteration, seconds); state 0 running, 1 converged, 2 stopped at max_iter."""
        st, conv, it, sec = C.c_int(), C.c_double(), C.c_uint(), C.c_double()
        _check(lib().dfm_batch_status(self.h, sys.h, C.byref(st), C.byref(conv), C.byref(it), C.byref(sec)))
        return st.value, conv.value, it.value, sec.value

    def info(self, sys):
        """(reliable updates, fell back to FP64)."""
        a, b = C.c_int(), C.c_int()
        _check(lib().dfm_batch_info(self.h, sys.h, C.byref(a), C.byref(b)))
        return a.value, bool(b.value)

    def remove(self, sys):
        _check(lib().dfm_batch_remove(self.h, sys.h))

    def close(self):
        if self.h:
            lib().dfm_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def newton_batch(ctx, systems, kind, load_scale, nsteps, max_newton=40, newton_tol=1e-6, precision=MIXED,
                 lin_tol=1e-6, lin_maxit=1000, glob=None):
    """dfm_newton_batch: returns (newton_iters [nsys, nsteps], pcg_iters [nsys], unconverged_pcg [nsys])."""
    n = len(systems)
    arr = (C.c_void_p * n)(*[s.h.value for s in systems])
    ls = _f64(load_scale)
    nit = np.zeros(n * nsteps, np.int32)
    lit = np.zeros(n, np.uint32)
    lun = np.zeros(n, np.int32)
    g = None if glob is None else _f64(glob)
    _check(lib().dfm_newton_batch(ctx.h, n, arr, kind, _p(g), nsteps, _p(ls), max_newton, C.c_double(newton_tol),
                                  precision, C.c_double(lin_tol), C.c_uint(lin_maxit), _p(nit), _p(lit), _p(lun)))
    return nit.reshape(n, nsteps), lit, lun


class Request(C.Structure):
    """dfm_request: one problem submitted to a Server (arrays must stay alive until submit returns)."""
    _fields_ = [("ndim", C.c_int), ("len", C.c_int), ("kind", C.c_int), ("nnode", C.c_int), ("nelem", C.c_int),
                ("coords", C.c_void_p), ("lnods", C.c_void_p), ("params", C.c_void_p), ("npar", C.c_int),
                ("pstride", C.c_int), ("glob", C.c_void_p), ("bc", C.c_void_p), ("nface", C.c_int),
                ("faces", C.c_void_p), ("trac", C.c_void_p), ("tol", C.c_double), ("max_iter", C.c_uint),
                ("method", C.c_int)]


REQ_QUEUED, REQ_SETUP, REQ_SOLVING, REQ_CONVERGED, REQ_MAXITER, REQ_FAILED = range(6)


EQN_SOLID2D, EQN_SOLID3D, EQN_SCALAR2D, EQN_SCALAR3D = range(4)


class Eqn:
    """An equation system (CEqnSystem_Solid2D, CEqn_Solid3D_Linear, CEqnSystem_Scalar2D or CEqn_Scalar3D) over the
    device linear systems: elements in groups (element arrays, elem_group per element), the boundary elements bnd
    (lines in 2D, triangles in 3D) fixed in component idof (-1: all) as AddFixElemAry fixes them."""

    def __init__(self, ctx, type_, coords, lnods, elem_group=None, ngroup=1, bnd=None, idof=-1):
        self.type = type_
        self.ndim = 3 if type_ in (EQN_SOLID3D, EQN_SCALAR3D) else 2
        self.len = {EQN_SOLID2D: 2, EQN_SOLID3D: 3}.get(type_, 1)
        coords, lnods = _f64(coords), _i32(lnods)
        self.nnode = coords.size // self.ndim
        eg = _i32(elem_group)
        self.h = C.c_void_p()
        _check(lib().dfm_eqn_create(ctx.h, type_, self.nnode, _p(coords), lnods.size // (self.ndim + 1), _p(lnods),
                                    _p(eg), ngroup, C.byref(self.h)))
        if bnd is not None:
            b = _i32(bnd)
            _check(lib().dfm_eqn_add_fix(self.h, b.size // self.ndim, self.ndim, _p(b), idof))

    def close(self):
        if self.h:
            lib().dfm_eqn_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_young_poisson(self, young, poisson, is_plane_stress=True, group=-1):
        _check(lib().dfm_eqn_set_young_poisson(self.h, group, C.c_double(young), C.c_double(poisson),
                                               1 if is_plane_stress else 0))

    def set_rho(self, rho, group=-1):
        _check(lib().dfm_eqn_set_rho(self.h, group, C.c_double(rho)))

    def set_gravitation(self, gx, gy, gz=0.0, group=-1):
        _check(lib().dfm_eqn_set_gravitation(self.h, group, C.c_double(gx), C.c_double(gy), C.c_double(gz)))

    def set_flag(self, which, on):
        """which: 0 geometrical nonlinearity, 1 stationary, 2 saved stiffness"""
        _check(lib().dfm_eqn_set_flag(self.h, which, 1 if on else 0))

    def set_scalar(self, which, val, group=-1):
        """which: 0 alpha, 1 capacity, 2 source"""
        _check(lib().dfm_eqn_set_scalar(self.h, which, group, C.c_double(val)))

    def set_time(self, dt, gamma=0.6, beta=-1.0):
        _check(lib().dfm_eqn_set_time(self.h, C.c_double(dt), C.c_double(gamma), C.c_double(beta)))

    def solve(self):
        _check(lib().dfm_eqn_solve(self.h))
        return True

    def itr(self):
        n = C.c_int()
        it = np.zeros(256, np.uint32)
        nr = np.zeros(256)
        _check(lib().dfm_eqn_itr_norm_res(self.h, C.byref(n), _p(it), _p(nr), 256))
        return [(int(it[k]), float(nr[k])) for k in range(min(n.value, 256))]

    def fields(self):
        m = self.nnode * self.len
        u, v, a = np.zeros(m), np.zeros(m), np.zeros(m)
        _check(lib().dfm_eqn_get_field(self.h, _p(u), _p(v), _p(a)))
        return u, v, a


class Server:
    """Continuous-batching solve server (dfm_server_*)."""

    def __init__(self, ctx, precision=MIXED, budget_bytes=64 << 30, max_active=64, nworkers=8):
        self.ctx = ctx
        self.h = C.c_void_p()
        _check(lib().dfm_server_create(ctx.h, precision, C.c_longlong(budget_bytes), max_active, nworkers,
                                       C.byref(self.h)))
        _live.add(self)

    def submit(self, ndim, length, kind, coords, lnods, params, bc, tol=1e-8, max_iter=20000, glob=None,
               faces=None, trac=None, method=0):
        """method: 0 PCG (batched), 1 PBiCGSTAB (solved by a setup worker beside the batch)"""
        coords, lnods, bc = _f64(coords), _i32(lnods), _i32(bc)
        params = _f64(params)
        npar, pstride = (params.size, 0) if params.ndim == 1 else (params.shape[1], params.shape[1])
        g = None if glob is None else _f64(glob)
        faces = None if faces is None else _i32(faces)
        trac = None if trac is None else _f64(trac)
        q = Request(ndim, length, kind, coords.size // ndim, lnods.size // (ndim + 1), coords.ctypes.data,
                    lnods.ctypes.data, params.ctypes.data, npar, pstride, None if g is None else g.ctypes.data,
                    bc.ctypes.data, 0 if faces is None else faces.size // 3,
                    None if faces is None else faces.ctypes.data, None if trac is None else trac.ctypes.data,
                    tol, max_iter, method)
        t = C.c_longlong()
        _check(lib().dfm_server_submit(self.h, C.byref(q), C.byref(t)))
        return t.value

    def progress(self, iterations=16, wait_ms=2):
        a, q = C.c_int(), C.c_int()
        _check(lib().dfm_server_progress(self.h, iterations, wait_ms, C.byref(a), C.byref(q)))
        return a.value, q.value

    def poll(self, ticket):
        """dict(state, latency_s, solve_s, iters, conv, fell_back)."""
        st, lat, sol, it, conv, fb = C.c_int(), C.c_double(), C.c_double(), C.c_uint(), C.c_double(), C.c_int()
        _check(lib().dfm_server_poll(self.h, C.c_longlong(ticket), C.byref(st), C.byref(lat), C.byref(sol),
                                     C.byref(it), C.byref(conv), C.byref(fb)))
        return dict(state=st.value, latency_s=lat.value, solve_s=sol.value, iters=it.value, conv=conv.value,
                    fell_back=bool(fb.value))

    def stats(self):
        out = (C.c_double * 12)()
        _check(lib().dfm_server_stats(self.h, out, 12))
        keys = ["steps", "s_steps", "s_rebuild", "rebuilds", "s_poll", "restarts", "s_idle", "s_setup", "setups",
                "s_check", "s_join", "s_progress"]
        return dict(zip(keys, list(out)))

    def true_residual(self, ticket):
        rel = C.c_double()
        _check(lib().dfm_server_true_residual(self.h, C.c_longlong(ticket), C.byref(rel)))
        return rel.value

    def result(self, ticket, ndof):
        u = np.empty(ndof)
        _check(lib().dfm_server_result(self.h, C.c_longlong(ticket), _p(u)))
        return u

    def release(self, ticket):
        _check(lib().dfm_server_release(self.h, C.c_longlong(ticket)))

    def close(self):
        if self.h:
            lib().dfm_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
