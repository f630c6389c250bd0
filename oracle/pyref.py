"""ctypes wrappers over the oracle libraries (test infrastructure only; see oracle/__init__.py)."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
CPU_REF_SO = os.path.join(HERE, "libcpu_ref.so")
REF_SO = os.path.join(HERE, "_ref", "libdfm_ref.so")

POISSON, SOLID2D, SOLID3D, STVK2D, STVK3D, DIFFUSION2D, SOLID2D_NEWMARK, SOLID3D_NEWMARK, SOLID2D_BEULER, STVK3D_NEWMARK, STVK2D_NEWMARK, STVK2D_BEULER = range(12)

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


_cr = None


def cpu_ref():
    global _cr
    if _cr is None:
        if not os.path.exists(CPU_REF_SO):
            raise RuntimeError(f"{CPU_REF_SO} is not built (make -C oracle)")
        _cr = C.CDLL(CPU_REF_SO)
        _cr.cr_pattern_count.restype = C.c_long
        _cr.cr_set_bc.restype = C.c_double
        _cr.cr_iluk_pattern.restype = C.c_long
        _cr.cr_iluk_pattern.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_long, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p]
        _cr.cr_ordering.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        _cr.cr_boundary_pattern.restype = C.c_long
        _cr.cr_make_residual_save.restype = C.c_double
        _cr.cr_pattern_dia.restype = C.c_long
        _cr.cr_make_residual_savedia.restype = C.c_double
    return _cr


ORDER_RCM, ORDER_RCM2, ORDER_AMD = 3, 4, 5


def ordering(kind, rowptr, colind):
    """COrdering_Blk's ordering (restated in cpu_ref.c) of a block pattern: new -> old as NewToOld reads it, or None
    where the reference asserts."""
    rp = np.ascontiguousarray(rowptr, dtype=np.uint32)
    ci = np.ascontiguousarray(colind, dtype=np.uint32)
    n = rp.size - 1
    out = np.zeros(max(n, 1), np.int32)
    ok = cpu_ref().cr_ordering(kind, n, rp.ctypes.data_as(C.c_void_p), ci.ctypes.data_as(C.c_void_p),
                               out.ctypes.data_as(C.c_void_p))
    return out[:n] if ok else None


class CpuSystem:
    """Block-CRS system assembled by the restatement."""

    def __init__(self, ndim, coords, lnods, len_):
        self.ndim, self.len = ndim, len_
        self.coords = _f64(coords)
        self.lnods = _i32(lnods)
        self.nnode = self.coords.size // ndim
        self.npoel = 3 if ndim == 2 else 4
        self.nelem = self.lnods.size // self.npoel
        lib = cpu_ref()
        self.rowptr = np.zeros(self.nnode + 1, dtype=np.uint32)
        ncrs = lib.cr_pattern_count(self.nnode, self.nelem, self.npoel, self.lnods.ctypes.data_as(C.c_void_p),
                                    self.rowptr.ctypes.data_as(C.c_void_p))
        self.colind = np.zeros(max(ncrs, 1), dtype=np.uint32)[:ncrs]
        lib.cr_pattern_fill(self.nnode, self.nelem, self.npoel, self.lnods.ctypes.data_as(C.c_void_p),
                            self.rowptr.ctypes.data_as(C.c_void_p), self.colind.ctypes.data_as(C.c_void_p))
        bs = len_ * len_
        self.crs = np.zeros(ncrs * bs)
        self.dia = np.zeros(self.nnode * bs)
        self.res = np.zeros(self.nnode * len_)

    def zero(self):
        self.crs[:] = 0.0
        self.dia[:] = 0.0
        self.res[:] = 0.0

    def assemble(self, kind, params, u=None, v=None, glob=(0.0, 0.0)):
        """params: (npar,) for all elements or (nelem, npar) per element."""
        params = _f64(params)
        if params.ndim == 1:
            npar, pstride = params.size, 0
        else:
            npar, pstride = params.shape[1], params.shape[1]
        glob = _f64(glob)
        u = None if u is None else _f64(u)
        v = None if v is None else _f64(v)
        p = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
        rc = cpu_ref().cr_assemble(kind, self.ndim, self.nnode, p(self.coords), self.nelem, p(self.lnods),
                                   p(params), npar, pstride, p(glob), p(u), p(v), self.len, p(self.rowptr),
                                   p(self.colind), p(self.crs), p(self.dia), p(self.res))
        assert rc == 0

    def assemble_save(self, kind, params, flags):
        """CLinearSystem_Save (restated): the *_SaveStiffMat element loop into the matrix, the boundary matrix
        (self.brp / self.bci / self.bval) and the force (self.force); FinalizeMarge's Dirichlet treatment of both.
        The residual stays zero."""
        params = _f64(params)
        pstride = 0 if params.ndim == 1 else params.shape[1]
        flags = _i32(flags)
        p = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
        lib = cpu_ref()
        n, bs = self.nnode, self.len * self.len
        nb = lib.cr_boundary_pattern(n, self.len, p(self.rowptr), p(self.colind), p(flags), None, None)
        self.brp = np.zeros(n + 1, np.uint32)
        bci = np.zeros(max(nb, 1), np.uint32)
        lib.cr_boundary_pattern(n, self.len, p(self.rowptr), p(self.colind), p(flags), p(self.brp), p(bci))
        self.bci = bci[:nb]
        self.bval = np.zeros(max(nb, 1) * bs)
        self.force = np.zeros(n * self.len)
        self.zero()
        rc = lib.cr_assemble_save(kind, self.ndim, n, p(self.coords), self.nelem, p(self.lnods), p(params), pstride,
                                  self.len, p(self.rowptr), p(self.colind), p(self.crs), p(self.dia), p(self.brp),
                                  p(bci), p(self.bval), p(self.force))
        assert rc == 0, rc
        self.bval = self.bval[: nb * bs]
        self.set_bc(flags)
        lib.cr_boundary_bc(n, self.len, p(self.brp), p(bci), p(self.bval), p(flags))
        self.flags = flags

    def make_residual(self, u):
        """MakeResidual (restated): returns (||r||, residual, update); the residual also becomes self.res."""
        u = _f64(u)
        res = np.zeros_like(u)
        upd = np.zeros_like(u)
        p = lambda a: a.ctypes.data_as(C.c_void_p)
        bci = self.bci if self.bci.size else np.zeros(1, np.uint32)
        bval = self.bval if self.bval.size else np.zeros(1)
        norm = cpu_ref().cr_make_residual_save(self.nnode, self.len, p(self.brp), p(bci), p(bval), p(self.force),
                                               p(u), p(res), p(upd))
        self.res = res.copy()
        return norm, res, upd

    def update_field_save(self, upd, u):
        """UpdateValueOfField of CLinearSystem_Save (restated): returns (update, field)."""
        upd, u = _f64(upd).copy(), _f64(u).copy()
        p = lambda a: a.ctypes.data_as(C.c_void_p)
        cpu_ref().cr_update_field_save(self.nnode, self.len, p(self.flags), p(upd), p(u))
        return upd, u

    def assemble_savedia(self, kind, params, flags, glob):
        """CLinearSystem_SaveDiaM_Newmark (kind DIFFUSION2D, params (rho, alpha, source)) or _NewmarkBeta (kind
        SOLID2D, (lambda, myu, rho, gx, gy)), restated: the element loop into the matrix, the boundary matrix
        (SetPatternDia: self.brp / self.bci / self.bval) and the force (self.force); FinalizeMarge: the matrix's
        Dirichlet treatment and the boundary matrix's fixed rows zeroed. glob = (dt, gamma, beta)."""
        params = _f64(params)
        pstride = 0 if params.ndim == 1 else params.shape[1]
        flags, glob = _i32(flags), _f64(glob)
        p = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
        lib = cpu_ref()
        n, bs = self.nnode, self.len * self.len
        nb = lib.cr_pattern_dia(n, p(self.rowptr), p(self.colind), None, None)
        self.brp = np.zeros(n + 1, np.uint32)
        self.bci = np.zeros(nb, np.uint32)
        lib.cr_pattern_dia(n, p(self.rowptr), p(self.colind), p(self.brp), p(self.bci))
        self.bval = np.zeros(nb * bs)
        self.force = np.zeros(n * self.len)
        self.zero()
        rc = lib.cr_assemble_savedia(kind, n, p(self.coords), self.nelem, p(self.lnods), p(params), pstride, p(glob),
                                     self.len, p(self.rowptr), p(self.colind), p(self.crs), p(self.dia), p(self.brp),
                                     p(self.bci), p(self.bval), p(self.force))
        assert rc == 0, rc
        self.set_bc(flags)
        lib.cr_boundary_rows(n, self.len, p(self.brp), p(self.bval), p(flags))
        self.flags, self.savedia = flags, (kind, glob)

    def make_residual_savedia(self, u, v, a=None):
        """MakeResidual of the lumped-mass saved systems (restated): returns (||r||, residual, update); the residual
        also becomes self.res."""
        kind, glob = self.savedia
        u, v = _f64(u), _f64(v)
        a = np.zeros_like(u) if a is None else _f64(a)
        res, upd = np.zeros_like(u), np.zeros_like(u)
        p = lambda x: x.ctypes.data_as(C.c_void_p)
        norm = cpu_ref().cr_make_residual_savedia(kind, self.nnode, self.len, p(self.brp), p(self.bci), p(self.bval),
                                                  p(self.force), p(glob), p(self.flags), p(u), p(v), p(a), p(res),
                                                  p(upd))
        self.res = res.copy()
        return norm, res, upd

    def update_field_savedia(self, upd, u, v, a=None):
        """UpdateValueOfField of the lumped-mass saved systems (restated): returns (update, u, v, a)."""
        kind, glob = self.savedia
        upd, u, v = _f64(upd).copy(), _f64(u).copy(), _f64(v).copy()
        a = np.zeros_like(u) if a is None else _f64(a).copy()
        p = lambda x: x.ctypes.data_as(C.c_void_p)
        cpu_ref().cr_update_field_savedia(kind, self.nnode, self.len, p(glob), p(self.flags), p(upd), p(u), p(v), p(a))
        return upd, u, v, a

    def assemble_newmark(self, params, u, v, a, glob):
        """Newmark-beta element loop (restated), 2D or 3D by the mesh: matrix and residual accumulate into
        self.crs / self.dia / self.res (zero them first). glob = (dt, gamma, beta, is_initial)."""
        params = _f64(params)
        pstride = 0 if params.ndim == 1 else params.shape[1]
        u, v, a, glob = _f64(u), _f64(v), _f64(a), _f64(glob)
        p = lambda x: x.ctypes.data_as(C.c_void_p)
        rc = cpu_ref().cr_assemble_newmark(self.ndim, self.nnode, p(self.coords), self.nelem, p(self.lnods),
                                           p(params), pstride, p(glob), p(u), p(v), p(a), p(self.rowptr),
                                           p(self.colind), p(self.crs), p(self.dia), p(self.res))
        assert rc == 0

    def assemble_stvk3d_newmark(self, params, u, v, a, glob):
        """Newmark-beta StVK element loop (restated; 2D or 3D by the mesh) into self.crs / self.dia / self.res (zero them first)."""
        params = _f64(params)
        pstride = 0 if params.ndim == 1 else params.shape[1]
        u, v, a, glob = _f64(u), _f64(v), _f64(a), _f64(glob)
        p = lambda x: x.ctypes.data_as(C.c_void_p)
        fn = cpu_ref().cr_assemble_stvk3d_newmark if self.ndim == 3 else cpu_ref().cr_assemble_stvk2d_newmark
        rc = fn(self.nnode, p(self.coords), self.nelem, p(self.lnods), p(params),
                                                  pstride, p(glob), p(u), p(v), p(a), p(self.rowptr), p(self.colind),
                                                  p(self.crs), p(self.dia), p(self.res))
        assert rc == 0

    def assemble_beuler(self, params, u, v, v0, dt, is_initial, stvk=False):
        """Backward-Euler 2D element loop (restated; linear elasticity, or St.Venant-Kirchhoff with stvk) into
        self.crs / self.dia / self.res (zero them first)."""
        params = _f64(params)
        pstride = 0 if params.ndim == 1 else params.shape[1]
        u, v, v0 = _f64(u), _f64(v), _f64(v0)
        p = lambda x: x.ctypes.data_as(C.c_void_p)
        fn = cpu_ref().cr_assemble_stvk2d_beuler if stvk else cpu_ref().cr_assemble_beuler
        rc = fn(self.nnode, p(self.coords), self.nelem, p(self.lnods), p(params), pstride,
                                          C.c_double(dt), 1 if is_initial else 0, p(u), p(v), p(v0), p(self.rowptr),
                                          p(self.colind), p(self.crs), p(self.dia), p(self.res))
        assert rc == 0

    @staticmethod
    def update_field_beuler(dt, is_initial, upd, u, v):
        """UpdateValueOfField_BackwardEular (restated): returns new (u, v)."""
        u, v, upd = _f64(u).copy(), _f64(v).copy(), _f64(upd)
        p = lambda x: x.ctypes.data_as(C.c_void_p)
        cpu_ref().cr_update_field_beuler(C.c_long(u.size), C.c_double(dt), 1 if is_initial else 0, p(upd), p(u), p(v))
        return u, v

    @staticmethod
    def update_field_newmark(gamma, beta, dt, is_initial, upd, u, v, a):
        """UpdateValueOfField_NewmarkBeta (restated): returns new (u, v, a)."""
        u, v, a, upd = _f64(u).copy(), _f64(v).copy(), _f64(a).copy(), _f64(upd)
        p = lambda x: x.ctypes.data_as(C.c_void_p)
        cpu_ref().cr_update_field_newmark(C.c_long(u.size), C.c_double(gamma), C.c_double(beta), C.c_double(dt),
                                          1 if is_initial else 0, p(upd), p(u), p(v), p(a))
        return u, v, a

    def traction(self, faces, trac):
        faces, trac = _i32(faces), _f64(trac)
        cpu_ref().cr_traction_loads(faces.size // 3, faces.ctypes.data_as(C.c_void_p),
                                    trac.ctypes.data_as(C.c_void_p), self.coords.ctypes.data_as(C.c_void_p),
                                    self.res.ctypes.data_as(C.c_void_p))

    def set_bc(self, flags):
        flags = _i32(flags)
        p = lambda a: a.ctypes.data_as(C.c_void_p)
        return cpu_ref().cr_set_bc(self.nnode, self.len, p(self.rowptr), p(self.colind), p(self.crs),
                                   p(self.dia), p(flags), p(self.res))

    def dense(self):
        """The assembled matrix as a dense array (for checks on small systems)."""
        n, b = self.nnode, self.len
        A = np.zeros((n * b, n * b))
        crs = self.crs.reshape(-1, b, b)
        dia = self.dia.reshape(-1, b, b)
        for i in range(n):
            A[i * b:(i + 1) * b, i * b:(i + 1) * b] = dia[i]
            for q in range(self.rowptr[i], self.rowptr[i + 1]):
                j = self.colind[q]
                A[i * b:(i + 1) * b, j * b:(j + 1) * b] = crs[q]
        return A

    def matvec(self, alpha, x, beta, y):
        y = _f64(y).copy()
        x = _f64(x)
        p = lambda a: a.ctypes.data_as(C.c_void_p)
        cpu_ref().cr_matvec(self.nnode, self.len, p(self.rowptr), p(self.colind), p(self.crs), p(self.dia),
                            C.c_double(alpha), p(x), C.c_double(beta), p(y))
        return y

    def ilu0(self, order=None):
        return self.ilu(order, 0)

    def ilu(self, order=None, lev=0):
        """ILU(lev) (lev -1: no fill limit) of the matrix under `order` (new -> old): the level-of-fill pattern
        of the ordered pattern, A's values scattered into it, DoILUDecomp. Entry levels in f["flev"]."""
        n, bs = self.nnode, self.len * self.len
        ncrs = int(self.rowptr[-1])
        order = None if order is None else _i32(order)
        p = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
        lib = cpu_ref()
        rp0 = np.zeros(n + 1, np.uint32)
        ci0 = np.zeros(max(ncrs, 1), np.uint32)
        di0 = np.zeros(n, np.uint32)
        lib.cr_ilu0_pattern(n, p(self.rowptr), p(self.colind), p(order), p(rp0), p(ci0), p(di0))
        cap = max(ncrs, 1)
        while True:
            f = {"frowptr": np.zeros(n + 1, np.uint32), "fcolind": np.zeros(cap, np.uint32),
                 "diaind": np.zeros(n, np.uint32), "flev": np.zeros(cap, np.int32)}
            got = lib.cr_iluk_pattern(n, p(rp0), p(ci0), lev, cap, p(f["frowptr"]), p(f["fcolind"]),
                                      p(f["diaind"]), p(f["flev"]))
            if got >= 0:
                break
            cap = max(2 * cap, -got)
        f["fcolind"], f["flev"] = f["fcolind"][:got], f["flev"][:got]
        f["fcrs"], f["fdia"] = np.zeros(max(got, 1) * bs)[: got * bs], np.zeros(n * bs)
        lib.cr_ilu0_setvalue(n, self.len, p(self.rowptr), p(self.colind), p(self.crs), p(self.dia), p(order),
                             p(f["frowptr"]), p(f["fcolind"]), p(f["fcrs"]), p(f["fdia"]))
        ok = lib.cr_ilu0_factor(n, self.len, p(f["frowptr"]), p(f["fcolind"]), p(f["diaind"]), p(f["fcrs"]),
                                p(f["fdia"]))
        f["ok"] = bool(ok)
        f["order"] = order
        return f

    def ilu_apply(self, f, r):
        v = _f64(r).copy()
        work = np.zeros_like(v)
        p = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
        cpu_ref().cr_ilu_solve(self.nnode, self.len, p(f["frowptr"]), p(f["fcolind"]), p(f["diaind"]),
                               p(f["fcrs"]), p(f["fdia"]), p(f["order"]), p(v), p(work))
        return v

    def _cr_system(self, f):
        if f is None:  # unpreconditioned solvers: no factor
            f = {"frowptr": None, "fcolind": None, "diaind": None, "fcrs": None, "fdia": None, "order": None}
        class S(C.Structure):
            _fields_ = [("n", C.c_int), ("len", C.c_int), ("rowptr", C.c_void_p), ("colind", C.c_void_p),
                        ("crs", C.c_void_p), ("dia", C.c_void_p), ("frowptr", C.c_void_p),
                        ("fcolind", C.c_void_p), ("diaind", C.c_void_p), ("fcrs", C.c_void_p),
                        ("fdia", C.c_void_p), ("order", C.c_void_p)]
        p = lambda a: None if a is None else a.ctypes.data
        return S(self.nnode, self.len, p(self.rowptr), p(self.colind), p(self.crs), p(self.dia),
                 p(f["frowptr"]), p(f["fcolind"]), p(f["diaind"]), p(f["fcrs"]), p(f["fdia"]), p(f["order"]))

    def solve(self, f, method="pcg", tol=1e-10, max_iter=1000, res=None):
        """Returns (ok, conv_ratio, iteration, update, residual_after)."""
        s = self._cr_system(f)
        r = (self.res if res is None else _f64(res)).copy()
        x = np.zeros_like(r)
        conv = C.c_double(tol)
        it = C.c_uint(max_iter)
        fn = {"pcg": cpu_ref().cr_solve_pcg, "pbicgstab": cpu_ref().cr_solve_pbicgstab, "cg": cpu_ref().cr_solve_cg,
              "bicgstab": cpu_ref().cr_solve_bicgstab, "mat_cg": cpu_ref().cr_solve_mat_cg,
              "mat_pcg": cpu_ref().cr_solve_mat_pcg, "mat_bicgstab": cpu_ref().cr_solve_mat_bicgstab,
              "mat_pbicgstab": cpu_ref().cr_solve_mat_pbicgstab}[method]
        ok = fn(C.byref(s), r.ctypes.data_as(C.c_void_p), x.ctypes.data_as(C.c_void_p), C.byref(conv),
                C.byref(it))
        return bool(ok), conv.value, it.value, x, r


_ref = None


def ref_available():
    return os.path.exists(REF_SO)


def reference():
    global _ref
    if _ref is None:
        if not os.path.exists(REF_SO):
            raise RuntimeError(f"{REF_SO} is not built (make -C oracle ref, where the reference sources exist)")
        _ref = C.CDLL(REF_SO)
        _ref.ref_sys_create.restype = C.c_void_p
        _ref.ref_sys_create2.restype = C.c_void_p
        _ref.ref_sys_make_residual.restype = C.c_double
        _ref.ref_sys_make_residual.argtypes = [C.c_void_p]
        for name in ["ref_sys_destroy", "ref_sys_ncrs", "ref_sys_pattern", "ref_sys_set_field", "ref_sys_get_field",
                     "ref_sys_set_bc", "ref_sys_assemble", "ref_sys_get_matrix", "ref_sys_get_vector",
                     "ref_sys_set_vector", "ref_sys_matvec", "ref_ilu0_factor", "ref_ilu_apply", "ref_sys_solve",
                     "ref_sys_update_field", "ref_ordering", "ref_mat_solve"]:
            getattr(_ref, name).argtypes = None
        _ref.ref_ilu_factor.restype = C.c_long
        _ref.ref_ilu_factor.argtypes = [C.c_void_p, C.c_int, C.c_long] + [C.c_void_p] * 5
        _ref.ref_sys_destroy.argtypes = [C.c_void_p]
        _ref.ref_sys_ncrs.argtypes = [C.c_void_p]
    return _ref


class RefSystem:
    """The reference DelFEM driven through its own API (oracle/ref_capi.cpp)."""

    def __init__(self, ndim, coords, lnods, len_, groups=None, save=False, dynamic=False, savedia=False):
        self.ndim, self.len = ndim, len_
        self.coords = _f64(coords)
        self.lnods = _i32(lnods)
        self.nnode = self.coords.size // ndim
        npoel = 3 if ndim == 2 else 4
        self.nelem = self.lnods.size // npoel
        self.groups = None if groups is None else _i32(groups)
        ngroup = 1 if groups is None else int(self.groups.max()) + 1
        self.ngroup = ngroup
        lib = reference()
        p = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
        self.h = C.c_void_p(lib.ref_sys_create2(ndim, self.nnode, p(self.coords), self.nelem, p(self.lnods), ngroup,
                                                p(self.groups), len_,
                                                (1 if save else 0) | (2 if dynamic else 0) | (4 if savedia else 0)))
        self.ncrs = lib.ref_sys_ncrs(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            reference().ref_sys_destroy(self.h)
            self.h = None

    def pattern(self):
        rp = np.zeros(self.nnode + 1, np.uint32)
        ci = np.zeros(max(self.ncrs, 1), np.uint32)
        reference().ref_sys_pattern(self.h, rp.ctypes.data_as(C.c_void_p), ci.ctypes.data_as(C.c_void_p))
        return rp, ci[: self.ncrs]

    def ordering(self, kind):
        """COrdering_Blk::MakeOrdering_RCM / _RCM2 / _AMD (kind 3 / 4 / 5) of the matrix, as NewToOld."""
        out = np.zeros(self.nnode, np.int32)
        reference().ref_ordering(self.h, kind, out.ctypes.data_as(C.c_void_p))
        return out

    def set_field(self, u=None, v=None):
        u = None if u is None else _f64(u)
        v = None if v is None else _f64(v)
        p = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
        reference().ref_sys_set_field(self.h, p(u), p(v))

    def get_field(self):
        u = np.zeros(self.nnode * self.len)
        reference().ref_sys_get_field(self.h, u.ctypes.data_as(C.c_void_p))
        return u

    def set_bc(self, flags):
        flags = _i32(flags)
        reference().ref_sys_set_bc(self.h, flags.ctypes.data_as(C.c_void_p))

    def assemble(self, kind, group_params, glob=(0.0, 0.0)):
        gp = _f64(group_params)
        if gp.ndim == 1:
            gp = gp[None, :]
        assert gp.shape[0] == self.ngroup
        glob = _f64(glob)
        rn = C.c_double(0.0)
        rc = reference().ref_sys_assemble(self.h, kind, gp.ctypes.data_as(C.c_void_p), gp.shape[1],
                                          glob.ctypes.data_as(C.c_void_p), C.byref(rn))
        assert rc == 0, rc
        return rn.value

    def matrix(self):
        bs = self.len * self.len
        crs = np.zeros(self.ncrs * bs)
        dia = np.zeros(self.nnode * bs)
        reference().ref_sys_get_matrix(self.h, crs.ctypes.data_as(C.c_void_p), dia.ctypes.data_as(C.c_void_p))
        return crs, dia

    def vector(self, iv):
        out = np.zeros(self.nnode * self.len)
        reference().ref_sys_get_vector(self.h, C.c_int(iv), out.ctypes.data_as(C.c_void_p))
        return out

    def set_vector(self, iv, vals):
        vals = _f64(vals)
        reference().ref_sys_set_vector(self.h, C.c_int(iv), vals.ctypes.data_as(C.c_void_p))

    def matvec(self, alpha, x, beta, y):
        x, y = _f64(x), _f64(y).copy()
        reference().ref_sys_matvec(self.h, C.c_double(alpha), x.ctypes.data_as(C.c_void_p), C.c_double(beta),
                                   y.ctypes.data_as(C.c_void_p))
        return y

    def ilu0(self, order=None):
        bs = self.len * self.len
        rp = np.zeros(self.nnode + 1, np.uint32)
        ci = np.zeros(max(self.ncrs, 1), np.uint32)
        crs = np.zeros(max(self.ncrs, 1) * bs)
        dia = np.zeros(self.nnode * bs)
        order = None if order is None else _i32(order)
        p = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
        ok = reference().ref_ilu0_factor(self.h, p(order), p(rp), p(ci), p(crs), p(dia))
        return bool(ok), rp, ci[: self.ncrs], crs[: self.ncrs * bs], dia

    def ilu(self, lev):
        """The reference's ILU(lev) factor in the natural order: (ok, rowptr, colind, crs, dia)."""
        bs = self.len * self.len
        cap = max(self.ncrs, 1)
        while True:
            rp = np.zeros(self.nnode + 1, np.uint32)
            ci = np.zeros(cap, np.uint32)
            crs = np.zeros(cap * bs)
            dia = np.zeros(self.nnode * bs)
            ok = C.c_int(0)
            p = lambda a: a.ctypes.data_as(C.c_void_p)
            got = reference().ref_ilu_factor(self.h, lev, cap, p(rp), p(ci), p(crs), p(dia), C.byref(ok))
            if got >= 0:
                return bool(ok.value), rp, ci[:got], crs[: got * bs], dia
            cap = -got

    def ilu_apply(self, r, lev=0, order=None):
        r = _f64(r)
        z = np.zeros_like(r)
        order = None if order is None else _i32(order)
        p = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
        ok = reference().ref_ilu_apply(self.h, C.c_int(lev), p(order), p(r), p(z))
        assert ok
        return z

    def solve_mat(self, method, tol=1e-10, max_iter=1000):
        """MatVec::Sol::Solve_CG / _PCG / _BiCGSTAB / _PBiCGSTAB (ILU(0), natural order) on the system."""
        conv = C.c_double(tol)
        it = C.c_uint(max_iter)
        code = {"cg": 0, "pcg": 1, "bicgstab": 2, "pbicgstab": 3}[method]
        ok = reference().ref_mat_solve(self.h, C.c_int(code), C.byref(conv), C.byref(it))
        return ok, conv.value, it.value

    def solve(self, method="pcg", tol=1e-10, max_iter=1000, lev=0, order=None):
        conv = C.c_double(tol)
        it = C.c_uint(max_iter)
        order = None if order is None else _i32(order)
        p = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
        code = {"pcg": 0, "pbicgstab": 1, "cg": 2, "bicgstab": 3}[method]
        ok = reference().ref_sys_solve(self.h, C.c_int(code), C.c_int(lev), p(order), C.byref(conv), C.byref(it))
        return ok, conv.value, it.value

    def update_field(self):
        reference().ref_sys_update_field(self.h)

    def set_acc(self, a=None):
        a = None if a is None else _f64(a)
        reference().ref_sys_set_acc(self.h, None if a is None else a.ctypes.data_as(C.c_void_p))

    def get_fields(self, acc=True):
        """(u, v, a) of the field; a is None (not read) when acc is False (a field made without ACCELERATION)."""
        out = [np.zeros(self.nnode * self.len) for _ in range(3 if acc else 2)]
        ptrs = [o.ctypes.data_as(C.c_void_p) for o in out] + ([] if acc else [None])
        reference().ref_sys_get_fields(self.h, *ptrs)
        return out if acc else out + [None]

    def set_velo_pre(self, v0=None):
        v0 = None if v0 is None else _f64(v0)
        reference().ref_sys_set_velo_pre(self.h, None if v0 is None else v0.ctypes.data_as(C.c_void_p))

    def update_field_beuler(self, dt, is_initial):
        reference().ref_sys_update_field_beuler(self.h, C.c_double(dt), 1 if is_initial else 0)

    def update_field_newmark(self, gamma, beta, dt, is_initial):
        reference().ref_sys_update_field_newmark(self.h, C.c_double(gamma), C.c_double(beta), C.c_double(dt),
                                                 1 if is_initial else 0)

    def assemble_save(self, kind, group_params):
        """CLinearSystem_Save: InitializeMarge, the *_SaveStiffMat loop, FinalizeMarge (system made with save)."""
        gp = _f64(group_params)
        if gp.ndim == 1:
            gp = gp[None, :]
        assert gp.shape[0] == self.ngroup
        rc = reference().ref_sys_assemble_save(self.h, kind, gp.ctypes.data_as(C.c_void_p), gp.shape[1])
        assert rc == 0, rc

    def make_residual(self):
        return reference().ref_sys_make_residual(self.h)

    def force(self):
        out = np.zeros(self.nnode * self.len)
        reference().ref_sys_get_force(self.h, out.ctypes.data_as(C.c_void_p))
        return out

    def update_field_save(self):
        reference().ref_sys_update_field_save(self.h)

    def assemble_savedia(self, kind, group_params, glob):
        """CLinearSystem_SaveDiaM_Newmark / _NewmarkBeta (system made with savedia): InitializeMarge,
        SetNewmarkParameter, the element loop per group, FinalizeMarge. glob = (dt, gamma, beta)."""
        gp = _f64(group_params)
        if gp.ndim == 1:
            gp = gp[None, :]
        assert gp.shape[0] == self.ngroup
        dt, gamma, beta = (float(x) for x in glob)
        rc = reference().ref_sys_assemble_savedia(self.h, kind, gp.ctypes.data_as(C.c_void_p), gp.shape[1],
                                                  C.c_double(dt), C.c_double(gamma), C.c_double(beta))
        assert rc == 0, rc

    def update_field_savedia(self):
        reference().ref_sys_update_field_savedia(self.h)

    def boundary(self):
        """The boundary matrix: (browptr, bcolind, values)."""
        lib = reference()
        lib.ref_sys_get_boundary.restype = C.c_long
        nb = lib.ref_sys_get_boundary(self.h, None, None, None)
        rp = np.zeros(self.nnode + 1, np.uint32)
        ci = np.zeros(max(nb, 1), np.uint32)
        val = np.zeros(max(nb, 1) * self.len * self.len)
        p = lambda a: a.ctypes.data_as(C.c_void_p)
        lib.ref_sys_get_boundary(self.h, p(rp), p(ci), p(val))
        return rp, ci[:nb], val[: nb * self.len * self.len]


# ---------------------------------------------------------------------------------------------------------
# Equation systems: CEqnSystem_Solid2D, CEqn_Solid3D_Linear, CEqnSystem_Scalar2D, CEqn_Scalar3D (include/delfem/
# eqnsys_solid.h, eqnsys_scalar.h; src/femeqn/eqnsys_solid.cpp, eqnsys_scalar.cpp). RefEqn drives the reference's
# own classes through oracle/ref_capi.cpp; EqnFlow restates their Solve flows over CpuSystem, and is what the GPU
# equation systems are tested against.

EQN_SOLID2D, EQN_SOLID3D, EQN_SCALAR2D, EQN_SCALAR3D = range(4)


def _eqn_lib():
    lib = reference()
    if not getattr(lib, "_eqn_ready", False):
        lib.ref_eqn_create.restype = C.c_void_p
        lib.ref_eqn_create.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                       C.c_int, C.c_void_p, C.c_int]
        lib.ref_eqn_destroy.argtypes = [C.c_void_p]
        lib.ref_eqn_set_young_poisson.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int]
        lib.ref_eqn_set_rho.argtypes = [C.c_void_p, C.c_int, C.c_double]
        lib.ref_eqn_set_gravitation.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double]
        lib.ref_eqn_set_flag.argtypes = [C.c_void_p, C.c_int, C.c_int]
        lib.ref_eqn_set_scalar.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double]
        lib.ref_eqn_set_time.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double]
        lib.ref_eqn_solve.argtypes = [C.c_void_p]
        lib.ref_eqn_itr.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        lib.ref_eqn_get_field.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib._eqn_ready = True
    return lib


def _eqn_dims(type_):
    ndim = 3 if type_ in (EQN_SOLID3D, EQN_SCALAR3D) else 2
    length = {EQN_SOLID2D: 2, EQN_SOLID3D: 3}.get(type_, 1)
    return ndim, length


class RefEqn:
    """The reference's equation system of `type_` on a mesh of element groups (one element array and CEqn_* per
    group, elem_group per element) whose boundary elements `bnd` (lines in 2D, triangles in 3D) are fixed with
    AddFixElemAry (component idof, -1 all)."""

    def __init__(self, type_, coords, lnods, elem_group=None, ngroup=1, bnd=None, idof=-1):
        self.type = type_
        self.ndim, self.len = _eqn_dims(type_)
        self.coords, self.lnods = _f64(coords), _i32(lnods)
        self.nnode = self.coords.size // self.ndim
        npoel = self.ndim + 1
        eg = None if elem_group is None else _i32(elem_group)
        bnd = np.zeros(0, np.int32) if bnd is None else _i32(bnd)
        nbno = self.ndim
        p = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
        self.h = C.c_void_p(_eqn_lib().ref_eqn_create(type_, self.nnode, p(self.coords), self.lnods.size // npoel,
                                                      p(self.lnods), ngroup, p(eg), bnd.size // nbno,
                                                      p(bnd) if bnd.size else None, idof))

    def __del__(self):
        if getattr(self, "h", None):
            _eqn_lib().ref_eqn_destroy(self.h)
            self.h = None

    def set_young_poisson(self, young, poisson, is_plane_stress=True, group=-1):
        _eqn_lib().ref_eqn_set_young_poisson(self.h, group, young, poisson, 1 if is_plane_stress else 0)

    def set_rho(self, rho, group=-1):
        _eqn_lib().ref_eqn_set_rho(self.h, group, rho)

    def set_gravitation(self, gx, gy, gz=0.0, group=-1):
        _eqn_lib().ref_eqn_set_gravitation(self.h, group, gx, gy, gz)

    def set_flag(self, which, on):
        """which: 0 geometrical nonlinearity, 1 stationary, 2 saved stiffness"""
        _eqn_lib().ref_eqn_set_flag(self.h, which, 1 if on else 0)

    def set_scalar(self, which, val, group=-1):
        """which: 0 alpha, 1 capacity, 2 source"""
        _eqn_lib().ref_eqn_set_scalar(self.h, which, group, val)

    def set_time(self, dt, gamma=0.6, beta=-1.0):
        _eqn_lib().ref_eqn_set_time(self.h, dt, gamma, beta)

    def solve(self):
        return bool(_eqn_lib().ref_eqn_solve(self.h))

    def itr(self):
        it = np.zeros(256, np.uint32)
        nr = np.zeros(256)
        n = _eqn_lib().ref_eqn_itr(self.h, it.ctypes.data_as(C.c_void_p), nr.ctypes.data_as(C.c_void_p), 256)
        return [(int(it[k]), float(nr[k])) for k in range(n)]

    def fields(self):
        m = self.nnode * self.len
        u, v, a = np.zeros(m), np.zeros(m), np.zeros(m)
        p = lambda x: x.ctypes.data_as(C.c_void_p)
        _eqn_lib().ref_eqn_get_field(self.h, p(u), p(v), p(a))
        return u, v, a


def _solid_lame(young, poisson, is_plane_stress):
    """CEqn_Solid2D::SetYoungPoisson (include/delfem/eqnsys_solid.h:190-196), including its flag: the plane-stress
    reduction of lambda applies when is_plane_stress is false."""
    lam = young * poisson / ((1.0 + poisson) * (1 - 2.0 * poisson))
    myu = young / (2.0 * (1.0 + poisson))
    if not is_plane_stress:
        lam = 2 * lam * myu / (lam + 2 * myu)
    return lam, myu


class EqnFlow:
    """The Solve flows of the reference's equation systems, restated over CpuSystem. The same constructor and
    setters as RefEqn. Kinds: 0 Poisson, 1/2 linear solid 2D/3D, 3/4 St.Venant-Kirchhoff 2D/3D, 5 diffusion."""

    def __init__(self, type_, coords, lnods, elem_group=None, ngroup=1, bnd=None, idof=-1):
        self.type = type_
        self.ndim, self.len = _eqn_dims(type_)
        npoel = self.ndim + 1
        coords, lnods = _f64(coords), _i32(lnods).reshape(-1, npoel)
        eg = np.zeros(lnods.shape[0], np.int32) if elem_group is None else _i32(elem_group)
        # the reference's element arrays: group by group, element order kept inside each group
        perm = np.argsort(eg, kind="stable")
        self.coords, self.lnods, self.group = coords, lnods[perm].ravel(), eg[perm]
        self.ngroup, self.nnode = ngroup, coords.size // self.ndim
        self.flags = np.zeros(self.nnode * self.len, np.int32)
        if bnd is not None:
            nodes = np.unique(_i32(bnd))
            fl = self.flags.reshape(-1, self.len)
            if idof < 0:
                fl[nodes, :] = 1
            else:
                fl[nodes, idof] = 1
        m = self.nnode * self.len
        self.u, self.v, self.a = np.zeros(m), np.zeros(m), np.zeros(m)
        self.dt, self.gamma, self.beta = 0.1, 0.6, 0.3025
        self.itr_ = []
        self.cs = None  # the linear system (pLS)
        self.f = None  # the preconditioner's factor (pPrec); solids ILU(1), scalars ILU(0)
        self.has_prec = False
        self.cleared_ls = self.cleared_prec = True
        self.save = False
        self.nonlin = False
        if type_ == EQN_SOLID2D:
            lam, myu = _solid_lame(1.0, 0.0, True)
            self.gp = [[lam, myu, 1.0, 0.0, 0.0] for _ in range(ngroup)]
            self.stationary = True
        elif type_ == EQN_SOLID3D:
            self.gp = [[0.0, 0.0, 1.0, 0.0, 0.0, 0.0]]
            self.stationary = False  # CEqn_Solid3D_Linear's constructor
        else:
            self.gp = [[1.0, 1.0, 0.0] for _ in range(ngroup)]  # alpha, capacity, source
            self.stationary = True
        self.lev = 1 if type_ in (EQN_SOLID2D, EQN_SOLID3D) else 0

    # setters: the values of the reference's setters, and which of them drop the linear system
    def _groups(self, group):
        return range(len(self.gp)) if group < 0 else [group]

    def set_young_poisson(self, young, poisson, is_plane_stress=True, group=-1):
        if self.type == EQN_SOLID3D:
            lam = young * poisson / ((1.0 + poisson) * (1 - 2.0 * poisson))
            myu = young / (2.0 * (1.0 + poisson))
        else:
            lam, myu = _solid_lame(young, poisson, is_plane_stress)
        for g in self._groups(group):
            self.gp[g][0], self.gp[g][1] = lam, myu
        self._clear_values(group)

    def set_rho(self, rho, group=-1):
        for g in self._groups(group):
            self.gp[g][2] = rho
        self._clear_values(group)

    def set_gravitation(self, gx, gy, gz=0.0, group=-1):
        for g in self._groups(group):
            self.gp[g][3], self.gp[g][4] = gx, gy
            if self.type == EQN_SOLID3D:
                self.gp[g][5] = gz
        self._clear_values(group)

    def set_scalar(self, which, val, group=-1):
        for g in self._groups(group):
            self.gp[g][which] = val
        self._clear_values(group)

    def _clear_values(self, group):
        # the system-level setters flag the values; SetEquation (one group) drops the linear system and the
        # preconditioner (CEqnSystem_Solid2D::SetEquation, eqnsys_solid.cpp)
        if group >= 0 and self.type in (EQN_SOLID2D, EQN_SCALAR2D):
            self._clear_ls()
        self.cleared_ls = self.cleared_prec = True

    def _clear_ls(self):
        self.cs, self.f, self.has_prec = None, None, False

    def set_flag(self, which, on):
        t = self.type
        if which == 0:  # geometrical nonlinearity
            if t == EQN_SOLID2D:
                self.nonlin = bool(on)
                if on:
                    self.save = False
                    self.cs = None
            elif t == EQN_SOLID3D:
                if on and not self.nonlin:
                    self.nonlin, self.save = True, False
                elif not on:
                    self.nonlin = False
            self.cleared_ls = self.cleared_prec = True
        elif which == 1:  # stationary
            if t == EQN_SOLID3D:
                if bool(on) != self.stationary:
                    self.stationary = bool(on)
                    if not on:
                        self.save = False
                    self.cleared_ls = self.cleared_prec = True
            else:
                self.stationary = bool(on)
                self._clear_ls()
        elif which == 2:  # saved stiffness
            if t == EQN_SOLID3D:
                if on and not (self.save or self.nonlin or not self.stationary):
                    self.save = True
                    self.cleared_ls = self.cleared_prec = True
                elif not on and self.save:
                    self.save = False
                    self.cleared_ls = self.cleared_prec = True
            else:
                self.save = bool(on) and not (t == EQN_SOLID2D and self.nonlin)
                self._clear_ls()

    def set_time(self, dt, gamma=0.6, beta=-1.0):
        """CEqnSystem::SetTimeIntegrationParameter (include/delfem/eqnsys.h:88-93)"""
        self.dt, self.gamma = dt, gamma
        self.beta = 0.25 * (self.gamma + 0.5) * (self.gamma + 0.5) if beta < 0.0 else beta

    # the linear system, the preconditioner and the element loops
    def _params(self):
        """per element, in the grouped element order"""
        if self.type in (EQN_SCALAR2D, EQN_SCALAR3D):
            gp = np.array(self.gp)
            if self.stationary or self.type == EQN_SCALAR3D:
                return np.stack([gp[self.group, 0], gp[self.group, 2]], 1)  # alpha, source
            return np.stack([gp[self.group, 1], gp[self.group, 0], gp[self.group, 2]], 1)  # capacity, alpha, source
        gp = np.array(self.gp)
        return gp[self.group if self.type == EQN_SOLID2D else np.zeros_like(self.group)]

    def _save_passes(self):
        """CEqn_Solid2D::AddLinSys_Save (src/femeqn/eqnsys_solid.cpp:425-432) passes no element array, so each
        group's equation assembles every element of the field with its own material: ngroup passes over all
        elements, group by group. Returns the number of passes."""
        reps = self.ngroup if (self.type == EQN_SOLID2D and self.save) else 1
        return reps

    def _init_ls(self):
        reps = self._save_passes()
        self.cs = CpuSystem(self.ndim, self.coords, np.tile(self.lnods, reps), self.len)
        self.f, self.has_prec = None, False

    def _make(self, initial):
        """MakeLinearSystem: InitializeMarge, the element loop of the mode, FinalizeMarge; returns ||r||"""
        if self.cs is None:
            self._init_ls()
        cs, t, prm = self.cs, self.type, self._params()
        self.cleared_ls = False
        if self.save:
            if not self.stationary:
                raise NotImplementedError("saved stiffness of a dynamic system (CLinearSystem_SaveDiaM_*)")
            kind = {EQN_SOLID2D: 1, EQN_SOLID3D: 2}.get(t, 0)
            if self._save_passes() > 1:
                prm = np.concatenate([np.tile(np.array(self.gp[g]), (prm.shape[0], 1)) for g in range(self.ngroup)])
            cs.assemble_save(kind, prm, self.flags)
            return 0.0
        cs.zero()
        solid = t in (EQN_SOLID2D, EQN_SOLID3D)
        if solid and not self.stationary:
            glob = (self.dt, self.gamma, self.beta, 1.0 if initial else 0.0)
            if self.nonlin:
                cs.assemble_stvk3d_newmark(prm, self.u, self.v, self.a, glob)
            else:
                cs.assemble_newmark(prm, self.u, self.v, self.a, glob)
        elif solid:
            kind = {(EQN_SOLID2D, False): 1, (EQN_SOLID3D, False): 2, (EQN_SOLID2D, True): 3,
                    (EQN_SOLID3D, True): 4}[(t, self.nonlin)]
            cs.assemble(kind, prm, self.u)
        elif t == EQN_SCALAR2D and not self.stationary:
            cs.assemble(5, prm, self.u, self.v, (self.dt, self.gamma))
        else:
            cs.assemble(0, prm, self.u)
        return cs.set_bc(self.flags)

    def _make_prec(self):
        self.f = self.cs.ilu(None, self.lev)
        self.has_prec = True
        self.cleared_prec = False

    def _pcg(self, res=None):
        ok, conv, it, x, _ = self.cs.solve(self.f, "pcg", 1.0e-6, 1000, res)
        self.itr_.append((it, conv))
        return x

    def _update(self, x, initial=True):
        if self.save:
            _, self.u = self.cs.update_field_save(x, self.u)
        elif self.stationary:
            self.u = self.u + x
        elif self.type in (EQN_SOLID2D, EQN_SOLID3D):
            self.u, self.v, self.a = CpuSystem.update_field_newmark(self.gamma, self.beta, self.dt, initial, x,
                                                                    self.u, self.v, self.a)
        else:
            # UpdateValueOfField_Newmark (src/femls/linearsystem_field.cpp:885-921), VELOCITY
            u = self.u + self.dt * self.v
            self.u = u + (self.gamma * self.dt) * x
            self.v = self.v + 1.0 * x

    def _residual_save(self):
        _, res, _ = self.cs.make_residual(self.u)
        return res

    def solve(self):
        t = self.type
        if t == EQN_SOLID2D or t == EQN_SOLID3D:
            return self._solve_solid(10 if t == EQN_SOLID2D else 40)
        if t == EQN_SCALAR3D:
            if self.cs is None:
                self._init_ls()
            self._make(True)
            self._make_prec()
            x = self._pcg()
            self.itr_ = self.itr_[-1:]
            self.u = self.u + x
            return True
        # CEqnSystem_Scalar2D::Solve (src/femeqn/eqnsys_scalar.cpp:294-359); its MakeLinearSystem also sets the
        # preconditioner's values
        res = None
        if self.cs is None or not self.has_prec:
            self._init_ls()
            self._make(True)
            self._make_prec()
        elif not self.save:
            self._make(True)
            self._make_prec()
        if self.save:
            if self.cleared_ls or self.cleared_prec:
                self._make(True)
                self._make_prec()
            res = self._residual_save()
        x = self._pcg(res)
        self.itr_ = self.itr_[-1:]
        self._update(x, True)
        return True

    def _solve_solid(self, max_newton):
        """CEqnSystem_Solid2D::Solve (src/femeqn/eqnsys_solid.cpp:671-761) and CEqn_Solid3D_Linear::Solve
        (:175-247): a Newton loop for geometrical nonlinearity, else one linear solve"""
        self.itr_ = []
        d3 = self.type == EQN_SOLID3D
        if self.nonlin:
            if self.cs is None or (d3 and not self.has_prec):
                self._init_ls()
            ini = 0.0
            for it in range(max_newton):
                norm = self._make(it == 0)
                self._make_prec()
                if it == 0:
                    if norm < 1.0e-20:
                        break
                    ini = norm
                if norm < ini * 1.0e-6:
                    break
                x = self._pcg()
                self._update(x, it == 0)
            return True
        res = None
        if d3:
            if self.cs is None or not self.has_prec:
                self._init_ls()
                self._make(True)
                self._make_prec()
            elif not self.save:
                self._make(True)
                self._make_prec()
            if self.save:
                res = self._residual_save()
        else:
            if self.cs is None:
                self._init_ls()
                self._make(True)
            elif not self.save:
                self._make(True)
            if self.save:
                if self.cleared_ls:
                    self._make(True)
                res = self._residual_save()
            if not self.has_prec or not self.save or self.cleared_prec:
                self._make_prec()
        x = self._pcg(res)
        self._update(x, True)
        return True

    def itr(self):
        return list(self.itr_)

    def fields(self):
        return self.u.copy(), self.v.copy(), self.a.copy()


# ---------------------------------------------------------------------------------------------------------------
# Multi-segment systems (LsSol::CLinearSystem with several segments, CPreconditioner_ILU over them).
# A system is described by plain arrays: nblk[i], lens[i] per segment and mats = {(i, j): dict(rowptr, colind, val
# [, dia])}; (i, i) is a diagonal segment (off-diagonal blocks in CRS, diagonal blocks in dia). Vectors and Dirichlet
# flags are concatenated over the segments.

def ms_offsets(nblk, lens):
    off = [0]
    for n, l in zip(nblk, lens):
        off.append(off[-1] + int(n) * int(l))
    return off


def _diaind(rowptr, colind):
    n = rowptr.size - 1
    out = np.zeros(n, np.uint32)
    for i in range(n):
        row = colind[rowptr[i]:rowptr[i + 1]]
        out[i] = rowptr[i] + int(np.searchsorted(row, i))
    return out


def ms_copy(mats):
    return {k: {f: np.array(v, copy=True) for f, v in m.items()} for k, m in mats.items()}


class _CrBlk(C.Structure):
    _fields_ = [("nrow", C.c_int), ("ncol", C.c_int), ("bi", C.c_int), ("bj", C.c_int), ("rowptr", C.c_void_p),
                ("colind", C.c_void_p), ("val", C.c_void_p), ("dia", C.c_void_p), ("diaind", C.c_void_p)]


class _CrMsys(C.Structure):
    _fields_ = [("nseg", C.c_int), ("nblk", C.c_int * 8), ("len", C.c_int * 8), ("off", C.c_long * 9),
                ("a", C.POINTER(_CrBlk) * 64), ("f", C.POINTER(_CrBlk) * 64)]


class CpuMultiSeg:
    """The restatement (cpu_ref.c, cr_ms_*) over a multi-segment system; keeps its own copy of the matrices."""

    def __init__(self, nblk, lens, mats):
        self.nblk, self.lens = [int(x) for x in nblk], [int(x) for x in lens]
        self.nseg = len(self.nblk)
        self.off = ms_offsets(self.nblk, self.lens)
        self.m = self.off[-1]
        self.mats = ms_copy(mats)
        for (i, j), a in self.mats.items():
            a["rowptr"] = np.ascontiguousarray(a["rowptr"], np.uint32)
            a["colind"] = np.ascontiguousarray(a["colind"], np.uint32)
            a["val"] = _f64(a["val"])
            if i == j:
                a["dia"] = _f64(a["dia"])
                a["diaind"] = _diaind(a["rowptr"], a["colind"])
        self.fac = {k: {"val": np.zeros_like(a["val"]), "dia": (np.zeros_like(a["dia"]) if "dia" in a else None)}
                    for k, a in self.mats.items()}
        self._keep = []
        self.s = _CrMsys()
        self.s.nseg = self.nseg
        for i in range(self.nseg):
            self.s.nblk[i], self.s.len[i] = self.nblk[i], self.lens[i]
        for i in range(self.nseg + 1):
            self.s.off[i] = self.off[i]
        ptr = lambda x: None if x is None else x.ctypes.data
        for (i, j), a in self.mats.items():
            for which, vals in (("a", a), ("f", self.fac[(i, j)])):
                b = _CrBlk(self.nblk[i], self.nblk[j], self.lens[i], self.lens[j], ptr(a["rowptr"]), ptr(a["colind"]),
                           ptr(vals["val"]), ptr(vals.get("dia")), ptr(a.get("diaind")))
                self._keep.append(b)
                getattr(self.s, which)[i * self.nseg + j] = C.pointer(b)
        lib = cpu_ref()
        lib.cr_ms_set_bc.restype = C.c_double

    def set_bc(self, flags, res):
        """FinalizeMarge: Dirichlet treatment of the matrices (in place); returns (norm, residual)."""
        fl = _i32(flags)
        r = _f64(res).copy()
        nrm = cpu_ref().cr_ms_set_bc(C.byref(self.s), fl.ctypes.data_as(C.c_void_p), r.ctypes.data_as(C.c_void_p))
        return nrm, r

    def matvec(self, alpha, x, beta, y):
        y = _f64(y).copy()
        cpu_ref().cr_ms_matvec(C.byref(self.s), C.c_double(alpha), _f64(x).ctypes.data_as(C.c_void_p),
                               C.c_double(beta), y.ctypes.data_as(C.c_void_p))
        return y

    def ilu0(self):
        return bool(cpu_ref().cr_ms_ilu0(C.byref(self.s)))

    def precond(self, v):
        v = _f64(v).copy()
        cpu_ref().cr_ms_precond(C.byref(self.s), v.ctypes.data_as(C.c_void_p))
        return v

    def solve(self, method, res, tol, max_iter):
        """Solve_PCG ("pcg") / Solve_PBiCGSTAB ("pbicgstab") with the ILU(0) of ilu0(): (ok, conv, iter, x, r)."""
        r = _f64(res).copy()
        x = np.zeros_like(r)
        conv, it = C.c_double(tol), C.c_uint(max_iter)
        ok = cpu_ref().cr_ms_solve(C.byref(self.s), 0 if method == "pcg" else 1, r.ctypes.data_as(C.c_void_p),
                                   x.ctypes.data_as(C.c_void_p), C.byref(conv), C.byref(it))
        return bool(ok), conv.value, it.value, x, r


def _ref_ms_lib():
    lib = reference()
    if not getattr(lib, "_ms_ready", False):
        lib.ref_ms_create.restype = C.c_void_p
        lib.ref_ms_hyper2d.restype = C.c_void_p
        lib.ref_ms_hyper2d.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p] + [C.c_double] * 5 + \
            [C.c_void_p]
        lib.ref_ms_destroy.argtypes = [C.c_void_p]
        lib.ref_ms_ncrs.restype = C.c_long
        lib.ref_ms_ncrs.argtypes = [C.c_void_p, C.c_int, C.c_int]
        lib.ref_ms_finalize.restype = C.c_double
        lib.ref_ms_finalize.argtypes = [C.c_void_p, C.c_void_p]
        lib.ref_ms_matvec.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_double, C.c_int]
        lib.ref_ms_dot.restype = C.c_double
        lib.ref_ms_dot.argtypes = [C.c_void_p, C.c_int, C.c_int]
        lib._ms_ready = True
    return lib


class RefMultiSeg:
    """The reference's own multi-segment CLinearSystem (oracle/ref_msys.cpp)."""

    def __init__(self, h):
        self.lib = _ref_ms_lib()
        self.h = C.c_void_p(h)
        nblk = np.zeros(8, np.int32)
        lens = np.zeros(8, np.int32)
        self.nseg = self.lib.ref_ms_info(self.h, nblk.ctypes.data_as(C.c_void_p), lens.ctypes.data_as(C.c_void_p))
        self.nblk = [int(x) for x in nblk[:self.nseg]]
        self.lens = [int(x) for x in lens[:self.nseg]]
        self.off = ms_offsets(self.nblk, self.lens)
        self.m = self.off[-1]

    @classmethod
    def create(cls, nblk, lens, patterns):
        """patterns: {(i, j): (rowptr, colind)}; (i, i) the off-diagonal pattern of a diagonal segment."""
        lib = _ref_ms_lib()
        nb, ln = _i32(nblk), _i32(lens)
        h = lib.ref_ms_create(len(nblk), nb.ctypes.data_as(C.c_void_p), ln.ctypes.data_as(C.c_void_p))
        self = cls(h)
        for (i, j), (rp, ci) in sorted(patterns.items()):
            rp = np.ascontiguousarray(rp, np.uint32)
            ci = np.ascontiguousarray(ci, np.uint32)
            lib.ref_ms_add_pattern(self.h, i, j, rp.ctypes.data_as(C.c_void_p), ci.ctypes.data_as(C.c_void_p))
        lib.ref_ms_initialize(self.h)
        return self

    @classmethod
    def hyper2d(cls, coords, lnods, fixed, c1, c2, rho, gx, gy):
        lib = _ref_ms_lib()
        co, ln, fx = _f64(coords), _i32(lnods), _i32(fixed)
        nrm = C.c_double(0.0)
        h = lib.ref_ms_hyper2d(co.size // 2, co.ctypes.data_as(C.c_void_p), ln.size // 3,
                               ln.ctypes.data_as(C.c_void_p), fx.ctypes.data_as(C.c_void_p), c1, c2, rho, gx, gy,
                               C.byref(nrm))
        self = cls(h)
        self.res_norm = nrm.value
        return self

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.ref_ms_destroy(self.h)
            self.h = None

    def pairs(self):
        return [(i, j) for i in range(self.nseg) for j in range(self.nseg) if self.lib.ref_ms_ncrs(self.h, i, j) >= 0]

    def matrices(self):
        """{(i, j): dict(rowptr, colind, val[, dia])} of the current values."""
        out = {}
        for (i, j) in self.pairs():
            ncrs = self.lib.ref_ms_ncrs(self.h, i, j)
            rp = np.zeros(self.nblk[i] + 1, np.uint32)
            ci = np.zeros(max(ncrs, 1), np.uint32)
            self.lib.ref_ms_get_pattern(self.h, i, j, rp.ctypes.data_as(C.c_void_p), ci.ctypes.data_as(C.c_void_p))
            bs = self.lens[i] * self.lens[j]
            val = np.zeros(max(ncrs, 1) * bs)
            dia = np.zeros(self.nblk[i] * bs) if i == j else None
            self.lib.ref_ms_get_values(self.h, i, j, val.ctypes.data_as(C.c_void_p),
                                       None if dia is None else dia.ctypes.data_as(C.c_void_p))
            m = {"rowptr": rp, "colind": ci[:ncrs], "val": val[:ncrs * bs]}
            if dia is not None:
                m["dia"] = dia
            out[(i, j)] = m
        return out

    def set_values(self, mats):
        for (i, j), m in mats.items():
            val = _f64(m["val"]) if m["val"].size else np.zeros(1)
            dia = _f64(m["dia"]) if i == j else None
            self.lib.ref_ms_set_values(self.h, i, j, val.ctypes.data_as(C.c_void_p),
                                       None if dia is None else dia.ctypes.data_as(C.c_void_p))

    def set_vector(self, iv, v):
        v = _f64(v)
        self.lib.ref_ms_set_vector(self.h, iv, v.ctypes.data_as(C.c_void_p))

    def vector(self, iv):
        v = np.zeros(self.m)
        self.lib.ref_ms_get_vector(self.h, iv, v.ctypes.data_as(C.c_void_p))
        return v

    def finalize(self, flags):
        fl = _i32(flags)
        return self.lib.ref_ms_finalize(self.h, fl.ctypes.data_as(C.c_void_p))

    def factor(self):
        """ILU(0) factor values as CPreconditioner_ILU computes them: (ok, {(i, j): val}, [dia per segment])."""
        pairs = self.pairs()
        ns = self.nseg
        vals = {k: np.zeros(max(self.lib.ref_ms_ncrs(self.h, *k), 1) * self.lens[k[0]] * self.lens[k[1]])
                for k in pairs}
        dias = [np.zeros(self.nblk[i] * self.lens[i] ** 2) for i in range(ns)]
        vp = (C.c_void_p * (ns * ns))()
        dp = (C.c_void_p * ns)()
        for (i, j), v in vals.items():
            vp[i * ns + j] = v.ctypes.data
        for i in range(ns):
            dp[i] = dias[i].ctypes.data
        ok = self.lib.ref_ms_factor(self.h, vp, dp)
        for k in pairs:
            vals[k] = vals[k][:self.lib.ref_ms_ncrs(self.h, *k) * self.lens[k[0]] * self.lens[k[1]]]
        return bool(ok), vals, dias

    def precond(self, v):
        self.set_vector(0, v)
        ok = self.lib.ref_ms_precond(self.h, 0)
        assert ok
        return self.vector(0)

    def matvec(self, alpha, x, beta, y):
        self.set_vector(0, x)
        self.set_vector(1, y)
        self.lib.ref_ms_matvec(self.h, alpha, 0, beta, 1)
        return self.vector(1)

    def solve(self, method, res, tol, max_iter):
        """(ok, conv, iter, x, r) of Solve_PCG / Solve_PBiCGSTAB from the residual res."""
        self.set_vector(-1, res)
        conv, it = C.c_double(tol), C.c_uint(max_iter)
        ok = self.lib.ref_ms_solve(self.h, 0 if method == "pcg" else 1, C.byref(conv), C.byref(it))
        return ok == 1, conv.value, it.value, self.vector(-2), self.vector(-1)
