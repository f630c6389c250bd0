"""Equation-system scenarios shared by the oracle pin (restated Solve flows against the reference's own
CEqnSystem classes) and the GPU tests (the library's equation systems against the restated flows). Each scenario
builds a mesh of element groups with a fixed boundary and drives an object with the RefEqn / EqnFlow / eqn
interface: setters, then a few Solve calls; it returns the number of Solve calls made."""
import numpy as np

import meshgen as M
from oracle.pyref import EQN_SCALAR2D, EQN_SCALAR3D, EQN_SOLID2D, EQN_SOLID3D


def mesh2d(nx=10, ny=4):
    co, tri = M.rect_tri(nx, ny, 1.0, 0.4, 51)
    t = tri.reshape(-1, 3)
    cx = co.reshape(-1, 2)[t, 0].mean(1)
    group = (cx > 0.5).astype(np.int32)  # two element arrays (materials)
    left = np.arange(ny + 1) * (nx + 1)
    bnd = np.stack([left[:-1], left[1:]], 1).ravel().astype(np.int32)  # the x = 0 edge as lines
    return co, tri, group, bnd


def mesh3d(nx=4, ny=3, nz=3):
    co, tet = M.box_tet(nx, ny, nz, 1.0, 0.6, 0.6, 52)
    bnd = M.box_face_tris(nx, ny, nz, 0, 0).ravel()  # the x = 0 face as triangles
    return co, tet, None, bnd


def _solid2d(obj_cls, mode):
    co, tri, group, bnd = mesh2d()
    e = obj_cls(EQN_SOLID2D, co, tri, group, 2, bnd, -1)
    e.set_young_poisson(1.0, 0.3, True, 0)   # plane strain lambda (the inverted flag)
    e.set_young_poisson(3.0, 0.25, False, 1)  # plane stress lambda
    e.set_gravitation(0.0, -0.02)
    return e


SCENARIOS = {}


def scenario(f):
    SCENARIOS[f.__name__] = f
    return f


@scenario
def solid2d_static(cls):
    e = _solid2d(cls, 0)
    e.set_flag(1, True)
    e.solve()
    e.set_gravitation(0.01, -0.03)
    e.solve()
    return e


@scenario
def solid2d_save(cls):
    e = _solid2d(cls, 0)
    e.set_flag(1, True)
    e.set_flag(2, True)
    e.solve()
    e.solve()  # the saved matrix, a new residual
    e.set_gravitation(0.0, -0.05)  # values cleared: the matrix is made again
    e.solve()
    return e


@scenario
def solid2d_nonlinear(cls):
    e = _solid2d(cls, 0)
    e.set_flag(1, True)
    e.set_flag(0, True)
    e.set_gravitation(0.0, -0.2)
    e.solve()
    return e


@scenario
def solid2d_dynamic(cls):
    e = _solid2d(cls, 0)
    e.set_flag(1, False)
    e.set_time(0.05)
    for _ in range(3):
        e.solve()
    return e


@scenario
def solid2d_dynamic_nonlinear(cls):
    e = _solid2d(cls, 0)
    e.set_flag(1, False)
    e.set_flag(0, True)
    e.set_gravitation(0.0, -0.2)
    e.set_time(0.05, 0.7)
    for _ in range(2):
        e.solve()
    return e


def _solid3d(cls):
    co, tet, _, bnd = mesh3d()
    e = cls(EQN_SOLID3D, co, tet, None, 1, bnd, -1)
    e.set_young_poisson(2.0, 0.3)
    e.set_gravitation(0.0, 0.0, -0.02)
    return e


@scenario
def solid3d_static(cls):
    e = _solid3d(cls)
    e.set_flag(1, True)
    e.solve()
    return e


@scenario
def solid3d_save(cls):
    e = _solid3d(cls)
    e.set_flag(1, True)
    e.set_flag(2, True)
    e.solve()
    e.solve()
    return e


@scenario
def solid3d_nonlinear(cls):
    e = _solid3d(cls)
    e.set_flag(1, True)
    e.set_flag(0, True)
    e.set_gravitation(0.0, 0.0, -0.3)
    e.solve()
    return e


@scenario
def solid3d_dynamic(cls):
    e = _solid3d(cls)  # not stationary: the constructor's default
    e.set_time(0.05)
    for _ in range(3):
        e.solve()
    return e


@scenario
def scalar2d_static(cls):
    co, tri, group, bnd = mesh2d()
    e = cls(EQN_SCALAR2D, co, tri, group, 2, bnd, -1)
    e.set_flag(1, True)
    e.set_scalar(0, 1.5, 0)
    e.set_scalar(0, 0.4, 1)
    e.set_scalar(2, 1.0)
    e.solve()
    return e


@scenario
def scalar2d_save(cls):
    co, tri, group, bnd = mesh2d()
    e = cls(EQN_SCALAR2D, co, tri, group, 2, bnd, -1)
    e.set_flag(1, True)
    e.set_flag(2, True)
    e.set_scalar(2, 1.0)
    e.solve()
    e.solve()
    return e


@scenario
def scalar2d_dynamic(cls):
    co, tri, group, bnd = mesh2d()
    e = cls(EQN_SCALAR2D, co, tri, group, 2, bnd, -1)
    e.set_flag(1, False)
    e.set_scalar(1, 2.0, 1)
    e.set_scalar(2, 1.0)
    e.set_time(0.02)
    for _ in range(3):
        e.solve()
    return e


@scenario
def scalar3d_static(cls):
    co, tet, _, bnd = mesh3d()
    e = cls(EQN_SCALAR3D, co, tet, None, 1, bnd, -1)
    e.set_scalar(0, 0.7)
    e.set_scalar(2, 1.0)
    e.solve()
    return e
