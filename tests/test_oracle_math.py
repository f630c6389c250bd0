"""What the mathematics fixes about the oracle, independently of the reference: closed-form element matrices,
rigid-body null spaces, zero row sums, symmetry, patch tests, the StVK tangent as the derivative of the
residual, exact factorisation of matrices whose ILU(0) has no dropped fill, PCG with an exact preconditioner,
and the consistent surface loads (the one oracle function with no reference counterpart)."""
import numpy as np
import pytest

import meshgen as M
from oracle import pyref as P


def dense(cpu):
    n, b = cpu.nnode, cpu.len
    A = np.zeros((n * b, n * b))
    for i in range(n):
        A[i * b:(i + 1) * b, i * b:(i + 1) * b] = cpu.dia[i * b * b:(i + 1) * b * b].reshape(b, b)
        for q in range(cpu.rowptr[i], cpu.rowptr[i + 1]):
            j = cpu.colind[q]
            A[i * b:(i + 1) * b, j * b:(j + 1) * b] = cpu.crs[q * b * b:(q + 1) * b * b].reshape(b, b)
    return A


def test_poisson_row_sums_zero_and_symmetry():
    co, tri = M.rect_tri(6, 5, 1.0, 1.0, 3)
    cpu = P.CpuSystem(2, co, tri, 1)
    cpu.assemble(P.POISSON, [1.7, 0.0])
    A = dense(cpu)
    assert np.allclose(A.sum(axis=1), 0.0, atol=1e-13)
    assert np.allclose(A, A.T, atol=1e-14)


def test_poisson_closed_form_single_triangle():
    co = np.array([0.0, 0.0, 2.0, 0.0, 0.0, 1.0])
    cpu = P.CpuSystem(2, co, np.array([0, 1, 2]), 1)
    cpu.assemble(P.POISSON, [3.0, 6.0])
    A = dense(cpu)
    # grad N: N0 = 1 - x/2 - y, N1 = x/2, N2 = y; area 1; K = alpha*area*G G^T
    G = np.array([[-0.5, -1.0], [0.5, 0.0], [0.0, 1.0]])
    assert np.allclose(A, 3.0 * 1.0 * G @ G.T, rtol=1e-15)
    assert np.allclose(cpu.res, 6.0 * 1.0 / 3.0)


def _rigid_modes_2d(co):
    x, y = co[0::2], co[1::2]
    t1 = np.stack([np.ones_like(x), np.zeros_like(x)], 1).ravel()
    t2 = np.stack([np.zeros_like(x), np.ones_like(x)], 1).ravel()
    r = np.stack([-y, x], 1).ravel()
    return [t1, t2, r]


def _rigid_modes_3d(co):
    p = co.reshape(-1, 3)
    modes = []
    for d in range(3):
        t = np.zeros_like(p)
        t[:, d] = 1
        modes.append(t.ravel())
    for (a, b) in [(0, 1), (1, 2), (2, 0)]:
        r = np.zeros_like(p)
        r[:, a] = -p[:, b]
        r[:, b] = p[:, a]
        modes.append(r.ravel())
    return modes


@pytest.mark.parametrize("ndim", [2, 3])
def test_linear_elasticity_rigid_null_space(ndim):
    if ndim == 2:
        co, ln = M.rect_tri(5, 4, 1.0, 0.5, 9)
        cpu = P.CpuSystem(2, co, ln, 2)
        cpu.assemble(P.SOLID2D, [0.4, 0.3, 0.0, 0.0, 0.0])
        modes = _rigid_modes_2d(co)
    else:
        co, ln = M.box_tet(3, 2, 2, 1.0, 0.8, 0.6, 9)
        cpu = P.CpuSystem(3, co, ln, 3)
        cpu.assemble(P.SOLID3D, [0.4, 0.3, 0.0, 0.0, 0.0, 0.0])
        modes = _rigid_modes_3d(co)
    A = dense(cpu)
    assert np.allclose(A, A.T, atol=1e-13)
    for mvec in modes:
        assert np.abs(A @ mvec).max() < 1e-12
    w = np.linalg.eigvalsh(A)
    assert np.sum(np.abs(w) < 1e-10) == len(modes)  # exactly the rigid modes


def test_tet_element_matrix_closed_form():
    """Ke = V B^T D B for a single tetrahedron (Voigt form), against the restatement."""
    co = np.array([0, 0, 0, 1.0, 0, 0, 0, 2.0, 0, 0, 0, 3.0])
    lam, mu = 0.7, 0.4
    cpu = P.CpuSystem(3, co, np.array([0, 1, 2, 3]), 3)
    cpu.assemble(P.SOLID3D, [lam, mu, 0.0, 0.0, 0.0, 0.0])
    A = dense(cpu)
    p = co.reshape(4, 3)
    T = np.hstack([np.ones((4, 1)), p])
    C = np.linalg.inv(T)
    dN = C[1:, :].T  # gradients of the barycentric functions, rows per node
    V = abs(np.linalg.det(T)) / 6
    B = np.zeros((6, 12))
    for a in range(4):
        gx, gy, gz = dN[a]
        B[:, 3 * a:3 * a + 3] = [[gx, 0, 0], [0, gy, 0], [0, 0, gz], [gy, gx, 0], [0, gz, gy], [gz, 0, gx]]
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[np.arange(3), np.arange(3)] += 2 * mu
    D[3:, 3:] = np.eye(3) * mu
    # the reference's TetVolume uses 0.1666666666666 instead of 1/6: a 4e-13 relative bias
    assert np.allclose(A, V * B.T @ D @ B, rtol=1e-11, atol=1e-12)


def test_patch_test_linear_field_zero_interior_residual():
    """A linear displacement field solves the homogeneous equations: interior residuals vanish."""
    co, ln = M.rect_tri(6, 6, 1.0, 1.0, 21)
    cpu = P.CpuSystem(2, co, ln, 2)
    x, y = co[0::2], co[1::2]
    u = np.stack([0.01 * x + 0.02 * y, -0.03 * x + 0.005 * y], 1).ravel()
    cpu.assemble(P.SOLID2D, [0.5, 0.4, 0.0, 0.0, 0.0], u=u)
    ix, iy = np.arange(49) % 7, np.arange(49) // 7
    inner = (ix > 0) & (ix < 6) & (iy > 0) & (iy < 6)
    assert np.abs(cpu.res.reshape(-1, 2)[inner]).max() < 1e-13


@pytest.mark.parametrize("ndim", [2, 3])
def test_stvk_tangent_is_residual_derivative(ndim):
    """K(u) = -d r(u) / d u (central differences, O(h^2)); at u = 0 StVK equals linear elasticity."""
    if ndim == 2:
        co, ln = M.rect_tri(3, 2, 1.0, 0.6, 5)
        kind, lin, par = P.STVK2D, P.SOLID2D, [0.5, 0.4, 1.0, 0.0, -0.1]
    else:
        co, ln = M.box_tet(2, 1, 1, 1.0, 0.6, 0.5, 5)
        kind, lin, par = P.STVK3D, P.SOLID3D, [0.5, 0.4, 1.0, 0.0, 0.0, -0.1]
    length = ndim
    n = co.size // ndim
    rng = np.random.default_rng(3)
    u = rng.uniform(-0.05, 0.05, n * length)
    cpu = P.CpuSystem(ndim, co, ln, length)
    cpu.assemble(kind, par, u=u)
    K = dense(cpu)
    h = 1e-6
    J = np.zeros_like(K)
    for k in range(n * length):
        e = np.zeros(n * length)
        e[k] = h
        cpu.zero()
        cpu.assemble(kind, par, u=u + e)
        rp = cpu.res.copy()
        cpu.zero()
        cpu.assemble(kind, par, u=u - e)
        J[:, k] = -(rp - cpu.res) / (2 * h)
    assert np.abs(K - J).max() < 1e-6 * np.abs(K).max()
    cpu.zero()
    cpu.assemble(kind, par, u=np.zeros(n * length))
    c2 = P.CpuSystem(ndim, co, ln, length)
    c2.assemble(lin, par, u=np.zeros(n * length))
    assert np.allclose(dense(cpu), dense(c2), rtol=1e-13, atol=1e-14)


def test_ilu0_exact_when_no_fill_is_dropped():
    """A single tetrahedron has a complete node graph, so ILU(0) drops no fill and equals the exact block LU:
    one application solves the system, and PCG converges at its first iteration (reported 0-based)."""
    co = np.array([0, 0, 0, 1.0, 0.1, 0, 0.2, 1.1, 0, 0.1, 0.3, 0.9])
    cpu = P.CpuSystem(3, co, np.array([0, 1, 2, 3]), 1)
    cpu.assemble(P.POISSON, [1.3, 1.0])
    flags = np.zeros(4, np.int32)
    flags[0] = 1
    cpu.set_bc(flags)
    A = dense(cpu)
    assert np.linalg.cond(A) < 1e3
    f = cpu.ilu0()
    b = np.random.default_rng(1).normal(size=4)
    z = cpu.ilu_apply(f, b)
    assert np.allclose(A @ z, b, rtol=1e-12, atol=1e-12)
    ok, conv, it, x, _ = cpu.solve(f, "pcg", 1e-12, 10, res=b)
    assert ok and it == 0
    assert np.allclose(A @ x, b, atol=1e-10)


def test_pcg_terminates_within_n_on_tiny_spd():
    co, tri = M.rect_tri(3, 3, 1.0, 1.0, 2)
    cpu = P.CpuSystem(2, co, tri, 1)
    cpu.assemble(P.POISSON, [1.0, 1.0])
    flags = np.zeros(16, np.int32)
    flags[[0, 3, 12, 15]] = 1
    cpu.set_bc(flags)
    order = np.random.default_rng(0).permutation(16).astype(np.int32)
    f = cpu.ilu0(order)
    ok, conv, it, x, _ = cpu.solve(f, "pcg", 1e-12, 16)
    A = dense(cpu)
    assert ok and it < 16
    assert np.allclose(A @ x, cpu.res, atol=1e-10)
    xs = np.linalg.solve(A, cpu.res)
    assert np.allclose(x, xs, rtol=1e-9, atol=1e-12)
    ok, conv, it, x2, _ = cpu.solve(f, "pbicgstab", 1e-12, 40)
    assert np.allclose(x2, xs, rtol=1e-8, atol=1e-11)


def test_dirichlet_rows_and_columns():
    co, tri = M.rect_tri(4, 3, 1.0, 1.0, 2)
    cpu = P.CpuSystem(2, co, tri, 2)
    cpu.assemble(P.SOLID2D, [0.5, 0.3, 1.0, 0.0, -1.0])
    flags = np.zeros((20, 2), np.int32)
    flags[0, :] = 1
    flags[5, 1] = 1
    cpu.set_bc(flags.ravel())
    A = dense(cpu)
    for k in np.flatnonzero(flags.ravel()):
        e = np.zeros(40)
        e[k] = 1
        assert np.array_equal(A[k], e) and np.array_equal(A[:, k], e)
        assert cpu.res[k] == 0.0


def test_traction_loads_closed_form():
    """Consistent P1 loads t*A/3 per node: they sum to t times the loaded area and split equally."""
    nx = 3
    co, tet = M.box_tet(nx, nx, nx, 1.0, 1.0, 1.0, 4, 0.0)
    faces = M.box_face_tris(nx, nx, nx, 2, 1)
    t = np.tile([0.1, -0.2, 0.3], faces.shape[0])
    cpu = P.CpuSystem(3, co, tet, 3)
    cpu.traction(faces.ravel(), t)
    tot = cpu.res.reshape(-1, 3).sum(axis=0)
    assert np.allclose(tot, [0.1, -0.2, 0.3], rtol=1e-14)
    # a corner node of the diagonal (shared by 2 triangles of one quad) gets 2 * A/3 * t
    a = 0.5 / nx / nx
    top_corner = (nx * (nx + 1) + 0) * (nx + 1) + 0 + 0  # node (0,0,nx)
    assert np.allclose(cpu.res.reshape(-1, 3)[top_corner], np.array([0.1, -0.2, 0.3]) * 2 * a / 3)


# ---- saved stiffness: what the mathematics fixes about MakeResidual and the saved solve --------------------
def _save_setup(ndim, length, kind, par, seed=3):
    co, ln = (M.rect_tri(7, 4, 1.0, 0.4, seed) if ndim == 2 else M.box_tet(3, 2, 2, 1.0, 0.7, 0.6, seed))
    x = co.reshape(-1, ndim)[:, 0]
    f = np.zeros((x.size, length), np.int32)
    f[x == x.min()] = 1
    f[x == x.max(), 0] = 1
    flags = f.ravel()
    rng = np.random.default_rng(seed)
    u = rng.uniform(-0.05, 0.05, flags.size)
    u[flags == 1] = rng.uniform(0.01, 0.1, int(flags.sum()))
    full = P.CpuSystem(ndim, co, ln, length)  # the full matrix, no Dirichlet treatment
    full.assemble(kind, par)
    cpu = P.CpuSystem(ndim, co, ln, length)
    cpu.assemble_save(kind, par, flags)
    return co, ln, flags, u, full, cpu


SAVE_MATH = [(2, 1, P.POISSON, [1.3, 0.8]), (2, 2, P.SOLID2D, [0.9, 0.4, 1.0, 0.1, -0.2]),
             (3, 3, P.SOLID3D, [0.6, 0.4, 1.0, 0.0, 0.2, -0.1])]


@pytest.mark.parametrize("ndim,length,kind,par", SAVE_MATH, ids=["poisson2d", "solid2d", "solid3d"])
def test_saved_residual_is_coupling_to_prescribed_values(ndim, length, kind, par):
    """MakeResidual: r_free = f_free - K[free, fixed] u_fixed and r_fixed = f_fixed, with K the full stiffness
    (a transposed coupling, a dropped diagonal block or a free column left in would break it)."""
    _, _, flags, u, full, cpu = _save_setup(ndim, length, kind, par)
    K = dense(full)
    fr, fx = flags == 0, flags == 1
    _, res, upd = cpu.make_residual(u)
    want = cpu.force.copy()
    want[fr] -= K[np.ix_(fr, fx)] @ u[fx]
    assert np.allclose(res, want, rtol=1e-13, atol=1e-15 * np.abs(cpu.force).max())
    assert np.array_equal(res[fx], cpu.force[fx])
    assert np.array_equal(upd, u)


@pytest.mark.parametrize("ndim,length,kind,par", SAVE_MATH, ids=["poisson2d", "solid2d", "solid3d"])
def test_saved_force_is_the_static_load(ndim, length, kind, par):
    """The saved force is the external load: the static loop's residual at u = 0 (same expression there)."""
    _, _, _, _, full, cpu = _save_setup(ndim, length, kind, par)
    assert np.allclose(cpu.force, full.res, rtol=1e-15, atol=0.0)
    # the saved matrix is the static one (2D elasticity writes its element matrix in another order: rounding)
    K_save, K = dense(cpu), dense(full)
    fr = np.flatnonzero(cpu.flags == 0)
    assert np.allclose(K_save[np.ix_(fr, fr)], K[np.ix_(fr, fr)], rtol=1e-14, atol=1e-14 * np.abs(K).max())


@pytest.mark.parametrize("ndim,length,kind,par", SAVE_MATH, ids=["poisson2d", "solid2d", "solid3d"])
def test_saved_solve_equals_incremental_solve(ndim, length, kind, par):
    """Solving K_ff u_f = f_f - K_fb u_b and setting u = update gives the field of the incremental formulation
    (r = f - K u, K du = r, u += du) and the direct dense solve."""
    co, ln, flags, u, full, cpu = _save_setup(ndim, length, kind, par)
    cpu.make_residual(u)
    f = cpu.ilu(None, 0)
    ok, _, _, x, _ = cpu.solve(f, "pcg", 1e-13, 2000)
    assert ok
    _, u_save = cpu.update_field_save(x, u)
    inc = P.CpuSystem(ndim, co, ln, length)
    inc.assemble(kind, par, u)
    inc.set_bc(flags)
    ok, _, _, du, _ = inc.solve(inc.ilu0(), "pcg", 1e-13, 2000)
    assert ok
    u_inc = u + du
    K = dense(full)
    fr, fx = flags == 0, flags == 1
    u_dir = u.copy()
    u_dir[fr] = np.linalg.solve(K[np.ix_(fr, fr)], cpu.force[fr] - K[np.ix_(fr, fx)] @ u[fx])
    scale = np.abs(u_dir).max()
    assert np.allclose(u_save, u_dir, rtol=0, atol=1e-9 * scale)
    assert np.allclose(u_inc, u_dir, rtol=0, atol=1e-9 * scale)
    assert np.array_equal(u_save[fx], u[fx])


# ---- Newmark-beta: what the mathematics fixes -------------------------------------------------------------------
def _nm_system(ndim, seed=4):
    co, ln = (M.rect_tri(6, 3, 1.0, 0.3, seed) if ndim == 2 else M.box_tet(3, 2, 2, 1.0, 0.6, 0.5, seed))
    return co, ln, co.size // ndim


def _nm_dense(co, ln, ndim, par, glob, u=None, v=None, a=None, flags=None):
    n = co.size // ndim
    z = np.zeros(n * ndim)
    cpu = P.CpuSystem(ndim, co, ln, ndim)
    cpu.assemble_newmark(par, z if u is None else u, z if v is None else v, z if a is None else a, glob)
    if flags is not None:
        cpu.set_bc(flags)
    return cpu, dense(cpu)


@pytest.mark.parametrize("ndim", [2, 3])
def test_newmark_mass_and_stiffness_parts(ndim):
    """beta = 0: the matrix is the consistent mass (rho w / 12 (2D) or / 20 (3D) per node pair, doubled on the
    diagonal; total mass rho * |domain| per direction); rho = 0, beta dt^2 = 1: it is the stiffness, entry for
    entry the saved loops' K; u = v = a = 0: the residual is the body force."""
    co, ln, n = _nm_system(ndim)
    g = [0.0, -0.3] if ndim == 2 else [0.1, 0.0, -0.3]
    rho = 1.7
    cpu, Mm = _nm_dense(co, ln, ndim, [0.8, 0.5, rho] + g, (0.1, 0.5, 0.0, 1.0))
    vol = sum(abs(np.linalg.det(np.column_stack([co.reshape(-1, ndim)[e[k]] - co.reshape(-1, ndim)[e[0]]
                                                    for k in range(1, ndim + 1)]))) / (2 if ndim == 2 else 6)
              for e in ln.reshape(-1, ndim + 1))
    for d in range(ndim):
        assert np.isclose(Mm[d::ndim, d::ndim].sum(), rho * vol, rtol=1e-12)
    assert np.allclose(Mm, Mm.T, atol=1e-15)
    e0 = ln.reshape(-1, ndim + 1)[0]
    one = P.CpuSystem(ndim, co, e0, ndim)
    one.assemble_newmark([0.8, 0.5, rho] + g, np.zeros(n * ndim), np.zeros(n * ndim), np.zeros(n * ndim),
                         (0.1, 0.5, 0.0, 1.0))
    M1 = dense(one)[np.ix_(np.repeat(e0, ndim) * ndim + np.tile(np.arange(ndim), ndim + 1),
                           np.repeat(e0, ndim) * ndim + np.tile(np.arange(ndim), ndim + 1))]
    w = abs(np.linalg.det(np.column_stack([co.reshape(-1, ndim)[e0[k]] - co.reshape(-1, ndim)[e0[0]]
                                           for k in range(1, ndim + 1)]))) / (2 if ndim == 2 else 6)
    want = np.kron(rho * w / (12 if ndim == 2 else 20) * (np.ones((ndim + 1, ndim + 1)) + np.eye(ndim + 1)),
                   np.eye(ndim))
    assert np.allclose(M1, want, rtol=1e-12)
    body = np.zeros(n * ndim)
    for e in ln.reshape(-1, ndim + 1):
        we = abs(np.linalg.det(np.column_stack([co.reshape(-1, ndim)[e[k]] - co.reshape(-1, ndim)[e[0]]
                                                for k in range(1, ndim + 1)]))) / (2 if ndim == 2 else 6)
        for k in e:
            body[k * ndim:(k + 1) * ndim] += we * rho * np.array(g) / (ndim + 1)
    assert np.allclose(cpu.res, body, rtol=1e-12, atol=1e-15)
    # stiffness part equals the saved loops' K bit for bit
    kk, K = _nm_dense(co, ln, ndim, [0.8, 0.5, 0.0] + g, (1.0, 0.5, 1.0, 1.0))
    sv = P.CpuSystem(ndim, co, ln, ndim)
    sv.assemble_save(P.SOLID2D if ndim == 2 else P.SOLID3D, [0.8, 0.5, 0.0] + g, np.zeros(n * ndim, np.int32))
    assert np.array_equal(kk.crs, sv.crs) and np.array_equal(kk.dia, sv.dia)


@pytest.mark.parametrize("ndim", [2, 3])
def test_newmark_average_acceleration_conserves_energy(ndim):
    """beta = 1/4, gamma = 1/2 on an undamped linear system: E = v'Mv/2 + u'Ku/2 is constant from step to step
    (consistent initial acceleration), and the update of UpdateValueOfField_NewmarkBeta with the solved increment
    satisfies M a + K u = 0 after every step."""
    co, ln, n = _nm_system(ndim, 6)
    m = n * ndim
    x = co.reshape(-1, ndim)
    flags = np.zeros((n, ndim), np.int32)
    flags[x[:, 0] == x[:, 0].min()] = 1
    flags = flags.ravel()
    par = [0.9, 0.6, 1.3] + [0.0] * ndim
    dt, gamma, beta = 0.02, 0.5, 0.25
    _, K = _nm_dense(co, ln, ndim, par[:2] + [0.0] + par[3:], (1.0, gamma, 1.0, 1.0))
    _, Mm = _nm_dense(co, ln, ndim, par, (dt, gamma, 0.0, 1.0))
    fr = flags == 0
    u = np.zeros(m)
    u[fr] = 0.01 * np.sin(3 * x.repeat(ndim, axis=0)[fr, 0])
    v = np.zeros(m)
    a = np.zeros(m)
    a[fr] = np.linalg.solve(Mm[np.ix_(fr, fr)], -(K[np.ix_(fr, fr)] @ u[fr]))
    energy = lambda u, v: 0.5 * v @ Mm @ v + 0.5 * u @ K @ u
    e0 = energy(u, v)
    for _ in range(20):
        cpu, A = _nm_dense(co, ln, ndim, par, (dt, gamma, beta, 1.0), u, v, a, flags)
        du = np.linalg.solve(A, cpu.res)
        u, v, a = P.CpuSystem.update_field_newmark(gamma, beta, dt, True, du, u, v, a)
        assert np.allclose((Mm @ a + K @ u)[fr], 0.0, atol=1e-10 * np.abs(K @ u).max())
        assert abs(energy(u, v) - e0) <= 1e-10 * e0
    assert np.abs(v).max() > 0


def test_beuler_step_is_implicit_euler():
    """The backward-Euler loop's system M dv = dt f - dt K u - dt^2 K v (is_initial) with u += dt v_new is one
    implicit Euler step of M v' = f - K u, u' = v: M (v1 - v) = dt (f - K u1) with u1 = u + dt v1."""
    co, ln = M.rect_tri(5, 3, 1.0, 0.3, 9)
    n = co.size // 2
    par = [0.7, 0.4, 1.3, 0.0, -0.2]
    rng = np.random.default_rng(1)
    u, v = rng.uniform(-0.01, 0.01, 2 * n), rng.uniform(-0.01, 0.01, 2 * n)
    dt = 0.05
    cpu = P.CpuSystem(2, co, ln, 2)
    cpu.assemble_beuler(par, u, v, np.zeros(2 * n), dt, True)
    A = dense(cpu)
    du = np.linalg.solve(A, cpu.res)
    u1, v1 = P.CpuSystem.update_field_beuler(dt, True, du, u, v)
    cm = P.CpuSystem(2, co, ln, 2)
    cm.assemble_newmark(par, np.zeros(2 * n), np.zeros(2 * n), np.zeros(2 * n), (dt, 0.5, 0.0, 1.0))
    Mm, f = dense(cm), cm.res
    ck = P.CpuSystem(2, co, ln, 2)
    ck.assemble_newmark(par[:2] + [0.0, 0.0, 0.0], np.zeros(2 * n), np.zeros(2 * n), np.zeros(2 * n), (1.0, 0.5, 1.0, 1.0))
    K = dense(ck)
    assert np.allclose(u1, u + dt * v1, rtol=0, atol=1e-15)
    lhs, rhs = Mm @ (v1 - v), dt * (f - K @ u1)
    assert np.allclose(lhs, rhs, rtol=0, atol=1e-12 * np.abs(rhs).max())


def test_stvk3d_newmark_small_strain_limit():
    """At vanishing displacement the dynamic StVK loop is the linear Newmark loop: same matrix, same residual (to
    rounding; the tangent and the linear stiffness are written differently)."""
    co, ln = M.box_tet(2, 2, 2, 1.0, 0.8, 0.6, 5)
    n = co.size // 3
    par = [0.6, 0.4, 1.2, 0.0, 0.1, -0.2]
    rng = np.random.default_rng(2)
    z = np.zeros(3 * n)
    v, a = rng.uniform(-0.1, 0.1, 3 * n), rng.uniform(-0.1, 0.1, 3 * n)
    glob = (0.05, 0.5, 0.25, 1.0)
    nl = P.CpuSystem(3, co, ln, 3)
    nl.assemble_stvk3d_newmark(par, z, v, a, glob)
    li = P.CpuSystem(3, co, ln, 3)
    li.assemble_newmark(par, z, v, a, glob)
    A, B = dense(nl), dense(li)
    assert np.allclose(A, B, rtol=0, atol=1e-13 * np.abs(B).max())
    assert np.allclose(nl.res, li.res, rtol=0, atol=1e-13 * np.abs(li.res).max())
