"""The oracle's critical time step and its Tchamwa-Wielgosz integrator,
pinned by closed forms rather than by restating their formulas:

- critical step (sld_updtss.f90:119-162 with elmlen.f90 ndime == 3 and
  sld_velsnd.f90:117-124): on a graded mesh of boxes h_e is the shortest
  edge and the sound speed sqrt((lam + 2 mu)/rho) (law 160) or
  sqrt((K + 4 mu/3)/rho) (law 102, sld_velsnd.f90:75-78); on a sheared
  parallelepiped h_e is the smallest distance between opposite faces,
  V / (largest face area); on a non-affine hex the lengths are those of the
  parallelepiped spanned by the mean edge vectors (elmlen takes the
  Jacobian at the element centre);
- Tchamwa-Wielgosz (mod_sld_solution_methods.f90:406-429,504-527): a body
  that falls from rest under gravity, with one direction prescribed at every
  node, follows v^n = (n - 1) dt g and u^n = dt^2 g ((n-1)(n-2)/2 + phi (n-1))
  on the free dofs (a^0 = 0, a^n = g after the first step), and keeps
  u = v = a = 0 on the prescribed ones.

Each check also runs with a mistake planted in the oracle (E for
lam + 2 mu, the largest length for the smallest, the 4 mu/3 dropped,
phi dt^2 a / 2, v and a kept on prescribed dofs, v^{n+1} in the displacement
update) and must then fail."""
import numpy as np
import pytest

from alya_sma_amd import problems
from oracle import oracle_py as O

NU = 0.33
RHO = problems.RHO_NITINOL


def _c_sma(E, nu=NU, rho=RHO):
    lam = nu * E / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    return np.sqrt((lam + 2 * mu) / rho)


def _graded_box():
    """2 x 3 x 2 boxes with unequal edges and seeded per-element E."""
    dx, dy, dz = np.array([1.0, 0.55]), np.array([0.7, 1.3, 0.9]), np.array([0.8, 0.45])
    nx, ny, nz = len(dx), len(dy), len(dz)
    coord, lnods = problems.box_mesh_np(nx, ny, nz, 1.0, 1.0, 1.0)
    xs, ys, zs = (np.concatenate([[0.0], np.cumsum(d)]) for d in (dx, dy, dz))
    i = np.rint(coord[:, 0] * nx).astype(int)
    j = np.rint(coord[:, 1] * ny).astype(int)
    k = np.rint(coord[:, 2] * nz).astype(int)
    coord = np.stack([xs[i], ys[j], zs[k]], axis=1)
    rng = np.random.default_rng(7)
    ne = lnods.shape[0]
    E = 60000.0 * (1.0 + 0.3 * rng.uniform(-1, 1, ne))
    prob = problems.bar_problem(nx=nx, ny=ny, nz=nz)
    prob.update(coord=coord, lnods=lnods, E_of_elem=E)
    prob["fixmask"] = np.zeros(coord.shape[0], np.uint8)
    prob["func_of_node"] = np.full(coord.shape[0], -1, np.int32)
    # edges of element (ei, ej, ek), numbered e = ei + nx (ej + ny ek)
    ee = np.arange(ne)
    ei, ej, ek = ee % nx, (ee // nx) % ny, ee // (nx * ny)
    hmin = np.minimum(np.minimum(dx[ei], dy[ej]), dz[ek])
    return prob, hmin


def _one_element(x8, props=None, law=None):
    prob = problems.bar_problem(nx=1, ny=1, nz=1)
    prob.update(coord=np.asarray(x8, float), lnods=np.arange(8, dtype=np.int32)[None])
    prob["fixmask"] = np.zeros(8, np.uint8)
    prob["func_of_node"] = np.full(8, -1, np.int32)
    prob.pop("E_of_elem", None)
    if props is not None:
        prob["props"] = np.asarray(props, float).reshape(1, 18)
    if law is not None:
        prob["law_of_mat"] = np.array([law], np.int32)
    return prob


SIGNS = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                  [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float)


def _rel(a, b):
    return abs(a - b) / abs(b)


def dt_box_error():
    prob, hmin = _graded_box()
    expect = np.min(hmin / _c_sma(prob["E_of_elem"]))
    return _rel(O.StepOracle(prob).critical_dt(), expect)


def dt_parallelepiped_error():
    # x(xi) = x0 + B xi: edge vectors e_k = 2 B[:, k]; the distance between
    # the faces xi_i = +-1 is |det(e)| / |e_j x e_k|
    B = np.array([[0.6, 0.25, -0.1], [0.05, 0.4, 0.15], [0.12, -0.08, 0.9]])
    x8 = np.array([1.0, 2.0, -0.5]) + SIGNS @ B.T
    e = 2.0 * B
    V = abs(np.linalg.det(e))
    heights = [V / np.linalg.norm(np.cross(e[:, (i + 1) % 3], e[:, (i + 2) % 3])) for i in range(3)]
    expect = min(heights) / _c_sma(60000.0)
    return _rel(O.StepOracle(_one_element(x8)).critical_dt(), expect)


def dt_trapezoid_error():
    # a non-affine hex: the top face shrunk and shifted.  elmlen uses the
    # Jacobian at the centre, whose columns are the mean edge vectors / 2
    x8 = (SIGNS + 1.0) / 2.0 * np.array([2.0, 1.0, 0.6])
    x8[4:, 0] = 0.3 + 0.6 * x8[4:, 0]
    x8[4:, 1] = 0.1 + 0.8 * x8[4:, 1]
    mean_edge = np.stack([x8[SIGNS[:, k] > 0].mean(0) - x8[SIGNS[:, k] < 0].mean(0)
                          for k in range(3)], axis=1)
    V = abs(np.linalg.det(mean_edge))
    heights = [V / np.linalg.norm(np.cross(mean_edge[:, (i + 1) % 3], mean_edge[:, (i + 2) % 3]))
               for i in range(3)]
    expect = min(heights) / _c_sma(60000.0)
    return _rel(O.StepOracle(_one_element(x8)).critical_dt(), expect)


def dt_neo_hookean_error():
    mu, K, rho = 0.03, 1.0, 1.0e-9
    x8 = (SIGNS + 1.0) / 2.0 * np.array([0.5, 0.8, 0.35])
    props = np.zeros(18)
    props[0], props[1] = mu, K
    prob = _one_element(x8, props=props, law=102)
    prob["rho"] = np.array([rho])
    expect = 0.35 / np.sqrt((K + 4.0 * mu / 3.0) / rho)
    return _rel(O.StepOracle(prob).critical_dt(), expect)


DT_CHECKS = [dt_box_error, dt_parallelepiped_error, dt_trapezoid_error]


@pytest.mark.parametrize("check", DT_CHECKS + [dt_neo_hookean_error])
def test_critical_dt_closed_form(check):
    assert check() < 1e-13


@pytest.mark.parametrize("mutation", [O.MUT_DT_YOUNG, O.MUT_DT_HMAX])
def test_planted_dt_mistakes_fail(mutation):
    with O.planted(mutation):
        errs = [c() for c in DT_CHECKS]
    assert max(errs) > 1e-3, errs


def test_planted_neo_hookean_sound_speed_fails():
    with O.planted(O.MUT_DT_NH_BULK):
        assert dt_neo_hookean_error() > 1e-3


# ---------------------------------------------------------------- TW scheme
G = np.array([-3000.0, 9810.0, 5000.0])


def _falling_box():
    prob = problems.cantilever_problem(nx=3, ny=2, nz=2, lx=1.5, ly=1.0, lz=1.0, nsteps=100,
                                       traction=0.0, E_noise=0.0)
    prob["fext0"], prob["fext_func"] = None, -1
    # z prescribed (to zero) at every node: the body stays unstrained
    prob["fixmask"] = np.full(prob["coord"].shape[0], 4, np.uint8)
    prob["func_of_node"] = np.full(prob["coord"].shape[0], -1, np.int32)
    prob["gravity"] = G.copy()
    # a dense body: the weights m g stand well above the rounding of the
    # internal forces of a rigid translation
    prob["rho"] = np.array([1.0e-3])
    return prob


def tw_fall_errors(n=30, phi=1.033):
    prob = _falling_box()
    orc = O.StepOracle(prob)
    dt = 0.5 * orc.critical_dt()
    orc.step(dt, scheme=1, phi=phi, n=n)
    g = G[:2]
    u_x = dt * dt * g * ((n - 1) * (n - 2) / 2 + phi * (n - 1))
    v_x = (n - 1) * dt * g
    # rigid translation: f_int is rounding only (F formed from X + u), so
    # the state follows the closed form to ~1e-11
    eu = np.max(np.abs(orc.u[:, :2] - u_x)) / np.max(np.abs(u_x))
    ev = np.max(np.abs(orc.v[:, :2] - v_x)) / np.max(np.abs(v_x))
    ea = np.max(np.abs(orc.a[:, :2] - g)) / np.max(np.abs(g))
    ez = np.max(np.abs(np.concatenate([orc.u[:, 2], orc.v[:, 2], orc.a[:, 2]])))
    return eu, ev, ea, ez


def test_tw_free_fall_closed_form():
    eu, ev, ea, ez = tw_fall_errors()
    assert eu < 1e-9 and ev < 1e-9 and ea < 1e-9, (eu, ev, ea)
    assert ez == 0.0


def test_tw_first_steps_from_rest():
    # n = 1: a^1 = g, v^1 = u^1 = 0; n = 2: v = dt g, u = phi dt^2 g
    prob = _falling_box()
    orc = O.StepOracle(prob)
    dt, phi = 0.5 * orc.critical_dt(), 1.2
    orc.step(dt, scheme=1, phi=phi, n=1)
    assert np.all(orc.u[:, :2] == 0.0) and np.all(orc.v[:, :2] == 0.0)
    np.testing.assert_allclose(orc.a[:, :2], np.tile(G[:2], (orc.npoin, 1)), rtol=1e-12)
    orc.step(dt, scheme=1, phi=phi, n=1)
    np.testing.assert_allclose(orc.v[:, :2], np.tile(dt * G[:2], (orc.npoin, 1)), rtol=1e-12)
    np.testing.assert_allclose(orc.u[:, :2], np.tile(phi * dt * dt * G[:2], (orc.npoin, 1)),
                               rtol=1e-12)


@pytest.mark.parametrize("mutation", [O.MUT_TW_HALF, O.MUT_TW_FIXVA, O.MUT_TW_VNEW])
def test_planted_tw_mistakes_fail(mutation):
    with O.planted(mutation):
        eu, ev, ea, ez = tw_fall_errors()
    assert max(eu, ev, ea, ez) > 1e-3, (eu, ev, ea, ez)
