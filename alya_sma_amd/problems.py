"""Seeded synthetic problems of the explicit SMA step.

Each builder returns a dict of plain numpy arrays -- the same arrays the
device library (``sld_dev_create``) and the CPU oracle take -- plus run
parameters.  Units: mm, MPa, tonne/mm^3, s.

Material parameters follow the reference's parco_sld(1:18) layout
(sld_reaphy.f90:477-490; meaning per sld_stress_model_160.f90:263-279):
[E, nu, eL, sigma_s_AS, sigma_f_AS, C_AS, Tf_AS, Ts_AS, sigma_n_AS (unused),
 sigma_s_SA, sigma_f_SA, C_SA, Tf_SA, Ts_SA, alpha, beta_AS, beta_SA, unused].
"""
import numpy as np

from . import _lib

# E = 60 GPa, nu = 0.33, eL = 0.05, AS start/finish 400/500 MPa,
# SA start/finish 250/150 MPa; temperature slopes zero (T is inert).
NITINOL = np.array([60000.0, 0.33, 0.05, 400.0, 500.0, 0, 0, 0, 0,
                    250.0, 150.0, 0, 0, 0, 0, 1.0, 1.0, 0], dtype=np.float64)
RHO_NITINOL = 6.45e-9   # 6450 kg/m^3 in tonne/mm^3


def box_mesh(nx, ny, nz, lx, ly, lz):
    """Structured HEX08 box through the library's host helper."""
    npoin = (nx + 1) * (ny + 1) * (nz + 1)
    coord = np.zeros((npoin, 3))
    lnods = np.zeros((nx * ny * nz, 8), dtype=np.int32)
    L = _lib.lib()
    _lib.check(L.sld_mesh_box(nx, ny, nz, lx, ly, lz, coord.ctypes.data, lnods.ctypes.data),
               "sld_mesh_box")
    return coord, lnods


def box_mesh_np(nx, ny, nz, lx, ly, lz):
    """Same numbering as sld_mesh_box, in numpy (no library needed)."""
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    px, py = nx + 1, ny + 1
    idx = lambda a, b, c: a + px * (b + py * c)
    coord = np.zeros(((nx + 1) * (ny + 1) * (nz + 1), 3))
    p = idx(i, j, k).ravel()
    coord[p, 0] = lx * i.ravel() / nx
    coord[p, 1] = ly * j.ravel() / ny
    coord[p, 2] = lz * k.ravel() / nz
    ei, ej, ek = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    e = (ei + nx * (ej + ny * ek)).ravel()
    ei, ej, ek = ei.ravel(), ej.ravel(), ek.ravel()
    lnods = np.zeros((nx * ny * nz, 8), dtype=np.int32)
    lnods[e] = np.stack([idx(ei, ej, ek), idx(ei + 1, ej, ek), idx(ei + 1, ej + 1, ek),
                         idx(ei, ej + 1, ek), idx(ei, ej, ek + 1), idx(ei + 1, ej, ek + 1),
                         idx(ei + 1, ej + 1, ek + 1), idx(ei, ej + 1, ek + 1)], axis=1)
    return coord, lnods


def smooth_table(t_end, values):
    """DISCRETE,SMOOTH load table: equally spaced (t, v) knots over [0, t_end].
    values: [npts][3] per-component amplitudes.  Rows (v_x, v_y, v_z, t)."""
    values = np.asarray(values, dtype=np.float64)
    n = values.shape[0]
    tab = np.zeros((n, 4))
    tab[:, :3] = values
    tab[:, 3] = np.linspace(0.0, t_end, n)
    return tab


# HEX08 faces as QUA04 node quadruples (local node order of the element:
# 1-4 on zeta = -1, 5-8 on zeta = +1, counter-clockwise)
HEX8_FACES = np.array([[0, 3, 2, 1], [4, 5, 6, 7], [0, 1, 5, 4],
                       [1, 2, 6, 5], [2, 3, 7, 6], [0, 4, 7, 3]], dtype=np.int64)


def boundary_faces(coord, lnods, select=None):
    """QUA04 boundary faces of a HEX08 mesh: faces that belong to one
    element only, each ordered so that t_s x t_t (the normal of the QUA04
    map, node 1 at (s, t) = (-1, -1), counter-clockwise) points out of its
    element.  select(centre [nf, 3]) -> bool mask keeps a subset.
    Returns (lnodb [nf, 4] int32, elem [nf] int32)."""
    lnods = np.asarray(lnods, dtype=np.int64)
    ne = lnods.shape[0]
    faces = lnods[:, HEX8_FACES].reshape(-1, 4)               # [6 ne, 4]
    key = np.sort(faces, axis=1)
    _, inv, cnt = np.unique(key, axis=0, return_inverse=True, return_counts=True)
    keep = cnt[inv.ravel()] == 1
    fb = faces[keep]
    el = np.repeat(np.arange(ne), 6)[keep]
    x = coord[fb]                                              # [nf, 4, 3]
    centre = x.mean(axis=1)
    if select is not None:
        m = np.asarray(select(centre), dtype=bool)
        fb, el, x, centre = fb[m], el[m], x[m], centre[m]
    # t_s x t_t at the face centre vs the direction out of the element
    ts = 0.25 * (-x[:, 0] + x[:, 1] + x[:, 2] - x[:, 3])
    tt = 0.25 * (-x[:, 0] - x[:, 1] + x[:, 2] + x[:, 3])
    out = centre - coord[lnods[el]].mean(axis=1)
    flip = np.einsum("ij,ij->i", np.cross(ts, tt), out) < 0.0
    fb[flip] = fb[flip][:, ::-1]
    return fb.astype(np.int32), el.astype(np.int32)


def bar_problem(nx=10, ny=2, nz=2, lx=10.0, ly=2.0, lz=2.0, strain=0.06,
                nsteps=2000, safety=0.5, props=NITINOL, rho=RHO_NITINOL, use_lib=False):
    """Uniaxial SMA bar (displacement-controlled ramp to `strain`, then unload).

    BCs as tests/uniaxial_test/sm154_d1C.sld.dat:80-107: u_x = 0 on x = 0,
    symmetry u_y = 0 on y = 0 and u_z = 0 on z = 0, prescribed u_x on x = lx
    through a DISCRETE,SMOOTH function 0 -> strain*lx -> 0 over nsteps steps.
    """
    coord, lnods = (box_mesh if use_lib else box_mesh_np)(nx, ny, nz, lx, ly, lz)
    npoin = coord.shape[0]
    fix = np.zeros(npoin, dtype=np.uint8)
    x, y, z = coord[:, 0], coord[:, 1], coord[:, 2]
    tol = 1e-9 * max(lx, ly, lz)
    fix[np.abs(x) < tol] |= 1
    fix[np.abs(y) < tol] |= 2
    fix[np.abs(z) < tol] |= 4
    tip = np.abs(x - lx) < tol
    fix[tip] |= 1
    func_of_node = np.full(npoin, -1, dtype=np.int32)
    func_of_node[tip] = 0
    prob = dict(coord=coord, lnods=lnods, props=np.asarray(props, np.float64).reshape(-1, 18),
                rho=np.array([rho], np.float64), fixmask=fix, func_of_node=func_of_node)
    # dt from the critical step of the mesh (uniform box: h = min edge)
    E, nu = prob["props"][0, 0], prob["props"][0, 1]
    lam = nu * E / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    c = np.sqrt((lam + 2 * mu) / rho)
    h = min(lx / nx, ly / ny, lz / nz)
    dt = safety * h / c
    t_end = nsteps * dt
    tab = smooth_table(t_end, [[0, 0, 0], [strain * lx, 0, 0], [0, 0, 0]])
    prob.update(func_table=tab[None], func_npts=np.array([3], np.int32),
                func_smooth=np.array([1], np.int32), fext0=None, fext_func=-1)
    prob["run"] = dict(dt=dt, nsteps=nsteps, safety=safety, scheme=0, phi=1.033)
    prob["name"] = f"bar {nx}x{ny}x{nz} hex8 over {lx}x{ly}x{lz} mm"
    return prob


def cantilever_problem(nx=400, ny=20, nz=20, lx=50.0, ly=2.5, lz=2.5, traction=500.0,
                       nsteps=5000, safety=0.5, seed=1, props=NITINOL, rho=RHO_NITINOL,
                       E_noise=0.02, use_lib=False, direction=1):
    """Cantilever clamped at x = 0 with a tip traction on x = lx in direction
    `direction` (0 = x, 1 = y) that is ramped up and released
    (DISCRETE,SMOOTH 0 -> 1 -> 0 over nsteps).  The per-element Young's
    modulus carries seeded uniform noise of +-E_noise."""
    coord, lnods = (box_mesh if use_lib else box_mesh_np)(nx, ny, nz, lx, ly, lz)
    npoin, nelem = coord.shape[0], lnods.shape[0]
    tol = 1e-9 * lx
    fix = np.zeros(npoin, dtype=np.uint8)
    fix[np.abs(coord[:, 0]) < tol] = 7
    func_of_node = np.full(npoin, -1, dtype=np.int32)
    # tip traction -> nodal forces by tributary area on the x = lx face
    tip = np.abs(coord[:, 0] - lx) < tol
    hy, hz = ly / ny, lz / nz
    wy = np.where((np.abs(coord[:, 1]) < tol) | (np.abs(coord[:, 1] - ly) < tol), 0.5, 1.0)
    wz = np.where((np.abs(coord[:, 2]) < tol) | (np.abs(coord[:, 2] - lz) < tol), 0.5, 1.0)
    fext0 = np.zeros((npoin, 3))
    fext0[tip, direction] = traction * hy * hz * (wy * wz)[tip]
    rng = np.random.default_rng(seed)
    P = np.asarray(props, np.float64).reshape(-1, 18)
    E_of_elem = P[0, 0] * (1.0 + E_noise * rng.uniform(-1.0, 1.0, nelem))
    prob = dict(coord=coord, lnods=lnods, props=P, rho=np.array([rho], np.float64),
                fixmask=fix, func_of_node=func_of_node, E_of_elem=E_of_elem, fext0=fext0)
    E, nu = E_of_elem.max(), P[0, 1]
    lam = nu * E / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    c = np.sqrt((lam + 2 * mu) / rho)
    dt = safety * min(lx / nx, hy, hz) / c
    t_end = nsteps * dt
    tab = smooth_table(t_end, [[0, 0, 0], [1, 0, 0], [0, 0, 0]])
    prob.update(func_table=tab[None], func_npts=np.array([3], np.int32),
                func_smooth=np.array([1], np.int32), fext_func=0)
    prob["run"] = dict(dt=dt, nsteps=nsteps, safety=safety, scheme=0, phi=1.033)
    prob["name"] = f"cantilever {nx}x{ny}x{nz} hex8 over {lx}x{ly}x{lz} mm"
    return prob


def batched_specimens(n=256, seed=3, median=8000, sigma=0.7, lo=2000, hi=60000, nsteps=2000,
                      width=1.0, safety=0.5, props=NITINOL, rho=RHO_NITINOL):
    """Many uniaxial SMA specimens of uneven size stepped as one mesh.

    Specimen s is a bar of cubic hex8 elements, n_e ~ LogNormal(ln median,
    sigma) truncated to [lo, hi] by redrawing, aspect ratio U[5, 40], cross
    section width x width mm; its own material row (E and the four
    transformation stresses scaled by factors U[0.9, 1.1]) and its own
    DISCRETE,SMOOTH function 0 -> amp*L -> 0 with amp ~ U[3%, 7%], prescribed
    on u_x at x = L, with the symmetry planes of the uniaxial bar (u_x = 0 on
    x = 0, u_y = 0 on y = 0, u_z = 0 on z = 0).  The specimens share no node;
    one time step (the smallest critical step times `safety`) serves all.
    Specimens are laid side by side in y (2 x width apart); that changes no
    result, only where a plot would put them."""
    rng = np.random.default_rng(seed)
    sizes = np.empty(n, dtype=np.int64)
    for s in range(n):
        while True:
            v = rng.lognormal(np.log(median), sigma)
            if lo <= v <= hi:
                break
        sizes[s] = int(round(v))
    aspect = rng.uniform(5.0, 40.0, n)
    fE = rng.uniform(0.9, 1.1, n)
    fS = rng.uniform(0.9, 1.1, n)
    amp = rng.uniform(0.03, 0.07, n)
    P0 = np.asarray(props, np.float64).reshape(18)
    coords, lnods, fixes, funcs, mats, dims = [], [], [], [], [], []
    P = np.zeros((n, 18))
    rhos = np.full(n, rho, dtype=np.float64)
    hmin_over_c = np.inf
    node0 = 0
    for s in range(n):
        ny = max(1, int(round((sizes[s] / aspect[s]) ** (1.0 / 3.0))))
        nx = max(1, int(round(sizes[s] / (ny * ny))))
        h = width / ny
        L = nx * h
        c, ln = box_mesh_np(nx, ny, ny, L, width, width)
        tol = 1e-9 * L
        fx = np.zeros(c.shape[0], dtype=np.uint8)
        fx[np.abs(c[:, 0]) < tol] |= 1
        fx[np.abs(c[:, 1]) < tol] |= 2
        fx[np.abs(c[:, 2]) < tol] |= 4
        tip = np.abs(c[:, 0] - L) < tol
        fx[tip] |= 1
        fn = np.full(c.shape[0], -1, dtype=np.int32)
        fn[tip] = s
        c = c.copy()
        c[:, 1] += 2.0 * width * s
        coords.append(c)
        lnods.append(ln + node0)
        fixes.append(fx)
        funcs.append(fn)
        mats.append(np.full(ln.shape[0], s, dtype=np.int32))
        dims.append((nx, ny, ny, L))
        node0 += c.shape[0]
        P[s] = P0
        P[s, 0] = P0[0] * fE[s]
        for k in (3, 4, 9, 10):
            P[s, k] = P0[k] * fS[s]
        E, nu = P[s, 0], P[s, 1]
        lam = nu * E / ((1 + nu) * (1 - 2 * nu))
        mu = E / (2 * (1 + nu))
        hmin_over_c = min(hmin_over_c, h / np.sqrt((lam + 2 * mu) / rho))
    dt = safety * hmin_over_c
    t_end = nsteps * dt
    tabs = np.stack([smooth_table(t_end, [[0, 0, 0], [amp[s] * dims[s][3], 0, 0], [0, 0, 0]])
                     for s in range(n)])
    prob = dict(coord=np.concatenate(coords), lnods=np.concatenate(lnods).astype(np.int32),
                props=P, rho=rhos, mat_of_elem=np.concatenate(mats), fixmask=np.concatenate(fixes),
                func_of_node=np.concatenate(funcs), func_table=tabs,
                func_npts=np.full(n, 3, np.int32), func_smooth=np.ones(n, np.int32),
                fext0=None, fext_func=-1)
    prob["run"] = dict(dt=dt, nsteps=nsteps, safety=safety, scheme=0, phi=1.033)
    prob["specimens"] = dict(elements=np.array([d[0] * d[1] * d[2] for d in dims]),
                             dims=np.array([d[:3] for d in dims]), length=np.array([d[3] for d in dims]),
                             amplitude=amp, E_factor=fE, stress_factor=fS)
    prob["name"] = (f"{n} SMA specimens, {prob['lnods'].shape[0]} hex8 in total "
                    f"({int(prob['specimens']['elements'].min())}-{int(prob['specimens']['elements'].max())} each)")
    return prob


# Tissue around the SMA tube: neo-Hookean (Ansys form) mu = 0.03 MPa, K = 1 MPa,
# density 1000 kg/m^3 (props[0] = mu, props[1] = K as parco_sld for NEOHO ANSYS).
TISSUE = np.array([0.03, 1.0] + [0.0] * 16, dtype=np.float64)
RHO_TISSUE = 1.0e-9
LAW_SMA, LAW_NEOHOOKEAN = 160, 102


def cylinder_problem(ntheta=360, nz=400, nr_tube=3, nr_tissue=39, r_in=4.0, r_tube=4.3,
                     r_out=10.0, length=40.0, crimp=0.4, nsteps=10000, safety=0.5,
                     props=NITINOL, rho=RHO_NITINOL):
    """SMA tube r in [r_in, r_tube] inside neo-Hookean tissue
    r in [r_tube, r_out], length `length`, structured hex8 in (r, theta, z)
    (nr_tube + nr_tissue) x ntheta x nz elements (default 42 x 360 x 400 =
    6.05 M), closed in theta.  The inner surface r = r_in is crimped by the
    prescribed radial displacement -crimp (mm) and released (DISCRETE,SMOOTH
    0 -> 1 -> 0 over nsteps); u_z = 0 on z = 0.  One load function per
    circumferential station carries that station's direction (cos, sin)."""
    nr = nr_tube + nr_tissue
    r = np.concatenate([np.linspace(r_in, r_tube, nr_tube + 1),
                        np.linspace(r_tube, r_out, nr_tissue + 1)[1:]])
    th = 2.0 * np.pi * np.arange(ntheta) / ntheta
    z = np.linspace(0.0, length, nz + 1)
    ir, it, iz = np.meshgrid(np.arange(nr + 1), np.arange(ntheta), np.arange(nz + 1), indexing="ij")
    node = lambda a, b, c: a + (nr + 1) * ((b % ntheta) + ntheta * c)
    npoin = (nr + 1) * ntheta * (nz + 1)
    coord = np.zeros((npoin, 3))
    p = node(ir, it, iz).ravel()
    coord[p, 0] = (r[ir] * np.cos(th[it])).ravel()
    coord[p, 1] = (r[ir] * np.sin(th[it])).ravel()
    coord[p, 2] = z[iz].ravel()
    # elements numbered like the nodes (r fastest, then theta, then z), so that
    # consecutive elements share nodes that sit close in memory
    er, et, ez = [x.ravel(order="F") for x in np.meshgrid(np.arange(nr), np.arange(ntheta),
                                                          np.arange(nz), indexing="ij")]
    # reference node order with (xi, eta, zeta) = (r, theta, z): positive volume
    lnods = np.stack([node(er, et, ez), node(er + 1, et, ez), node(er + 1, et + 1, ez),
                      node(er, et + 1, ez), node(er, et, ez + 1), node(er + 1, et, ez + 1),
                      node(er + 1, et + 1, ez + 1), node(er, et + 1, ez + 1)], axis=1).astype(np.int32)
    mat_of_elem = (er >= nr_tube).astype(np.int32)          # 0 = SMA tube, 1 = tissue
    fix = np.zeros(npoin, dtype=np.uint8)
    fix[np.abs(coord[:, 2]) < 1e-9 * length] |= 4
    inner = np.zeros(npoin, dtype=bool)
    inner[node(np.zeros_like(it[0]), it[0], iz[0]).ravel()] = True
    fix[inner] |= 3
    func_of_node = np.full(npoin, -1, dtype=np.int32)
    stat = np.zeros(npoin, dtype=np.int64)
    stat[p] = it.ravel()
    func_of_node[inner] = stat[inner]
    P = np.stack([np.asarray(props, np.float64).reshape(18), TISSUE])
    rhos = np.array([rho, RHO_TISSUE])
    # critical step: the SMA tube's smallest edge at the SMA wave speed (the
    # tissue's speed is ~100x lower)
    E, nu = P[0, 0], P[0, 1]
    c_sma = np.sqrt((nu * E / ((1 + nu) * (1 - 2 * nu)) + E / (1 + nu)) / rho)
    h = min((r_tube - r_in) / nr_tube, 2 * r_in * np.sin(np.pi / ntheta), length / nz)
    dt = safety * h / c_sma
    t_end = nsteps * dt
    tabs = np.stack([smooth_table(t_end, [[0, 0, 0], [-crimp * np.cos(a), -crimp * np.sin(a), 0],
                                          [0, 0, 0]]) for a in th])
    prob = dict(coord=coord, lnods=lnods, props=P, rho=rhos, mat_of_elem=mat_of_elem,
                law_of_mat=np.array([LAW_SMA, LAW_NEOHOOKEAN], np.int32), fixmask=fix,
                func_of_node=func_of_node, func_table=tabs, func_npts=np.full(ntheta, 3, np.int32),
                func_smooth=np.ones(ntheta, np.int32), fext0=None, fext_func=-1)
    prob["run"] = dict(dt=dt, nsteps=nsteps, safety=safety, scheme=0, phi=1.033)
    prob["name"] = (f"SMA tube in neo-Hookean tissue, {nr}x{ntheta}x{nz} = {lnods.shape[0]} hex8 "
                    f"(tube {nr_tube} layers, r {r_in}-{r_tube}-{r_out} mm, L {length} mm)")
    return prob
