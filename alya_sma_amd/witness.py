"""Witness points of a solidz run (sld_outwit.f90).

A witness point is a coordinate where nodal results are interpolated with
the shape functions of the element that contains it: DISPX..Z, VELOX..Z,
ACCEX..Z, the reactions FRXIX..Z (-frxid_sld, sld_outwit.f90:184-211), the
Cauchy stress SIGXX, SIGYY, SIGZZ, SIGYZ, SIGXZ, SIGXY of the nodal projection
caust_sld (sld_calculate_pp_tensors.f90, sld_calculate_pp_tensors_ielem.f90):
per Gauss point sigma = F P^T / det F, lumped onto the nodes with the weights
N_a gpvol and divided by the lumped volume sum_g N_a gpvol; the Green strain
EPSXX..EPSXY (green_sld, E = (F^T F - I)/2) and the logarithmic strain
LEPXX..LEPXY (lepsi_sld, the eigen form of log(F F^T)/2,
sld_calculate_pp_tensors_ielem.f90:125-215) by the same projection, and the
von Mises stress SEQVM of the projected Cauchy stress (seqvm_sld,
sld_calculate_pp_tensors.f90:136-141).  The points come
from the WITNESS_POINTS section of the case's .ker.dat (ker_readat.f90:
3832-3870), the variables from the WITNESS_POINTS list of its .sld.dat.
Host-side post-processing of the device's fields; HEX08 meshes.
"""
import numpy as np

GPOS = 0.577350269189626
# reference HEX08 node signs (xi, eta, zeta) and Gauss points (xi outer, zeta inner)
NODE_SIGNS = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                       [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=np.float64)
GP_SIGNS = np.array([[(g >> 2) & 1, (g >> 1) & 1, g & 1] for g in range(8)], dtype=np.float64) * 2 - 1

VARIABLES = {"DISPX": ("u", 0), "DISPY": ("u", 1), "DISPZ": ("u", 2),
             "VELOX": ("v", 0), "VELOY": ("v", 1), "VELOZ": ("v", 2),
             "ACCEX": ("a", 0), "ACCEY": ("a", 1), "ACCEZ": ("a", 2),
             "FRXIX": ("frx", 0), "FRXIY": ("frx", 1), "FRXIZ": ("frx", 2),
             "SIGXX": ("sig", 0), "SIGYY": ("sig", 1), "SIGZZ": ("sig", 2),
             "SIGYZ": ("sig", 3), "SIGXZ": ("sig", 4), "SIGXY": ("sig", 5),
             "EPSXX": ("eps", 0), "EPSYY": ("eps", 1), "EPSZZ": ("eps", 2),
             "EPSYZ": ("eps", 3), "EPSXZ": ("eps", 4), "EPSXY": ("eps", 5),
             "LEPXX": ("lep", 0), "LEPYY": ("lep", 1), "LEPZZ": ("lep", 2),
             "LEPYZ": ("lep", 3), "LEPXZ": ("lep", 4), "LEPXY": ("lep", 5),
             "SEQVM": ("vm", 0)}


def shape(xi):
    """Trilinear N [..., 8] and dN/dxi [..., 8, 3] at natural coordinates xi [..., 3]."""
    xi = np.asarray(xi, dtype=np.float64)
    f = 1.0 + NODE_SIGNS * xi[..., None, :]                  # [..., 8, 3]
    N = 0.125 * f[..., 0] * f[..., 1] * f[..., 2]
    dN = np.empty(xi.shape[:-1] + (8, 3))
    dN[..., 0] = 0.125 * NODE_SIGNS[:, 0] * f[..., 1] * f[..., 2]
    dN[..., 1] = 0.125 * NODE_SIGNS[:, 1] * f[..., 0] * f[..., 2]
    dN[..., 2] = 0.125 * NODE_SIGNS[:, 2] * f[..., 0] * f[..., 1]
    return N, dN


def locate(coord, lnods, points, tol=1e-8):
    """Element [nw] (-1 when outside the mesh) and natural coordinates [nw, 3]
    of each point: Newton on x(xi) = sum_a N_a(xi) x_a over the elements whose
    bounding box holds the point."""
    coord = np.asarray(coord, dtype=np.float64)
    X = coord[np.asarray(lnods)]                               # [ne, 8, 3]
    lo, hi = X.min(axis=1), X.max(axis=1)
    pad = 1e-9 * max(1.0, float(np.abs(coord).max()))
    elem = np.full(len(points), -1, dtype=np.int64)
    xis = np.zeros((len(points), 3))
    for w, p in enumerate(np.asarray(points, dtype=np.float64)):
        cand = np.nonzero(np.all((p >= lo - pad) & (p <= hi + pad), axis=1))[0]
        for e in cand:
            xi = np.zeros(3)
            for _ in range(30):
                N, dN = shape(xi)
                r = N @ X[e] - p
                J = X[e].T @ dN                                # dx/dxi
                step = np.linalg.solve(J, r)
                xi = xi - step
                if np.abs(step).max() < 1e-14:
                    break
            if np.abs(xi).max() <= 1.0 + tol:
                elem[w], xis[w] = e, xi
                break
    return elem, xis


def gp_kinematics(coord, lnods, u):
    """Per element and Gauss point: gpvol = w det J [ne, 8], N [8, 8] and
    F = I + sum_a u_a (x) dN_a/dX [ne, 8, 3, 3] (reference configuration)."""
    X = np.asarray(coord, dtype=np.float64)[lnods]
    U = np.asarray(u, dtype=np.float64)[lnods]
    N, dN = shape(GP_SIGNS * GPOS)                             # [8 gp, 8], [8 gp, 8, 3]
    J = np.einsum("ean,gak->egnk", X, dN)
    gpvol = np.linalg.det(J)                                   # weights are 1
    Ji = np.linalg.inv(J)
    gpcar = np.einsum("gak,egkn->egan", dN, Ji)               # dN_a/dX_n
    F = np.eye(3) + np.einsum("eai,egaj->egij", U, gpcar)
    return gpvol, N, F


def _project(coord, lnods, gpvol, N, tens):
    """Nodal projection of a symmetric tensor per Gauss point tens [ne, 8,
    3, 3]: Voigt (xx, yy, zz, yz, xz, xy) lumped with N_a gpvol, divided by
    the lumped volume."""
    voigt = np.stack([tens[..., 0, 0], tens[..., 1, 1], tens[..., 2, 2],
                      tens[..., 1, 2], tens[..., 0, 2], tens[..., 0, 1]], axis=-1)
    wgt = N[None, :, :] * gpvol[:, :, None]                    # [ne, 8 gp, 8 node]
    npoin = len(coord)
    out = np.zeros((npoin, 6))
    vmass = np.zeros(npoin)
    np.add.at(out, lnods, np.einsum("ega,egv->eav", wgt, voigt))
    np.add.at(vmass, lnods, wgt.sum(axis=1))
    return out / vmass[:, None]


def nodal_strains(coord, lnods, u):
    """(green, lepsi) [npoin, 6] each: the nodal projections of the Green
    strain (F^T F - I)/2 and of the logarithmic strain log(F F^T)/2."""
    lnods = np.asarray(lnods)
    gpvol, N, F = gp_kinematics(coord, lnods, u)
    green = 0.5 * (np.einsum("egki,egkj->egij", F, F) - np.eye(3))
    w, v = np.linalg.eigh(np.einsum("egik,egjk->egij", F, F))
    lep = np.einsum("egik,egk,egjk->egij", v, 0.5 * np.log(w), v)
    return _project(coord, lnods, gpvol, N, green), _project(coord, lnods, gpvol, N, lep)


def von_mises(caust):
    """[npoin, 1] von Mises stress of the nodal Cauchy stress caust [npoin, 6]
    (xx, yy, zz, yz, xz, xy)."""
    s = np.asarray(caust)
    return np.sqrt(0.5 * ((s[:, 0] - s[:, 1]) ** 2 + (s[:, 1] - s[:, 2]) ** 2
                          + (s[:, 2] - s[:, 0]) ** 2
                          + 6.0 * (s[:, 3] ** 2 + s[:, 4] ** 2 + s[:, 5] ** 2)))[:, None]


def nodal_cauchy(coord, lnods, u, pk1):
    """caust [npoin, 6] (xx, yy, zz, yz, xz, xy) from the first Piola stress
    pk1 [ne, 8, 9] (column-major per Gauss point, the device's layout)."""
    lnods = np.asarray(lnods)
    gpvol, N, F = gp_kinematics(coord, lnods, u)
    P = np.asarray(pk1, dtype=np.float64).reshape(-1, 8, 3, 3).transpose(0, 1, 3, 2)
    sig = np.einsum("egij,egkj->egik", F, P) / np.linalg.det(F)[..., None, None]
    return _project(coord, lnods, gpvol, N, sig)


def values(lnods, elem, xi, fields, variables):
    """[nw, nvar] values of the named variables at located points; fields
    maps "u", "v", "a", "frx", "sig", "eps", "lep", "vm" to nodal arrays.  Points
    outside the mesh give NaN."""
    lnods = np.asarray(lnods)
    out = np.full((len(elem), len(variables)), np.nan)
    N, _ = shape(xi)
    for w, e in enumerate(elem):
        if e < 0:
            continue
        nodes = lnods[e]
        for j, name in enumerate(variables):
            key, comp = VARIABLES[name]
            out[w, j] = N[w] @ fields[key][nodes, comp]
    return out
