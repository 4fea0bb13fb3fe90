"""Oracle: Hex8 geometry, strain fields, internal force / tangent assembly (FP64).

Test infrastructure only (see ``oracle/__init__.py``).

Restates:

* structured box mesh          ``src/fem/mesh/structured.py:36-80``
* Hex8 shape functions         ``src/fem/elements/lagrange.py:50-76``
* tensor Gauss rule            ``src/fem/quadrature/rules.py:24-33`` (zeta fastest)
* 3-D B matrix                 ``src/fem/assembly/elasticity.py:160-183``
* quadrature geometry cache    ``src/fem/assembly/nonlinear_solid.py:391-425``
* strain / strain-rate fields  ``nonlinear_solid.py:428-454``
* force + tangent assembly     ``nonlinear_solid.py:457-493`` (COO -> CSR)
* consistent / lumped mass     ``src/fem/assembly/elasticity.py:50-77,109-141``

Pinning (tests/test_oracle_hex8.py): uniform-strain reaction equals
(lambda + 2G) eps at 1e-12 (``tests/unit/fem/test_nonlinear_solid.py:29-77``),
affine patch test, sum of w detJ = volume, element stiffness symmetric with
exactly six rigid-body null vectors, lumped mass sums to rho V.
"""

from __future__ import annotations

import itertools

import numpy as np
from numpy.polynomial.legendre import leggauss
from scipy.sparse import coo_matrix

SIGNS = np.asarray(
    [
        [-1.0, -1.0, -1.0],
        [1.0, -1.0, -1.0],
        [-1.0, 1.0, -1.0],
        [1.0, 1.0, -1.0],
        [-1.0, -1.0, 1.0],
        [1.0, -1.0, 1.0],
        [-1.0, 1.0, 1.0],
        [1.0, 1.0, 1.0],
    ]
)


def structured_box(lengths, divisions):
    """``build_structured_hex8_box``: node id = i + (nx+1)(j + (ny+1)k); element i fastest."""
    nx, ny, nz = divisions
    lx, ly, lz = lengths
    xs = np.linspace(0.0, lx, nx + 1)
    ys = np.linspace(0.0, ly, ny + 1)
    zs = np.linspace(0.0, lz, nz + 1)
    kk, jj, ii = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    nodes = np.stack([xs[ii.ravel()], ys[jj.ravel()], zs[kk.ravel()]], axis=1)
    ek, ej, ei = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ei, ej, ek = ei.ravel(), ej.ravel(), ek.ravel()

    def nid(i, j, k):
        return i + (nx + 1) * (j + (ny + 1) * k)

    elements = np.stack(
        [nid(ei + s[0], ej + s[1], ek + s[2]) for s in ((SIGNS + 1) // 2).astype(int)], axis=1
    ).astype(np.int64)
    return nodes, elements


def gauss_rule(order: int = 2):
    """``_tensor_product_rule(3, order)``: itertools.product, first coordinate slowest."""
    p1, w1 = leggauss(order)
    pts = np.asarray(list(itertools.product(p1, repeat=3)), dtype=np.float64)
    wts = np.asarray([np.prod(v) for v in itertools.product(w1, repeat=3)], dtype=np.float64)
    return pts, wts


def shape_functions(points):
    """``_hex8_shape_functions``: values (nq, 8) and reference gradients (nq, 8, 3)."""
    f = 1.0 + points[:, None, :] * SIGNS[None, :, :]
    values = 0.125 * np.prod(f, axis=2)
    grads = np.zeros((points.shape[0], 8, 3))
    grads[:, :, 0] = 0.125 * SIGNS[None, :, 0] * f[:, :, 1] * f[:, :, 2]
    grads[:, :, 1] = 0.125 * SIGNS[None, :, 1] * f[:, :, 0] * f[:, :, 2]
    grads[:, :, 2] = 0.125 * SIGNS[None, :, 2] * f[:, :, 0] * f[:, :, 1]
    return values, grads


def b_matrix(grads_phys):
    """``_build_three_dimensional_b_matrix`` batched over leading axes (..., 8, 3) -> (..., 6, 24)."""
    shp = grads_phys.shape[:-2]
    b = np.zeros(shp + (6, 24))
    gx, gy, gz = grads_phys[..., 0], grads_phys[..., 1], grads_phys[..., 2]
    b[..., 0, 0::3] = gx
    b[..., 1, 1::3] = gy
    b[..., 2, 2::3] = gz
    b[..., 3, 1::3] = gz
    b[..., 3, 2::3] = gy
    b[..., 4, 0::3] = gz
    b[..., 4, 2::3] = gx
    b[..., 5, 0::3] = gy
    b[..., 5, 1::3] = gx
    return b


def geometry_cache(nodes, elements, order: int = 2):
    """``_build_hex8_quadrature_geometry``: B (n_el, nq, 6, 24) and w*detJ (n_el, nq)."""
    pts, wts = gauss_rule(order)
    _, gref = shape_functions(pts)
    x = nodes[elements]  # (ne, 8, 3)
    jac = np.einsum("eai,qaj->eqij", x, gref)  # J = X^T dN
    det = np.linalg.det(jac)
    if np.any(det <= 0.0):
        raise ValueError("Hex8 nonlinear assembly requires positive Jacobian.")
    inv = np.linalg.inv(jac)
    gphys = np.einsum("qaj,eqjk->eqak", gref, inv)
    return b_matrix(gphys), det * wts[None, :]


def element_dofs(elements):
    return (3 * elements[:, :, None] + np.arange(3)[None, None, :]).reshape(elements.shape[0], 24)


def strain_fields(elements, b, u, u_prev, dt):
    """``_compute_hex8_strain_fields``: strains, strain rates (n_el*nq, 6)."""
    dofs = element_dofs(elements)
    ue = u[dofs]
    upe = u_prev[dofs]
    eps = np.matmul(b, ue[:, None, :, None])[..., 0].reshape(-1, 6)
    rate = (np.matmul(b, (ue - upe)[:, None, :, None])[..., 0] / dt).reshape(-1, 6)
    return eps, rate


def internal_force(elements, b, jw, stress, n_dof):
    """Force part of ``_assemble_force_and_tangent`` (element-ordered accumulation)."""
    ne, nq = jw.shape
    s = stress.reshape(ne, nq, 6)
    fe = np.sum(np.matmul(np.swapaxes(b, -1, -2), (s * jw[..., None])[..., None])[..., 0], axis=1)
    f = np.zeros(n_dof)
    np.add.at(f, element_dofs(elements).ravel(), fe.ravel())
    return f


def tangent_matrix(elements, b, jw, tangents, n_dof):
    """Tangent part of ``_assemble_force_and_tangent``: COO list -> CSR."""
    ne, nq = jw.shape
    c = tangents.reshape(ne, nq, 6, 6)
    bt = np.swapaxes(b, -1, -2)
    ke = np.einsum("eqab,eq->eab", np.matmul(np.matmul(bt, c), b), jw)
    dofs = element_dofs(elements)
    rows = np.repeat(dofs[:, :, None], 24, axis=2).ravel()
    cols = np.repeat(dofs[:, None, :], 24, axis=1).ravel()
    return coo_matrix((ke.ravel(), (rows, cols)), shape=(n_dof, n_dof)).tocsr()


def element_stiffness(b_e, jw_e, c):
    """Single-element K_e = sum_q B^T C B w (for null-space pins)."""
    return np.einsum("qia,ij,qjb,q->ab", b_e, c, b_e, jw_e)


def lumped_mass(nodes, elements, density, order: int = 2):
    """``assemble_lumped_mass_vector``: row sums of the consistent mass (per dof).

    ``density`` is a scalar or a per-element array.
    """
    pts, wts = gauss_rule(order)
    vals, gref = shape_functions(pts)
    x = nodes[elements]
    jac = np.einsum("eai,qaj->eqij", x, gref)
    det = np.linalg.det(jac)
    rho = np.broadcast_to(np.asarray(density, dtype=np.float64), (elements.shape[0],))
    # row sum of rho N^T N w detJ over a node: rho * N_a * sum_b N_b * w detJ = rho N_a w detJ
    me = np.einsum("e,qa,qb,eq->ea", rho, vals, vals, det * wts[None, :])
    m_node = np.zeros(nodes.shape[0])
    np.add.at(m_node, elements.ravel(), me.ravel())
    return np.repeat(m_node, 3)


def consistent_mass(nodes, elements, density, order: int = 2):
    """``assemble_consistent_mass_matrix`` (``assembly/elasticity.py:50-70,109-126``):
    M_e = sum_q rho N^T N w detJ with the bundle material's density, 3 dofs per
    node; CSR (3 n_nodes)^2."""
    pts, wts = gauss_rule(order)
    vals, gref = shape_functions(pts)
    x = nodes[elements]
    jac = np.einsum("eai,qaj->eqij", x, gref)
    det = np.linalg.det(jac)
    me = float(density) * np.einsum("qa,qb,eq->eab", vals, vals, det * wts[None, :])
    ne = elements.shape[0]
    dofs = element_dofs(elements)  # (ne, 24) node-major 3 a + c
    blk = np.zeros((ne, 24, 24))
    for c in range(3):
        blk[:, c::3, c::3] = me
    rows = np.repeat(dofs[:, :, None], 24, axis=2).ravel()
    cols = np.repeat(dofs[:, None, :], 24, axis=1).ravel()
    n_dof = 3 * nodes.shape[0]
    return coo_matrix((blk.ravel(), (rows, cols)), shape=(n_dof, n_dof)).tocsr()
