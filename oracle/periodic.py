"""Oracle: exact periodic constraint for structured Hex8 boxes (FP64).

Test infrastructure only (see ``oracle/__init__.py``).

Restates ``src/fem/rve/periodic.py``:
* class ids  (i mod nx) + nx((j mod ny) + ny(k mod nz))       ``:20-47``
* elimination operator T (0/1, reference class = class of node 0)  ``:50-104``
* opposite face node sets                                      ``:107-133``
* periodic fluctuation error                                   ``:136-150``
* affine map u = A eps (1/2 on engineering shear)              ``:153-169``

Pinning (tests/test_oracle_rve.py): class count = nx ny nz, rank T = 3(C-1),
T^T T diagonal with the class multiplicities, A eps reproduces x . eps for
every node (affine consistency at 1e-14), as ``tests/unit/fem/test_rve.py:36-63``.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.sparse import coo_matrix, csr_matrix


@dataclass
class Constraint:
    transformation: csr_matrix
    affine: np.ndarray
    class_ids: np.ndarray
    representatives: np.ndarray
    reference_class: int
    divisions: tuple
    faces: tuple


def structured_class_ids(divisions):
    nx, ny, nz = divisions
    k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    return ((i % nx) + nx * ((j % ny) + ny * (k % nz))).ravel().astype(np.int64)


def face_nodes(divisions):
    nx, ny, nz = divisions
    k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    idx = (i.ravel(), j.ravel(), k.ravel())
    ids = np.arange(i.size, dtype=np.int64)
    return tuple(
        (ids[idx[a] == 0], ids[idx[a] == divisions[a]]) for a in range(3)
    )


def affine_matrix(nodes):
    rel = nodes - np.min(nodes, axis=0)
    n = nodes.shape[0]
    a = np.zeros((3 * n, 6))
    x, y, z = rel[:, 0], rel[:, 1], rel[:, 2]
    a[0::3, 0] = x
    a[0::3, 4] = 0.5 * z
    a[0::3, 5] = 0.5 * y
    a[1::3, 1] = y
    a[1::3, 3] = 0.5 * z
    a[1::3, 5] = 0.5 * x
    a[2::3, 2] = z
    a[2::3, 3] = 0.5 * y
    a[2::3, 4] = 0.5 * x
    return a


def structured_constraint(nodes, divisions) -> Constraint:
    """``build_structured_hex8_periodic_constraint`` (``src/fem/rve/periodic.py:20-47``)."""
    return class_constraint(nodes, structured_class_ids(divisions), divisions, face_nodes(divisions))


def class_constraint(nodes, class_ids, divisions, faces) -> Constraint:
    """``build_periodic_constraint_from_classes`` (``src/fem/rve/periodic.py:50-104``):
    classes renumbered by ``np.unique``, each class's representative its first node,
    the reference class that of the first x- face node (its fluctuation is
    eliminated), T the 0/1 map from the other classes' fluctuations to the nodes."""
    cls_raw = np.asarray(class_ids, dtype=np.int64)
    uniq, cls = np.unique(cls_raw, return_inverse=True)
    cls = cls.astype(np.int64)
    count = uniq.size
    reps = np.full(count, -1, dtype=np.int64)
    order = np.arange(cls.size)[::-1]
    reps[cls[order]] = order  # first occurrence wins
    ref = int(cls[faces[0][0][0]])
    reduced = np.cumsum(np.arange(count) != ref) - 1  # class -> reduced id
    keep = cls != ref
    nodes_k = np.flatnonzero(keep)
    rows = (3 * nodes_k[:, None] + np.arange(3)[None, :]).ravel()
    cols = (3 * reduced[cls[nodes_k]][:, None] + np.arange(3)[None, :]).ravel()
    t = coo_matrix(
        (np.ones(rows.size), (rows, cols)), shape=(3 * nodes.shape[0], 3 * (count - 1))
    ).tocsr()
    return Constraint(t, affine_matrix(nodes), cls, reps, ref, tuple(divisions), faces)


def fluctuation_error(u, macro_strain, con: Constraint) -> float:
    fl = (u - con.affine @ macro_strain).reshape(-1, 3)
    mism = fl - fl[con.representatives[con.class_ids]]
    return float(np.max(np.linalg.norm(mism, axis=1)))


def face_reaction_error(f, con: Constraint) -> float:
    """``_compute_face_reaction_error`` (``src/fem/rve/solver.py:464-477``)."""
    nf = f.reshape(-1, 3)
    out = []
    for minus, plus in con.faces:
        a = np.sum(nf[minus], axis=0)
        b = np.sum(nf[plus], axis=0)
        den = max(float(np.linalg.norm(a) + np.linalg.norm(b)), 1.0)
        out.append(float(np.linalg.norm(a + b)) / den)
    return max(out)


def class_reaction_error(f, con: Constraint) -> float:
    """``_compute_periodic_class_reaction_error`` (``solver.py:480-494``)."""
    nf = f.reshape(-1, 3)
    scale = max(float(np.linalg.norm(f)), 1.0)
    sums = np.zeros((con.representatives.size, 3))
    np.add.at(sums, con.class_ids, nf)
    return float(np.max(np.linalg.norm(sums, axis=1)) / scale)
