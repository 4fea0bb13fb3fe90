"""CPU check of the spectral reference-medium preconditioner's algebra.

The GPU PCG applies z = P K0^-1 (r - (sum r) delta_0) through two packed
complex FFTs (x + i y and z components), splitting the component spectra by
conjugate symmetry per mode pair (k, -k) and pinning class 0 by a translation.
This restates that algebra with NumPy FFTs and compares it with the dense
inverse of the pinned reduced stiffness T^T K0 T assembled by the oracle.
"""

import numpy as np
import pytest

from oracle import hex8, j2vp as OJ, periodic


def _symbol(b, jw, c0, nx, ny, nz):
    ke = hex8.element_stiffness(b, jw, c0).reshape(8, 3, 8, 3)
    S = np.zeros((3, 3, 3, 3, 3))
    for a in range(8):
        for bb in range(8):
            o = ((bb & 1) - (a & 1), ((bb >> 1) & 1) - ((a >> 1) & 1), ((bb >> 2) & 1) - ((a >> 2) & 1))
            S[o[0] + 1, o[1] + 1, o[2] + 1] += ke[a, :, bb, :]
    kinv = np.zeros((nz, ny, nx, 3, 3), complex)
    offs = [(ox, oy, oz) for ox in (-1, 0, 1) for oy in (-1, 0, 1) for oz in (-1, 0, 1)]
    for kz in range(nz):
        for ky in range(ny):
            for kx in range(nx):
                if kx == ky == kz == 0:
                    continue
                th = 2 * np.pi * np.array([kx / nx, ky / ny, kz / nz])
                A = sum(S[o[0] + 1, o[1] + 1, o[2] + 1] * np.exp(1j * (th @ o)) for o in offs)
                kinv[kz, ky, kx] = np.linalg.inv(A)
    return kinv


@pytest.mark.parametrize("div", [(4, 3, 5), (2, 2, 2), (5, 4, 3)])
def test_packed_spectral_inverse_matches_dense(div):
    nx, ny, nz = div
    nodes, el = hex8.structured_box((1.0, 0.9, 1.3), div)
    b, jw = hex8.geometry_cache(nodes, el)
    c0 = OJ.elasticity_matrix(30e9, 0.22)
    kf = hex8.tangent_matrix(el, b, jw, np.broadcast_to(c0, (el.shape[0] * 8, 6, 6)), 3 * nodes.shape[0])
    t = periodic.structured_constraint(nodes, div).transformation
    kred = (t.T @ kf @ t).toarray()
    n = nx * ny * nz
    kinv = _symbol(b[0], jw[0], c0, nx, ny, nz)
    rng = np.random.default_rng(0)
    r = rng.normal(size=(n, 3))
    r[0] = 0.0
    ua = np.fft.fftn((r[:, 0] + 1j * r[:, 1]).reshape(nz, ny, nx))
    ub = np.fft.fftn((r[:, 2] + 0j).reshape(nz, ny, nx))
    sr = r.sum(axis=0)
    va = np.zeros_like(ua)
    vb = np.zeros_like(ub)
    rz = 0.0
    for kz in range(nz):
        for ky in range(ny):
            for kx in range(nx):
                m = (-kz % nz, -ky % ny, -kx % nx)
                a, am = ua[kz, ky, kx], ua[m]
                spec = np.array([(a + np.conj(am)) / 2 - sr[0], (a - np.conj(am)) / 2j - sr[1],
                                 ub[kz, ky, kx] - sr[2]])
                z = kinv[kz, ky, kx] @ spec / n
                rz += np.real(np.conj(spec) @ z)
                va[kz, ky, kx] = z[0] + 1j * z[1]
                vb[kz, ky, kx] = z[2]
    za = np.fft.ifftn(va) * n
    zb = np.fft.ifftn(vb) * n
    z = np.stack([za.real.ravel(), za.imag.ravel(), zb.real.ravel()], 1)
    z -= z[0]
    dense = np.linalg.solve(kred, r[1:].ravel())
    assert np.max(np.abs(z[1:].ravel() - dense)) <= 1e-12 * np.max(np.abs(dense))
    assert rz == pytest.approx(r[1:].ravel() @ dense, rel=1e-12)
    assert np.max(np.abs(zb.imag)) <= 1e-12 * np.max(np.abs(zb.real))
