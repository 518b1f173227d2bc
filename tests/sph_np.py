"""Brute-force FP64 NumPy statements of reference formulas, for tests.

Independent of the oracle's data structures: O(N^2) pair sets instead of cell
grids and capped lists, dense operators instead of matrix-free sweeps.  Each
function cites the reference expression it states.  Used only by tests/.
"""
from __future__ import annotations

import numpy as np


def kernel_w(r, h):
    """cubic_kernel, SPH/main.py:859-880 (Monaghan's spline, sigma = 8 / (pi h^3))."""
    q = np.asarray(r, dtype=np.float64) / h
    k = 8.0 / (np.pi * h ** 3)
    w = np.where(q <= 0.5, k * (6 * q ** 3 - 6 * q ** 2 + 1), 2 * k * (1 - q) ** 3)
    return np.where(q <= 1.0, w, 0.0)


def kernel_dw(r, h):
    """d/dr of kernel_w: the scalar of cubic_kernel_derivative, :883-905."""
    q = np.asarray(r, dtype=np.float64) / h
    k = 8.0 / (np.pi * h ** 3)
    d = np.where(q <= 0.5, 6 * k * q * (3 * q - 2) / h, -6 * k * (1 - q) ** 2 / h)
    return np.where((q <= 1.0) & (np.asarray(r) > 1e-5), d, 0.0)


def kernel_grad(rv, h):
    """cubic_kernel_derivative(r) = dW/dr * r / |r| (:883-905); rv (..., 3)."""
    rv = np.asarray(rv, dtype=np.float64)
    rn = np.linalg.norm(rv, axis=-1)
    safe = np.where(rn > 0, rn, 1.0)
    return (kernel_dw(rn, h) / safe)[..., None] * rv


def pairs(x, h):
    """Directed neighbor pairs (i, j), j != i, |x_i - x_j| < h evaluated in FP32
    as the reference does (store_neighbors, :845)."""
    x32 = np.asarray(x, dtype=np.float32)
    d = x32[:, None, :] - x32[None, :, :]
    r = np.sqrt(np.sum(d * d, axis=-1, dtype=np.float32), dtype=np.float32)
    mask = r < np.float32(h)
    np.fill_diagonal(mask, False)
    return mask


class Scene:
    """Positions, velocities and kinds (0 fluid, 1 static, 2 elastic) with the
    reference's derived constants (Parameter, :598-621)."""

    def __init__(self, x, v, kind, r=0.01, rho0=1000.0, nu=0.01, bnu=0.0):
        self.x = np.asarray(x, dtype=np.float64)
        self.v = np.asarray(v, dtype=np.float64)
        self.kind = np.asarray(kind)
        self.h = 4 * r
        self.V0 = 0.8 * (2 * r) ** 3
        self.m = self.V0 * 1000.0  # initialize_particles, :743
        self.rho0 = rho0
        self.nu = nu
        self.bnu = bnu
        self.nb = pairs(self.x, self.h)
        self.fluid = self.kind == 0
        self.solid = ~self.fluid
        self.dyn_solid = self.kind == 2
        self.rv = self.x[:, None, :] - self.x[None, :, :]  # r_ij = x_i - x_j
        self.rn = np.linalg.norm(self.rv, axis=-1)
        self.W = np.where(self.nb, kernel_w(self.rn, self.h), 0.0)
        self.gW = np.where(self.nb[..., None], kernel_grad(self.rv, self.h), 0.0)
        self.vol = self._volumes()
        self.rho = self._densities()

    def _volumes(self):
        """V0 for fluid; 3 / (W0 + sum over solid neighbors W) for solids
        (compute_*_boundary_volume, :1010-1035)."""
        W0 = kernel_w(0.0, self.h)
        vol = np.full(len(self.x), self.V0)
        ws = (self.W * self.solid[None, :]).sum(axis=1)
        vol[self.solid] = 3.0 / (W0 + ws[self.solid])
        return vol

    def _densities(self):
        """rho_i = rho0 (V_i W0 + sum_j V_j W_ij), every neighbor (:962-984)."""
        W0 = kernel_w(0.0, self.h)
        return self.rho0 * (self.vol * W0 + self.W @ self.vol)

    def viscosity_acc(self):
        """ViscosityModel.compute_viscosity_force_task, :1427-1460: the fluid term,
        the boundary term when boundary_viscosity != 0, and its reaction
        a_j -= f_v on dynamic solids j."""
        d = 10.0  # 2 (dim + 2)
        eps = 0.01 * self.h ** 2
        vij = self.v[:, None, :] - self.v[None, :, :]
        vxy = np.sum(vij * self.rv, axis=-1)
        base = vxy / (self.rn ** 2 + eps)
        a = np.zeros_like(self.x)
        ff = self.fluid[:, None] & self.fluid[None, :] & self.nb
        coef = np.where(ff, d * self.nu * (self.m / self.rho)[None, :] * base, 0.0)
        a += np.einsum("ij,ijk->ik", coef, self.gW)
        if self.bnu != 0.0:
            fs = self.fluid[:, None] & self.solid[None, :] & self.nb
            rho_i = np.where(self.fluid, self.rho, 1.0)
            cb = np.where(fs, d * self.bnu * (self.rho0 * self.vol[None, :] / rho_i[:, None]) * base, 0.0)
            fv = cb[..., None] * self.gW  # f_v of pair (fluid i, solid j)
            a += fv.sum(axis=1)
            react = -(fv * self.dyn_solid[None, :, None]).sum(axis=0)
            a += react
        a[~self.fluid & ~self.dyn_solid] = 0.0
        return a

    def weiler_matrix(self, dt, mu):
        """Dense 3nf x 3nf Weiler 2018 operator over the fluid particles in
        index order: A v = v - dt / rho_i * sum_{fluid j} 10 mu_ij (m_j / rho_j)
        ((v_i - v_j) . r_ij) / (|r_ij|^2 + 0.01 h^2) gradW_ij, mu_ij the mean of
        the two dynamic viscosities (mu = rho0 * kinematic)."""
        idx = np.flatnonzero(self.fluid)
        nf = len(idx)
        eps = 0.01 * self.h ** 2
        A = np.eye(3 * nf)
        for a_, i in enumerate(idx):
            blk_ii = np.zeros((3, 3))
            for b_, j in enumerate(idx):
                if not self.nb[i, j]:
                    continue
                muij = 0.5 * (mu[i] + mu[j])
                c = 10.0 * muij * (self.m / self.rho[j]) / (self.rn[i, j] ** 2 + eps)
                # (v_i - v_j) . r  times gradW: outer(gradW, r)
                blk = dt / self.rho[i] * c * np.outer(self.gW[i, j], self.rv[i, j])
                blk_ii += blk
                A[3 * a_:3 * a_ + 3, 3 * b_:3 * b_ + 3] += blk
            A[3 * a_:3 * a_ + 3, 3 * a_:3 * a_ + 3] -= blk_ii
        return A


def lattice(n, spacing=0.02, origin=0.2):
    """n^3 (or nx x ny x nz) points of a cubic lattice, x slowest."""
    ns = (n, n, n) if np.isscalar(n) else tuple(n)
    axes = [np.arange(k) * spacing + origin for k in ns]
    return np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3)
