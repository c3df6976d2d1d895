"""Seeded synthetic right-hand sides for the benchmark workloads.

Generator (the same for every workload):
  rng = numpy.random.Generator(PCG64(seed)), seed = 20260000 + 100*index + field
  (fields u=0, v=1, w=2, p=3; index 0 = CPU_33, the 33^2 x 16 CPU case, 1 = SOLVE_129,
  2 = SOLVE_257, 3 = STEP_129, the pseudo-spectral step, 4 = ADJOINT_129);
  g1, g2 ~ N(0,1) of shape (M, N+1, N+1) drawn in the order (mode, q, p);
  a_pq^(k) = (g1 + i g2) exp(-0.15 (p+q)) [times exp(-0.2 |k|) for STEP_129];
  f_k(y_j, z_l) = sum_pq a_pq T_p(y_j) T_q(z_l) at the CGL nodes y_j = -cos(pi j/N)
  (pseudomat.f90:83-85); the k=0 mode keeps only its real part.
Dirichlet fields get their boundary overwritten with 0 (helm3d.f90:48-60 does
the same); Neumann solves read only the interior.

STEP_129's velocity (duct_velocity) is the laminar square-duct profile plus a
divergence-free perturbation curl(A), A_c = (1-y^2)^2 (1-z^2)^2 h_c with h_c a
seeded Chebyshev series, coefficients (g1 + i g2) exp(-0.2 (p + q + |k|)),
scaled to 5% of the laminar maximum; seeds seed_for(3, c) for A_x, A_y, A_z.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

FIELDS = ("u", "v", "w", "p")
BOUNDS = "dddn"


@dataclass(frozen=True)
class Config:
    index: int
    m: int
    nmodes: int
    nfields: int
    alambda_dir: float = 1000.0   # u, v, w: (D^2 - (1000 + k^2)) u = f
    alambda_neu: float = 0.0      # p: pressure Poisson, xlambda = k^2
    alp: float = 1.0              # Lx = 2 pi -> k_x = m
    mode_decay: float = 0.0


CONFIGS = {
    0: Config(0, 33, 16, 1),
    1: Config(1, 129, 64, 4),
    2: Config(2, 257, 256, 4),
    3: Config(3, 129, 64, 4, mode_decay=0.2),
    4: Config(4, 129, 64, 4),
}
# The workloads by what they are (the index is only the seed key of seed_for):
CPU_33 = CONFIGS[0]        # 33^2 CGL x 16 modes, one Dirichlet field: the CPU-runnable case
SOLVE_129 = CONFIGS[1]     # 129^2 x 64 modes x u, v, w, p
SOLVE_257 = CONFIGS[2]     # 257^2 x 256 modes x u, v, w, p: the benchmarked solve stage
STEP_129 = CONFIGS[3]      # 129^2 x 64 modes, mgalx 192: the pseudo-spectral step
ADJOINT_129 = CONFIGS[4]   # 129^2 x 64 modes: the adjoint (gradient) checks


def cheb_nodes_matrix(m: int, deg: int) -> np.ndarray:
    """T_p(y_j) for y_j = -cos(pi j/N): (-1)^p cos(p pi j / N), shape (m, deg+1)."""
    N = m - 1
    j = np.arange(m)[:, None]
    p = np.arange(deg + 1)[None, :]
    phase = (p * j) % (2 * N)                    # exact integer reduction
    sign = np.where(p % 2 == 0, 1.0, -1.0)
    return sign * np.cos(np.pi * phase / N)


def seed_for(index: int, field: int) -> int:
    return 20260000 + 100 * int(index) + int(field)


def rhs_field(cfg: Config, field: int, m: int | None = None, nmodes: int | None = None) -> np.ndarray:
    """Complex128 (nmodes, m, m) [mode, z, y] right-hand side of one field."""
    m = cfg.m if m is None else m
    M = cfg.nmodes if nmodes is None else nmodes
    N = m - 1
    rng = np.random.Generator(np.random.PCG64(seed_for(cfg.index, field)))
    g1 = rng.standard_normal((M, N + 1, N + 1))
    g2 = rng.standard_normal((M, N + 1, N + 1))
    q = np.arange(N + 1)[:, None]
    p = np.arange(N + 1)[None, :]
    decay = np.exp(-0.15 * (p + q))              # indexed [q, p]
    a = (g1 + 1j * g2) * decay[None, :, :]
    if cfg.mode_decay:
        a *= np.exp(-cfg.mode_decay * np.arange(M))[:, None, None]
    T = cheb_nodes_matrix(m, N)                  # [node, degree]
    # f[k, z_l, y_j] = sum_q sum_p T[l, q] a[k, q, p] T[j, p]
    f = np.einsum("lq,kqp,jp->klj", T, a, T, optimize=True)
    f[0] = f[0].real
    if BOUNDS[field] == "d":
        f[:, 0, :] = 0
        f[:, -1, :] = 0
        f[:, :, 0] = 0
        f[:, :, -1] = 0
    return np.ascontiguousarray(f)


def alambda_of(cfg: Config, field: int) -> float:
    return cfg.alambda_dir if BOUNDS[field] == "d" else cfg.alambda_neu


def laminar_duct(m: int) -> np.ndarray:
    """Laminar square-duct profile U(z, y): (D_y^2 + D_z^2) U = -1, U = 0 on the walls,
    scaled to max U = 1 (the flow the reference's dpval table describes, operators.F:2486-2507)."""
    from .basis import collocation_matrix
    _, DC = collocation_matrix(m)
    D2 = (DC @ DC)[1:-1, 1:-1]
    lam, P = np.linalg.eig(D2)
    lam, P = lam.real, P.real
    Pi = np.linalg.inv(P)
    F = -np.ones((m - 2, m - 2))
    G = (Pi @ F @ Pi.T) / (lam[:, None] + lam[None, :])
    U = np.zeros((m, m))
    U[1:-1, 1:-1] = P @ G @ P.T
    return U / U.max()


def duct_velocity(cfg: Config, m: int | None = None, nmodes: int | None = None, amp: float = 0.05,
                  kdeg: int = 24):
    """u, v, w complex128 (nmodes, m, m) [mode, z, y] of the pseudo-spectral step (STEP_129)."""
    from numpy.polynomial import chebyshev as C
    m = cfg.m if m is None else m
    M = cfg.nmodes if nmodes is None else nmodes
    K = min(kdeg, m - 5)                         # g(y) h(y) stays of degree <= m-1
    x = -np.cos(np.pi * np.arange(m) / (m - 1))
    T = cheb_nodes_matrix(m, K - 1)              # T_p(x_j), [node, p]
    dT = np.stack([C.chebval(x, C.chebder(np.eye(K)[p])) for p in range(K)], axis=1)
    g, dg = (1 - x ** 2) ** 2, -4.0 * x * (1 - x ** 2)
    k = np.arange(M)
    decay = np.exp(-0.2 * (np.arange(K)[:, None] + np.arange(K)[None, :]))
    A, Ay, Az = [], [], []
    for c in range(3):
        rng = np.random.Generator(np.random.PCG64(seed_for(cfg.index, c)))
        a = (rng.standard_normal((M, K, K)) + 1j * rng.standard_normal((M, K, K))) * decay[None]
        a *= np.exp(-0.2 * k)[:, None, None]
        a[0] = a[0].real
        H = np.einsum("lq,kqp,jp->klj", T, a, T, optimize=True)       # [k, z, y]
        Hy = np.einsum("lq,kqp,jp->klj", T, a, dT, optimize=True)
        Hz = np.einsum("lq,kqp,jp->klj", dT, a, T, optimize=True)
        gz, gy = g[None, :, None], g[None, None, :]
        A.append(gz * gy * H)
        Ay.append(gz * (dg[None, None, :] * H + gy * Hy))
        Az.append(gy * (dg[None, :, None] * H + gz * Hz))
    ikx = (1j * cfg.alp * k)[:, None, None]
    up = Ay[2] - Az[1]                            # u' = dA_z/dy - dA_y/dz
    vp = Az[0] - ikx * A[2]                       # v' = dA_x/dz - dA_z/dx
    wp = ikx * A[1] - Ay[0]                       # w' = dA_y/dx - dA_x/dy
    eps = amp / max(np.abs(up).max(), np.abs(vp).max(), np.abs(wp).max())
    u, v, w = eps * up, eps * vp, eps * wp
    u[0] += laminar_duct(m)
    for f in (u, v, w):
        f[0] = f[0].real
    return [np.ascontiguousarray(f) for f in (u, v, w)]
