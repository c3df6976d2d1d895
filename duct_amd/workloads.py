"""Seeded synthetic right-hand sides for the benchmark configurations.

Generator (the same for every configuration):
  rng = numpy.random.Generator(PCG64(seed)), seed = 20260000 + 100*config + field
  (fields u=0, v=1, w=2, p=3; config 0 = 33^2 x 16 CPU case, 1 = 129^2 x 64,
  2 = 257^2 x 256, 3 = pseudo-spectral step, 4 = adjoint);
  g1, g2 ~ N(0,1) of shape (M, N+1, N+1) drawn in the order (mode, q, p);
  a_pq^(k) = (g1 + i g2) exp(-0.15 (p+q)) [times exp(-0.2 |k|) for config 3];
  f_k(y_j, z_l) = sum_pq a_pq T_p(y_j) T_q(z_l) at the CGL nodes y_j = -cos(pi j/N)
  (pseudomat.f90:83-85); the k=0 mode keeps only its real part.
Dirichlet fields get their boundary overwritten with 0 (helm3d.f90:48-60 does
the same); Neumann solves read only the interior.
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


def cheb_nodes_matrix(m: int, deg: int) -> np.ndarray:
    """T_p(y_j) for y_j = -cos(pi j/N): (-1)^p cos(p pi j / N), shape (m, deg+1)."""
    N = m - 1
    j = np.arange(m)[:, None]
    p = np.arange(deg + 1)[None, :]
    phase = (p * j) % (2 * N)                    # exact integer reduction
    sign = np.where(p % 2 == 0, 1.0, -1.0)
    return sign * np.cos(np.pi * phase / N)


def seed_for(config: int, field: int) -> int:
    return 20260000 + 100 * int(config) + int(field)


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
