"""Synthetic workloads of the upwind path, fully determined by committed seeds.

Each workload is a 3-D periodic box of lengths (1, 1, 1):
  * dt = (cfl * dx) / (|v0| + |v1| + |v2|) in FP64, left to right, with
    dx = min_j L_j/N_j (cubic grids here), computed once and handed to the
    GPU and the oracle as the same double;
  * random draws use SplitMix64 in counter form, u = (z >> 11) * 2^-53
    (the same generator the library's init kernels and the oracle use).

Workloads (named by what they are):
  sphere128      128^3, 10 steps, v=(1,1,1), CFL 0.5, sphere indicator
                 (|x-0.5|^2 < 0.15^2) -- the CPU-runnable FP64 case.
  sphere_noise512 512^3, 100 steps, seeded mixed-sign v in [-1,1]^3,
                 sphere + 0.01*U noise -- the FP64 headline workload.
  gauss1024      1024^3, 200 steps, v=(0.7,-0.4,0.2), exp(-|x-c|^2/0.005)
                 with c_j = 0.3+0.4u_j -- mixed precision.
  misfit256      256^3, T=50: gradient of 0.5*sum((f_T - f_target)^2).
  misfit_large   1024x1024x768, T=400, 8 Gaussian bubbles
                 a*exp(-|x-c|^2/w^2) -- checkpointed adjoint.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Tuple

MASK64 = (1 << 64) - 1


def splitmix64(seed: int, i: int) -> int:
    z = (seed + (i + 1) * 0x9E3779B97F4A7C15) & MASK64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


def uniform(seed: int, i: int) -> float:
    return (splitmix64(seed, i) >> 11) * 2.0 ** -53


def cfl_dt(num_cells, velocity, lengths=(1.0, 1.0, 1.0), cfl: float = 0.5) -> float:
    dx = min(L / n for L, n in zip(lengths, num_cells))
    s = 0.0
    for v in velocity:
        s = s + abs(v)
    return (cfl * dx) / s


def mixed_sign_velocity(seed: int) -> Tuple[float, float, float]:
    """v_j = 2u-1 from consecutive draws, redrawn until the signs are mixed."""
    k = 0
    while True:
        v = tuple(2.0 * uniform(seed, k + j) - 1.0 for j in range(3))
        k += 3
        if any(x < 0 for x in v) and any(x > 0 for x in v):
            return v


# init kinds, as in fdb_init_kind
INIT_DELTA, INIT_SPHERE, INIT_GAUSS, INIT_GAUSS_SUM, INIT_UNIFORM = range(5)


@dataclass(frozen=True)
class Workload:
    name: str
    num_cells: Tuple[int, int, int]
    steps: int
    velocity: Tuple[float, float, float]
    init_kind: int
    init_params: Tuple[float, ...]
    init_seed: int = 0
    lengths: Tuple[float, float, float] = (1.0, 1.0, 1.0)
    cfl: float = 0.5
    dt_velocity: Tuple[float, ...] | None = None  # velocity that fixes dt (adjoint)
    extra: dict = field(default_factory=dict, compare=False, hash=False)

    @property
    def dt(self) -> float:
        return cfl_dt(self.num_cells, self.dt_velocity or self.velocity, self.lengths, self.cfl)

    @property
    def ntot(self) -> int:
        n0, n1, n2 = self.num_cells
        return n0 * n1 * n2


SPHERE = (0.5, 0.5, 0.5, 0.15, 0.0)


def sphere128() -> Workload:
    return Workload("sphere128", (128, 128, 128), 10, (1.0, 1.0, 1.0), INIT_SPHERE, SPHERE)


def sphere_noise512(n: int = 512, steps: int = 100) -> Workload:
    return Workload(f"sphere_noise{n}", (n, n, n), steps, mixed_sign_velocity(2), INIT_SPHERE,
                    (0.5, 0.5, 0.5, 0.15, 0.01), init_seed=12345)


def gauss1024(n: int = 1024, steps: int = 200) -> Workload:
    c = tuple(0.3 + 0.4 * uniform(3, j) for j in range(3))
    return Workload(f"gauss{n}", (n, n, n), steps, (0.7, -0.4, 0.2), INIT_GAUSS,
                    c + (0.005, 1.0))


def misfit256(n: int = 256, steps: int = 50) -> Workload:
    """Evaluation point of the gradient workload; the target run is
    misfit_target(); both use dt from sum|v| = 1.6."""
    return Workload(f"misfit{n}", (n, n, n), steps, (0.9, 0.25, -0.45), INIT_SPHERE,
                    (0.5, 0.5, 0.5, 0.15, 0.01), init_seed=777,
                    dt_velocity=(0.8, 0.3, -0.5))


def misfit_target(n: int = 256, steps: int = 50) -> Workload:
    return Workload(f"misfit_target{n}", (n, n, n), steps, (0.8, 0.3, -0.5), INIT_SPHERE, SPHERE,
                    dt_velocity=(0.8, 0.3, -0.5))


def gauss_sum_params(seed: int = 5, m: int = 8) -> Tuple[float, ...]:
    """{m, (cx, cy, cz, w, a) x m}: centres in [0.2,0.8], w in [0.02,0.1],
    a in [0.5,1], for sum_k a_k exp(-|x-c_k|^2 / w_k^2)."""
    p = [float(m)]
    k = 0
    for _ in range(m):
        cx, cy, cz = (0.2 + 0.6 * uniform(seed, k + j) for j in range(3))
        w = 0.02 + 0.08 * uniform(seed, k + 3)
        a = 0.5 + 0.5 * uniform(seed, k + 4)
        k += 5
        p += [cx, cy, cz, w, a]
    return tuple(p)


def misfit_large(shape=(1024, 1024, 768), steps: int = 400) -> Workload:
    return Workload("misfit_large", tuple(shape), steps, (0.9, 0.25, -0.45), INIT_GAUSS_SUM,
                    gauss_sum_params(), dt_velocity=(0.8, 0.3, -0.5))
