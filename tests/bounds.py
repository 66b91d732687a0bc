"""Error bounds the GPU tests use to compare two solutions of one discrete
Poisson problem that each stop at a residual tolerance.

For the 7-point Laplacian Lap_h with ghost-node Neumann faces, two iterates
u_a, u_b of one problem satisfy Lap_h (u_a - u_b) = r_b - r_a, so
|u_a - u_b|inf <= ||Lap_h^-1||inf (|r_a|inf + |r_b|inf).  ||Lap_h^-1||inf =
max w with -Lap_h w = 1 (Lap_h is an M-matrix up to sign), and a parabola
along one axis bounds w: t (L - t) / 2 (Dirichlet on both faces of that axis,
max L^2 / 8) or t (2L - t) / 2 (Dirichlet on one face, Neumann on the other,
max L^2 / 2) solves the one-axis problem exactly, ghost nodes included, is
>= 0 on every other Dirichlet face and satisfies every other homogeneous
Neumann face, so w <= its maximum (discrete maximum principle).
"""
import math

import numpy as np

W, E, S, N, B, T = 3, 2, 1, 0, 5, 4  # Location_t order N,S,E,W,T,B


def linv_bound(dims, h, bc):
    """Upper bound of ||Lap_h^-1||inf for a grid of `dims` points, spacing h,
    face types bc ("N"/"D" per face in N,S,E,W,T,B order)."""
    best = math.inf
    for ax, (lo, hi) in enumerate(((W, E), (S, N), (B, T))):
        length = (dims[ax] - 1) * h
        dlo, dhi = bc[lo] == "D", bc[hi] == "D"
        if dlo and dhi:
            best = min(best, length * length / 8.0)
        elif dlo or dhi:
            best = min(best, length * length / 2.0)
    if not math.isfinite(best):
        raise ValueError("no Dirichlet face: the problem is singular")
    return best


def gap_bound(dims, h, bc, rel_a, rel_b, fnorm, df=0.0, umax=1.0):
    """Largest |u_a - u_b|inf allowed for two solutions with relative
    residuals rel_a, rel_b (||r|| / ||f||inf, ||f||inf = fnorm) of right-hand
    sides that differ by at most df, plus 8 ulp of max |u| for the rounding
    of the iterates to FP64; 25 % margin on the residual terms."""
    B_ = linv_bound(dims, h, bc)
    return 1.25 * B_ * ((rel_a + rel_b) * fnorm + df) + 8.0 * np.spacing(max(umax, 1e-300))
