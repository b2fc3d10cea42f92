"""Independent numpy formulation of the PM path, for tiny sizes.

Written from the mathematical definitions (periodic hat-function CIC with
stagger, numpy's pocketfft for the transforms, ``np.roll`` central differences),
not from the oracle's loops, so that agreement between the two pins index,
sign and layout mistakes in either.  Arrays are indexed ``[z, y, x]``.
Used by ``make_golden.py`` to write the committed fixtures and by the tests.
"""
from __future__ import annotations

import numpy as np


def hat_deposit(pos, N, q, stagger=(0.0, 0.0, 0.0), ngp=0):
    """sum_p q_p prod_d hat(pos_d*N - n_d - s_d) over periodic images; [z,y,x].
    Axes in the bit mask ``ngp`` take the nearest-grid-point kernel (the indicator
    of |pos_d*N - n_d - s_d| <= 1/2, upper node on a tie) instead of the hat."""
    out = np.zeros((N, N, N))
    t = pos * N - np.asarray(stagger)[None, :]
    n0 = np.floor(t).astype(np.int64)
    f = t - n0
    for d in range(3):
        if (ngp >> d) & 1:
            n0[:, d] = np.floor(t[:, d] + 0.5).astype(np.int64)
            f[:, d] = 0.0
    for i in (0, 1):
        for j in (0, 1):
            for k in (0, 1):
                w = (f[:, 0] if i else 1 - f[:, 0]) * (f[:, 1] if j else 1 - f[:, 1]) * \
                    (f[:, 2] if k else 1 - f[:, 2])
                x = (n0[:, 0] + i) % N
                y = (n0[:, 1] + j) % N
                z = (n0[:, 2] + k) % N
                np.add.at(out, (z, y, x), q * w)
    return out


def deposits(pos, vel, mass, N):
    """scalar rho, staggered vector B (3), VecVec T (6, symmetric order)."""
    inv_vol = float(N) ** 3
    m = mass * inv_vol
    rho = hat_deposit(pos, N, m)
    B = np.stack([hat_deposit(pos, N, m * vel[:, a], tuple(0.5 if d == a else 0.0 for d in range(3)))
                  for a in range(3)])
    pairs = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    T = []
    for a, b in pairs:
        st = tuple(0.0 if a == b else (0.5 if d in (a, b) else 0.0) for d in range(3))
        T.append(hat_deposit(pos, N, m * vel[:, a] * vel[:, b], st))
    return rho, B, np.stack(T)


def green_eigen(N):
    """K^2 = sum_i N^2 (2 - 2 cos(2 pi k_i / N)) on the rfftn grid [kz, ky, kx]"""
    k = np.arange(N)
    s = N * N * (2.0 - 2.0 * np.cos(2.0 * np.pi * k / N))
    sx = s[: N // 2 + 1]
    return s[:, None, None] + s[None, :, None] + sx[None, None, :]


def poisson(rho, N):
    d = rho - rho.mean()
    K = np.fft.rfftn(d)
    K2 = green_eigen(N)
    K2[0, 0, 0] = 1.0
    phik = -K / K2
    phik[0, 0, 0] = 0.0
    return np.fft.irfftn(phik, s=(N, N, N), axes=(0, 1, 2))


def force(phi, N):
    """F_d = -(phi(n+e_d) - phi(n-e_d)) N/2; returns [3, z, y, x] with d = x, y, z"""
    F = []
    for ax in (2, 1, 0):  # x, y, z axes of a [z,y,x] array
        F.append(-(np.roll(phi, -1, axis=ax) - np.roll(phi, 1, axis=ax)) * N / 2.0)
    return np.stack(F)


def gather(pos, F, N):
    t = pos * N
    n0 = np.floor(t).astype(np.int64)
    f = t - n0
    acc = np.zeros_like(pos)
    for i in (0, 1):
        for j in (0, 1):
            for k in (0, 1):
                w = (f[:, 0] if i else 1 - f[:, 0]) * (f[:, 1] if j else 1 - f[:, 1]) * \
                    (f[:, 2] if k else 1 - f[:, 2])
                x = (n0[:, 0] + i) % N
                y = (n0[:, 1] + j) % N
                z = (n0[:, 2] + k) % N
                for d in range(3):
                    acc[:, d] += w * F[d][z, y, x]
    return acc


def pm_step(pos, vel, mass, N, dt_kick, dt_drift):
    rho = hat_deposit(pos, N, mass * float(N) ** 3)
    phi = poisson(rho, N)
    F = force(phi, N)
    acc = gather(pos, F, N)
    vel2 = vel + dt_kick * acc
    pos2 = pos + dt_drift * vel2
    pos2 = np.where(pos2 < 0, pos2 + 1.0, pos2)
    pos2 = np.where(pos2 >= 1.0, pos2 - 1.0, pos2)
    return {"rho": rho, "phi": phi, "acc": acc, "vel": vel2, "pos": pos2}


def deposits_ngp(pos, vel, mass, N):
    """the CIC-NGP hybrids: B (3), component a on the link grid along a with the
    nearest-grid-point kernel there and CIC elsewhere; T (6, symmetric order), T_aa
    CIC on the vertices, T_ab (a<b) NGP on the links along a and b, CIC on the third."""
    m = mass * float(N) ** 3
    B = np.stack([hat_deposit(pos, N, m * vel[:, a], tuple(0.5 if d == a else 0.0 for d in range(3)),
                              ngp=1 << a) for a in range(3)])
    pairs = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    T = []
    for a, b in pairs:
        if a == b:
            T.append(hat_deposit(pos, N, m * vel[:, a] * vel[:, a]))
        else:
            st = tuple(0.5 if d in (a, b) else 0.0 for d in range(3))
            T.append(hat_deposit(pos, N, m * vel[:, a] * vel[:, b], st, ngp=(1 << a) | (1 << b)))
    return B, np.stack(T)
