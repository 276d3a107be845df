"""Exact infinite-temperature FIDs of Sz-conserving two-spin Hamiltonians by sector diagonalisation.

F(t) = Tr(X(t) X) / (N 2^N), X = sum_s X_s, X(t) = e^{iHt} X e^{-iHt}: the quantity the engine's
readout estimates (sum of the single-site X coefficients over N, engine/src/tpgp_rev.c:311-388,
engine/src/cli_rev.c:576-581).

Method (test infrastructure, independent of the engine and of the oracle):
* H = sum over pairs of zz Z_iZ_j + xx X_iX_j + yy Y_iY_j with xx == yy per pair (the secular
  dipolar coupling of every benchmark model) conserves the number k of down spins, so it is
  block diagonal over sectors of dimension C(N, k); each block is real symmetric (ZZ diagonal,
  XX + YY a flip-flop of amplitude xx + yy) and is diagonalised with LAPACK.
* X couples sector k only to k +- 1, so Tr(X(t) X) = 2 sum_k sum_{a in k+1, b in k}
  |<a|X|b>|^2 cos((E_a - E_b) t).
* The global flip prod_s X_s commutes with H and with X and maps sector k to N - k, so the
  pair (k, k+1) contributes as much as (N-1-k, N-k): only pairs k <= (N-1)/2 are computed.
"""
from __future__ import annotations

import itertools
import math

import numpy as np


def _pair_couplings(n, terms):
    zz, ff = {}, {}
    xx, yy = {}, {}
    for (i, j, qi, qj, h) in terms:
        if i == j or not (0 <= i < n and 0 <= j < n):
            raise ValueError(f"bad sites {(i, j)}")
        if qi != qj or qi not in (1, 2, 3):
            raise ValueError("only ZZ, XX and YY couplings conserve Sz")
        key = (min(i, j), max(i, j))
        tab = {1: xx, 2: yy, 3: zz}[qi]
        tab[key] = tab.get(key, 0.0) + h
    for key in set(xx) | set(yy):
        if xx.get(key, 0.0) != yy.get(key, 0.0):
            raise ValueError(f"pair {key}: XX != YY, Sz not conserved")
        ff[key] = xx.get(key, 0.0) + yy.get(key, 0.0)
    return zz, ff


def sector_states(n, k):
    """Basis states with k set bits (bit s set: site s spin down), ascending."""
    out = np.zeros(math.comb(n, k), np.int64)
    for idx, sites in enumerate(itertools.combinations(range(n), k)):
        v = 0
        for s in sites:
            v |= 1 << s
        out[idx] = v
    out.sort()
    return out


def sector_hamiltonian(n, states, zz, ff):
    d = len(states)
    Hk = np.zeros((d, d))
    bits = (states[:, None] >> np.arange(n)[None, :]) & 1
    z = 1.0 - 2.0 * bits
    diag = np.zeros(d)
    for (i, j), h in zz.items():
        diag += h * z[:, i] * z[:, j]
    Hk[np.arange(d), np.arange(d)] = diag
    for (i, j), f in ff.items():
        if f == 0.0:
            continue
        src = np.nonzero(bits[:, i] != bits[:, j])[0]
        dst = np.searchsorted(states, states[src] ^ ((1 << i) | (1 << j)))
        Hk[dst, src] += f
    return Hk


def _raise_apply(n, lo_states, hi_states, V_lo):
    """sum_s sigma^-_s (k -> k+1 flips) applied to the columns of V_lo: (d_hi, ncols)."""
    out = np.zeros((len(hi_states), V_lo.shape[1]))
    for s in range(n):
        src = np.nonzero(((lo_states >> s) & 1) == 0)[0]
        dst = np.searchsorted(hi_states, lo_states[src] | (1 << s))
        out[dst] += V_lo[src]
    return out


def sector_fid(n, terms, times, progress=None):
    """F(t) at every t in `times` (float64 array)."""
    zz, ff = _pair_couplings(n, terms)
    t = np.asarray(times, np.float64)
    total = np.zeros(len(t))
    kmax = (n - 1) // 2
    prev = None
    for k in range(0, kmax + 2):
        states = sector_states(n, k)
        w, V = np.linalg.eigh(sector_hamiltonian(n, states, zz, ff))
        if progress:
            progress(f"sector {k}: dim {len(states)}")
        if prev is not None:
            ps, pw, pV = prev
            M = V.T @ _raise_apply(n, ps, states, pV)  # <a in k | X | b in k-1>
            W = M * M
            del M
            ca, sa = np.cos(np.outer(w, t)), np.sin(np.outer(w, t))
            cb, sb = np.cos(np.outer(pw, t)), np.sin(np.outer(pw, t))
            pair = np.sum(ca * (W @ cb) + sa * (W @ sb), axis=0)
            weight = 1.0 if (k - 1) == n - 1 - (k - 1) else 2.0  # mirror pair (N-k, N-k+1)
            total += 2.0 * weight * pair
            del W
        prev = (states, w, V)
    return total / (n * 2.0 ** n)


def second_moment(n, terms):
    """M2 = -F''(0) = Tr(C^dagger C) / (N 2^N) with C = [H, X], from the Pauli algebra alone:
    [P_i Q_j, X_i] = [P, X]_i Q_j with [Z, X] = 2iY, [Y, X] = -2iZ, [X, X] = 0, so C is a sum of
    Pauli strings and the trace is the sum of their squared total coefficients."""
    strings = {}
    for (i, j, qi, qj, h) in terms:
        for site, other in ((i, j), (j, i)):
            q = qi if site == i else qj
            qo = qj if site == i else qi
            if q == 1:
                continue  # X_site commutes with X_site
            # [h P_i P_j, X_site] = h [P_site, X_site] P_other: the product symbol at `site`
            sym = {2: 3, 3: 2}[q]
            key = tuple(sorted(((site, sym), (other, qo))))
            sign = 1.0 if q == 3 else -1.0  # [Z, X] = 2iY, [Y, X] = -2iZ
            strings[key] = strings.get(key, 0.0) + 2.0 * sign * h
    return sum(v * v for v in strings.values()) / n
