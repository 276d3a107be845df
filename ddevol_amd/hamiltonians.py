"""Term tables of the benchmark models, in the reference CLI's order.

A term is (i, j, qi, qj, h) with Pauli symbols I=0, X=1, Y=2, Z=3 (engine/include/tpgp_rev.h:30-35).

* ``edge`` follows engine/src/cli_rev.c:125-132: the secular dipolar coupling d(2ZZ-XX-YY)/4
  becomes ZZ with h=d/2, XX with h=-d/4, YY with h=-d/4, in that order; ``append`` skips h == 0
  (cli_rev.c:108-124).
* ``ring`` / ``j1j2`` follow cli_rev.c:235-245 (for each i: bond (i, i+1), then (i, i+2)).
* ``cubic`` follows cli_rev.c:183-207 (periodic lattice, nearest and face-diagonal shells).
* ``InputPCG`` restates engine/include/pcg_inputs.hpp:7-37 (PCG XSL-RR 128/64 with the frozen
  states of seeds 7 and 20260711).

The remaining builders define the benchmark workloads that have no reference driver (powder
cluster, range-4 ring, serving rings); their inputs are seeded and fully determined here.
"""
from __future__ import annotations

import math

MASK64 = (1 << 64) - 1
MASK128 = (1 << 128) - 1
_PCG_MULT = 47026247687942121848144207491837523525
_PCG_STATES = {
    7: (208745520555909116978795849195383758904, 261136684632268670825940853076396136793),
    20260711: (233554721566408405256262523697011956744, 211610054476143376060500860209735615305),
}


class InputPCG:
    """pcg_inputs.hpp:7-37.  Only the two frozen streams exist."""

    def __init__(self, seed: int):
        if seed not in _PCG_STATES:
            raise ValueError("only input streams 7 and 20260711 are defined")
        self.state, self.increment = _PCG_STATES[seed]

    def next(self) -> int:
        self.state = (self.state * _PCG_MULT + self.increment) & MASK128
        x = ((self.state >> 64) ^ self.state) & MASK64
        rot = self.state >> 122
        return ((x >> rot) | (x << ((-rot) & 63))) & MASK64

    def uniform(self) -> float:
        return (self.next() >> 11) * 2.0 ** -53


def edge(terms, i, j, d, model="secular"):
    def append(a, b, h):
        if h != 0:
            terms.append((i, j, a, b, h))

    if model == "ising":
        append(3, 3, d / 4)
    elif model == "xy":
        append(1, 1, d / 4)
        append(2, 2, d / 4)
    else:
        append(3, 3, d / 2)
        append(1, 1, -d / 4)
        append(2, 2, -d / 4)
    return terms


def ring(n, model="secular"):
    t = []
    for i in range(n):
        edge(t, i, (i + 1) % n, 1.0, model)
    return t


def j1j2(n, j2):
    """j2: a float (uniform) or a sequence of per-bond values J2_i for bond (i, i+2)."""
    t = []
    for i in range(n):
        edge(t, i, (i + 1) % n, 1.0)
        edge(t, i, (i + 2) % n, j2[i] if hasattr(j2, "__len__") else j2)
    return t


def edges(edge_list, model="secular"):
    t = []
    for (i, j, d) in edge_list:
        edge(t, i, j, d, model)
    return t


def cubic(L):
    """Periodic [001] cubic lattice of L^3 sites, r^2 in {1, 2} (cli_rev.c:183-207)."""
    n = L ** 3
    t = []
    for i in range(n):
        for j in range(i + 1, n):
            a = (i % L, (i // L) % L, i // (L * L))
            b = (j % L, (j // L) % L, j // (L * L))
            r2 = 0
            dz = 0
            for k in range(3):
                d = b[k] - a[k]
                if d > L // 2:
                    d -= L
                if d < -(L // 2):
                    d += L
                r2 += d * d
                if k == 2:
                    dz = d
            if r2 in (1, 2):
                coupling = (1 - 3.0 * dz * dz / r2) / (r2 * math.sqrt(float(r2)))
                edge(t, i, j, coupling)
    return t


# ---- benchmark workloads -----------------------------------------------------------------
def j1j2_n16_couplings(n=16):
    """J2_i = 0.1 + 0.3 u_i with u_i the first n uniforms of InputPCG(7)."""
    g = InputPCG(7)
    return [0.1 + 0.3 * g.uniform() for _ in range(n)]


def j1j2_n16_terms(n=16):
    return j1j2(n, j1j2_n16_couplings(n))


def powder_orientations(count=512):
    """Unit vectors uniform on the sphere from InputPCG(20260711): z = 2u-1, phi = 2 pi u'."""
    g = InputPCG(20260711)
    out = []
    for _ in range(count):
        z = 2.0 * g.uniform() - 1.0
        phi = 2.0 * math.pi * g.uniform()
        r = math.sqrt(max(0.0, 1.0 - z * z))
        out.append((r * math.cos(phi), r * math.sin(phi), z))
    return out


def cluster_pairs(L=3, r2max=3):
    """Open simple-cubic L^3 cluster, site x + L y + L^2 z; pairs i<j with r^2 <= r2max."""
    n = L ** 3
    pos = [(s % L, (s // L) % L, s // (L * L)) for s in range(n)]
    pairs = []
    for i in range(n):
        for j in range(i + 1, n):
            d = [pos[j][k] - pos[i][k] for k in range(3)]
            r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2]
            if r2 <= r2max:
                pairs.append((i, j, d, r2))
    return pairs


def powder_terms(orientation, L=3, r2max=3):
    """b_ij = (1 - 3 cos^2 theta_ij) / r^3 for field direction `orientation`."""
    t = []
    for (i, j, d, r2) in cluster_pairs(L, r2max):
        r = math.sqrt(r2)
        cos_t = (d[0] * orientation[0] + d[1] * orientation[1] + d[2] * orientation[2]) / r
        b = (1.0 - 3.0 * cos_t * cos_t) / (r2 * r)
        edge(t, i, j, b)
    return t


def range_ring(n=64, rmax=4):
    """Ring with 1/r^3 couplings for r = 1..rmax, order: for i, for r."""
    t = []
    for i in range(n):
        for r in range(1, rmax + 1):
            edge(t, i, (i + r) % n, 1.0 / r ** 3)
    return t


def serving_ring(n, nn):
    """Ring with nearest-neighbour couplings nn[i] and next-nearest 1/2^3."""
    t = []
    for i in range(n):
        edge(t, i, (i + 1) % n, nn[i])
        edge(t, i, (i + 2) % n, 0.125)
    return t


def global_x(n):
    """Initial observable sum_s X_s with coefficients exactly 1 (cli_rev.c:514-523)."""
    return [(s, 1, 1.0) for s in range(n)]
