"""
Pure-Python loop restatement of the D4 pair and triple sums for molecules
with a handful of atoms (TEST INFRASTRUCTURE ONLY).

The three-body sum is written over *unordered* triples {a, b, c} with the
per-atom partition that the reference's ordered sum implies
(``dispersion/threebody.py:147-163``: atom i collects e/6 for every ordered
(j, k) with r_ij and r_jk inside the cutoff, r_ik unscreened):

    E_a += e/6 [bc]([ab] + [ac]),  and cyclically for b and c,

with [xy] = 1 if r_xy^2 <= cutoff^2. Agreement with ``dense`` therefore
checks the mask semantics independently of the dense tensor code.
"""

from __future__ import annotations

import math
from itertools import combinations


def _dist(p, q):
    return math.sqrt(sum((a - b) ** 2 for a, b in zip(p, q)))


def cn_d4(numbers, positions, rcov, en, cutoff=30.0, kcn=7.5,
          k4=4.10451, k5=19.08857, k6=2 * 11.28174**2):
    n = len(numbers)
    cn = [0.0] * n
    for i in range(n):
        for j in range(n):
            if i == j or numbers[i] == 0 or numbers[j] == 0:
                continue
            r = _dist(positions[i], positions[j])
            if r > cutoff:
                continue
            rc = rcov[numbers[i]] + rcov[numbers[j]]
            den = k4 * math.exp(-((abs(en[numbers[i]] - en[numbers[j]]) + k5) ** 2) / k6)
            cn[i] += den * 0.5 * (1.0 + math.erf(-kcn * (r / rc - 1.0)))
    return cn


def cn_eeq(numbers, positions, rcov, cutoff=25.0, kcn=7.5, cn_max=8.0):
    n = len(numbers)
    out = []
    for i in range(n):
        s = 0.0
        for j in range(n):
            if i == j or numbers[i] == 0 or numbers[j] == 0:
                continue
            r = _dist(positions[i], positions[j])
            if r <= cutoff:
                rc = rcov[numbers[i]] + rcov[numbers[j]]
                s += 0.5 * (1.0 + math.erf(-kcn * (r / rc - 1.0)))
        out.append(math.log(1 + math.exp(cn_max)) - math.log(1 + math.exp(cn_max - s)))
    return out


def _solve(a, b):
    """Gaussian elimination with partial pivoting on lists."""
    n = len(b)
    m = [row[:] + [b[i]] for i, row in enumerate(a)]
    for k in range(n):
        p = max(range(k, n), key=lambda r: abs(m[r][k]))
        m[k], m[p] = m[p], m[k]
        for r in range(k + 1, n):
            f = m[r][k] / m[k][k]
            for c in range(k, n + 1):
                m[r][c] -= f * m[k][c]
    x = [0.0] * n
    for k in range(n - 1, -1, -1):
        x[k] = (m[k][n] - sum(m[k][c] * x[c] for c in range(k + 1, n))) / m[k][k]
    return x


def eeq_charges(numbers, positions, total_charge, rcov, chi, eta, kcn, rad,
                cutoff=25.0):
    """Bordered EEQ system assembled element by element and solved by
    pivoted Gaussian elimination (real atoms only)."""
    idx = [i for i, z in enumerate(numbers) if z != 0]
    z = [numbers[i] for i in idx]
    p = [positions[i] for i in idx]
    cn = cn_eeq(z, p, rcov, cutoff)
    n = len(z)
    a = [[0.0] * (n + 1) for _ in range(n + 1)]
    b = [0.0] * (n + 1)
    for i in range(n):
        a[i][i] = eta[z[i]] + math.sqrt(2.0 / math.pi) / rad[z[i]]
        a[i][n] = a[n][i] = 1.0
        b[i] = -chi[z[i]] + kcn[z[i]] * math.sqrt(max(cn[i], 2.220446049250313e-16))
        for j in range(n):
            if i != j:
                r = _dist(p[i], p[j])
                g = 1.0 / math.sqrt(rad[z[i]] ** 2 + rad[z[j]] ** 2)
                a[i][j] = math.erf(g * r) / r
    b[n] = float(total_charge)
    x = _solve(a, b)
    q = [0.0] * len(numbers)
    for k, i in enumerate(idx):
        q[i] = x[k]
    return q


def twobody(numbers, positions, c6, r4r2, s6, s8, a1, a2, cutoff=60.0):
    n = len(numbers)
    e = [0.0] * n
    for i in range(n):
        for j in range(n):
            if i == j or numbers[i] == 0 or numbers[j] == 0:
                continue
            r = _dist(positions[i], positions[j])
            if r > cutoff:
                continue
            qq = 3.0 * r4r2[numbers[i]] * r4r2[numbers[j]]
            r0 = a1 * math.sqrt(qq) + a2
            t6 = 1.0 / (r**6 + r0**6)
            t8 = 1.0 / (r**8 + r0**8)
            e[i] -= 0.5 * (s6 * c6[i][j] * t6 + s8 * c6[i][j] * qq * t8)
    return e


def threebody(numbers, positions, c6_0, r4r2, s9, a1, a2, alp=16.0,
              cutoff=40.0):
    n = len(numbers)
    e = [0.0] * n
    c2 = cutoff * cutoff

    def r0(i, j):
        return a1 * math.sqrt(3.0 * r4r2[numbers[i]] * r4r2[numbers[j]]) + a2

    for a, b, c in combinations(range(n), 3):
        if 0 in (numbers[a], numbers[b], numbers[c]):
            continue
        rab2 = sum((x - y) ** 2 for x, y in zip(positions[a], positions[b]))
        rac2 = sum((x - y) ** 2 for x, y in zip(positions[a], positions[c]))
        rbc2 = sum((x - y) ** 2 for x, y in zip(positions[b], positions[c]))
        ab, ac, bc = rab2 <= c2, rac2 <= c2, rbc2 <= c2
        rbar = math.sqrt(rab2 * rac2 * rbc2)
        r0bar = r0(a, b) * r0(a, c) * r0(b, c)
        f = 1.0 / (1.0 + 6.0 * (r0bar / rbar) ** (alp / 3.0))
        s = (rab2 + rbc2 - rac2) * (rab2 - rbc2 + rac2) * (-rab2 + rbc2 + rac2)
        ang = 0.375 * s / rbar**5 + 1.0 / rbar**3
        c9 = math.sqrt(abs(c6_0[a][b] * c6_0[a][c] * c6_0[b][c]))
        en = s9 * c9 * ang * f / 6.0
        e[a] += en * bc * (ab + ac)
        e[b] += en * ac * (ab + bc)
        e[c] += en * ab * (ac + bc)
    return e
