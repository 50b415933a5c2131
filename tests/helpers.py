"""Shared helpers of the parity tests."""
import numpy as np


def bicgstab_spread(orc, csr, b, eps, nruns=20, itmax=1000):
    """Iteration counts of the oracle's BiCGSTAB when b is perturbed at the
    rounding level (b*(1 + k*2^-52)): the spread any other summation order
    can land in (min, max over nruns perturbations)."""
    its = []
    for k in range(nruns):
        bk = b * (1.0 + k * 2.0 ** -52)
        its.append(orc.krylov("BICGSTAB", csr, bk, eps=eps, itmax=itmax)[1])
    return min(its), max(its)


def in_spread(it, spread, slack=2):
    """it within the rounding spread widened by a quarter of its width + slack
    (the spread is sampled, not exhaustive)."""
    lo, hi = spread
    m = 0.25 * (hi - lo) + slack
    return lo - m <= it <= hi + m


def tri_csr(orc, n, lower, unit, seed, density=0.15, integer=False):
    """Random triangular CSR (1-based, columns ascending): strictly triangular
    entries plus, unless unit, a dominant diagonal (last entry of a lower
    row, first of an upper row)."""
    rng = np.random.default_rng(seed)
    ia, ja, va = [], [], []
    for i in range(n):
        cols = [j for j in range(n) if (j < i if lower else j > i) and rng.random() < density]
        vals = (rng.integers(-2, 3, len(cols)).astype(float) if integer
                else rng.uniform(-1.0, 1.0, len(cols)))
        for j, v in zip(cols, vals):
            ia.append(i + 1); ja.append(j + 1); va.append(v)
        if not unit:
            ia.append(i + 1); ja.append(i + 1)
            va.append(float(rng.choice([-1.0, 1.0])) if integer
                      else (1.0 + np.abs(vals).sum()) * rng.choice([-1.0, 1.0]))
    coo = orc.fix_coo(orc.COO(n, n, np.array(ia, np.int32), np.array(ja, np.int32),
                              np.array(va)))
    return orc.csr_from_coo(coo)


def dense(irp, ja, val, n):
    A = np.zeros((n, n))
    for i in range(n):
        for k in range(irp[i] - 1, irp[i + 1] - 1):
            A[i, ja[k] - 1] += val[k]
    return A
