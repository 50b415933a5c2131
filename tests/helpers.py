"""Shared helpers of the parity tests."""


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
