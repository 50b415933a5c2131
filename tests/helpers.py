"""Shared helpers of the parity tests."""


def bicgstab_spread(orc, csr, b, eps, nruns=10, itmax=1000):
    """Iteration counts of the oracle's BiCGSTAB when b is perturbed at the
    rounding level (b*(1 + k*2^-52)): the spread any other summation order
    can land in."""
    its = []
    for k in range(nruns):
        bk = b * (1.0 + k * 2.0 ** -52)
        its.append(orc.krylov("BICGSTAB", csr, bk, eps=eps, itmax=itmax)[1])
    return min(its), max(its)
