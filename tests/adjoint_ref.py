"""CPU references for the adjoint route, built on the oracle's KKR matrices.

The reference has no Lloyd term or derivative of tau (it never forms
d ln det M / dE), so these quantities are pinned by independent numerical
routes on the oracle's M(E):
  - dM/dE by the four-point complex difference (M is holomorphic in E)
      M'(E) ~ [M(E+h) - M(E-h) - i (M(E+ih) - M(E-ih))] / (4h),  error O(h^4);
  - Tr(M^-1 dM/dE) from that derivative with LAPACK;
  - d ln det M / dE by the central difference of LAPACK's log-determinant
    (a second, independent route to the same number);
  - the tau00 gradient by central differences of tau00 along store directions.
"""
import numpy as np
import scipy.linalg as sla

from oracle import oracle as O


def oracle_m(system, a, ie, ispin=0):
    at = system.atoms[a]
    return O.build_kkr(system.maxlmax, at.LIZPos, at.LIZlmax, at.LIZStoreIdx, system.tmatStore,
                       system.blkSizeTmatStore, ie, system.prel(ie), system.n_spin_cant,
                       system.n_spin_pola, ispin)


def oracle_t0(system, a, ie, ispin=0):
    at = system.atoms[a]
    return O.t0(system.maxlmax, int(at.LIZStoreIdx[0]), system.tmatStore,
                system.blkSizeTmatStore, ie, system.n_spin_cant, system.n_spin_pola, ispin)


def m_at(system, a, energies, ispin=0):
    """Oracle KKR matrices of atom a at arbitrary energies (t rebuilt there)."""
    s2 = system.at_energies(energies)
    return [oracle_m(s2, a, k, ispin) for k in range(len(energies))]


def dm_de(system, a, e, ispin=0, h=1e-3):
    """Four-point complex difference of M at energy e."""
    mp, mm, mip, mim = m_at(system, a, [e + h, e - h, e + 1j * h, e - 1j * h], ispin)
    return (mp - mm - 1j * (mip - mim)) / (4.0 * h)


def lloyd_trace(system, a, e, ispin=0, h=1e-3):
    """Tr(M^-1 dM/dE) with LAPACK on the oracle's matrices."""
    m = m_at(system, a, [e], ispin)[0]
    d = dm_de(system, a, e, ispin, h)
    lu = sla.lu_factor(m)
    return complex(np.trace(sla.lu_solve(lu, d)))


def logdet(m):
    """ln det m as (log|det|, arg det)."""
    lu, piv = sla.lu_factor(m)
    dg = np.diag(lu)
    sign = (-1) ** int(np.sum(piv != np.arange(len(piv))))
    return float(np.sum(np.log(np.abs(dg)))), float(np.angle(sign * np.prod(dg / np.abs(dg))))


def lloyd_logdet_fd(system, a, e, ispin=0, h=1e-6):
    """Central difference of ln det M(E) (phase difference wrapped to (-pi, pi])."""
    mp, mm = m_at(system, a, [e + h, e - h], ispin)
    (lp, pp), (lm, pm) = logdet(mp), logdet(mm)
    dph = (pp - pm + np.pi) % (2 * np.pi) - np.pi
    return complex((lp - lm) / (2 * h), dph / (2 * h))


def site_t(system, store_col, ie, lmax_site, ispin=0):
    """t of one site packed as the assembly packs it (leading kkri block of
    each spin block)."""
    nsc = system.n_spin_cant
    kkrsz = (system.maxlmax + 1) ** 2
    full = O.t0(system.maxlmax, int(store_col), system.tmatStore, system.blkSizeTmatStore, ie,
                nsc, system.n_spin_pola, ispin)
    kkri = (lmax_site + 1) ** 2
    idx = [rs * kkrsz + ri for rs in range(nsc) for ri in range(kkri)]
    return full[np.ix_(idx, idx)]


def tau_full(system, a, ie, ispin=0):
    """tau = M^-1 T, T the block diagonal of the LIZ sites' t-matrices."""
    at = system.atoms[a]
    m = oracle_m(system, a, ie, ispin)
    blocks = [site_t(system, at.LIZStoreIdx[s], ie, int(at.LIZlmax[s]), ispin)
              for s in range(at.numLIZ)]
    T = sla.block_diag(*blocks)
    return sla.solve(m, T)
