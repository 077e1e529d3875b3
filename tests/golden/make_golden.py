"""Write tests/golden/c1_golden.npz from the reference's own leaf routines.

Run in a tree where oracle/_ref was built from the reference sources
(`make -C oracle ref`):  python tests/golden/make_golden.py

Contents (bcc a=5.42 bohr, 2-atom cell, LIZ 15 sites, lmax=3, 31-point
semicircle contour -0.5..0.7 Ry, seeded phase shifts of lsms_amd.configs):
  m_ref[n]   KKR matrix of atom m_atom at contour point m_energy_index[n],
             assembled block by block from the reference's makegij_/setgij_
             (MultipleScattering/makegij_c.f, setgij.f) with the t packing of
             calculateTauMatrix.cpp:155-183;
  tau00[a,e] solver-level tau00 of every (atom, energy): LAPACK zgetrf/zgetrs
             (scipy's LAPACK) on those matrices, as solveTau00zgetrf does
             (linearSolvers_CPU.cpp:98-123).
"""
import os
import sys

import numpy as np
import scipy.linalg as sl

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from lsms_amd import configs as CF  # noqa: E402
from oracle import ref_leaf as R  # noqa: E402


def ref_assemble(s, a, ie):
    at = s.atoms[a]
    kkrsz = (s.maxlmax + 1) ** 2
    blk = s.blkSizeTmatStore
    p = s.prel(ie)
    kk = [(l + 1) ** 2 for l in at.LIZlmax]
    off = np.concatenate([[0], np.cumsum(kk)])
    M = np.eye(off[-1], dtype=complex)
    cg = R.cgaunt(s.maxlmax)
    il, i1 = R.ifacts(s.maxlmax)
    for i in range(at.numLIZ):
        T = s.tmatStore[ie * blk:ie * blk + kkrsz * kkrsz, at.LIZStoreIdx[i]].reshape(
            kkrsz, kkrsz, order="F")
        k1 = kk[i]
        for j in range(at.numLIZ):
            if i != j:
                g = R.makegij(int(at.LIZlmax[i]), int(at.LIZlmax[j]), s.maxlmax, p,
                              at.LIZPos[:, i] - at.LIZPos[:, j], cg, il, i1)
                M[off[i]:off[i + 1], off[j]:off[j + 1]] = -T[:k1, :k1] @ R.setgij(g, 1, s.energies[ie])
    return M


def main():
    assert R.available(), "build oracle/_ref first (make -C oracle ref)"
    s = CF.config_c1()
    na, ne = len(s.atoms), len(s.energies)
    tau = np.zeros((na, ne, 16, 16), dtype=complex)
    mats = {}
    for a in range(na):
        for ie in range(ne):
            M = ref_assemble(s, a, ie)
            if a == 0 and ie == 15:
                mats[ie] = M.copy()
            t0 = s.tmatStore[ie * s.blkSizeTmatStore:ie * s.blkSizeTmatStore + 256,
                             s.atoms[a].LIZStoreIdx[0]].reshape(16, 16, order="F")
            lu, piv, info = sl.lapack.zgetrf(M)
            assert info == 0
            rhs = np.zeros((M.shape[0], 16), complex)
            rhs[:16] = t0
            x, info = sl.lapack.zgetrs(lu, piv, rhs)
            tau[a, ie] = x[:16]
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c1_golden.npz")
    np.savez_compressed(out, m_atom=np.int64(0), m_energy_index=np.array([15]),
                        m_ref=np.array([mats[15]]), tau00=tau, energies=s.energies,
                        deltas=s.deltas)
    print("wrote", out)


if __name__ == "__main__":
    main()
