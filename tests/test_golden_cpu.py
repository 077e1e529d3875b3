"""The oracle reproduces the committed fixtures of tests/golden (CPU only).

The fixtures were written from the reference's own makegij_/setgij_ and LAPACK
(tests/golden/make_golden.py), so this pins the oracle where the reference
tree is absent.
"""
import os

import numpy as np

from lsms_amd import configs as CF
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "c1_golden.npz")


def test_fixture_inputs_match_config():
    g = np.load(GOLD, allow_pickle=False)
    s = CF.config_c1()
    assert np.array_equal(g["energies"], s.energies)
    assert np.array_equal(g["deltas"], s.deltas)


def test_oracle_reproduces_golden_matrix():
    g = np.load(GOLD, allow_pickle=False)
    s = CF.config_c1()
    at = s.atoms[int(g["m_atom"])]
    for n, ie in enumerate(g["m_energy_index"]):
        M = O.build_kkr(3, at.LIZPos, at.LIZlmax, at.LIZStoreIdx, s.tmatStore, s.blkSizeTmatStore,
                        int(ie), s.prel(int(ie)))
        ref = g["m_ref"][n]
        assert np.abs(M - ref).max() <= 1e-14 * np.abs(ref).max()


def test_oracle_reproduces_golden_tau00():
    g = np.load(GOLD, allow_pickle=False)
    s = CF.config_c1()
    worst = 0.0
    for a in range(len(s.atoms)):
        at = s.atoms[a]
        for ie in range(0, len(s.energies), 5):
            M = O.build_kkr(3, at.LIZPos, at.LIZlmax, at.LIZStoreIdx, s.tmatStore,
                            s.blkSizeTmatStore, ie, s.prel(ie))
            t0 = O.t0(3, int(at.LIZStoreIdx[0]), s.tmatStore, s.blkSizeTmatStore, ie)
            tau, info = O.solve_tau00_zgetrf(M, t0)
            ref = g["tau00"][a, ie]
            worst = max(worst, np.linalg.norm(tau - ref) / np.linalg.norm(ref))
    assert worst < 1e-12
