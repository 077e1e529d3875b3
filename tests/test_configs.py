"""Synthetic workload generators (CPU only)."""
import numpy as np

from lsms_amd import configs as CF
from oracle import oracle as O


def test_c1_liz_is_two_bcc_shells():
    s = CF.config_c1()
    assert len(s.atoms) == 2
    for at in s.atoms:
        assert at.numLIZ == 15 and at.nrmat == 240
        d = np.sqrt((at.LIZPos ** 2).sum(axis=0))
        assert d[0] == 0.0 and np.all(np.diff(d) >= -1e-12)
        assert np.allclose(d[1:9], 5.42 * np.sqrt(3) / 2) and np.allclose(d[9:], 5.42)


def test_c2_liz_is_135_fcc_sites():
    s = CF.config_c2(natoms=3)
    for at in s.atoms:
        assert at.numLIZ == 135 and at.nrmat == 2160
    # shell counts of fcc up to a*sqrt(7/2)
    d = np.sqrt((s.atoms[0].LIZPos ** 2).sum(axis=0)) / 6.82
    shells = np.round(2 * d * d).astype(int)
    counts = [int(np.sum(shells == n)) for n in range(8)]
    assert counts == [1, 12, 6, 24, 12, 24, 8, 48]


def test_contour_matches_oracle():
    e, w = CF.contour_semicircle(-0.5, 0.7, 31)
    e2, w2 = O.congauss(-0.5, 0.7, 31)
    assert np.abs(e - e2).max() < 1e-14 and np.abs(w - w2).max() < 1e-14


def test_tmat_store_layout():
    s = CF.config_c1()
    blk = s.blkSizeTmatStore
    assert blk == 32 * 32
    ie, st = 4, 1
    t = s.tmatStore[ie * blk:ie * blk + 256, st].reshape(16, 16, order="F")
    kap = CF.prel_of(s.energies[ie])
    for l in range(4):
        d = s.deltas[st, 0, l]
        for m in range(2 * l + 1):
            k = l * l + m
            assert abs(t[k, k] - (-np.sin(d) * np.exp(1j * d) / kap)) < 1e-15
    assert np.count_nonzero(t - np.diag(np.diag(t))) == 0


def test_prel_matches_oracle():
    for e in (-0.5 + 0.01j, 0.3 + 0.2j):
        assert abs(complex(CF.prel_of(e)) - O.prel(e)) < 1e-15
