"""Energy-contour consumers of the tau00 path (integrated densities of states).

The synthetic workloads have no radial wave functions, so the reference's
Green-function consumer is taken with unit radial integrals:
  - multiple-scattering DOS (MultipleScattering/mdosms_c.f:34-50,
    dos = -(sum zz*w1 - zj)/pi with w1 = tau00 - t, here zz = identity and
    zj = 0), integrated along the contour:
        N_ms = -(1/pi) Im sum_k w_k Tr[tau00(E_k) - t0(E_k)];
  - the Lloyd integrated DOS of the LIZ cluster, which the reference lacks:
        N_L = -(1/pi) Im sum_k w_k Tr(M^-1 dM/dE)(E_k)
    (d ln det M / dE from the adjoint route).
w_k are the contour weights (Misc/congauss_c.f:93); energies off the contour
carry weight 0.
"""
import numpy as np


def idos_ms(tau00, t0, weights):
    """-(1/pi) Im sum_k w_k Tr[tau00_k - t0_k] for blocks (ne, k, k)."""
    tr = np.trace(np.asarray(tau00) - np.asarray(t0), axis1=-2, axis2=-1)
    return float(-np.imag(np.sum(np.asarray(weights) * tr)) / np.pi)


def idos_lloyd(lloyd_values, weights):
    """-(1/pi) Im sum_k w_k Tr(M^-1 dM/dE)(E_k)."""
    return float(-np.imag(np.sum(np.asarray(weights) * np.asarray(lloyd_values))) / np.pi)
