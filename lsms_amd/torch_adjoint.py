"""torch.autograd binding of the adjoint route (lsmsx_tau00_adjoint).

Tau00Function maps a t-matrix store (complex128 tensor, the layout of
LocalTypeInfo::tmatStore) to the solver-level tau00 blocks of a batch of
(atom, spin, energy slot) solves.  Its backward returns the vector-Jacobian
product with respect to every store entry, computed on the GPU from an
explicit inverse of each KKR matrix.  PyTorch's convention for complex
holomorphic maps is grad_in = J^H grad_out, which is the convention of the
C entry point.  tau00 is holomorphic in t, so no conjugate-linear part exists.

The Lloyd term Tr(M^-1 dM/dE) = d ln det M / dE is returned by `lloyd` (not
differentiable).

The native library is the only compute path; there is no CPU fallback.
"""
import numpy as np
import torch


class _Batch:
    """The fixed, non-tensor arguments of a batch."""

    def __init__(self, engine, system, atom_ids, iies, ispins=None):
        self.engine = engine
        self.system = system
        self.atom_ids = np.asarray(atom_ids, dtype=np.int32)
        self.iies = np.asarray(iies, dtype=np.int32)
        self.ispins = None if ispins is None else np.asarray(ispins, dtype=np.int32)
        self.energies = np.asarray(system.energies, dtype=np.complex128)[self.iies]
        self.kc = system.kkrsz_ns

    def upload(self, store):
        s = self.system
        self.engine.upload_tmats(store, s.blkSizeTmatStore, len(s.energies))


class Tau00Function(torch.autograd.Function):
    """tau00 = f(tmat_store) for one batch; see the module docstring."""

    @staticmethod
    def forward(ctx, tmat_store, batch):
        store = tmat_store.detach().cpu().numpy()
        batch.upload(store)
        tau, _, _, info = batch.engine.tau00_adjoint(batch.atom_ids, batch.iies, batch.energies,
                                                     batch.kc, ispins=batch.ispins, lloyd=False)
        if np.any(info != 0):
            raise RuntimeError(f"singular KKR matrix in the batch (info={info.tolist()})")
        ctx.batch = batch
        ctx.save_for_backward(tmat_store)
        return torch.from_numpy(tau)

    @staticmethod
    def backward(ctx, grad_tau):
        (tmat_store,) = ctx.saved_tensors
        batch = ctx.batch
        store = tmat_store.detach().cpu().numpy()
        batch.upload(store)
        _, _, grad, _ = batch.engine.tau00_adjoint(
            batch.atom_ids, batch.iies, batch.energies, batch.kc, ispins=batch.ispins,
            lloyd=False, tau00_bar=grad_tau.detach().cpu().numpy(), store_shape=store.shape)
        return torch.from_numpy(np.asarray(grad)), None


def tau00(tmat_store, engine, system, atom_ids, iies, ispins=None):
    """Differentiable tau00 (n, k, k) of a batch, with `engine` a capi.Context
    holding the system's tables and LIZs."""
    return Tau00Function.apply(tmat_store, _Batch(engine, system, atom_ids, iies, ispins))


def lloyd(engine, system, atom_ids, iies, ispins=None):
    """Tr(M^-1 dM/dE) per solve (complex), with dt/dE of the system's store."""
    kc = system.kkrsz_ns
    e = np.asarray(system.energies, dtype=np.complex128)[np.asarray(iies)]
    engine.upload_dtmats(system.dtmat_store(), system.blkSizeTmatStore, len(system.energies))
    _, ll, _, info = engine.tau00_adjoint(atom_ids, iies, e, kc, ispins=ispins, lloyd=True)
    return ll, info


__all__ = ["Tau00Function", "tau00", "lloyd"]
