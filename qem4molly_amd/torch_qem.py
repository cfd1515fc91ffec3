"""PyTorch-ROCm binding: QEM energy and forces as one differentiable op.

    from qem4molly_amd.torch_qem import qem_energy_forces
    E, F = qem_energy_forces(ctx, xyz, q, k_idx, gamma=gamma)   # cuda float64

forward  runs qem_forces_energy_dev on torch's current stream: forces, E,
         dE/dq and dE/dgamma in one call (in mixed precision one pair
         quadrature for both), on the tensors' device pointers; nothing is
         copied to the host except the k indices and the scalar results.
backward (no reference counterpart; the reference has forces only):
  d/dxyz : -gE F  +  J_F^T gF, with J_F^T gF taken as -H gF from the
           Hessian-vector product kernel (qem_hvp_dev): the force Jacobian is
           symmetric for the conservative force (QEM_SEM_CORRECTED, up to the
           motion of the quadrature limits, an effect at the accuracy level);
  d/dq   : gE dE/dq (the per-atom potential); a force cotangent gF with
           respect to q is not provided (raises);
  d/dgamma: gE (dE/dgamma_1, dE/dgamma_2); a force cotangent gF with
           respect to gamma is not provided either (raises, rather than
           returning the energy term alone).  `gamma` must equal the
           context's contrasts; it is only the handle autograd differentiates.
With keep_sums (the default when xyz requires grad) the forward keeps the
pair quadrature's tangent node sums (qem_forces_energy_save_dev, 224 B per
pair in mixed precision) and the backward's HVP combines them
(qem_hvp_saved_dev) instead of running the quadrature again: about 1.3 ms
more forward and 2.9 ms less backward at N = 100k.  The context keeps one set
of sums; when another forces + energy call on the same context has replaced
them before this backward runs, the HVP recomputes the quadrature.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib as L

if L._lib is not None and not L._loaded_with_torch:
    raise ImportError("libqem_hip.so was loaded before torch; import torch (or "
                      "qem4molly_amd.torch_qem) first so that both use one HIP runtime")


def _stream():
    """torch's current stream for the library: its handle, or, for torch's
    default stream (handle 0), the library's marker for HIP's null stream
    (a NULL stream argument would select the context's own stream, whose
    work torch's default stream does not wait for)."""
    h = torch.cuda.current_stream().cuda_stream
    return h if h else L.STREAM_NULL


class QEMEnergyForces(torch.autograd.Function):
    @staticmethod
    def forward(fctx, xyz, q, gamma, qctx, k_idx, keep_sums=False):
        if xyz.dtype != torch.float64 or q.dtype != torch.float64 or not xyz.is_cuda:
            raise TypeError("xyz and q must be float64 tensors on a ROCm device")
        n = xyz.shape[0]
        x = xyz.detach().contiguous()
        qq = q.detach().contiguous()
        F = torch.empty_like(x)
        phi = torch.empty_like(qq)
        s = _stream()
        E, dg = qctx.forces_energy_dev(n, x.data_ptr(), qq.data_ptr(), k_idx, F.data_ptr(),
                                       phi.data_ptr(), s, save=keep_sums)
        fctx.sums_gen = qctx.sums_gen if keep_sums else None
        fctx.save_for_backward(x, qq, F, phi)
        fctx.qctx = qctx
        fctx.k_idx = np.array(k_idx, dtype=np.int64)
        fctx.dg = dg
        return x.new_tensor(E), F

    @staticmethod
    def backward(fctx, gE, gF):
        x, qq, F, phi = fctx.saved_tensors
        n = x.shape[0]
        gx = gq = gg = None
        if gE is None:
            gE = x.new_zeros(())
        if fctx.needs_input_grad[0]:
            gx = -gE * F
            if gF is not None and bool(torch.any(gF != 0)):
                hv = torch.empty_like(x)
                gFc = gF.contiguous()
                fctx.qctx.hvp_dev(n, x.data_ptr(), qq.data_ptr(), fctx.k_idx, gFc.data_ptr(),
                                  hv.data_ptr(), _stream(),
                                  saved=fctx.sums_gen == fctx.qctx.sums_gen)
                gx = gx - hv
        if fctx.needs_input_grad[1]:
            if gF is not None and bool(torch.any(gF != 0)):
                raise NotImplementedError(
                    "gradient of the forces with respect to the charges is not provided")
            gq = gE * phi
        if fctx.needs_input_grad[2]:
            if gF is not None and bool(torch.any(gF != 0)):
                raise NotImplementedError(
                    "gradient of the forces with respect to gamma is not provided")
            gg = gE * x.new_tensor(fctx.dg)
        return gx, gq, gg, None, None, None


def qem_energy_forces(qctx: L.Context, xyz, q, k_idx, gamma=None, keep_sums=None):
    """(E, F) for one electrostatic evaluation; differentiable (see module).
    keep_sums (default: xyz.requires_grad under grad mode) keeps the pair
    quadrature's tangent sums for the backward's HVP."""
    p = qctx.params
    if keep_sums is None:
        keep_sums = bool(xyz.requires_grad and torch.is_grad_enabled())
    if gamma is None:
        gamma = xyz.new_tensor([p.gamma_1, p.gamma_2])
    else:
        g = gamma.detach().cpu().numpy()
        if not np.allclose(g, [p.gamma_1, p.gamma_2], rtol=0, atol=0):
            raise ValueError("gamma must equal the context's gamma_1, gamma_2")
    return QEMEnergyForces.apply(xyz, q, gamma, qctx, k_idx, bool(keep_sums))


def qem_forces(qctx: L.Context, xyz, q, k_idx=None, walls=(), full_sum=False, out=None):
    """Forces only (not differentiable): QEM short + long for the batch k_idx
    plus the given Substrate_LJ walls, on torch's current stream, or with
    full_sum the deterministic sum over the whole K set (weights
    exp(-k^2/(4 alpha))) in place of the batch.  xyz (N x 3) and q (N) are
    float64 tensors on the device; out (N x 3) is reused when given."""
    if xyz.dtype != torch.float64 or q.dtype != torch.float64 or not xyz.is_cuda:
        raise TypeError("xyz and q must be float64 tensors on a ROCm device")
    x = xyz.contiguous()
    qq = q.contiguous()
    n = x.shape[0]
    F = torch.empty_like(x) if out is None else out
    s = _stream()
    if full_sum:
        if not hasattr(qctx, "_kset_dev"):
            K, _, _ = qctx.kset()
            qctx._kset_dev = K
        qctx.short_forces_dev(n, x.data_ptr(), qq.data_ptr(), F.data_ptr(), s)
        qctx.long_terms_dev(n, x.data_ptr(), qq.data_ptr(), qctx._kset_dev, -1.0, L.LONG_ALL,
                            F.data_ptr(), accumulate=True, stream=s)
        for w in walls:
            qctx.substrate_lj_dev(n, w, x.data_ptr(), F.data_ptr(), accumulate=True, stream=s)
    else:
        qctx.general_forces_dev(n, x.data_ptr(), qq.data_ptr(), k_idx, list(walls),
                                F.data_ptr(), s)
    return F
