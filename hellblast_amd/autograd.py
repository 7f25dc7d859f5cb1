"""Differentiable SHTs: torch.autograd.Function over the adjoint kernels.

No reference counterpart (the reference has no gradients).  The backward of
each transform is the exact transpose, run by the same HIP kernels with the
adjoint flags (the lambda recursion is recomputed, never stored):
    alm2map:      grad_alm  = hbx_alm2map_adjoint(grad_map)
                  (unweighted analysis; m = 0 rows real, m > 0 rows doubled)
    map2alm:      grad_map  = hbx_map2alm_adjoint(grad_alm)
                  (synthesis of grad_alm with m > 0 halved, times omega_r =
                  2 pi w_r / nphi_r)
and likewise for the spin pair.  Complex gradients follow PyTorch's
convention grad = dL/dRe + i dL/dIm.  Backward functions are themselves the
forward transforms' transposes, so second derivatives go through the same
pair of Functions.
"""
import torch


def _cast(t, dtype):
    return t if t.dtype == dtype else t.to(dtype)


class _Alm2Map(torch.autograd.Function):
    @staticmethod
    def forward(ctx, alm, plan, dtype):
        ctx.plan, ctx.in_dtype = plan, alm.dtype
        return plan._alm2map_raw(alm.detach(), dtype)

    @staticmethod
    def backward(ctx, g):
        return _cast(_Alm2MapAdjoint.apply(g, ctx.plan), ctx.in_dtype), None, None


class _Alm2MapAdjoint(torch.autograd.Function):
    @staticmethod
    def forward(ctx, maps, plan):
        ctx.plan, ctx.in_dtype = plan, maps.dtype
        return plan._alm2map_adjoint_raw(maps.detach().contiguous())

    @staticmethod
    def backward(ctx, g):
        return _cast(_Alm2Map.apply(g, ctx.plan, ctx.in_dtype), ctx.in_dtype), None


class _Map2Alm(torch.autograd.Function):
    @staticmethod
    def forward(ctx, maps, plan):
        ctx.plan, ctx.in_dtype = plan, maps.dtype
        return plan._map2alm_raw(maps.detach())

    @staticmethod
    def backward(ctx, g):
        return _cast(_Map2AlmAdjoint.apply(g, ctx.plan, ctx.in_dtype), ctx.in_dtype), None


class _Map2AlmAdjoint(torch.autograd.Function):
    @staticmethod
    def forward(ctx, alm, plan, dtype):
        ctx.plan, ctx.in_dtype = plan, alm.dtype
        return plan._map2alm_adjoint_raw(alm.detach().contiguous(), dtype)

    @staticmethod
    def backward(ctx, g):
        return _cast(_Map2Alm.apply(g, ctx.plan), ctx.in_dtype), None, None


class _Alm2MapSpin(torch.autograd.Function):
    @staticmethod
    def forward(ctx, alm, plan, dtype):
        ctx.plan, ctx.in_dtype = plan, alm.dtype
        return plan._alm2map_spin_raw(alm.detach(), dtype)

    @staticmethod
    def backward(ctx, g):
        return _cast(_Alm2MapSpinAdjoint.apply(g, ctx.plan), ctx.in_dtype), None, None


class _Alm2MapSpinAdjoint(torch.autograd.Function):
    @staticmethod
    def forward(ctx, maps, plan):
        ctx.plan, ctx.in_dtype = plan, maps.dtype
        return plan._alm2map_spin_adjoint_raw(maps.detach().contiguous())

    @staticmethod
    def backward(ctx, g):
        return _cast(_Alm2MapSpin.apply(g, ctx.plan, ctx.in_dtype), ctx.in_dtype), None


class _Map2AlmSpin(torch.autograd.Function):
    @staticmethod
    def forward(ctx, maps, plan):
        ctx.plan, ctx.in_dtype = plan, maps.dtype
        return plan._map2alm_spin_raw(maps.detach())

    @staticmethod
    def backward(ctx, g):
        return _cast(_Map2AlmSpinAdjoint.apply(g, ctx.plan, ctx.in_dtype), ctx.in_dtype), None


class _Map2AlmSpinAdjoint(torch.autograd.Function):
    @staticmethod
    def forward(ctx, alm, plan, dtype):
        ctx.plan, ctx.in_dtype = plan, alm.dtype
        return plan._map2alm_spin_adjoint_raw(alm.detach().contiguous(), dtype)

    @staticmethod
    def backward(ctx, g):
        return _cast(_Map2AlmSpin.apply(g, ctx.plan), ctx.in_dtype), None, None


def alm2map(alm, plan, dtype=torch.float64):
    """Differentiable ``plan.alm2map``: [..., nalm] complex -> [..., npix] real."""
    return _Alm2Map.apply(alm, plan, dtype)


def map2alm(maps, plan):
    """Differentiable ``plan.map2alm``: [..., npix] real -> [..., nalm] complex."""
    return _Map2Alm.apply(maps, plan)


def alm2map_spin(alm, plan, dtype=torch.float64):
    """Differentiable ``plan.alm2map_spin``: [..., 2, nalm] (E, B) -> [..., 2, npix] (Q, U)."""
    return _Alm2MapSpin.apply(alm, plan, dtype)


def map2alm_spin(maps, plan):
    """Differentiable ``plan.map2alm_spin``: [..., 2, npix] (Q, U) -> [..., 2, nalm] (E, B)."""
    return _Map2AlmSpin.apply(maps, plan)
