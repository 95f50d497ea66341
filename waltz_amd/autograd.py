"""Differentiable CDF 9/7 transforms for torch.autograd (no reference
counterpart). Forward passes run the codec's own HIP lifting kernels (so
``idwt`` reproduces the decoder's reconstruction bit for bit before repairs);
backward passes run the HIP adjoint kernels (exact transposes of the same
linear maps, computed in the tensor's precision).

    c = waltz_amd.autograd.dwt(field).detach().requires_grad_()
    x_hat = waltz_amd.autograd.idwt(c)                       # IDWT
    loss = torch.nn.functional.mse_loss(x_hat, target)
    loss.backward()                                          # c.grad = IDWT^T(dL/dx_hat)

Shapes are (z, y, x) with x fastest; float32 or float64 device tensors. The
level counts default to the codec's own choice for the shape.
"""
import torch

from . import stages


def dwt_levels(n):
    """Levels of one axis group (include/utils/Blob.hpp:22-29): halve (rounding
    up) while the extent exceeds 8, at most 5 times."""
    levels = 0
    while n > 8 and levels < 5:
        n -= n // 2
        levels += 1
    return levels


def default_levels(shape):
    """(levels_xy, levels_z, dyadic) the codec uses for a (z, y, x) shape."""
    shape = tuple(int(v) for v in shape)
    if len(shape) == 2:
        nz, (ny, nx) = 1, shape
    elif len(shape) == 3:
        nz, ny, nx = shape
    else:
        raise ValueError("expected a (y, x) or (z, y, x) shape")
    lxy = dwt_levels(min(nx, ny))
    lz = dwt_levels(nz)
    return lxy, lz, lxy == lz


def _levels(t, levels):
    if t.dim() != 3:
        raise ValueError("expected a (z, y, x) tensor")
    return default_levels(t.shape) if levels is None else tuple(levels)


class _Forward(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, lxy, lz, dyadic):
        ctx.geometry = (lxy, lz, dyadic)
        return stages.transform(x.detach(), lxy, lz, dyadic, inverse=False)

    @staticmethod
    def backward(ctx, grad):
        lxy, lz, dyadic = ctx.geometry
        return stages.transform_adjoint(grad, lxy, lz, dyadic, of_inverse=False), None, None, None


class _Inverse(torch.autograd.Function):
    @staticmethod
    def forward(ctx, c, lxy, lz, dyadic):
        ctx.geometry = (lxy, lz, dyadic)
        return stages.transform(c.detach(), lxy, lz, dyadic, inverse=True)

    @staticmethod
    def backward(ctx, grad):
        lxy, lz, dyadic = ctx.geometry
        return stages.transform_adjoint(grad, lxy, lz, dyadic, of_inverse=True), None, None, None


def dwt(x, levels=None):
    """Coefficients of a field (the codec's forward transform); gradient DWT^T."""
    lxy, lz, dyadic = _levels(x, levels)
    return _Forward.apply(x, lxy, lz, bool(dyadic))


def idwt(c, levels=None):
    """Field of a coefficient volume (the decoder's inverse transform); gradient IDWT^T."""
    lxy, lz, dyadic = _levels(c, levels)
    return _Inverse.apply(c, lxy, lz, bool(dyadic))


def reconstruct(q, quantum, levels=None):
    """dequant∘IDWT: IDWT(quantum * q) for signed integer-valued magnitudes q
    (a float tensor; its gradient is quantum * IDWT^T(dL/dx))."""
    return idwt(q * quantum, levels)
