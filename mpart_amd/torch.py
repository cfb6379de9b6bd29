"""torch integration, mirroring the reference's torch wrappers
(bindings/python/package/torch_helpers.py:33-115 and torch.py:33-107).

``MpartTorchAutograd`` keeps the reference semantics: ``forward`` returns the
map output (bitwise equal to ``Evaluate``) and optionally the log-determinant;
``backward`` returns per-point input gradients and coefficient gradients summed
over the points.  Unlike the reference (which copies to host tensors), all work
stays on the GPU: the coefficient gradient of both outputs is one fused
forward+adjoint launch (``ConditionalMap.CoeffVJP``) whose quadrature nodes are
recomputed rather than stored, and the input gradient of both outputs is
another (``ConditionalMap.InputVJPImpl``) in place of the reference's
Gradient and LogDeterminantInputGrad calls.
"""
import torch

from . import ConditionalMap


class MpartTorchAutograd(torch.autograd.Function):

    @staticmethod
    def forward(ctx, input, coeffs, f, return_logdet):
        """input is inputDim x N (a strided view is accepted), on a GPU."""
        ctx.save_for_backward(input, coeffs)
        ctx.f = f
        if coeffs is not None:
            cd = coeffs.detach().to(torch.float64).contiguous()
            ctx.coeffs_dev = cd
            f.WrapCoeffs(cd)
        x = input.detach().to(torch.float64)
        out = torch.empty((f.outputDim, x.shape[1]), dtype=torch.float64, device=x.device)
        f.EvaluateImpl(x, out)
        if return_logdet:
            ld = torch.empty(x.shape[1], dtype=torch.float64, device=x.device)
            f.LogDeterminantImpl(x, ld)
            return out.to(input.dtype), ld.to(input.dtype)
        return out.to(input.dtype)

    @staticmethod
    def backward(ctx, output_sens, logdet_sens=None):
        input, coeffs = ctx.saved_tensors
        f = ctx.f
        if coeffs is not None:
            f.WrapCoeffs(ctx.coeffs_dev)
        x = input.detach().to(torch.float64)
        sens = output_sens.to(torch.float64) if output_sens is not None else None
        lsens = logdet_sens.to(torch.float64).contiguous() if logdet_sens is not None else None

        want_c = coeffs is not None and coeffs.requires_grad
        grad = None
        if input.requires_grad:
            grad = torch.zeros((f.inputDim, x.shape[1]), dtype=torch.float64, device=x.device)
        coeff_grad = None
        if want_c:
            coeff_grad = torch.zeros(f.numCoeffs, dtype=torch.float64, device=x.device)
        if sens is not None or lsens is not None:
            if want_c:
                # both gradients from one fused pass
                f.VJPImpl(x, sens, lsens, coeff_grad, grad)
            elif grad is not None:
                f.InputVJPImpl(x, sens, lsens, grad)
        if grad is not None:
            grad = grad.to(input.dtype)
        if coeff_grad is not None:
            coeff_grad = coeff_grad.to(coeffs.dtype)
        return grad, coeff_grad, None, None


class TorchParameterizedFunctionBase(torch.nn.Module):
    """(torch.py:5-30)"""

    def __init__(self, f, store_coeffs=True, dtype=torch.double):
        super().__init__()
        self.f = f
        if store_coeffs:
            self.coeffs = torch.nn.Parameter(torch.tensor(f.CoeffMap(), dtype=dtype, device="cuda"))
        else:
            self.coeffs = None

    def forward(self, x, coeffs=None):
        if coeffs is None:
            if self.coeffs is None:
                raise RuntimeError('Must either set store_coeffs=True in constructor or pass coeffs to forward function.')
            coeffs = self.coeffs
        return MpartTorchAutograd.apply(x.T, coeffs, self.f, False).T


class TorchConditionalMapBase(torch.nn.Module):
    """(torch.py:33-107).  ``x`` is N x inputDim."""

    def __init__(self, f, store_coeffs=True, return_logdet=False, dtype=torch.double):
        super().__init__()
        self.return_logdet = return_logdet
        self.f = f
        if store_coeffs:
            self.coeffs = torch.nn.Parameter(torch.tensor(f.CoeffMap(), dtype=dtype, device="cuda"))
        else:
            self.coeffs = None

    def forward(self, x, coeffs=None):
        if coeffs is None:
            if self.coeffs is None:
                raise RuntimeError('Must either set store_coeffs=True in constructor or pass coeffs to forward function.')
            coeffs = self.coeffs
        self._check_shapes(x, None, coeffs)
        if self.return_logdet:
            y, logdet = MpartTorchAutograd.apply(x.T, coeffs, self.f, True)
            return y.T, logdet
        return MpartTorchAutograd.apply(x.T, coeffs, self.f, False).T

    def inverse(self, x, r, coeffs=None):
        c = (self.coeffs if coeffs is None else coeffs).detach().to(torch.float64).contiguous()
        self.f.WrapCoeffs(c)
        x_dbl = x.to(torch.float64)
        r_dbl = r.to(torch.float64)
        if (x.shape[1] != self.f.inputDim) and ((x.shape[1] + r.shape[1]) == self.f.inputDim):
            x_dbl = torch.hstack([x_dbl, r_dbl])
        self._check_shapes(x_dbl, r_dbl, c)
        out = torch.zeros((self.f.outputDim, x_dbl.shape[0]), dtype=torch.float64, device=x_dbl.device)
        extra = self.f.inputDim - self.f.outputDim
        if self.f.NumComponents() > 1 or extra > 0 or self.f.outputDim > 1:
            x1 = x_dbl[:, :extra].T
        else:
            x1 = x_dbl.T
        self.f.InverseImpl(x1, r_dbl.T, out)
        return out.T.to(x.dtype)

    def _check_shapes(self, x, r, coeffs):
        if x is not None and x.shape[1] != self.f.inputDim:
            raise ValueError(f'Input to map has wrong shape.  x has {x.shape[1]} columns but map expects input '
                             f'dimension of {self.f.inputDim}.')
        if r is not None and r.shape[1] != self.f.outputDim:
            raise ValueError(f'Reference input to map inverse has wrong shape.  r has {r.shape[1]} columns but map '
                             f'expects input dimension of {self.f.outputDim}.')
        if coeffs is not None and coeffs.shape[0] != self.f.numCoeffs:
            raise ValueError(f'Specified coefficients have wrong shape. coeffs input has length {coeffs.shape[0]} '
                             f'but map has {self.f.numCoeffs} coefficients.')


def kl_loss(f, x, coeffs):
    """mean(0.5 |T(x)|^2 - logdet(x)) through the autograd Function (x: inputDim x N view)."""
    y, ld = MpartTorchAutograd.apply(x, coeffs, f, True)
    return (0.5 * (y * y).sum(0) - ld).mean()


class KLObjective(torch.autograd.Function):
    """The KL objective mean(0.5 |T(x)|^2 - log det grad T(x)) of
    FastGaussianReverseKLObjective (src/MapObjective.cpp:146-234) as one autograd
    node: forward runs the fused forward+adjoint pass
    (ConditionalMap.KLValueAndGradsImpl) and keeps the coefficient gradient
    and, when the points require a gradient, the per-point gradient
    (Gradient(x, T) - LogDeterminantInputGrad(x)) / N that
    MpartTorchAutograd's backward would form from separate calls and T; both
    come from the same quadrature.  backward scales them by the incoming
    sensitivity."""

    @staticmethod
    def forward(ctx, x, coeffs, f):
        cd = coeffs.detach().to(torch.float64).contiguous()
        f.WrapCoeffs(cd)
        xd = x.detach().to(torch.float64)
        loss = torch.zeros(1, dtype=torch.float64, device=x.device)
        grad = torch.zeros(f.numCoeffs, dtype=torch.float64, device=x.device)
        xgrad = None
        if ctx.needs_input_grad[0]:
            xgrad = torch.empty((f.inputDim, xd.shape[1]), dtype=torch.float64, device=x.device)
        # one pass for the loss and both gradients
        f.KLValueAndGradsImpl(xd, loss, grad, xgrad)
        ctx.save_for_backward(grad, xgrad)
        ctx.x_dtype = x.dtype
        return loss[0].to(coeffs.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        grad, xgrad = ctx.saved_tensors
        gx = None
        if xgrad is not None:
            gx = (grad_out.to(torch.float64) * xgrad).to(ctx.x_dtype)
        gc = (grad_out * grad).to(grad_out.dtype) if ctx.needs_input_grad[1] else None
        return gx, gc, None


def kl_objective(f, x, coeffs):
    """Fused KL objective; x is inputDim x N (a strided view is accepted)."""
    return KLObjective.apply(x, coeffs, f)


__all__ = ["MpartTorchAutograd", "TorchParameterizedFunctionBase", "TorchConditionalMapBase", "kl_loss",
           "KLObjective", "kl_objective", "ConditionalMap"]
