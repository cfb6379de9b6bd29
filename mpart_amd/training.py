"""Map training against a standard-normal reference with torch.optim.LBFGS.

The role of the reference's TrainMap with FastGaussianReverseKLObjective
(src/TrainMap.cpp:111-216, src/MapObjective.cpp:146-234), driven through torch:
each objective evaluation is one fused forward+adjoint launch
(``mpart_amd.torch.KLObjective``).
"""
import torch

from .torch import kl_objective


def train_kl_lbfgs(f, x, steps=50, coeffs0=None, history_size=20, tolerance_grad=1e-12,
                   tolerance_change=1e-14, line_search_fn="strong_wolfe"):
    """Runs `steps` L-BFGS iterations on mean(0.5 |T(x)|^2 - logdet) over the
    coefficients of map `f`.

    x: inputDim x N float64 cuda tensor (or a strided view of one).
    Returns (coefficients, list of objective values, one per evaluation).
    The map's coefficients are set to the result.
    """
    c0 = torch.as_tensor(f.CoeffMap() if coeffs0 is None else coeffs0, dtype=torch.float64, device=x.device)
    coeffs = torch.nn.Parameter(c0.clone())
    opt = torch.optim.LBFGS([coeffs], lr=1.0, max_iter=steps, max_eval=int(steps * 1.25) + 1,
                            history_size=history_size, tolerance_grad=tolerance_grad,
                            tolerance_change=tolerance_change, line_search_fn=line_search_fn)
    history = []

    def closure():
        opt.zero_grad()
        loss = kl_objective(f, x, coeffs)
        loss.backward()
        history.append(float(loss.detach()))
        return loss

    opt.step(closure)
    out = coeffs.detach().clone()
    f.SetCoeffs(out)
    return out, history
