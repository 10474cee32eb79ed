"""torch.autograd wrapper of the adjoint sweep: the detector response of a
steady-state solve as a differentiable function of the per-cell absorption
field.

    siga = torch.tensor(..., dtype=torch.float64, requires_grad=True)  # [nz, ny, nx]
    R = detector_response(siga, problem, resp)
    R.backward()          # siga.grad = dR/dsiga from the reverse sweeps

The total cross section of group g is sigt_cell[g] = siga + sigs[g] (the
problem's material-0 scattering cross section), so dR/dsiga = sum_g dR/dt_xs[g].
The forward solve and the adjoint run on the HIP device through libsnap_hip;
torch only carries the values and the gradient.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import torch

from .problems import Problem
from .solver import Solver


class DetectorResponse(torch.autograd.Function):
    @staticmethod
    def forward(ctx, siga: torch.Tensor, problem: Problem, resp: np.ndarray, precision: str, tol):
        a = siga.detach().to("cpu", torch.float64).numpy()
        sigs = problem.sigs[0]
        sigt_cell = a[None, :, :, :] + sigs[:, None, None, None]
        p = dataclasses.replace(problem, sigt_cell=np.ascontiguousarray(sigt_cell))
        with Solver(p, precision="mixed" if precision == "mixed" else "fp64") as s:
            st = s.run()
            if not st["converged"]:
                raise RuntimeError("forward solve did not converge")
            grad, R, _ = s.gradient(resp, precision=precision, tol=tol)
        ctx.save_for_backward(torch.from_numpy(grad.sum(axis=0)).to(siga.device))
        return torch.tensor(R, dtype=torch.float64, device=siga.device)

    @staticmethod
    def backward(ctx, g_out):
        (g,) = ctx.saved_tensors
        return g_out * g, None, None, None, None


def detector_response(siga: torch.Tensor, problem: Problem, resp: np.ndarray,
                      precision: str = "fp64", tol=None) -> torch.Tensor:
    """R = sum(resp * flux0) of `problem` with per-cell absorption `siga`."""
    return DetectorResponse.apply(siga, problem, resp, precision, tol)
