"""The gradient entry point behind torch.autograd (a gradcheck harness only).

DDAObjective wraps ddafft_gradient as a torch.autograd.Function of the
refractive index (Re m, Im m) of one material: forward returns the weighted
objective (Cext, Cabs), backward the adjoint gradient.  torch.autograd.gradcheck
then compares it with central differences of the forward solves.  The
forward runs in float32 on the GPU, so the step is 1e-3 and the tolerance is
written in the test (2e-2 relative): float32 solve noise over the step.
"""

import numpy as np
import pytest

import ddafft

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

DIMS, KD = (8, 8, 16), 0.5


class DDAObjective(torch.autograd.Function):
    @staticmethod
    def forward(actx, m, weights, occ):
        mc = complex(float(m[0]), float(m[1]))
        kinc = np.array([0, 0, 1.0], np.float32)
        einc = np.array([[1, 0, 0], [0, 1, 0]], np.complex64)
        alpha = np.array([[0, ddafft.polarizability(1, KD, mc)]] * 2)
        dadm = np.array([[0, ddafft.polarizability_deriv(1, KD, mc)]] * 2)
        W = weights.detach().cpu().numpy().astype(np.float32)[None]
        with ddafft.Context(DIMS, KD, nrhs_max=2) as ctx:
            ctx.set_target(occ, [0, 1])
            res = ctx.gradient("bicgstab", alpha, dadm, [1.0, mc], kinc, einc, W, ndirs=0, max_step=400,
                               err_max=1e-7)
        actx.save_for_backward(torch.tensor(res["grad_m"][0, 1], dtype=torch.float64))
        return torch.tensor(float(res["f"][0]), dtype=torch.float64)

    @staticmethod
    def backward(actx, gout):
        (g,) = actx.saved_tensors
        return gout * g, None, None


def test_gradcheck_material_index():
    occ, _ = ddafft.target_ellipsoid(DIMS, 1.0, 2.0)
    for w in ([1.0, 0.0], [0.0, 1.0], [0.7, -0.4]):
        print("case gradcheck weights", w, flush=True)
        m = torch.tensor([1.5, 0.01], dtype=torch.float64, requires_grad=True)
        weights = torch.tensor(w, dtype=torch.float64)
        f = DDAObjective.apply(m, weights, occ)
        assert torch.autograd.gradcheck(lambda mm: DDAObjective.apply(mm, weights, occ), (m,), eps=1e-3,
                                        atol=2e-2 * abs(float(f)) + 1e-6, rtol=2e-2)
