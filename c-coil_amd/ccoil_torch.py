"""torch.autograd.Function over the field path: B at points as a differentiable function of every coil's
(current, R, a, b, px, py, pz, alpha, beta), with ccoil_field_gradient (the adjoint kernel) as its backward.

The quadrature orders of each coil (its PrecisionArguments) are frozen when the model is built, as the reference's
integer planning is piecewise constant in the geometry; J = N I / (a b) is re-derived from the parameters on every
call. Parameters and points are float64 tensors on the context's device.

The library's kernels run on torch's current stream (the context is bound to it on every call), so they are ordered
after the work that produced the points and the incoming gradient without a device-wide synchronisation. The
parameter values are copied to the host once per forward (n_coils x 9 doubles), because the coil constants are
formed there.
"""
import numpy as np
import torch

import ccoil_abi as abi

PARAM_NAMES = abi.GRAD_TERMS   # current, R, a, b, px, py, pz, alpha, beta


class CoilFieldModel:
    """Frozen part of a coil set: method (fast/slow), quadrature orders and turns, from C ABI rows."""

    def __init__(self, rows, context, precision=abi.FP64):
        self.rows = [dict(r) for r in rows]
        self.context = context
        self.precision = precision

    def initial_params(self, device="cuda"):
        p = [[r["current"], r["R"], r["a"], r["b"], *r["pos"], r["alpha"], r["beta"]] for r in self.rows]
        return torch.tensor(p, dtype=torch.float64, device=device)

    def coilset(self, params):
        """CoilSet for the given (n_coils, 9) parameter values (host copy)."""
        v = params.detach().to("cpu", torch.float64).numpy()
        if v.shape != (len(self.rows), len(PARAM_NAMES)):
            raise ValueError(f"params must be ({len(self.rows)}, {len(PARAM_NAMES)})")
        I, R, a, b = v[:, 0], v[:, 1], v[:, 2], v[:, 3]
        N = np.array([r["turns"] for r in self.rows])
        return abi.CoilSet([r["use_fast"] for r in self.rows], [r["n_ang"] for r in self.rows],
                           [r["n_thk"] for r in self.rows], R, a, b, N * I / (a * b), v[:, 4:7].reshape(-1),
                           v[:, 7], v[:, 8], current=I, turns=N, n_len=[r["n_len"] for r in self.rows])

    def B(self, params, points):
        return CoilFieldB.apply(params, points, self)


class CoilFieldB(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, points, model):
        pts = points.detach().contiguous()
        cs = model.coilset(params)
        B = torch.empty((pts.shape[0], 3), dtype=torch.float64, device=pts.device)
        model.context.set_stream(torch.cuda.current_stream(pts.device).cuda_stream)
        model.context.fields(cs, pts, abi.B, model.precision, abi.DEVICE_POINTERS, out={"B": B})
        ctx.model, ctx.cs = model, cs
        ctx.save_for_backward(pts)
        return B

    @staticmethod
    def backward(ctx, gB):
        (pts,) = ctx.saved_tensors
        g = gB.contiguous().to(torch.float64)
        grad = torch.empty((ctx.cs.n, len(PARAM_NAMES)), dtype=torch.float64, device=pts.device)
        ctx.model.context.set_stream(torch.cuda.current_stream(pts.device).cuda_stream)
        ctx.model.context.field_gradient(ctx.cs, pts, g, ctx.model.precision, abi.DEVICE_POINTERS, out=grad)
        return grad, None, None


def design_loss(B, points, target_bz=0.01, radial_weight=1e-3):
    """mean over points of (B_z - target)^2 + w |B_r|^2, B_r = B_x cos(phi) + B_y sin(phi)."""
    phi = torch.atan2(points[:, 1], points[:, 0])
    br = B[:, 0] * torch.cos(phi) + B[:, 1] * torch.sin(phi)
    return torch.mean((B[:, 2] - target_bz) ** 2 + radial_weight * br ** 2)
