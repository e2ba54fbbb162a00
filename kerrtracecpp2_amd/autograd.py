"""torch.autograd over the C ABI: the disk-intensity image as a differentiable function of the
black-hole and camera parameters p = (a, Q, Lambda, theta_obs [deg], r_obs).

forward  -> kt_hip_render_intensity (binary64 DOPRI5 on the GPU)
backward -> kt_hip_render_adjoint with the incoming image gradient as per-pixel weights: the
            hand-written reverse pass over checkpointed steps returns dL/dp.

    import torch, kerrtracecpp2_amd as kt
    from kerrtracecpp2_amd.autograd import KerrIntensity
    ctx = kt.Context(0)
    p = torch.tensor([0.7, 0.2, 0.0, 60.0, 40.0], dtype=torch.float64, requires_grad=True)
    img = KerrIntensity.apply(p, ctx, dict(width=256, height=256, fov=30.0, disk_out=12.0), opts)
    loss = ((img - target) ** 2).sum(); loss.backward()     # p.grad: dL/dp
"""
import numpy as np
import torch

from . import make_frame


def _frame(p, spec):
    a, Q, Lam, theta, r_obs = (float(v) for v in p)
    return make_frame(spec["width"], spec["height"], a=a, Q=Q, Lambda=Lam, theta=theta, r_obs=r_obs,
                      fov=spec.get("fov", 30.0), disk_out=spec.get("disk_out", 12.0), phi=spec.get("phi", 0.0),
                      M=spec.get("M", 1.0))


class KerrIntensity(torch.autograd.Function):
    """img[H, W] = I(p); p is a float64 tensor of 5 parameters."""

    @staticmethod
    def forward(ctx, p, kt_ctx, spec, opts):
        pv = p.detach().cpu().numpy().astype(np.float64)
        img, _ = kt_ctx.render_intensity(_frame(pv, spec), opts)
        ctx.kt = (kt_ctx, spec, opts, pv)
        return torch.from_numpy(img).to(dtype=torch.float64, device=p.device)

    @staticmethod
    def backward(ctx, grad_img):
        kt_ctx, spec, opts, pv = ctx.kt
        w = grad_img.detach().to(torch.float64).cpu().numpy()
        out = kt_ctx.render_adjoint(_frame(pv, spec), opts, weights=w)
        g = torch.from_numpy(out["grad"]).to(dtype=grad_img.dtype, device=grad_img.device)
        return g, None, None, None
