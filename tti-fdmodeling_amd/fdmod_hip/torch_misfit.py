"""torch.autograd.Function over the libfdmod_hip misfit gradient (all three forms).

J(vp, eps, delta, theta) = 1/2 sum (d(m) - obs)^2 for one shot, where d(m) is the
modelled data of a fdmod_hip.Modeling context (source and receivers set) and the
absorbing layers are those of the context's last set_model.  The forward pass calls
fdmod_update_model and fdmod_gradient (one run from rest plus the checkpointed
adjoint); backward scales the stored gradient by the incoming gradient.  All device
work is in libfdmod_hip; torch only carries the tensors.
"""
import numpy as np
import torch

__all__ = ["TTIMisfit", "misfit"]


def _host(t):
    if isinstance(t, torch.Tensor):
        return t.detach().to("cpu", torch.float32).contiguous().numpy()
    return np.ascontiguousarray(t, dtype=np.float32)


class TTIMisfit(torch.autograd.Function):
    """Model tensors are (nx, nz) in the model-file layout (iz fastest); theta in
    degrees.  obs maps "vx", "vz", "p" to (ntrace, ntsample) arrays (absent keys are
    not in the misfit).  Returns J as a float64 0-d tensor on vp's device."""

    @staticmethod
    def forward(ctx, vp, eps, delta, theta, modeling, vs, rho, obs, ncheckpoints=0, storage="f32", history=False):
        modeling.update_model(_host(vp), _host(vs), _host(rho), _host(eps), _host(delta), _host(theta))
        obs_h = {k: _host(v) for k, v in obs.items()}
        if any(ctx.needs_input_grad[:4]):
            J, g = modeling.gradient(obs_h.get("vx"), obs_h.get("vz"), obs_h.get("p"), ncheckpoints, storage, history)
            ctx.save_for_backward(torch.from_numpy(g))
        else:
            modeling.run()
            J = 0.0
            for comp, key in ((1, "vx"), (3, "vz"), (4, "p")):
                if key in obs_h:
                    r = modeling.traces(comp).astype(np.float64) - obs_h[key]
                    J += 0.5 * float((r * r).sum())
            ctx.save_for_backward(torch.empty(0))
        ctx.meta = [(t.dtype, t.device) for t in (vp, eps, delta, theta)]
        return torch.tensor(J, dtype=torch.float64, device=vp.device)

    @staticmethod
    def backward(ctx, gJ):
        (g,) = ctx.saved_tensors
        s = gJ.detach().to("cpu", torch.float64)
        grads = []
        for i, (dt, dev) in enumerate(ctx.meta):
            grads.append((s * g[i]).to(dt).to(dev) if ctx.needs_input_grad[i] else None)
        return (*grads, None, None, None, None, None, None, None)


def misfit(vp, eps, delta, theta, modeling, vs, rho, obs, ncheckpoints=0, storage="f32", history=False):
    """Differentiable L2 trace misfit of one shot; see TTIMisfit.  ncheckpoints, storage and
    history select how the gradient keeps the forward wavefield (Modeling.gradient)."""
    return TTIMisfit.apply(vp, eps, delta, theta, modeling, vs, rho, obs, ncheckpoints, storage, history)
