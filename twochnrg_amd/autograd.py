"""torch.autograd face of the derivative pass (a test harness, not a training
path).

NRGObservables maps the Anderson parameters (U, Gamma, ed) to observables of
the fixed-truncation OneChQS NRG_main loop on the device: <n_d> and T*chi at
the record N_out, with T = D_N / betabar.  Its forward runs the derivative
pass (nrg_onechqs_jvp) with the three unit tangents, so the full Jacobian
comes out of one device run; backward contracts the incoming gradient with
it.  With three inputs and several outputs, three tangents cost less than
one reverse sweep per output.  The Kept set of every step is frozen at the
first evaluation's (kept_arrays of a plain run), as the derivative assumes.
"""
import numpy as np
import torch

from . import Context, kept_arrays

_OUTPUTS = ("nd", "Tchi")


def _observables(steps, tans, n_out):
    s = next(x for x in steps if x["N"] == n_out)
    t = next(x for x in tans if x["N"] == n_out)
    # opavg[0] = <n_d>; suscep[3] = T chi at the record's own betabar (nc = 4)
    val = np.array([s["opavg"][0], s["suscep"][3]])
    jac = np.stack([t["dopavg"][:, 0], t["dsuscep"][:, 3]])   # (outputs, 3)
    return val, jac


class NRGObservables(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, run):
        p = params.detach().cpu().double().numpy()
        kw = dict(run["kw"], U=float(p[0]), Gamma=float(p[1]), ed=float(p[2]))
        steps, tans = run["ctx"].onechqs_jvp(kept=run["kept"], dpar=np.eye(3), **kw)
        val, jac = _observables(steps, tans, run["n_out"])
        ctx.save_for_backward(torch.from_numpy(jac))
        return torch.from_numpy(val).to(params.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        (jac,) = ctx.saved_tensors
        return grad_out.double() @ jac.to(grad_out.device), None


def make_run(ctx: Context, n_out, base_params, **kw):
    """Freeze the Kept sets of a plain run at base_params = (U, Gamma, ed)."""
    kw = dict(kw, nsitesmax=max(kw.get("nsitesmax", n_out + 1), n_out + 1))
    U, G, ed = base_params
    steps, _ = ctx.run_onechqs(U=U, Gamma=G, ed=ed, **kw)
    return {"ctx": ctx, "kw": kw, "kept": kept_arrays(steps), "n_out": n_out}


def observables(params, run):
    """(<n_d>, T chi) at record N_out as a differentiable function of params."""
    return NRGObservables.apply(params, run)
