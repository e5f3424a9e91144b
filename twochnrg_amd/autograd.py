"""torch.autograd face of the derivative pass (a test harness, not a training
path).

NRGSpectral maps (U, Gamma, ed) to the DM-NRG spectral function of c_d on an
omega grid: each omega is evaluated by the one shell whose CalcSpec_ManyOmegas
range holds it (omega / D_N in (Lambda^0.75, Lambda^1.25], the lowest omegas by
the last shell), A(omega) = A_N(omega / D_N) / D_N as that driver records it.

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


def calc_dn(Lambda, N):
    """CalcDN (z = 1): D_N = (1 + 1/Lambda)/2 Lambda^{-(N-1)/2}."""
    return 0.5 * (1.0 + 1.0 / Lambda) * Lambda ** (-(N - 1) / 2.0)


def omega_shells(omega, Lambda, nshells):
    """(shell, omegabar, D_N) per omega: the shell whose ManyOmegas range
    (Lambda^0.75, Lambda^1.25] holds |omega| / D_N; omegas above shell 0's
    range go to shell 0, those below the last shell's to the last shell."""
    out = []
    for w in np.atleast_1d(omega):
        N = nshells - 1
        for n in range(nshells):
            if abs(w) / calc_dn(Lambda, n) > Lambda ** 0.75:
                N = n
                break
        d = calc_dn(Lambda, N)
        out.append((N, w / d, d))
    return out


class NRGSpectral(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, run):
        p = params.detach().cpu().double().numpy()
        kw = dict(run["kw"], U=float(p[0]), Gamma=float(p[1]), ed=float(p[2]))
        grid = run["grid"]
        with run["ctx"].dm(kept=run["kept"], dpar=np.eye(3), **kw) as d:
            sh = [g[0] for g in grid]
            wb = [g[1] for g in grid]
            dn = np.array([g[2] for g in grid])
            val = d.spec(sh, wb) / dn
            jac = d.spec_jvp(sh, wb) / dn[:, None]
        ctx.save_for_backward(torch.from_numpy(jac))
        return torch.from_numpy(val).to(params.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        (jac,) = ctx.saved_tensors
        return grad_out.double() @ jac.to(grad_out.device), None


def make_spectral_run(ctx: Context, base_params, omega, **kw):
    """Freeze the Kept sets of a calcdens-1 chain at base_params and map omega to
    shells for NRGSpectral.  kw: the DM-NRG parameters of Context.dm."""
    U, G, ed = base_params
    steps, _ = ctx.run_onechqs(U=U, Gamma=G, ed=ed, Lambda=kw["Lambda"], nsitesmax=kw["nsitesmax"],
                               ncut=kw["ncut"], calcdens=1, disc_scheme=kw.get("disc_scheme", 0))
    with ctx.dm(U=U, Gamma=G, ed=ed, **kw) as d:
        nsh = d.spec_nshell_max
    return {"ctx": ctx, "kw": kw, "kept": kept_arrays(steps), "grid": omega_shells(omega, kw["Lambda"], nsh)}


def spectral(params, run):
    """A(omega) on run's grid as a differentiable function of params."""
    return NRGSpectral.apply(params, run)


def make_run(ctx: Context, n_out, base_params, **kw):
    """Freeze the Kept sets of a plain run at base_params = (U, Gamma, ed)."""
    kw = dict(kw, nsitesmax=max(kw.get("nsitesmax", n_out + 1), n_out + 1))
    U, G, ed = base_params
    steps, _ = ctx.run_onechqs(U=U, Gamma=G, ed=ed, **kw)
    return {"ctx": ctx, "kw": kw, "kept": kept_arrays(steps), "n_out": n_out}


def observables(params, run):
    """(<n_d>, T chi) at record N_out as a differentiable function of params."""
    return NRGObservables.apply(params, run)
