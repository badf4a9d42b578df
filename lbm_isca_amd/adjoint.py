"""Sensitivity of a response to the forcing and the initial state.

The reference has no adjoint.  A response that is linear in the states,

    J_e = sum_k w_k sum_l w_l sum_{j,i} W_{j,i} psi_l(j, i; step k),

gets its gradient with respect to member e's spectral initial state and
temperature forcing from one backward sweep of the hand-written transposed
kernels, for every member at once:

* the tangent-linear model (``Config(tangent_linear=True)``: strictly linear
  tendencies, ``ps' = ps_bar pi'``) is exactly linear, so
  ``J_e = <g_x, x0_e> + <g_f, f_e>`` and the sweep needs no trajectory;
* the default model keeps u'u'/2 + v'v'/2 and the exact ps' <-> pi' pair, so
  its gradient is taken about the trajectory of the run: the run keeps the
  whole state every K steps (``plan.set_checkpoints(K)``, 48 by default
  here), and the sweep recomputes one K-step segment at a time from its
  checkpoint, storing the per-step u, v and ps' it needs, before running
  the transposed kernels about them backward through that segment.

``LinearResponse`` wraps a forward run and the sweep as a
``torch.autograd.Function``; complex gradients follow torch's convention
``dJ/dRe + i dJ/dIm``.
"""

import numpy as np

from . import _lib  # noqa: F401  (the library must be present)


def box_weights(lat_deg, lon_deg, gw, lat_range, lon_range):
    """Area-mean weights of a latitude/longitude box on the Gaussian grid.

    The Gaussian weight ``gw_j`` is the area of latitude band ``j``; every
    longitude of a band has the same area.  Returns (J, I) weights summing
    to one over the grid points inside the box (edges included).
    """
    lat = np.asarray(lat_deg, dtype=np.float64)
    lon = np.mod(np.asarray(lon_deg, dtype=np.float64), 360.0)
    inlat = (lat >= lat_range[0] - 1e-9) & (lat <= lat_range[1] + 1e-9)
    lo, hi = lon_range[0] % 360.0, lon_range[1] % 360.0
    if lo <= hi:
        inlon = (lon >= lo - 1e-9) & (lon <= hi + 1e-9)
    else:
        inlon = (lon >= lo - 1e-9) | (lon <= hi + 1e-9)
    w = np.outer(np.where(inlat, np.asarray(gw, dtype=np.float64), 0.0), inlon.astype(np.float64))
    tot = w.sum()
    if tot <= 0:
        raise ValueError("the box contains no grid point")
    return w / tot


def sigma_weights(sig, target):
    """Weights of the full levels that interpolate linearly in sigma to
    ``target`` (levels ordered surface first)."""
    sig = np.asarray(sig, dtype=np.float64)
    w = np.zeros(sig.size)
    if target >= sig[0]:
        w[0] = 1.0
        return w
    if target <= sig[-1]:
        w[-1] = 1.0
        return w
    k = int(np.nonzero(sig >= target)[0][-1])
    f = (sig[k] - target) / (sig[k] - sig[k + 1])
    w[k], w[k + 1] = 1.0 - f, f
    return w


def step_weights(nsteps, inistp, steps_per_day, day0, day1):
    """Mean over the states after the full steps whose model time lies in
    [day0, day1]; index k of the result is the run's step k (the
    ``inistp`` start-up half steps come first and get no weight)."""
    w = np.zeros(int(inistp) + int(nsteps))
    k = np.arange(1, int(nsteps) + 1)
    t = k / float(steps_per_day)
    sel = (t >= day0 - 1e-9) & (t <= day1 + 1e-9)
    if not sel.any():
        raise ValueError("no step falls in the averaging window")
    w[int(inistp) + k[sel] - 1] = 1.0 / sel.sum()
    return w


class StreamfunctionResponse:
    """Area mean of the streamfunction at one sigma over a box, averaged over
    a window of days; installs itself on a plan with :meth:`apply`."""

    def __init__(self, plan, lat_deg, lon_deg, gw, sig, sigma, lat_range, lon_range,
                 day0, day1, nsteps, steps_per_day, inistp):
        self.wgrid = box_weights(lat_deg, lon_deg, gw, lat_range, lon_range)
        self.wlev = sigma_weights(sig, sigma)
        self.wstep = step_weights(nsteps, inistp, steps_per_day, day0, day1)
        self.plan = plan
        self.apply()

    def apply(self):
        self.plan.set_response(self.wgrid, self.wlev, self.wstep)


def _autograd_function():
    import torch

    class LinearResponse(torch.autograd.Function):
        """J = LinearResponse.apply(plan, nsteps, use_graph, frc, vor, div, t, q, pi).

        ``plan`` is a :class:`lbm_isca_amd.plan.Plan` (tangent-linear, or
        the default model with checkpoints, which ``response`` turns on)
        with a response installed; ``frc`` (E|1, J, I, L) float64 K/s or
        None; the spectral initial state (E, M, M, L) / (E, M, M) complex128
        or None.  ``nsteps`` <= 0 runs each member for its own length
        (``plan.set_lengths``).  Returns J (E,) float64.  The backward sweep
        refuses a plan whose response, lengths, forcing or basic state
        changed after the forward run (the gradient would belong to another
        function).
        """

        @staticmethod
        def forward(ctx, plan, nsteps, use_graph, frc, vor, div, t, q, pi):
            plan.set_forcing(frc)
            plan.set_spectral_state(vor, div, t, q, pi)
            plan.run(nsteps, use_graph=use_graph)
            J = torch.as_tensor(plan.get_response(), dtype=torch.float64, device="cuda")
            ctx.plan, ctx.nsteps, ctx.use_graph = plan, nsteps, use_graph
            ctx.generation = plan.generation
            ctx.frc_shape = None if frc is None else tuple(frc.shape)
            ctx.present = [x is not None for x in (vor, div, t, q, pi)]
            return J

        @staticmethod
        def backward(ctx, gJ):
            plan = ctx.plan
            if plan.generation != ctx.generation:
                raise RuntimeError("the plan's response, lengths, forcing or basic state changed "
                                   "after the forward run; run it again before the backward")
            gx, gf = plan.adjoint_run(ctx.nsteps, use_graph=ctx.use_graph)
            w = gJ.to(torch.float64)
            E = plan.nmem
            grads = []
            for g, present in zip(gx, ctx.present):
                grads.append(g * w.view((E,) + (1,) * (g.dim() - 1)) if present else None)
            gfrc = None
            if ctx.frc_shape is not None:
                gfrc = gf * w.view(E, 1, 1, 1)
                if len(ctx.frc_shape) == 3 or ctx.frc_shape[0] == 1:
                    # one forcing shared by every member
                    gfrc = gfrc.sum(0).reshape(ctx.frc_shape)
            return (None, None, None, gfrc, *grads)

    return LinearResponse


_FN = None


CHECKPOINT_EVERY = 48


def response(plan, nsteps, frc=None, vor=None, div=None, t=None, q=None, pi=None,
             use_graph=True, checkpoint_every=CHECKPOINT_EVERY):
    """J (E,) of the installed response, differentiable with respect to
    ``frc`` and the spectral initial state through the adjoint sweep: of the
    tangent-linear model, or of the default model about its trajectory
    (checkpoints every ``checkpoint_every`` steps)."""
    global _FN
    if _FN is None:
        _FN = _autograd_function()
    if not plan.tangent_linear:
        plan.set_checkpoints(checkpoint_every)
    return _FN.apply(plan, int(nsteps), bool(use_graph), frc, vor, div, t, q, pi)


def linear_response(plan, nsteps, frc=None, vor=None, div=None, t=None, q=None, pi=None,
                    use_graph=True):
    """``response`` of a tangent-linear plan."""
    if not plan.tangent_linear:
        raise RuntimeError("linear_response needs a plan built with tangent_linear=True "
                           "(use response() for the default model)")
    return response(plan, nsteps, frc, vor, div, t, q, pi, use_graph)
