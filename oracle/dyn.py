"""Oracle: basic-state terms and explicit linearised tendencies (numpy FP64).
TEST INFRASTRUCTURE -- see ``oracle/__init__.py``.

Restates ``pylbm_isca/dynamics.py:51-271``.  Perturbation grid fields carry
a member axis ``(E, J, I, L)`` (surface fields ``(E, J, I)``); the basic
state is a single state broadcast over members (``ensemble.py:3-7``).
"""

import numpy as np

from .vgrid import KAPPA, R_DRY, compute_tbar

OMEGA = 7.2921150e-5
EPSVT = 461.50 / 287.04 - 1.0


def column_sum(x, dsig):
    """S[k] = sum_{l>=k} dsig_l x_l with S[L] = 0 (``dynamics.py:96-102``)."""
    w = x * dsig
    out = np.zeros(x.shape[:-1] + (x.shape[-1] + 1,))
    out[..., :-1] = np.cumsum(w[..., ::-1], axis=-1)[..., ::-1]
    return out


def sigma_dot(ssum, sigm):
    """sdot_k = sigm_k S_0 - S_k, zero at both ends (``dynamics.py:105-114``)."""
    sd = sigm * ssum[..., :1] - ssum
    sd[..., 0] = 0.0
    sd[..., -1] = 0.0
    return sd


def vert_advect(f, sdot, dsig):
    """Second-centred advective-form vertical advection (``dynamics.py:117-143``):
    (F_{k+1} - F_k - f_k (sdot_{k+1} - sdot_k)) / dsig_k,
    F_k = sdot_k (f_k + f_{k-1}) / 2 for interior half levels, 0 at the ends.
    """
    kmax = f.shape[-1]
    shape = np.broadcast_shapes(f.shape[:-1], sdot.shape[:-1])
    flux = np.zeros(shape + (kmax + 1,))
    flux[..., 1:kmax] = sdot[..., 1:kmax] * 0.5 * (f[..., 1:kmax] + f[..., :kmax - 1])
    dsd = sdot[..., 1:] - sdot[..., :-1]
    return (flux[..., 1:] - flux[..., :-1] - f * dsd) / dsig


class BasicTerms:
    """Fields derived from the time-constant basic state (``dynamics.py:54-93``).

    ``bs`` has ``u, v, t, q`` of shape (J, I, L) and ``ps`` (J, I) in hPa.
    Builds the reference-temperature operators on ``vg`` as the reference does.
    """

    def __init__(self, sht, vg, bs):
        e = lambda x: np.asarray(x, dtype=np.float64)[None]  # noqa: E731
        self.u, self.v, self.t, self.q = e(bs.u), e(bs.v), e(bs.t), e(bs.q)
        self.ps = e(bs.ps)                                        # (1, J, I)
        self.coriol = (2.0 * OMEGA * sht.mu)[None, :, None, None]
        self.vor_s, self.div_s = sht.uv_to_vordiv(self.u, self.v)
        self.vor = sht.wave_to_grid(self.vor_s)
        self.div = sht.wave_to_grid(self.div_s)
        self.lnps = np.log(self.ps)
        pix, piy = sht.hgrad(sht.grid_to_wave(self.lnps[..., None]))
        self.pix, self.piy = pix, piy                             # (1, J, I, 1)
        self.tv = self.t * (1.0 + EPSVT * self.q)
        self.padv = self.u * pix + self.v * piy
        self.ssum = column_sum(self.div + self.padv, vg.dsig)
        self.sdot = sigma_dot(self.ssum, vg.sigm)
        self.x5 = vg.w3 * self.ssum[..., 1:] + vg.alpha * self.div + vg.wP * self.padv
        vg.set_reference_temperature(compute_tbar(self.tv[0]))


def tendencies(sht, vg, bst, st, frc_t=None, strictly_linear=False,
               frc_vor=None, frc_div=None, frc_ps=None, frc_q=None):
    """Explicit spectral tendencies (vor, div, t, pi, q) and diagnostics.

    ``st`` has grid fields u, v, t, q (E, J, I, L), ps (E, J, I) and ps_bar
    (J, I).  Reference: ``dynamics.py:156-271``.
    """
    dsig, sigm = vg.dsig, vg.sigm
    tref = vg.tbar
    up, vp, tp, qp = st.u, st.v, st.t, st.q

    vor_s, div_s = sht.uv_to_vordiv(up, vp)
    vor = sht.wave_to_grid(vor_s)
    div = sht.wave_to_grid(div_s)

    pi = st.lnps_pert()                                           # (E, J, I)
    pix, piy = sht.hgrad(sht.grid_to_wave(pi[..., None]))         # (E, J, I, 1)

    padv = up * bst.pix + bst.u * pix + vp * bst.piy + bst.v * piy
    ssum = column_sum(div + padv, dsig)
    sdot = sigma_dot(ssum, sigm)
    gpid = -ssum[..., 0]
    tv = tp + EPSVT * (tp * bst.q + bst.t * qp)

    # continuity (dynamics.py:188-191)
    gt_pi = gpid + (div * dsig).sum(axis=-1)
    if frc_ps is not None:
        gt_pi = gt_pi + frc_ps
    wt_pi = sht.grid_to_wave(gt_pi[..., None])[..., 0]

    # temperature (dynamics.py:197-219)
    x5p = vg.alpha * div + vg.wP * padv + vg.w3 * ssum[..., 1:]
    adiab = -KAPPA * (tv * bst.x5 + bst.tv * x5p)
    tdev = bst.t - tref
    xu = up * tdev + bst.u * tp
    xv = vp * tdev + bst.v * tp
    hadv = tp * bst.div + tdev * div
    vadv = vert_advect(tp, bst.sdot, dsig) + vert_advect(bst.t, sdot, dsig)
    gt_t = adiab + hadv + vadv + div @ vg.tmtx.T
    if frc_t is not None:
        gt_t = gt_t + frc_t
    wt_t = sht.neg_divergence(xu, xv) + sht.grid_to_wave(gt_t)

    # momentum (dynamics.py:224-251)
    absvor = bst.vor + bst.coriol
    fu = vp * absvor + bst.v * vor
    fv = -up * absvor - bst.u * vor
    rcpg = R_DRY * vg.cpg
    tvdev = bst.tv - tref
    fu = fu - rcpg * (tv * bst.pix + tvdev * pix)
    fv = fv - rcpg * (tv * bst.piy + tvdev * piy)
    fu = fu + vert_advect(up, bst.sdot, dsig) + vert_advect(bst.u, sdot, dsig)
    fv = fv + vert_advect(vp, bst.sdot, dsig) + vert_advect(bst.v, sdot, dsig)
    ke = vg.hydrostatic(tv - tp) + up * bst.u + vp * bst.v
    if not strictly_linear:
        ke = ke + 0.5 * (up * up + vp * vp)
    wt_vor, wt_div = sht.uv_to_vordiv(fu, fv)
    wt_div = wt_div - sht.laplacian(sht.grid_to_wave(ke))
    if frc_vor is not None:
        wt_vor = wt_vor + sht.grid_to_wave(frc_vor)
    if frc_div is not None:
        wt_div = wt_div + sht.grid_to_wave(frc_div)

    # humidity (dynamics.py:256-262)
    yu = up * bst.q + bst.u * qp
    yv = vp * bst.q + bst.v * qp
    g = qp * bst.div + bst.q * div
    g = g + vert_advect(qp, bst.sdot, dsig) + vert_advect(bst.q, sdot, dsig)
    if frc_q is not None:
        g = g + frc_q
    wt_q = sht.neg_divergence(yu, yv) + sht.grid_to_wave(g)

    # omega diagnostic (dynamics.py:267)
    omega = -vg.sig * (x5p * bst.ps[..., None] + bst.x5 * st.ps[..., None])
    diag = {"vor": vor, "div": div, "vor_s": vor_s, "div_s": div_s,
            "omega": omega, "sdot": sdot, "pi": pi}
    return (wt_vor, wt_div, wt_t, wt_pi, wt_q), diag
