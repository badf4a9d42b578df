"""Builders shared by the oracle tests and the HIP parity tests."""

import numpy as np

import cases
from oracle import BasicTerms, GridState, OracleConfig, OracleModel, SHT, VerticalOps


def oracle_setup(sigm=None):
    sht = SHT()
    vg = VerticalOps() if sigm is None else VerticalOps(sigm)
    bs = cases.jet_basic_state(sht.lat, sht.lon, vg.sig)
    bst = BasicTerms(sht, vg, bs)
    return sht, vg, bs, bst


def grid_from_spectra(sht, spectra):
    """(psi, chi, t, q, pi) spectra -> grid u, v, t, q (1, J, I, L), pi (1, J, I)."""
    psi, chi, t, q, pi = (x[None] for x in spectra)
    u, v = sht.psichi_to_uv(psi, chi)
    return (u, v, sht.wave_to_grid(t), sht.wave_to_grid(q),
            sht.wave_to_grid(pi[..., None])[..., 0])


def oracle_state(sht, bs, nlev, seed):
    u, v, t, q, pi = grid_from_spectra(sht, cases.perturbation_spectra(nlev, seed))
    st = GridState(bs.ps, 1, nlev)
    st.u, st.v, st.t, st.q = u, v, t, q
    st.set_from_lnps(pi)
    return st


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.sqrt((np.abs(b) ** 2).sum())
    return float(np.sqrt((np.abs(a - b) ** 2).sum()) / (den if den > 0 else 1.0))


def oracle_response(bs, nmem, spec, frc, nsteps, wgrid, wlev, wstep, sigm=None):
    """J_e = sum_k wstep[k] sum_l wlev[l] sum_ji wgrid psi_l of the oracle's
    linear-mode run from a spectral state (both time levels), with psi the
    inverse Laplacian of the new vorticity synthesised on the grid."""
    m = OracleModel(bs, nmem, OracleConfig(linear=True), sigm=sigm)
    m.ga.load_spectral(m.sht, spec)
    m.gb.load_spectral(m.sht, spec)
    c = m.cfg
    J = np.zeros(nmem)
    for k in range(c.inistp + nsteps):
        oinit = k < c.inistp
        dt = c.delt / (2.0 ** (c.inistp - k)) if oinit else c.delt
        new = m.step(dt, oinit, frc)
        if k < len(wstep) and wstep[k] != 0.0:
            psi = m.sht.wave_to_grid(m.sht.inverse_laplacian(new["vor"]))
            J += wstep[k] * np.einsum("ejil,ji,l->e", psi, wgrid, wlev)
    return J
