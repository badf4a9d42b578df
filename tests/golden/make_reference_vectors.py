#!/usr/bin/env python3
"""Write reference-generated vectors for the parts of the hot path that no
closed form, invariant or reference-held file pins: the tendencies and the
semi-implicit step on a non-trivial (moist, zonally asymmetric, baroclinic)
basic state, and a short multi-step run through start-up, Robert-Asselin and
the exact log-pressure round trip.

Run once, in a container where the reference checkout exists:

    python tests/golden/make_reference_vectors.py /root/reference

It imports the reference's ``pylbm_isca`` package from that path, evaluates
it on the inputs of ``tests/golden/cases.py`` and writes
``tests/golden/reference_vectors.npz`` (inputs are rebuilt from the seeds,
only outputs and sample indices are stored).  Nothing else in this
repository imports the reference.
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import cases  # noqa: E402


def main(ref_root):
    sys.path.insert(0, ref_root)
    from pylbm_isca import Config, LBM, Spectral, VerticalGrid
    from pylbm_isca.dynamics import BasicStateTerms, Dynamics
    from pylbm_isca.forcing import SteadyForcing
    from pylbm_isca.state import State

    out = {}
    sp = Spectral(device="cpu")
    vg = VerticalGrid(device="cpu")
    L = vg.kmax
    lv = list(cases.SAMPLE_LEVELS)

    # -- vertical operators (stretched L20) -------------------------------
    for name in ("alpha", "sig", "cpg", "w3", "wP", "dsigm", "ww"):
        out["vg_" + name] = np.asarray(getattr(vg, name))

    bs = cases.jet_basic_state(sp.lat, sp.lon, vg.sig)
    bst = BasicStateTerms(sp, vg, bs)
    out["bs_tbar"] = np.asarray(vg.tbar)
    out["bs_tmtx"] = np.asarray(vg.tmtx)
    for name in ("x5", "sdot", "ssum", "padv", "vor_g", "div"):
        arr = np.asarray(getattr(bst, name))
        idx = cases.sample_index(arr.shape, 512, seed=11)
        out["bs_%s_idx" % name] = idx
        out["bs_%s_val" % name] = arr.ravel()[idx]
    out["bs_pix"] = np.asarray(bst.pix)
    out["bs_piy"] = np.asarray(bst.piy)

    def make_state(seed):
        psi, chi, t, q, pi = cases.perturbation_spectra(L, seed)
        u, v = sp.psichi_to_uv(psi, chi)
        st = State.zeros(np.asarray(bs.ps), L, xp=np)
        st.u, st.v = np.asarray(u), np.asarray(v)
        st.t = np.asarray(sp.wave_to_grid(t))
        st.q = np.asarray(sp.wave_to_grid(q))
        st.set_from_lnps(np.asarray(sp.wave_to_grid(pi)))
        return st

    st1 = make_state(0)
    st2 = make_state(1)
    names = ("vor", "div", "t", "pi", "q")
    for lin in (False, True):
        dyn = Dynamics(sp, vg, bst, strictly_linear=lin)
        tend = dyn.tendencies(st1)
        tag = "lin" if lin else "full"
        for nm, x in zip(names, tend):
            x = np.asarray(x)
            out["tend_%s_%s" % (tag, nm)] = x if x.ndim == 2 else x[:, :, lv]
            out["tend_%s_%s_norm" % (tag, nm)] = np.sqrt((np.abs(x) ** 2).sum(axis=(0, 1)))
        if not lin:
            om = np.asarray(dyn.diag["omega"])
            idx = cases.sample_index(om.shape, 512, seed=12)
            out["diag_omega_idx"] = idx
            out["diag_omega_val"] = om.ravel()[idx]
            tend_full = tend

    # -- one semi-implicit step with damping + vertical diffusion ----------
    model = LBM(bs, sp, vg, Config())
    prev = st2.to_spectral(sp)
    for dt in (1200.0, 300.0):
        new = model.integ.step(dt, prev, tend_full)
        for nm in names:
            x = np.asarray(new[nm])
            key = "step%d_%s" % (int(dt), nm)
            out[key] = x if x.ndim == 2 else x[:, :, lv]
            out[key + "_norm"] = np.sqrt((np.abs(x) ** 2).sum(axis=(0, 1)))
        out["semi_cond_%d" % int(dt)] = np.asarray(model.semi.cond)

    # -- short runs: raw (non-band-limited) initial T, heating, jet --------
    rng = np.random.default_rng(5)
    t_raw = 0.05 * rng.standard_normal((sp.nlat, sp.nlon, L))
    heat = cases.heating(sp.lat, sp.lon, vg.sig, lat0=10.0, lon0=150.0)
    frc = SteadyForcing.zeros(sp.nlat, sp.nlon, L)
    frc.t = heat
    for nsteps in (1, 10, 72):
        m = LBM(bs, sp, vg, Config(ndays=nsteps / 72.0))
        m.set_initial_perturbation(u=st1.u, v=st1.v, t=st1.t + t_raw, q=st1.q, ps=st1.ps)
        m.run(steady_forcing=frc, verbose=False)
        for f in ("u", "v", "t", "q", "ps"):
            arr = np.asarray(getattr(m.ga, f))
            idx = cases.sample_index(arr.shape, 256, seed=20 + nsteps)
            out["run%d_%s_idx" % (nsteps, f)] = idx
            out["run%d_%s_val" % (nsteps, f)] = arr.ravel()[idx]
            out["run%d_%s_norm" % (nsteps, f)] = np.sqrt((arr ** 2).sum())
    out["run_t_raw_seed"] = np.array(5)

    path = os.path.join(HERE, "reference_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
