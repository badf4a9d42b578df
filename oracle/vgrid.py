"""Oracle: Simmons-Burridge pure-sigma vertical operators (numpy FP64).
TEST INFRASTRUCTURE -- see ``oracle/__init__.py``.

Restates ``pylbm_isca/vertical.py:59-216``.  Index k=0 is the bottom layer;
half level k lies below full level k (``sigm[k] > sig[k] > sigm[k+1]``).
"""

import numpy as np

R_DRY = 287.04
KAPPA = 2.0 / 7.0

# Stretched L20 half levels of the reference (pylbm_isca/constants.py:92-97).
SIGM_L20 = np.array([1.000, 0.990, 0.970, 0.930, 0.870, 0.790, 0.700, 0.600,
                     0.500, 0.410, 0.330, 0.260, 0.200, 0.150, 0.100, 0.070,
                     0.050, 0.040, 0.030, 0.020, 0.000])


class VerticalOps:
    """alpha, full-level sigma, log-pressure weights, cpg/w3/wP, dsigm, WW.

    Reference: ``pylbm_isca/vertical.py:59-129``.
    """

    def __init__(self, sigm=SIGM_L20):
        self.sigm = np.asarray(sigm, dtype=np.float64)
        self.kmax = self.sigm.size - 1
        bot, top = self.sigm[:-1], self.sigm[1:]
        self.dsig = bot - top
        interior = top > 0.0
        safe_top = np.where(interior, top, 1.0)
        ratio = np.where(interior, top * (np.log(bot) - np.log(safe_top)) / self.dsig, 0.0)
        self.alpha = 1.0 - ratio
        self.sig = bot * np.exp(-self.alpha)
        self.dlog1 = self.alpha.copy()
        self.dlog2 = np.where(interior, np.log(self.sig) - np.log(safe_top), 0.0)
        self.dlog3 = np.where(interior, np.log(bot) - np.log(safe_top), 0.0)
        self.cpg = (bot * self.dlog1 + top * self.dlog2) / self.dsig
        self.w3 = self.dlog3 / self.dsig
        self.wP = self.alpha - self.cpg
        k = self.kmax
        self.dsigm = np.empty(k + 1)
        self.dsigm[0] = self.sigm[0] - self.sig[0]
        self.dsigm[1:k] = self.sig[:-1] - self.sig[1:]
        self.dsigm[k] = self.sig[-1] - self.sigm[-1]
        # phi[k] = sum_{l<k} R dlog3[l] Tv[l] + R alpha[k] Tv[k]
        self.ww = np.tril(np.ones((k, k)), -1) * (R_DRY * self.dlog3)[None, :]
        self.ww[np.arange(k), np.arange(k)] = R_DRY * self.alpha
        self.tbar = self.gg = self.tmtx = self.whmtx = self.gcmtx = None

    def set_reference_temperature(self, tbar):
        """gg = R Tbar cpg; tmtx (upper triangular, diag kappa Tbar alpha);
        whmtx = WW tmtx; gcmtx = gg (x) dsig.  Reference: ``vertical.py:139-193``.
        """
        tbar = np.asarray(tbar, dtype=np.float64).copy()
        k = self.kmax
        self.tbar = tbar
        self.gg = R_DRY * tbar * self.cpg
        tm = np.zeros((k, k))
        for r in range(k):
            tm[r, r] = KAPPA * tbar[r] * self.alpha[r]
            tm[r, r + 1:] = KAPPA * tbar[r] * self.w3[r] * self.dsig[r + 1:]
        self.tmtx = tm
        self.whmtx = self.ww @ tm
        self.gcmtx = np.outer(self.gg, self.dsig)

    def hydrostatic(self, tv):
        """phi = Tv @ WW^T over the last axis (``vertical.py:196-203``)."""
        return tv @ self.ww.T


def compute_tbar(tv):
    """Plain grid mean over (lat, lon) of a (J, I, L) field (``vertical.py:206-216``)."""
    nlat, nlon, _ = tv.shape
    return tv.sum(axis=(0, 1)) / (nlat * nlon)
