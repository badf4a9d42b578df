"""Oracle: Gaussian grid, associated Legendre tables and spherical-harmonic
transforms (numpy FP64).  TEST INFRASTRUCTURE -- see ``oracle/__init__.py``.

Restates ``pylbm_isca/spectral.py`` of the reference:

* Gaussian grid (``spectral.py:57-68``): Gauss-Legendre nodes mu sorted north
  first, weights halved so they sum to one, ``gwdel = gw/(a(1-mu^2))``.
* Fully normalised P_n^m without Condon-Shortley phase and
  H_n^m = (1-mu^2) dP_n^m/dmu (``spectral.py:93-154``).
* Transforms (``spectral.py:159-276``): rFFT/nlon truncated to m<=N, per-m
  Legendre contraction over latitude; the inverse zero-fills m>N, multiplies
  by nlon and takes the real inverse FFT (the imaginary parts of m=0 and of
  the Nyquist mode are discarded by the real inverse).

Array convention here: every grid field carries an explicit leading member
axis and a trailing column axis, ``(E, J, I, K)``; spectral fields are
``(E, M, M, K)`` complex indexed ``[m, n]`` with zeros for n<m.  The
reference folds the member axis into the column axis (``ensemble.py:55-72``);
the arithmetic per column is identical.
"""

import numpy as np


def gauss_grid(nlat, refine=False):
    """North-first Gauss-Legendre nodes and weights summing to one.

    Reference: ``pylbm_isca/spectral.py:57-68`` (numpy ``leggauss``).
    ``refine=True`` (not the reference's) polishes the nodes by Newton steps
    on P_J and recomputes w = 2/((1-x^2) P_J'(x)^2) in extended precision:
    ``leggauss`` weights carry relative errors growing with J (6e-11 at J =
    192), which the device's weights do not.
    """
    mu, w = np.polynomial.legendre.leggauss(nlat)
    if refine:
        x = mu.astype(np.longdouble)
        for _ in range(3):
            p0, p1 = np.ones_like(x), x.copy()
            for k in range(2, nlat + 1):
                p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
            dp = nlat * (x * p1 - p0) / (x * x - 1)
            x = x - p1 / dp
        p0, p1 = np.ones_like(x), x.copy()
        for k in range(2, nlat + 1):
            p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
        dp = nlat * (x * p1 - p0) / (x * x - 1)
        wl = 2 / ((1 - x * x) * dp * dp)
        x = (x - x[::-1]) / 2
        wl = (wl + wl[::-1]) / 2
        wl = wl * (2 / wl.sum())
        mu, w = x.astype(np.float64), wl.astype(np.float64)
    order = np.argsort(-mu)
    return mu[order], w[order] / 2.0


def legendre_tables(nmax, mu):
    """P[m, n, j] for n <= nmax+1 and H[m, n, j] for n <= nmax, plus eps.

    Recurrences of ``pylbm_isca/spectral.py:102-138``:
    eps(n,m) = sqrt((n^2-m^2)/(4n^2-1));
    P_m^m = P_{m-1}^{m-1} sqrt((2m+1)/(2m)) cos(lat); P_{m+1}^m = sqrt(2m+3) mu P_m^m;
    P_n^m = (mu P_{n-1}^m - eps(n-1,m) P_{n-2}^m) / eps(n,m);
    H_n^m = -n eps(n+1,m) P_{n+1}^m + (n+1) eps(n,m) P_{n-1}^m.
    """
    nlat = mu.size
    nx = nmax + 2
    coslat = np.sqrt(1.0 - mu * mu)
    eps = np.zeros((nmax + 1, nx))
    for m in range(nmax + 1):
        n = np.arange(m + 1, nx, dtype=np.float64)
        eps[m, m + 1:] = np.sqrt((n * n - m * m) / (4.0 * n * n - 1.0))
    p = np.zeros((nmax + 1, nx, nlat))
    p[0, 0] = 1.0
    for m in range(1, nmax + 1):
        p[m, m] = p[m - 1, m - 1] * np.sqrt((2.0 * m + 1.0) / (2.0 * m)) * coslat
    for m in range(nmax + 1):
        p[m, m + 1] = p[m, m] * np.sqrt(2.0 * m + 3.0) * mu
        for n in range(m + 2, nx):
            p[m, n] = (mu * p[m, n - 1] - eps[m, n - 1] * p[m, n - 2]) / eps[m, n]
    h = np.zeros((nmax + 1, nmax + 1, nlat))
    for m in range(nmax + 1):
        for n in range(m, nmax + 1):
            acc = -n * eps[m, n + 1] * p[m, n + 1]
            if n - 1 >= m:
                acc = acc + (n + 1.0) * eps[m, n] * p[m, n - 1]
            h[m, n] = acc
    return p[:, :nmax + 1, :].copy(), h, eps


class SHT:
    """Spherical-harmonic transform pair on the Gaussian grid.

    Reference: ``pylbm_isca/spectral.py:43-289`` (class ``Spectral``).
    """

    def __init__(self, nlon=128, nlat=64, nmax=42, radius=6.371e6, refine_gauss=False):
        self.nlon, self.nlat, self.nmax = int(nlon), int(nlat), int(nmax)
        self.a = float(radius)
        self.mu, self.gw = gauss_grid(self.nlat, refine_gauss)
        self.lat = np.arcsin(self.mu)
        self.coslat = np.sqrt(1.0 - self.mu ** 2)
        self.lon = np.arange(self.nlon) * (2.0 * np.pi / self.nlon)
        self.gwdel = self.gw / self.a / (1.0 - self.mu ** 2)
        self.pnm, self.hnm, self.eps = legendre_tables(self.nmax, self.mu)
        n = np.arange(self.nmax + 1, dtype=np.float64)
        self.nn1 = n * (n + 1.0)
        self.laplace = -self.nn1 / self.a ** 2
        self.inv_laplace = np.zeros_like(n)
        self.inv_laplace[1:] = -self.a ** 2 / self.nn1[1:]
        self.im = 1j * np.arange(self.nmax + 1, dtype=np.float64)
        mm, nn = np.meshgrid(np.arange(self.nmax + 1), np.arange(self.nmax + 1),
                             indexing="ij")
        self.mask = nn >= mm
        # pre-weighted analysis tables, [m, n, j]
        self.p_gw = self.pnm * self.gw
        self.p_gwdel = self.pnm * self.gwdel
        self.h_gwdel = self.hnm * self.gwdel

    @property
    def lat_deg(self):
        return np.degrees(self.lat)

    @property
    def lon_deg(self):
        return np.degrees(self.lon)

    # -- Fourier stage (spectral.py:159-172) --------------------------------
    def fourier(self, f):
        """(E, J, I, K) real -> (E, M, J, K) complex, rfft/nlon, m <= N."""
        fm = np.fft.rfft(f, axis=2) / self.nlon
        return np.moveaxis(fm[:, :, :self.nmax + 1, :], 2, 1)

    def fourier_inv(self, fm):
        """(E, M, J, K) complex -> (E, J, I, K) real; zero-fill, *nlon, irfft."""
        e, m, j, k = fm.shape
        full = np.zeros((e, j, self.nlon // 2 + 1, k), dtype=np.complex128)
        full[:, :, :m, :] = np.moveaxis(fm, 1, 2)
        return np.fft.irfft(full * self.nlon, n=self.nlon, axis=2)

    # -- Legendre stage -------------------------------------------------------
    @staticmethod
    def _contract(table, x):
        # Fold the member axis into the column axis (ensemble.py:55-72), one
        # batched matmul over m (spectral.py:190,196), unfold.
        e, m, _, k = x.shape
        y = np.moveaxis(x, 0, 2).reshape(m, x.shape[2], e * k)
        out = np.matmul(table, y)
        return np.moveaxis(out.reshape(m, out.shape[1], e, k), 2, 0)

    def _ana(self, table, fm):
        # table (M, N, J) real; fm (E, M, J, K) -> (E, M, N, K)
        return self._contract(table, fm)

    def _syn(self, table, xi):
        # transposed table (M, J, N); xi (E, M, N, K) -> (E, M, J, K)
        return self._contract(np.ascontiguousarray(np.swapaxes(table, 1, 2)), xi)

    def _imk(self, x):
        return self.im[None, :, None, None] * x

    # -- public transforms (spectral.py:186-276) ----------------------------
    def grid_to_wave(self, f):
        return self._ana(self.p_gw, self.fourier(f))

    def wave_to_grid(self, xi):
        return self.fourier_inv(self._syn(self.pnm, xi))

    def laplacian(self, xi):
        return xi * self.laplace[None, None, :, None]

    def inverse_laplacian(self, xi):
        return xi * self.inv_laplace[None, None, :, None]

    def hgrad(self, xi):
        scale = 1.0 / (self.a * self.coslat)[None, :, None, None]
        gx = self.fourier_inv(self._syn(self.pnm, self._imk(xi))) * scale
        gy = self.fourier_inv(self._syn(self.hnm, xi)) * scale
        return gx, gy

    def uv_to_vordiv(self, u, v):
        c = self.coslat[None, :, None, None]
        um = self.fourier(u * c)
        vm = self.fourier(v * c)
        vor = self._ana(self.p_gwdel, self._imk(vm)) + self._ana(self.h_gwdel, um)
        div = self._ana(self.p_gwdel, self._imk(um)) - self._ana(self.h_gwdel, vm)
        return vor, div

    def neg_divergence(self, fx, fy):
        return -self.uv_to_vordiv(fx, fy)[1]

    def psichi_to_uv(self, psi, chi):
        um = self._syn(self.pnm, self._imk(chi)) - self._syn(self.hnm, psi)
        vm = self._syn(self.pnm, self._imk(psi)) + self._syn(self.hnm, chi)
        scale = 1.0 / (self.a * self.coslat)[None, :, None, None]
        return self.fourier_inv(um) * scale, self.fourier_inv(vm) * scale

    def vordiv_to_uv(self, vor, div):
        return self.psichi_to_uv(self.inverse_laplacian(vor),
                                 self.inverse_laplacian(div))
