"""Shared inputs of the tests."""
import numpy as np

from gturbo_amd.params import Params, TWO_PI


def params_for(n, alpv=2.0, rko=4, truncate=1, nu=None, mu=0.1, alpnu=4.0, alpmu=0.0,
               dt=None, famp=1.0, ring=None, rtrunc=0.6666666666666666667, xlx=TWO_PI,
               xly=None):
    if nu is None:
        nu = (n / 256.0) ** -8 * 1e-16 if alpnu == 4.0 else 1e-3
    if dt is None:
        dt = 0.25 / n
    if ring is None:
        kf = max(n // 16, 2)
        ring = (kf - 1.0, kf + 1.0)
    return Params(n=n, xlx=xlx, xly=xlx if xly is None else xly, nu=nu, alpnu=alpnu, mu=mu, alpmu=alpmu, alpv=alpv,
                  dt=dt, rko=rko, npas=10, iout=10**9, imix=10**9, rtrunc=rtrunc,
                  truncate=truncate, famp=famp, k1f=ring[0], k2f=ring[1])


def random_spectral(n, seed, truncated=True, p=None, hermitian=True):
    """Reference-layout (N, N+2) float64 array holding a random half-spectrum
    (complex (N, N/2+1) viewed as interleaved doubles)."""
    rng = np.random.default_rng(seed)
    nh = n // 2 + 1
    q = rng.standard_normal((n, nh)) + 1j * rng.standard_normal((n, nh))
    kyi = np.fft.fftfreq(n, 1.0 / n)
    kx = np.arange(nh)
    KX, KY = np.meshgrid(kx, kyi)
    k = np.sqrt(KX ** 2 + KY ** 2)
    # amplitude ~ 16/N keeps the field O(1) and the step stable at dt = 0.25/N
    q *= np.exp(-(k / (n / 6.0)) ** 2) * (16.0 / n)
    if hermitian:
        half = n // 2
        q[half + 1:, 0] = np.conj(q[1:half, 0][::-1])
        q[0, 0] = q[0, 0].real
        q[half, 0] = q[half, 0].real
        q[:, -1] = 0.0
        q[half, :] = 0.0
    if truncated:
        cut = (KY ** 2 / (n / 2) ** 2 + KX ** 2 / (n / 2) ** 2) > (2.0 / 3.0) ** 2
        q[cut] = 0.0
        q[0, 0] = 0.0
    return np.ascontiguousarray(q).view(np.float64).reshape(n, n + 2).copy()


def random_physical(n, seed):
    rng = np.random.default_rng(seed)
    z = np.zeros((n, n + 2))
    z[:, :n] = rng.standard_normal((n, n))
    return z


def as_complex(z):
    n = z.shape[0]
    return np.ascontiguousarray(z).view(np.complex128).reshape(n, n // 2 + 1)
