"""Initial conditions (the reference's generator, gen_noise.py called by
src/reset.sh:36, is not part of the reference).

Random phases from numpy's PCG64 (``numpy.random.default_rng(seed)``), shell
amplitudes |q(k)|^2 proportional to k**alpv * S(k) / k with
S(k) = k**p * exp(-(k/k0)**2) so that the energy spectrum follows S, Hermitian
symmetry on kx = 0, zero mean, the 2/3 truncation of src/trunc.f90, scaled to
total energy E = sum fac |q|^2 / k**alpv (the ee of src/diagnostics.f90:150-156),
then transformed to physical space and rounded to float32 -- the format the
reference reads (src/read.f90:10-19).
"""
from __future__ import annotations

import numpy as np

from .params import InitialSpectrum, Params


def spectral_field(p: Params, ic: InitialSpectrum) -> np.ndarray:
    """Half-spectrum q(ky, kx) as complex128 array (N, N/2+1)."""
    n = p.n
    dk = 2.0 * 3.14159265358979323846 / p.xlx
    kx = np.arange(n // 2 + 1) * dk
    kyi = np.fft.fftfreq(n, 1.0 / n)
    ky = kyi * (2.0 * 3.14159265358979323846 / p.xly)
    KX, KY = np.meshgrid(kx, ky)
    k2 = KX ** 2 + KY ** 2
    k = np.sqrt(k2)
    rng = np.random.default_rng(ic.seed)
    phase = rng.uniform(0.0, 2.0 * np.pi, size=k.shape)
    with np.errstate(divide="ignore", invalid="ignore"):
        shape = k ** ic.p * np.exp(-(k / ic.k0) ** 2)
        amp2 = np.where(k > 0, k ** p.alpv * shape / k, 0.0)
    q = np.sqrt(amp2) * np.exp(1j * phase)
    # kx = 0: q(0, -ky) = conj q(0, ky)
    half = n // 2
    q[half + 1:, 0] = np.conj(q[1:half, 0][::-1])
    q[0, 0] = 0.0
    q[half, :] = 0.0
    q[:, -1] = 0.0
    # src/trunc.f90:13-15 (applied whatever the truncate flag, as the
    # reference always starts from a truncated state when truncate = 1)
    if p.truncate == 1:
        kxm = (n // 2 * dk) ** 2
        kym = (n // 2 * (2.0 * 3.14159265358979323846 / p.xly)) ** 2
        cut = (KY ** 2 / kym + KX ** 2 / kxm) > p.rtrunc ** 2
        q[cut] = 0.0
    fac = np.ones(k.shape)
    fac[:, 0] = 0.5
    fac[:, -1] = 0.5
    with np.errstate(divide="ignore", invalid="ignore"):
        ka = np.where(k2 > 0, k2 ** (p.alpv / 2.0), 1.0)
    e = float(np.sum(fac * np.abs(q) ** 2 / ka))
    if e > 0:
        q *= np.sqrt(ic.energy / e)
    return q


def physical_field(p: Params, ic: InitialSpectrum) -> np.ndarray:
    """float32 physical field w[y][x] (Fortran w(x, y))."""
    q = spectral_field(p, ic)
    w = np.fft.irfft2(q, s=(p.n, p.n)) * (p.n * p.n)
    return w.astype(np.float32)
