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

import argparse
import os
import shutil
import sys

import numpy as np

from . import io
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


def make_runs(params_path: str, n: int, copies: int, prefix: str = "", ic_seed: int = 1,
              forcing_seed: int = 481929, p: float = 4.0, k0: float = 8.0,
              energy: float = 0.5, same_forcing_seed: bool = False) -> list:
    """What src/reset.sh does before submitting its jobs: run directories
    {prefix}1run ... {prefix}{copies}run, each with a copy of params.dat, its
    own initial field (seed ic_seed + k) in fields/w.001, files/seed.001 (the
    forcing seed, forcing_seed + k unless same_forcing_seed) and curframe.dat
    "1 0".  Returns the directories."""
    par = Params.read(params_path, n)
    dirs = []
    for k in range(copies):
        d = f"{prefix}{k + 1}run"
        w = physical_field(par, InitialSpectrum(seed=ic_seed + k, p=p, k0=k0, energy=energy))
        os.makedirs(os.path.join(d, "files"), exist_ok=True)
        os.makedirs(os.path.join(d, "fields"), exist_ok=True)
        dst = os.path.join(d, "params.dat")
        if not (os.path.exists(dst) and os.path.samefile(params_path, dst)):
            shutil.copyfile(params_path, dst)  # the file itself, as cp -r does
        io.write_curframe(os.path.join(d, "curframe.dat"), 1, 0.0)
        io.write_seed(os.path.join(d, "files", "seed.001"),
                      forcing_seed + (0 if same_forcing_seed else k))
        io.write_field(os.path.join(d, "fields", "w.001"), w)
        dirs.append(d)
    return dirs


def main(argv=None):
    ap = argparse.ArgumentParser(
        description="Lay out run directories with random initial fields (src/reset.sh with "
                    "its gen_noise.py).")
    ap.add_argument("params", help="the params.dat to copy into every directory")
    ap.add_argument("--n", type=int, default=8192, help="grid size (the driver's argument)")
    ap.add_argument("--copies", type=int, default=1)
    ap.add_argument("--prefix", default="", help="directory name prefix ({prefix}1run, ...)")
    ap.add_argument("--ic-seed", type=int, default=1, help="seed of the first directory's field")
    ap.add_argument("--forcing-seed", type=int, default=481929, help="src/seed.0")
    ap.add_argument("--same-forcing-seed", action="store_true")
    ap.add_argument("--slope", type=float, default=4.0, help="p of k**p exp(-(k/k0)**2)")
    ap.add_argument("--k0", type=float, default=8.0)
    ap.add_argument("--energy", type=float, default=0.5)
    a = ap.parse_args(argv)
    for d in make_runs(a.params, a.n, a.copies, a.prefix, a.ic_seed, a.forcing_seed, a.slope,
                       a.k0, a.energy, a.same_forcing_seed):
        print(d)
    return 0


if __name__ == "__main__":
    sys.exit(main())
