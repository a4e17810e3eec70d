"""Run parameters: the params.dat file of the reference and the named
workloads this package is measured on.

params.dat holds six list-directed records (src/inieuler.f90:13-20):

    xlx xly
    nu alpnu mu alpmu alpv
    dt rko
    npas iout imix
    rtrunc truncate
    famp k1f k2f

The grid size N is not in the file: the reference fixes it at compile time
(src/modules.f90:4); here it is a field of :class:`Params`.
"""
from __future__ import annotations

import dataclasses
import re

import numpy as np

TWO_PI = 6.2831853071795864769  # the literal of src/params.dat:1


@dataclasses.dataclass
class Params:
    n: int = 256
    xlx: float = TWO_PI
    xly: float = TWO_PI
    nu: float = 1.0e-5
    alpnu: float = 1.0
    mu: float = 0.0
    alpmu: float = 0.0
    alpv: float = 2.0
    dt: float = 2.0e-5
    rko: int = 4
    npas: int = 600000
    iout: int = 600000
    imix: int = 5000
    rtrunc: float = 0.6666666666666666667
    truncate: int = 1
    famp: float = 1.0
    k1f: float = 63.0
    k2f: float = 65.0

    def as_array(self) -> np.ndarray:
        """Values in params.dat order (the layout the C entry points and the
        oracle take)."""
        return np.array([self.xlx, self.xly, self.nu, self.alpnu, self.mu,
                         self.alpmu, self.alpv, self.dt, self.rko, self.npas,
                         self.iout, self.imix, self.rtrunc, self.truncate,
                         self.famp, self.k1f, self.k2f], dtype=np.float64)

    def replace(self, **kw) -> "Params":
        return dataclasses.replace(self, **kw)

    # params.dat --------------------------------------------------------
    def to_text(self) -> str:
        def d(x):
            return repr(float(x)).replace("e", "d") if "e" in repr(float(x)) else repr(float(x)) + "d0"
        lines = [
            f"{d(self.xlx)}\t{d(self.xly)}",
            f"{d(self.nu)}\t{d(self.alpnu)}\t{d(self.mu)}\t{d(self.alpmu)}\t{d(self.alpv)}",
            f"{d(self.dt)}\t{int(self.rko)}",
            f"{int(self.npas)}\t{int(self.iout)}\t{int(self.imix)}",
            f"{d(self.rtrunc)}\t{int(self.truncate)}",
            f"{d(self.famp)}\t{d(self.k1f)}\t{d(self.k2f)}",
            "",
            "xlx   yly",
            "nu    alpnu\tmu\talpmu\talpv",
            "dt\t\trko",
            "npas\tiout  imix",
            "rtrunc\ttruncate",
            "famp  fk1\tfk2",
        ]
        return "\n".join(lines) + "\n"

    def write(self, path: str) -> None:
        with open(path, "w") as f:
            f.write(self.to_text())

    @classmethod
    def read(cls, path: str, n: int) -> "Params":
        with open(path) as f:
            return cls.parse(f.read(), n)

    @classmethod
    def parse(cls, text: str, n: int) -> "Params":
        """List-directed read of the six records (src/inieuler.f90:14-19).
        Each READ takes the values it needs, continuing onto following lines
        if a line runs short, and discards the rest of its last line."""
        counts = [2, 5, 2, 3, 2, 3]
        lines = text.splitlines()
        vals = []
        li = 0
        for cnt in counts:
            got = []
            while len(got) < cnt:
                if li >= len(lines):
                    raise ValueError("params.dat: end of file inside a record")
                toks = [t for t in re.split(r"[\s,]+", lines[li].strip()) if t]
                li += 1
                for t in toks:
                    if len(got) < cnt:
                        got.append(float(t.replace("d", "e").replace("D", "e")))
            vals.append(got)
        (xlx, xly), (nu, alpnu, mu, alpmu, alpv), (dt, rko), (npas, iout, imix), \
            (rtrunc, truncate), (famp, k1f, k2f) = vals
        return cls(n=int(n), xlx=xlx, xly=xly, nu=nu, alpnu=alpnu, mu=mu,
                   alpmu=alpmu, alpv=alpv, dt=dt, rko=int(rko), npas=int(npas),
                   iout=int(iout), imix=int(imix), rtrunc=rtrunc,
                   truncate=int(truncate), famp=famp, k1f=k1f, k2f=k2f)


@dataclasses.dataclass
class InitialSpectrum:
    """Initial-condition recipe: random phases from ``seed`` and a shell
    spectrum k**p * exp(-(k/k0)**2), scaled to total energy ``energy``."""
    seed: int = 0
    p: float = 4.0
    k0: float = 8.0
    energy: float = 0.5


@dataclasses.dataclass
class Workload:
    name: str
    params: Params
    ic: InitialSpectrum
    forcing_seed: int = 481929  # src/seed.0


def _ring(n):
    kf = n // 16
    return float(kf - 1), float(kf + 1)


def workloads() -> dict:
    """The measured workloads.  'Injection rate 1.0' maps to famp = 1.0, the
    reference's forcing amplitude (src/params.dat:6); the enstrophy input is
    then nf*famp**2 (src/iniforcing.f90:27)."""
    w = {}
    w["euler256"] = Workload(
        "euler256: N=256 alpha=2 RK4 dt=1e-3 1000 steps, nu=1e-16 k^8, mu=0.1, ring (15,17]",
        Params(n=256, nu=1e-16, alpnu=4.0, mu=0.1, alpmu=0.0, alpv=2.0, dt=1e-3,
               rko=4, npas=1000, iout=10**9, imix=10**9, famp=1.0, k1f=15.0, k2f=17.0),
        InitialSpectrum(seed=0, p=4.0, k0=8.0, energy=0.5))
    w["sqg1024"] = Workload(
        "sqg1024: N=1024 alpha=1 RK4 dt=2e-4 2000 steps, nu=1e-22 k^8, mu=0.05, ring (31,33]",
        Params(n=1024, nu=1e-22, alpnu=4.0, mu=0.05, alpmu=0.0, alpv=1.0, dt=2e-4,
               rko=4, npas=2000, iout=10**9, imix=100, famp=1.0, k1f=31.0, k2f=33.0),
        InitialSpectrum(seed=1, p=2.0, k0=16.0, energy=0.5))
    for rko in (4, 2):
        w[f"euler4096_rk{rko}"] = Workload(
            f"euler4096_rk{rko}: N=4096 alpha=2 RK{rko} dt=5e-5 500 steps, nu=1e-28 k^8, mu=0.02, ring (63,65]",
            Params(n=4096, nu=1e-28, alpnu=4.0, mu=0.02, alpmu=0.0, alpv=2.0, dt=5e-5,
                   rko=rko, npas=500, iout=10**9, imix=10**9, famp=1.0, k1f=63.0, k2f=65.0),
            InitialSpectrum(seed=2, p=4.0, k0=32.0, energy=0.5))
    w["adjoint512"] = Workload(
        "adjoint512: N=512 alpha=1 RK4 dt=5e-4 200 steps, forcing off",
        Params(n=512, nu=1e-16 * (512 / 256) ** -8, alpnu=4.0, mu=0.1, alpmu=0.0,
               alpv=1.0, dt=5e-4, rko=4, npas=200, iout=10**9, imix=10**9,
               famp=0.0, k1f=31.0, k2f=33.0),
        InitialSpectrum(seed=200, p=2.0, k0=16.0, energy=0.5))
    return w


def ensemble_members() -> list:
    """64 members: 32 at 256^2, 24 at 512^2, 8 at 1024^2.  Member m has
    seed 100+m, alpha cycling through {0.5, 1, 2}, ring (N/16-1, N/16+1],
    nu = (N/256)**-8 * 1e-16 (k^8), mu = 0.1, dt = 0.25/N, 1000 steps.
    The forcing amplitude is not part of the workload's definition: famp =
    0.1 keeps every member stable over the 1000 steps (at famp = 1 the
    alpha = 1/2 members blow up near step 500 at this dt, in the oracle as
    on the GPU)."""
    out = []
    sizes = [256] * 32 + [512] * 24 + [1024] * 8
    for m, n in enumerate(sizes):
        k1, k2 = _ring(n)
        alpv = (0.5, 1.0, 2.0)[m % 3]
        p = Params(n=n, nu=(n / 256.0) ** -8 * 1e-16, alpnu=4.0, mu=0.1, alpmu=0.0,
                   alpv=alpv, dt=0.25 / n, rko=4, npas=1000, iout=10**9,
                   imix=10**9, famp=0.1, k1f=k1, k2f=k2)
        ic = InitialSpectrum(seed=100 + m, p=4.0 if alpv == 2.0 else 2.0,
                             k0=n / 32.0, energy=0.5)
        out.append(Workload(f"member{m}: N={n} alpha={alpv}", p, ic, 481929 + m))
    return out


def kept_half_width(p: Params) -> int:
    """Largest |k| index kept by the 2/3 rule along one axis (0 based):
    the biggest i with (i*dk)^2/kmax^2 <= rtrunc^2 (src/trunc.f90:13-15)."""
    if p.truncate != 1:
        return p.n // 2
    dk = 2.0 * 3.14159265358979323846 / p.xlx
    kmax2 = ((p.n // 2) * dk) ** 2
    r2 = p.rtrunc * p.rtrunc
    m = 0
    for i in range(p.n // 2 + 1):
        if ((i * dk) * (i * dk)) / kmax2 <= r2:
            m = i
    return m


__all__ = ["Params", "InitialSpectrum", "Workload", "workloads",
           "ensemble_members", "kept_half_width", "TWO_PI"]
