"""gturbo_amd.runs: several run directories driven as the drop-in driver
(gturbo_amd/fortran/gturbo_run) drives each, batched in one ensemble or one
by one, over consecutive segments chained through curframe.dat (the
reference's job arrays, src/reset.sh:12-77, src/jobscript:14-19).  Every
file the driver writes -- diagnostics text, frames, FP64 states, seeds,
curframe -- must come out the same: binary files byte for byte, text lines
character for character, seeds equal, times equal to the last bit the
driver's list-directed output keeps."""
import os
import shutil
import subprocess

import pytest

from gturbo_amd import io, runs
from gturbo_amd.params import InitialSpectrum, Params

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "gturbo_amd", "fortran", "gturbo_run")
N = 128


def _dirs(base):
    from gturbo_amd import ic
    out = []
    for k, alpv in enumerate((2.0, 1.0, 0.5)):
        p = Params(n=N, nu=(N / 256.0) ** -8 * 1e-16, alpnu=4.0, mu=0.1, alpmu=0.0, alpv=alpv,
                   dt=0.25 / N, rko=4, npas=12, iout=5, imix=4, famp=0.1, k1f=7.0, k2f=9.0)
        w0 = ic.physical_field(p, InitialSpectrum(seed=30 + k, p=4.0, k0=N / 16, energy=0.5))
        d = os.path.join(base, f"{k + 1}run")
        io.make_run_dir(d, p, w0, 481929 + 17 * k, ifr=1, t=0.0)
        out.append(d)
    return out


def _copy(dirs, base):
    os.makedirs(base)
    res = []
    for d in dirs:
        t = os.path.join(base, os.path.basename(d))
        shutil.copytree(d, t)
        res.append(t)
    return res


def _files(d):
    out = []
    for r, _, fs in os.walk(d):
        out += [os.path.relpath(os.path.join(r, f), d) for f in fs]
    return sorted(out)


def _same_outputs(a, b):
    fa, fb = _files(a), _files(b)
    assert fa == fb
    for rel in fa:
        pa, pb = os.path.join(a, rel), os.path.join(b, rel)
        name = os.path.basename(rel)
        if rel == "curframe.dat":
            ia, ta = io.read_curframe(pa)
            ib, tb = io.read_curframe(pb)
            assert ia == ib and abs(ta - tb) <= 1e-15 * abs(ta), (ta, tb)
        elif name.startswith("seed."):
            assert io.read_seed(pa) == io.read_seed(pb), rel
        elif name == "params.dat":
            continue
        elif rel.startswith("files"):
            with open(pa) as x, open(pb) as y:
                assert x.read().splitlines() == y.read().splitlines(), rel
        else:
            with open(pa, "rb") as x, open(pb, "rb") as y:
                assert x.read() == y.read(), rel


@pytest.mark.parametrize("exact", [False, True])
def test_run_dirs_write_what_the_driver_writes(tmp_path, exact):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (run __graft_entry__.build())")
    base = _dirs(str(tmp_path / "template"))
    ref = _copy(base, str(tmp_path / "driver"))
    bat = _copy(base, str(tmp_path / "batched"))
    solo = _copy(base, str(tmp_path / "solo"))
    for d in ref:  # two chained driver runs per directory
        for _ in range(2):
            args = [EXE, str(N)] + (["exact"] if exact else [])
            r = subprocess.run(args, cwd=d, capture_output=True, text=True, timeout=120)
            assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    logs = []
    runs.run_dirs(bat, n=N, segments=2, exact=exact, batch=True, log=logs.append)
    print(logs)
    assert all("batched" in line for line in logs)
    runs.run_dirs(solo, n=N, segments=2, exact=exact, batch=False, log=logs.append)
    for a, b, c in zip(ref, bat, solo):
        assert "fields/w.003" in _files(a) and "fields/w.002.002" in _files(a)
        _same_outputs(a, b)
        _same_outputs(a, c)
