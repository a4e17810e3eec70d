"""The Fortran driver (gturbo_amd/fortran/main.f90, the drop-in for
src/main.f90) run end to end on the GPU through ISO_C_BINDING, in a run
directory laid out like the reference's: params.dat, curframe.dat,
files/seed.NNN, fields/w.NNN.  Its output files are checked against the
oracle run on the same float32 input field and seed."""
import os
import subprocess

import numpy as np
import pytest

from conftest import rel_l2
from gturbo_amd import io
from gturbo_amd.params import InitialSpectrum, Params

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "gturbo_amd", "fortran", "gturbo_run")


def _setup(tmp_path, p, seed=481929):
    from gturbo_amd import ic
    w0 = ic.physical_field(p, InitialSpectrum(seed=3, p=4.0, k0=p.n / 16, energy=0.5))
    io.make_run_dir(str(tmp_path), p, w0, seed, ifr=1, t=0.0)
    return w0


def _run(tmp_path, n):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (run __graft_entry__.build())")
    r = subprocess.run([EXE, str(n)], cwd=str(tmp_path), capture_output=True, text=True,
                       timeout=120)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stderr
    return r.stdout


@pytest.mark.parametrize("rko", [4, 2])
def test_fortran_driver_matches_oracle(oracle_mod, tmp_path, rko):
    n = 64
    p = Params(n=n, nu=1e-12, alpnu=4.0, mu=0.1, alpmu=0.0, alpv=2.0, dt=2e-3, rko=rko,
               npas=20, iout=10, imix=5, famp=1.0, k1f=3.0, k2f=5.0)
    w0 = _setup(tmp_path, p)
    out = _run(tmp_path, n)
    assert "Mean time per iteration" in out

    o = oracle_mod.Oracle(p)
    z = io.field_to_ref(w0)
    diag = []
    seed = 481929
    o.fft_dir(z)
    o.trunc(z)
    snap = None
    for ipas in range(1, p.npas + 1):
        (o.step4 if rko == 4 else o.step)(z)
        if ipas % p.imix == 0:
            znl = z.copy()
            o.nlt(znl)
            sp, sz = o.spectrum(z)
            fle, flz = o.fluxes(z, znl)
            diag.append((ipas, sp, sz, fle, flz, o.glob(z)))
        if ipas % p.iout == 0 and ipas != p.npas:
            snap = z.copy()
            o.fft_inv(snap)
        seed = o.forcing(z, seed)
    o.fft_inv(z)

    # final field, seed and frame counter (src/main.f90:140-154,182)
    w_final = io.read_field(os.path.join(tmp_path, "fields", "w.002"), n)
    expect = z[:, :n].astype(np.float32)
    assert np.abs(w_final - expect).max() <= 2.5e-7 * np.abs(expect).max()
    assert io.read_seed(os.path.join(tmp_path, "files", "seed.002")) == seed
    ifr, t = io.read_curframe(os.path.join(tmp_path, "curframe.dat"))
    assert ifr == 2 and t == pytest.approx(p.npas * p.dt)
    # mid-run snapshot w.001.001 (src/main.f90:119-133)
    w_snap = io.read_field(os.path.join(tmp_path, "fields", "w.001.001"), n)
    assert np.abs(w_snap - snap[:, :n].astype(np.float32)).max() <= 2.5e-7 * np.abs(snap).max()

    # diagnostics text files, full precision columns
    glob = np.loadtxt(os.path.join(tmp_path, "files", "global.001"))
    spec = np.loadtxt(os.path.join(tmp_path, "files", "spectra.001")).reshape(-1, n // 2, 3)
    flux = np.loadtxt(os.path.join(tmp_path, "files", "fluxes.001")).reshape(-1, n // 2, 3)
    assert glob.shape == (len(diag), 9) and spec.shape[0] == flux.shape[0] == len(diag)
    for d, (ipas, sp, sz, fle, flz, g) in enumerate(diag):
        assert glob[d, 0] == pytest.approx(ipas * p.dt)
        assert np.allclose(glob[d, [2, 1, 3, 4, 6, 7]], g, rtol=1e-9, atol=0)
        # inpz, inpe of the library's iniforcing (src/iniforcing.f90:27-28,
        # src/diagnostics.f90:166)
        assert glob[d, 5] == pytest.approx(o.inpz, rel=1e-14)
        assert glob[d, 8] == pytest.approx(o.inpe, rel=1e-14)
        assert rel_l2(spec[d, :, 1], sp) < 1e-9 and rel_l2(spec[d, :, 2], sz) < 1e-9
        assert np.abs(flux[d, :, 2] - flz).max() <= 1e-9 * np.abs(flz).max()
        assert np.abs(flux[d, :, 1] - fle).max() <= 1e-9 * np.abs(fle).max()


def test_fortran_driver_resume_chain(tmp_path):
    """Two chained 10-step segments resume from the frame the first one wrote
    (curframe.dat, files/seed.002, fields/w.002), as the SLURM chain of
    src/jobscript:14 does; the float32 restart makes the result close to, not
    identical with, one 20-step segment."""
    n = 64
    base = Params(n=n, nu=1e-12, alpnu=4.0, mu=0.1, alpmu=0.0, alpv=2.0, dt=2e-3, rko=4,
                  npas=20, iout=10**6, imix=10**6, famp=1.0, k1f=3.0, k2f=5.0)
    one = tmp_path / "one"
    two = tmp_path / "two"
    _setup(one, base)
    _run(one, n)
    _setup(two, base.replace(npas=10))
    _run(two, n)
    _run(two, n)
    assert io.read_curframe(str(two / "curframe.dat"))[0] == 3
    a = io.read_field(str(one / "fields" / "w.002"), n)
    b = io.read_field(str(two / "fields" / "w.003"), n)
    assert io.read_seed(str(one / "files" / "seed.002")) == io.read_seed(str(two / "files" / "seed.003"))
    assert rel_l2(b, a) < 1e-5


def test_fortran_initial_condition_inversion(tmp_path):
    """gturbo_invert (Fortran, through adjoint_forward/adjoint_backward)
    lowers the misfit of a perturbed first guess to an observation made by a
    forward run from the true initial field."""
    from gturbo_amd.adjoint import Adjoint
    n = 64
    p = Params(n=n, nu=1e-12, alpnu=4.0, mu=0.1, alpmu=0.0, alpv=1.0, dt=2e-3, rko=4,
               npas=20, iout=10**6, imix=10**6, famp=0.0, k1f=3.0, k2f=5.0)
    from gturbo_amd import ic
    truth = ic.physical_field(p, InitialSpectrum(seed=21, p=2.0, k0=4.0, energy=0.5))
    other = ic.physical_field(p, InitialSpectrum(seed=22, p=2.0, k0=4.0, energy=0.5))
    guess = (truth + 0.2 * other).astype(np.float32)
    with Adjoint(p, p.npas, kchk=5) as adj:
        obs, _ = adj.forward(truth.astype(np.float64), 481929)
    io.make_run_dir(str(tmp_path), p, guess, 481929, ifr=1, t=0.0)
    io.write_field(str(tmp_path / "fields" / "obs.001"), obs.astype(np.float32))
    exe = os.path.join(ROOT, "gturbo_amd", "fortran", "gturbo_invert")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (run __graft_entry__.build())")
    r = subprocess.run([exe, str(n), "6"], cwd=str(tmp_path), capture_output=True, text=True,
                       timeout=120)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stderr
    losses = [float(line.split()[3]) for line in r.stdout.splitlines()
              if line.strip().startswith("iteration")]
    assert len(losses) >= 4
    assert all(b < a for a, b in zip(losses, losses[1:]))
    assert losses[-1] < 0.5 * losses[0]
    assert os.path.exists(tmp_path / "fields" / "winv.001")


def test_fortran_fp64_checkpoint_resume_is_exact(tmp_path):
    """Two chained 10-step segments resumed from the FP64 state fields/z.002
    ("exact") end bit-identical to one 20-step segment: same FP64 state,
    same float32 frame, same seed."""
    n = 64
    base = Params(n=n, nu=1e-12, alpnu=4.0, mu=0.1, alpmu=0.0, alpv=1.0, dt=2e-3, rko=4,
                  npas=20, iout=10**6, imix=10**6, famp=1.0, k1f=3.0, k2f=5.0)
    one = tmp_path / "one"
    two = tmp_path / "two"
    _setup(one, base)
    _run(one, n)
    _setup(two, base.replace(npas=10))
    _run(two, n)
    r = subprocess.run([EXE, str(n), "exact"], cwd=str(two), capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Resuming from the FP64 state" in r.stdout
    za = np.fromfile(one / "fields" / "z.002", dtype=np.uint8)
    zb = np.fromfile(two / "fields" / "z.003", dtype=np.uint8)
    assert za.size == zb.size and np.array_equal(za, zb)
    wa = io.read_field(str(one / "fields" / "w.002"), n)
    wb = io.read_field(str(two / "fields" / "w.003"), n)
    assert np.array_equal(wa, wb)
    assert io.read_seed(str(one / "files" / "seed.002")) == io.read_seed(str(two / "files" / "seed.003"))
