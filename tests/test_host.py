"""Host-side logic that needs no GPU: params.dat parsing, the field/seed/
curframe file formats, the initial-condition generator, and the C ABI
library's exported symbols (no compute calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

from gturbo_amd import io, params
from gturbo_amd.params import Params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def test_params_dat_of_the_reference():
    p = Params.read(os.path.join(GOLD, "params.dat"), n=8192)
    assert p.xlx == pytest.approx(2 * np.pi) and p.xly == p.xlx
    assert (p.nu, p.alpnu, p.mu, p.alpmu, p.alpv) == (1e-5, 1.0, 0.0, 0.0, 2.0)
    assert (p.dt, p.rko) == (2e-5, 4)
    assert (p.npas, p.iout, p.imix) == (600000, 600000, 5000)
    assert p.rtrunc == pytest.approx(2 / 3) and p.truncate == 1
    assert (p.famp, p.k1f, p.k2f) == (1.0, 63.0, 65.0)


def test_params_round_trip_and_short_lines(tmp_path):
    p = params.workloads()["sqg1024"].params
    f = tmp_path / "params.dat"
    p.write(str(f))
    q = Params.read(str(f), n=p.n)
    assert q == p
    # list-directed input continues onto the next line when one runs short
    txt = "6.28d0\n6.28d0\n1d-5 1 0 0 2\n1d-3 2\n10 10 10\n0.5 1\n1 2 3\n"
    r = Params.parse(txt, 16)
    assert (r.xlx, r.xly, r.alpv, r.rko, r.famp) == (6.28, 6.28, 2.0, 2, 1.0)
    with pytest.raises(ValueError):
        Params.parse("1 2\n3\n", 16)


def test_seed_and_curframe_files(tmp_path):
    assert io.read_seed(os.path.join(GOLD, "seed.0")) == 481929
    assert io.read_curframe(os.path.join(GOLD, "startframe.dat")) == (0, 0.0)
    io.write_seed(str(tmp_path / "s"), 1074099382)
    assert io.read_seed(str(tmp_path / "s")) == 1074099382
    io.write_curframe(str(tmp_path / "c"), 3, 0.125)
    assert io.read_curframe(str(tmp_path / "c")) == (3, 0.125)


def test_field_record_format(tmp_path):
    n = 16
    w = np.arange(n * n, dtype=np.float32).reshape(n, n)
    f = tmp_path / "w.001"
    io.write_field(str(f), w)
    raw = f.read_bytes()
    # Fortran unformatted sequential: 4-byte length, payload, 4-byte length
    assert len(raw) == 4 * n * n + 8
    assert int.from_bytes(raw[:4], "little") == 4 * n * n == int.from_bytes(raw[-4:], "little")
    # w(i,j) with i (x) fastest: the first values are row y=0
    assert np.frombuffer(raw[4:12], "<f4").tolist() == [0.0, 1.0]
    assert np.array_equal(io.read_field(str(f), n), w)
    z = io.field_to_ref(w)
    assert z.shape == (n, n + 2) and np.all(z[:, n:] == 0)
    assert np.array_equal(io.ref_to_field(z), w)


def test_truncated_record_is_an_error(tmp_path):
    f = tmp_path / "w.bad"
    f.write_bytes((64).to_bytes(4, "little") + b"\0" * 10)
    with pytest.raises(ValueError):
        io.read_field(str(f), 4)


def test_initial_condition_properties():
    from gturbo_amd import ic
    w = params.workloads()["euler256"]
    q = ic.spectral_field(w.params, w.ic)
    n = w.params.n
    # Hermitian on kx = 0, zero mean, truncated, energy as requested
    assert q[0, 0] == 0
    assert np.allclose(q[n - 5, 0], np.conj(q[5, 0]))
    kx = np.arange(n // 2 + 1)
    ky = np.fft.fftfreq(n, 1.0 / n)
    k2 = np.add.outer(ky ** 2, kx ** 2)
    assert np.all(q[k2 > (n / 3) ** 2 + 1] == 0)
    fac = np.ones(q.shape); fac[:, 0] = fac[:, -1] = 0.5
    kk = np.where(k2 > 0, k2, 1.0)
    assert np.sum(fac * np.abs(q) ** 2 / kk) == pytest.approx(0.5, rel=1e-12)
    f = ic.physical_field(w.params, w.ic)
    assert f.dtype == np.float32 and f.shape == (n, n)
    # deterministic in the seed
    assert np.array_equal(f, ic.physical_field(w.params, w.ic))


def test_workload_table():
    w = params.workloads()
    assert w["euler4096_rk4"].params.n == 4096 and w["euler4096_rk4"].params.rko == 4
    assert w["euler4096_rk2"].params.rko == 2
    m = params.ensemble_members()
    assert len(m) == 64 and sum(x.params.n == 256 for x in m) == 32
    assert [x.params.alpv for x in m[:3]] == [0.5, 1.0, 2.0]
    assert params.kept_half_width(w["euler256"].params) == 85


def _header_symbols():
    txt = open(os.path.join(ROOT, "include", "gturbo.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gt_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from gturbo_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    L = ctypes.CDLL(_lib.LIB_PATH)
    syms = _header_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing
    # the Python binding declares exactly the header's entry points
    assert sorted(_lib.SIGNATURES) == syms


def test_library_without_gpu_fails_cleanly():
    """On a host with no GPU gt_create reports GT_ERR_NOGPU; nothing falls back
    to a CPU path."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from gturbo_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    from gturbo_amd.solver import Solver
    with pytest.raises(_lib.GturboError, match="error 5"):
        Solver(params.workloads()["euler256"].params)


def test_make_runs_lays_out_run_directories(tmp_path, monkeypatch):
    """gturbo_amd.ic.make_runs (src/reset.sh with its gen_noise.py): every
    directory gets the params.dat, its own float32 field at the requested
    energy, its forcing seed and curframe "1 0"."""
    from gturbo_amd import ic, io
    from gturbo_amd.params import Params
    monkeypatch.chdir(tmp_path)
    n = 32
    Params(n=n).write("params.dat")
    dirs = ic.make_runs("params.dat", n, 3, prefix="e", ic_seed=5, forcing_seed=100)
    assert dirs == ["e1run", "e2run", "e3run"]
    fields = []
    for k, d in enumerate(dirs):
        assert open(os.path.join(d, "params.dat")).read() == open("params.dat").read()
        assert io.read_curframe(os.path.join(d, "curframe.dat")) == (1, 0.0)
        assert io.read_seed(os.path.join(d, "files", "seed.001")) == 100 + k
        w = io.read_field(os.path.join(d, "fields", "w.001"), n)
        assert w.dtype == np.float32 and w.shape == (n, n)
        # energy of the float32 field: sum fac |q|^2 / k^alpha with q = fft/N^2
        p = Params.read("params.dat", n)
        q = np.fft.rfft2(w.astype(np.float64)) / n ** 2
        kx = np.arange(n // 2 + 1)
        ky = np.fft.fftfreq(n, 1.0 / n)
        k2 = kx[None, :] ** 2 + ky[:, None] ** 2
        fac = np.ones(k2.shape)
        fac[:, 0] = fac[:, -1] = 0.5
        with np.errstate(divide="ignore"):
            e = np.sum(np.where(k2 > 0, fac * np.abs(q) ** 2 / k2 ** (p.alpv / 2), 0.0))
        assert abs(e - 0.5) < 1e-4
        fields.append(w)
    assert not np.array_equal(fields[0], fields[1])


def test_field_arrays_are_checked_before_their_address_is_taken():
    """The library reads and writes a field as row-major (N, N+2) float64
    memory; an array that is not exactly that must be refused, not misread."""
    from gturbo_amd.solver import _field_addr
    n = 16
    good = np.zeros((n, n + 2))
    assert _field_addr(good, n) == good.ctypes.data
    assert _field_addr(good.reshape(-1), n) == good.ctypes.data  # same memory, flat
    assert _field_addr(12345, n) == 12345                          # raw address as is
    bad = [np.zeros((n, n + 2), np.float32),                       # wrong dtype
           np.asfortranarray(np.zeros((n, n + 2))),                # column-major
           np.zeros((n, 2 * (n + 2)))[:, ::2],                     # strided view
           np.zeros((n + 2, n + 2)),                               # wrong size
           np.zeros((n, n + 2), np.complex128)]                    # same bytes x 2
    for a in bad:
        with pytest.raises(TypeError):
            _field_addr(a, n)
