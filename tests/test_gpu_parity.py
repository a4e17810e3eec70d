"""Parity of the HIP path (through the C ABI) with the CPU oracle on the same
seeded inputs.  Bars: integer/index results bit-exact; FP64 fields within the
tolerances written in each test (1e-10 relative L2 for the vorticity after a
run, 1e-9 for spectra and fluxes, tighter for single operators)."""
import numpy as np
import pytest

from conftest import rel_l2
from helpers import params_for, random_physical, random_spectral, as_complex

pytestmark = pytest.mark.gpu

SIZES = [16, 32, 64, 128, 256, 512]


def _pair(oracle_mod, p):
    from gturbo_amd.solver import Solver
    return Solver(p), oracle_mod.Oracle(p)


def _params_fft(n):
    # at 8192 the default ring (N/16 +- 1) holds more than NFORC modes
    return params_for(n, ring=(63.0, 65.0) if n >= 8192 else None)


def _present(s, *arrays):
    for a in arrays:
        s.enter_data(a)


@pytest.mark.parametrize("n", SIZES + [1024, 4096, 8192])
def test_fft_dir_matches_oracle(oracle_mod, n):
    p = _params_fft(n)
    s, o = _pair(oracle_mod, p)
    z = random_physical(n, 10 + n)
    zg = z.copy()
    s.enter_data(zg)
    s.fft_dir(zg)
    s.update_host(zg)
    o.fft_dir(z)
    assert rel_l2(zg, z) < 1e-14


@pytest.mark.parametrize("n", SIZES + [1024, 4096, 8192])
def test_fft_inv_matches_oracle(oracle_mod, n):
    p = _params_fft(n)
    s, o = _pair(oracle_mod, p)
    z = random_spectral(n, 20 + n, truncated=False)
    zg = z.copy()
    s.enter_data(zg)
    s.fft_inv(zg)
    s.update_host(zg)
    o.fft_inv(z)
    assert rel_l2(zg, z) < 1e-14


@pytest.mark.parametrize("n", [16, 64, 256])
def test_trunc_and_produ(oracle_mod, n):
    p = params_for(n)
    s, o = _pair(oracle_mod, p)
    z = random_spectral(n, 3, truncated=False)
    zg = z.copy()
    s.enter_data(zg)
    s.trunc(zg)
    s.update_host(zg)
    o.trunc(z)
    assert np.array_equal(zg, z)
    q, u, v = (random_physical(n, k) for k in (1, 2, 3))
    qg, ug, vg = q.copy(), u.copy(), v.copy()
    _present(s, qg, ug, vg)
    s.produ(qg, ug, vg)
    s.update_host(ug)
    s.update_host(vg)
    o.produ(q, u, v)
    assert np.array_equal(ug, u) and np.array_equal(vg, v)


@pytest.mark.parametrize("n", SIZES + [1024])
@pytest.mark.parametrize("alpv", [0.5, 1.0, 2.0])
def test_nlt_matches_oracle(oracle_mod, n, alpv):
    p = params_for(n, alpv=alpv)
    s, o = _pair(oracle_mod, p)
    z = random_spectral(n, 7 * n + int(alpv * 10))
    zg, ug, vg = z.copy(), s.field(), s.field()
    _present(s, zg, ug, vg)
    s.nlt(zg, ug, vg)
    s.update_host(zg)
    o.nlt(z)
    assert rel_l2(zg, z) < 1e-12


@pytest.mark.parametrize("n", [32, 128])
def test_nlt_untruncated_input(oracle_mod, n):
    """A spectrum with modes outside the 2/3 rectangle takes the full-rectangle
    path and still matches."""
    p = params_for(n, alpv=1.0)
    s, o = _pair(oracle_mod, p)
    z = random_spectral(n, 5, truncated=False)
    zg, ug, vg = z.copy(), s.field(), s.field()
    _present(s, zg, ug, vg)
    s.nlt(zg, ug, vg)
    s.update_host(zg)
    o.nlt(z)
    assert rel_l2(zg, z) < 1e-12


@pytest.mark.parametrize("n", [16, 64, 256])
@pytest.mark.parametrize("truncate", [0, 1])
def test_nlt_truncate_flag(oracle_mod, n, truncate):
    p = params_for(n, alpv=2.0, truncate=truncate)
    s, o = _pair(oracle_mod, p)
    z = random_spectral(n, 11, truncated=bool(truncate))
    zg, ug, vg = z.copy(), s.field(), s.field()
    _present(s, zg, ug, vg)
    s.nlt(zg, ug, vg)
    s.update_host(zg)
    o.nlt(z)
    assert rel_l2(zg, z) < 1e-12


@pytest.mark.parametrize("n", [16, 64, 256, 512])
@pytest.mark.parametrize("rko", [2, 4])
def test_step_matches_oracle(oracle_mod, n, rko):
    p = params_for(n, alpv=1.0, rko=rko)
    s, o = _pair(oracle_mod, p)
    z = random_spectral(n, 100 + n)
    zg = z.copy()
    scratch = [s.field() for _ in range(4)]
    _present(s, zg, *scratch)
    for _ in range(3):
        if rko == 4:
            s.step4(zg, *scratch)
            o.step4(z)
        else:
            s.step(zg, scratch[0], scratch[2], scratch[3])
            o.step(z)
    s.update_host(zg)
    assert rel_l2(zg, z) < 1e-12


@pytest.mark.parametrize("n", [32, 128])
def test_split_rk2_matches_oracle(oracle_mod, n):
    """alpnu = 1, alpmu = 0, rko = 2 selects the split integrating factor
    emkx*emky (src/inieuler.f90:24-36, src/rkstep.f90:21-27)."""
    p = params_for(n, alpv=2.0, rko=2, alpnu=1.0, nu=1e-3, alpmu=0.0)
    s, o = _pair(oracle_mod, p)
    assert s.split == 1 and o.split == 1
    z = random_spectral(n, 9)
    zg = z.copy()
    scratch = [s.field() for _ in range(3)]
    _present(s, zg, *scratch)
    for _ in range(3):
        s.step(zg, *scratch)
        o.step(z)
    s.update_host(zg)
    assert rel_l2(zg, z) < 1e-12


@pytest.mark.parametrize("n", [16, 64, 256, 1024])
@pytest.mark.parametrize("dt", [None, 3.1e-4])
def test_forcing_bit_exact(oracle_mod, n, dt):
    """dt = 3.1e-4 makes sqrt(dt) a non-power of two, so a contracted
    multiply-add would show up as a last-bit difference."""
    p = params_for(n, dt=dt)
    s, o = _pair(oracle_mod, p)
    assert s.nf == o.nf and s.nf > 0
    z = random_spectral(n, 4)
    zg = z.copy()
    s.enter_data(zg)
    seed_g = seed_o = 481929
    for _ in range(5):
        seed_g = s.forcing(zg, seed_g)
        seed_o = o.forcing(z, seed_o)
    s.update_host(zg)
    assert seed_g == seed_o
    bad = np.argwhere(zg != z)
    if len(bad):
        zc, gc = as_complex(z), as_complex(zg)
        cb = np.argwhere(gc != zc)[:8]
        print("mismatches (ky_index, kx):", [tuple(x) for x in cb],
              [(zc[tuple(x)], gc[tuple(x)]) for x in cb])
    assert len(bad) == 0


@pytest.mark.parametrize("n,ring,alpv,famp", [
    (256, (15.0, 17.0), 2.0, 1.0), (512, (31.0, 33.0), 1.0, 0.7),
    (1024, (31.0, 33.0), 1.0, 1.0), (4096, (63.0, 65.0), 2.0, 1.0),
    (8192, (63.0, 65.0), 0.5, 1.3)])
def test_iniforcing_tables_match_oracle(oracle_mod, n, ring, alpv, famp):
    """The library's src/iniforcing.f90 set-up: the forced modes in the
    reference's order (ikxf, ikyf, ff), nf, and the injection sums inpz =
    sum famp^2 and inpe = sum famp^2/k^alpha (iniforcing.f90:27-28), which
    feed files/global.NNN and the Zb/Eb line."""
    p = params_for(n, ring=ring, alpv=alpv, famp=famp)
    s, o = _pair(oracle_mod, p)
    assert s.nf == o.nf and s.nf > 0
    ikx, iky, ff = s.forcing_modes()
    oikx, oiky, off = o.forcing_modes()
    assert np.array_equal(ikx, oikx) and np.array_equal(iky, oiky)
    assert np.array_equal(ff, off)
    assert s.inpz == o.inpz and s.inpe == o.inpe  # same sums in the same order
    s.close()


def test_iniforcing_overflow_is_nforc_error():
    """More forced modes than NFORC = 2048 (src/modules.f90:41) stops the
    reference with 'NFORC too small !' (src/iniforcing.f90:22); gt_create
    returns GT_ERR_NFORC with that message and no context."""
    import ctypes as C
    from gturbo_amd import _lib
    L = _lib.load()
    p = params_for(4096, ring=(200.0, 210.0))  # about 6,400 modes in the ring
    arr = np.ascontiguousarray(p.as_array(), np.float64)
    h = C.c_void_p(1)
    rc = L.gt_create(4096, arr.ctypes.data, C.byref(h))
    assert rc == 4  # GT_ERR_NFORC
    assert L.gt_last_error().decode() == "NFORC too small !"
    assert not h.value
    p = params_for(4096, ring=(63.0, 65.0))  # the library is usable afterwards
    arr = np.ascontiguousarray(p.as_array(), np.float64)
    assert L.gt_create(4096, arr.ctypes.data, C.byref(h)) == 0
    assert L.gt_destroy(h) == 0


@pytest.mark.parametrize("n", [16, 256, 4096])
def test_shell_index_bit_exact(oracle_mod, n):
    p = params_for(n)
    s, o = _pair(oracle_mod, p)
    assert np.array_equal(s.shell_index(), o.shell_index())


@pytest.mark.parametrize("n", [16, 64, 256, 1024])
@pytest.mark.parametrize("alpv", [1.0, 2.0])
def test_diagnostics_match_oracle(oracle_mod, n, alpv):
    p = params_for(n, alpv=alpv)
    s, o = _pair(oracle_mod, p)
    z = random_spectral(n, 55 + n)
    znl = z.copy()
    o.nlt(znl)
    zg, ng = z.copy(), znl.copy()
    _present(s, zg, ng)
    sp, sz = s.spectrum(zg)
    osp, osz = o.spectrum(z)
    assert rel_l2(sp, osp) < 1e-12 and rel_l2(sz, osz) < 1e-12
    fle, flz = s.fluxes(zg, ng)
    ofle, oflz = o.fluxes(z, znl)
    scale_e = np.abs(ofle).max() + 1e-300
    scale_z = np.abs(oflz).max() + 1e-300
    assert np.abs(fle - ofle).max() / scale_e < 1e-9
    assert np.abs(flz - oflz).max() / scale_z < 1e-9
    g, og = s.glob(zg), o.glob(z)
    assert np.allclose(g, og, rtol=1e-12, atol=0)


@pytest.mark.parametrize("n,xlx,xly,rtrunc,alpmu,alpnu,alpv,steps", [
    (128, 4 * np.pi, 4 * np.pi, 2 / 3, 0.0, 4.0, 2.0, 3),   # larger box: dk = 1/2
    (128, 2 * np.pi, np.pi, 2 / 3, 0.0, 4.0, 1.0, 3),       # anisotropic box
    (256, 2 * np.pi, 2 * np.pi, 0.5, -1.0, 2.0, 2.0, 3),    # smaller radius, drag mu k^-2
    (256, 2 * np.pi, 2 * np.pi, 0.9, 0.0, 4.0, 0.5, 3),     # past 2/3: full rectangle
    (4096, 2 * np.pi, 2 * np.pi, 0.9, 0.0, 4.0, 2.0, 1),    # the same at 4096
    (4096, 2 * np.pi, 2 * np.pi, 0.5, 0.0, 4.0, 2.0, 1),    # in-place y path, smaller rect
])
def test_other_parameters_steps_and_diagnostics(oracle_mod, n, xlx, xly, rtrunc, alpmu,
                                                alpnu, alpv, steps):
    """Domain sizes other than 2 pi (square and not), truncation radii
    other than 2/3, a drag exponent other than 0 and dissipation k^4: RK4
    steps with forcing, then spectrum, fluxes and glob, against the oracle."""
    p = params_for(n, alpv=alpv, alpnu=alpnu, alpmu=alpmu, rtrunc=rtrunc, xlx=xlx, xly=xly)
    s, o = _pair(oracle_mod, p)
    assert s.nf == o.nf and s.nf > 0
    z = random_spectral(n, 90 + n, truncated=rtrunc <= 2 / 3 + 1e-12)
    o.trunc(z)
    zg = z.copy()
    scratch = [s.field() for _ in range(4)]
    _present(s, zg, *scratch)
    seed_g = seed_o = 481929
    print(n, xlx, xly, rtrunc, alpmu, flush=True)
    for _ in range(steps):
        s.step4(zg, *scratch)
        seed_g = s.forcing(zg, seed_g)
        o.step4(z)
        seed_o = o.forcing(z, seed_o)
    s.update_host(zg)
    assert seed_g == seed_o
    err = rel_l2(zg, z)
    print(f"  {steps} steps rel L2 = {err:.3e}", flush=True)
    assert err < 1e-12
    znl = z.copy()
    o.nlt(znl)
    ng = znl.copy()
    _present(s, ng)
    sp, sz = s.spectrum(zg)
    osp, osz = o.spectrum(z)
    assert rel_l2(sp, osp) < 1e-12 and rel_l2(sz, osz) < 1e-12
    fle, flz = s.fluxes(zg, ng)
    ofle, oflz = o.fluxes(z, znl)
    assert np.abs(fle - ofle).max() / (np.abs(ofle).max() + 1e-300) < 1e-9
    assert np.abs(flz - oflz).max() / (np.abs(oflz).max() + 1e-300) < 1e-9
    assert np.allclose(s.glob(zg), o.glob(z), rtol=1e-12, atol=0)


def test_run_euler256_matches_oracle(oracle_mod):
    """The 256^2 Euler workload (alpha=2, RK4, dt=1e-3, nu=1e-16 k^8, mu=0.1,
    ring (15,17]) over its full 1000 steps from the same float32 initial field
    and seed: vorticity within 1e-10 relative L2 of the oracle."""
    from gturbo_amd import ic, io, workloads
    w = workloads()["euler256"]
    p = w.params
    s, o = _pair(oracle_mod, p)
    w0 = ic.physical_field(p, w.ic)
    zg = io.field_to_ref(w0)
    zo = zg.copy()
    seed_g, _ = s.run(zg, w.forcing_seed, p.npas)
    seed_o, _, _ = o.run(zo, w.forcing_seed, p.npas)
    assert seed_g == seed_o
    err = rel_l2(zg[:, :p.n], zo[:, :p.n])
    print(f"euler256 1000-step rel L2 = {err:.3e}")
    assert err < 1e-10


@pytest.mark.parametrize("rko", [4, 2])
def test_headline_workload_steps_match_oracle(oracle_mod, rko):
    """The 4096^2 headline workload (alpha=2, dt=5e-5, nu=1e-28 k^8, mu=0.02,
    ring (63,65], seed-2 initial field) for 6 steps with forcing, at its full
    size: vorticity within 1e-12 relative L2 of the oracle, seed bit-exact."""
    from gturbo_amd import ic, io, workloads
    w = workloads()[f"euler4096_rk{rko}"]
    p = w.params
    s, o = _pair(oracle_mod, p)
    zg = io.field_to_ref(ic.physical_field(p, w.ic))
    zo = zg.copy()
    print("4096^2 RK%d, 6 steps" % rko, flush=True)
    seed_g, _ = s.run(zg, w.forcing_seed, 6)
    seed_o, _, _ = o.run(zo, w.forcing_seed, 6)
    assert seed_g == seed_o
    err = rel_l2(zg[:, :p.n], zo[:, :p.n])
    print(f"euler4096 RK{rko} 6-step rel L2 = {err:.3e}")
    assert err < 1e-12


@pytest.mark.parametrize("alpv,truncate", [(2.0, 1), (1.0, 1), (2.0, 0)])
def test_4096_nlt_and_steps_on_each_transform_path(oracle_mod, alpv, truncate):
    """At 4096^2 the x pass always runs the in-place transforms, and the y
    passes do for alpha = 2 on the 2/3 rectangle (rows kept in transform
    order between the passes) and otherwise run the Stockham ones: nlt, one
    RK4 step and one RK2 step on each combination."""
    n = 4096
    for rko in (4, 2):
        p = params_for(n, alpv=alpv, truncate=truncate, rko=rko)
        s, o = _pair(oracle_mod, p)
        z = random_spectral(n, 31 + rko, truncated=bool(truncate))
        zg, ug, vg = z.copy(), s.field(), s.field()
        _present(s, zg, ug, vg)
        print("4096 alpha", alpv, "truncate", truncate, "rko", rko, flush=True)
        if rko == 4:  # nlt once, on its own
            zn = z.copy()
            zng = zn.copy()
            _present(s, zng)
            s.nlt(zng, ug, vg)
            s.update_host(zng)
            o.nlt(zn)
            assert rel_l2(zng, zn) < 1e-12
        scratch = [s.field() for _ in range(4)]
        _present(s, *scratch)
        if rko == 4:
            s.step4(zg, *scratch)
            o.step4(z)
        else:
            s.step(zg, scratch[0], scratch[2], scratch[3])
            o.step(z)
        s.update_host(zg)
        err = rel_l2(zg, z)
        print(f"  step rel L2 = {err:.3e}", flush=True)
        assert err < 1e-12
        s.close()


@pytest.mark.parametrize("alpv,truncate", [(2.0, 1), (1.0, 1), (2.0, 0)])
def test_8192_nlt_and_steps(oracle_mod, alpv, truncate):
    """N = 8192, the reference's compile-time default (src/modules.f90): nlt,
    one RK4 step and one RK2 step with forcing against the oracle.  With the
    2/3 rule the x pass transforms one row per workgroup (the 4096-point
    transform of its even and odd samples) and the y passes run the in-place
    transforms; without it (the whole half-spectrum kept) row pairs and the
    Stockham transforms."""
    n = 8192
    for rko in (4, 2):
        p = params_for(n, alpv=alpv, rko=rko, truncate=truncate,
                       ring=(63.0, 65.0))  # nf within NFORC
        s, o = _pair(oracle_mod, p)
        z = random_spectral(n, 41 + rko, truncated=bool(truncate))
        zg = z.copy()
        u, v = s.field(), s.field()
        scratch = [s.field() for _ in range(4)]
        _present(s, zg, u, v, *scratch)
        print("8192 alpha", alpv, "truncate", truncate, "rko", rko, flush=True)
        if rko == 4:
            zn, zng = z.copy(), z.copy()
            _present(s, zng)
            s.nlt(zng, u, v)
            s.update_host(zng)
            o.nlt(zn)
            assert rel_l2(zng, zn) < 1e-12
            s.exit_data(zng, copyout=False)
            s.step4(zg, *scratch)
            o.step4(z)
        else:
            s.step(zg, scratch[0], scratch[2], scratch[3])
            o.step(z)
        seed_g = s.forcing(zg, 481929)
        seed_o = o.forcing(z, 481929)
        s.update_host(zg)
        assert seed_g == seed_o
        err = rel_l2(zg, z)
        print(f"  step rel L2 = {err:.3e}", flush=True)
        assert err < 1e-12
        s.close()


def test_sqg1024_workload_with_diagnostics_matches_oracle(oracle_mod):
    """The 1024^2 SQG workload (alpha=1, RK4, dt=2e-4, nu=1e-22 k^8,
    mu=0.05, ring (31,33], seed-1 initial field, spectra and fluxes every 100
    steps) over its first 1000 steps: vorticity within 1e-10, spectra, fluxes
    and global sums within 1e-9 of the oracle.  (Its 2000 steps are past the
    flow's predictability for FP64 rounding: two runs on the GPU from initial
    fields 6e-16 apart separate to 1e-10 near step 1350 and 1e-4 at step 2000,
    growing about 8x per 100 steps late in the run, as the GPU and the oracle
    do; 1000 steps keep that separation near 1e-12.)"""
    from gturbo_amd import ic, io, workloads
    w = workloads()["sqg1024"]
    p = w.params
    s, o = _pair(oracle_mod, p)
    zg = io.field_to_ref(ic.physical_field(p, w.ic))
    zo = zg.copy()
    got = []
    nsteps = 1000
    print(f"sqg1024, {nsteps} steps", flush=True)
    seed_g, _ = s.run(zg, w.forcing_seed, nsteps, diag=lambda *d: got.append(d))
    seed_o, _, diag = o.run(zo, w.forcing_seed, nsteps, with_diag=True)
    assert seed_g == seed_o
    err = rel_l2(zg[:, :p.n], zo[:, :p.n])
    print(f"sqg1024 {nsteps}-step rel L2 = {err:.3e}")
    assert err < 1e-10
    nb = p.n // 2
    assert len(got) == diag.shape[0] == nsteps // p.imix
    for (ipas, sp, sz, fle, flz, g), row in zip(got, diag):
        osp, osz, ofle, oflz = (row[i * nb:(i + 1) * nb] for i in range(4))
        assert rel_l2(sp, osp) < 1e-9 and rel_l2(sz, osz) < 1e-9
        assert np.abs(fle - ofle).max() <= 1e-9 * np.abs(ofle).max()
        assert np.abs(flz - oflz).max() <= 1e-9 * np.abs(oflz).max()
        assert np.allclose(g, row[4 * nb:], rtol=1e-9, atol=0)


@pytest.mark.parametrize("n,rko", [(64, 4), (256, 4), (256, 2), (4096, 4)])
def test_step_boundary_sequences_match_oracle(oracle_mod, n, rko):
    """gt_step/gt_step4 leave their last stage queued; forcing(z) and the
    next step on the same arrays fuse it with the next step's first pass,
    anything else runs it first.  Two states of one context through the
    sequences a caller can make -- step, forcing, step (fused), step without
    forcing between (fused, unforced), a step on the other state (the queued
    end runs first), forcing of the other state, diagnostics and a host copy
    in between -- against the oracle doing the same operations."""
    p = params_for(n, alpv=2.0 if n == 4096 else 1.0, rko=rko)
    s, o = _pair(oracle_mod, p)
    A = random_spectral(n, 301)
    B = random_spectral(n, 302)
    Ag, Bg = A.copy(), B.copy()
    wa = [s.field() for _ in range(4)]
    wb = [s.field() for _ in range(4)]
    _present(s, Ag, Bg, *wa, *wb)

    def gstep(z, w):
        if rko == 4:
            s.step4(z, *w)
        else:
            s.step(z, w[0], w[2], w[3])

    def ostep(z):
        (o.step4 if rko == 4 else o.step)(z)

    sa = sb = 481929
    oa = ob = 481929
    # A: step, forcing, step (fused), forcing, step, step (fused without forcing)
    for forced in (True, True, False):
        gstep(Ag, wa)
        ostep(A)
        if forced:
            sa = s.forcing(Ag, sa)
            oa = o.forcing(A, oa)
    gstep(Ag, wa)
    ostep(A)
    # B between A's step and A's forcing: A's queued end runs first
    gstep(Bg, wb)
    ostep(B)
    sa = s.forcing(Ag, sa)
    oa = o.forcing(A, oa)
    sb = s.forcing(Bg, sb)
    ob = o.forcing(B, ob)
    # diagnostics between a step and its forcing, as the driver does
    gstep(Ag, wa)
    ostep(A)
    g = s.glob(Ag)
    assert np.allclose(g, o.glob(A), rtol=1e-12, atol=0)
    sa = s.forcing(Ag, sa)
    oa = o.forcing(A, oa)
    gstep(Ag, wa)  # fused with the forcing above? no: the glob ran the end first
    ostep(A)
    s.update_host(Ag)
    sa = s.forcing(Ag, sa)  # after a host copy: queued end already run
    oa = o.forcing(A, oa)
    gstep(Bg, wb)
    ostep(B)
    sb = s.forcing(Bg, sb)
    ob = o.forcing(B, ob)
    gstep(Bg, wb)  # fused with B's forcing
    ostep(B)
    s.update_host(Ag)
    s.update_host(Bg)
    assert (sa, sb) == (oa, ob)
    ea, eb = rel_l2(Ag, A), rel_l2(Bg, B)
    print(f"n={n} rk{rko}: A {ea:.2e} B {eb:.2e}", flush=True)
    assert ea < 1e-12 and eb < 1e-12
    s.close()
