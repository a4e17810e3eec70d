"""Pins the CPU oracle (oracle/gturbo_oracle.c) before it is trusted as the
checker of the HIP path.  Every check comes from something other than the
oracle's own formulas: the DFT definition, numpy's pocketfft, closed forms in
physical space, conservation laws, convergence order, the LCG's closed form,
exact rational arithmetic, or an independent enumeration.  Each test names
the reference file it pins."""
import math
from fractions import Fraction

import numpy as np
import pytest

from conftest import rel_l2
from helpers import as_complex, params_for, random_physical, random_spectral


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


def mk(O, n, **kw):
    return O.Oracle(params_for(n, **kw))


# --------------------------------------------------------------------------
# rann / forcing: src/forcing.f90:40-51
# --------------------------------------------------------------------------

def lcg_closed_form(seed0, n):
    """seed_n = (a^n seed_0 + (a^n - 1)/(a - 1)) mod 2^31 for
    seed <- (69069 seed + 1) mod 2^32 masked to 31 bits (the mask makes the
    recurrence exact modulo 2^31)."""
    a, m = 69069, 2 ** 31
    an = pow(a, n, m * (a - 1))
    return (an * seed0 + (an - 1) // (a - 1)) % m


def test_rann_matches_lcg_closed_form(O):
    s = 481929
    for n in range(1, 2001):
        r, s = O.rann(s)
        assert s == lcg_closed_form(481929, n)
        assert r == np.float32(s / 2147483647.0)


def test_rann_first_draw_from_reference_seed(O):
    # src/seed.0 holds 481929; the first draw was checked with amdflang on a
    # standalone copy of the LCG expression (integer wrap as compiled).
    r, s = O.rann(481929)
    assert s == 1074099382


def test_forcing_closed_form_on_zero_field(O):
    n = 64
    o = mk(O, n, dt=3.1e-4)
    z = o.field()
    seed = 481929
    o.forcing(z, seed)
    zc = as_complex(z)
    ikx, iky, ff = o.forcing_modes()
    pi2 = 2.0 * 3.14159265358979          # src/forcing.f90:10
    amp = math.sqrt(3.1e-4) * 1.0
    s = seed
    for k in range(o.nf):
        r, s = O.rann(s)
        te = pi2 * float(r)
        i, j = ikx[k] - 1, iky[k] - 1
        assert abs(zc[j, i].real - amp * math.cos(te)) < 1e-17
        assert abs(zc[j, i].imag - amp * math.sin(te)) < 1e-17
        if i == 0:  # Hermitian partner, src/forcing.f90:28-31
            jc = (n - j) % n
            assert zc[jc, 0] == np.conj(zc[j, 0])


# --------------------------------------------------------------------------
# transforms: src/fft_dir.f90, src/fft_inv.f90 (cuFFT D2Z/Z2D semantics)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("n", [4, 8, 16])
def test_fft_dir_is_the_dft(O, n):
    o = mk(O, n) if n >= 16 else O.Oracle(params_for(16).replace(n=n))
    z = random_physical(n, 1)
    q = z[:, :n].copy()
    o.fft_dir(z)
    x = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(x, x) / n)
    dft = W @ q @ W.T / (n * n)   # [ky][kx]
    assert np.abs(as_complex(z) - dft[:, : n // 2 + 1]).max() < 1e-15


@pytest.mark.parametrize("n", [32, 64, 256])
def test_fft_dir_matches_numpy(O, n):
    o = mk(O, n)
    z = random_physical(n, 2)
    ref = np.fft.rfft2(z[:, :n]) / n / n
    o.fft_dir(z)
    assert np.abs(as_complex(z) - ref).max() < 1e-16 * n


@pytest.mark.parametrize("n", [32, 64, 256])
def test_fft_inv_matches_numpy(O, n):
    o = mk(O, n)
    z = random_spectral(n, 3, truncated=False)
    ref = np.fft.irfft2(as_complex(z), s=(n, n)) * n * n
    o.fft_inv(z)
    assert np.abs(z[:, :n] - ref).max() < 1e-13 * np.abs(ref).max()


def test_single_mode_closed_form(O):
    n = 32
    o = mk(O, n)
    x = 2 * np.pi * np.arange(n) / n
    X, Y = np.meshgrid(x, x)            # X[j][i] = x_i, Y[j][i] = y_j
    z = np.zeros((n, n + 2))
    z[:, :n] = np.cos(3 * X + 5 * Y) + 2.0 * np.sin(-4 * X + 7 * Y)
    o.fft_dir(z)
    zc = as_complex(z)
    expect = np.zeros_like(zc)
    expect[5, 3] = 0.5                  # cos(3x+5y): 1/2 at (kx=3, ky=5)
    expect[n - 7, 4] = 1j               # 2 sin(-4x+7y) = i e^{i(4x-7y)} - ...
    assert np.abs(zc - expect).max() < 1e-15


@pytest.mark.parametrize("n", [16, 128])
def test_round_trip_and_parseval(O, n):
    o = mk(O, n)
    z = random_physical(n, 4)
    q = z[:, :n].copy()
    o.fft_dir(z)
    zc = as_complex(z)
    fac = np.full(zc.shape, 2.0)
    fac[:, 0] = fac[:, -1] = 1.0
    assert abs(np.sum(fac * np.abs(zc) ** 2) - np.mean(q * q)) < 1e-13 * np.mean(q * q)
    o.fft_inv(z)
    assert np.abs(z[:, :n] - q).max() < 1e-13


# --------------------------------------------------------------------------
# defk / trunc / iniforcing
# --------------------------------------------------------------------------

def test_defk_tables(O):
    n = 32
    o = mk(O, n, nu=1e-3, mu=0.2, alpnu=2.0, alpmu=0.0, dt=1e-2)
    d = o.defk()
    assert list(d["jc"]) == [1] + [n + 2 - j for j in range(2, n + 1)]
    assert d["fkx"][5] == 5.0 and d["fky"][n - 1] == -1.0 and d["fky"][n // 2] == n / 2
    emk = d["emk"]
    assert emk[0, 0] == 0.0
    # depends on |k| only, symmetric in ky <-> -ky and kx <-> ky
    for kx, ky in [(1, 2), (3, 4), (0, 5), (7, 1)]:
        a = emk[ky, kx]
        assert a == emk[(n - ky) % n, kx] == emk[kx, ky]
    # monotone decreasing with |k| and equal to exp(-dt/2 mu) at k -> 0 limit
    k2 = np.add.outer(d["fkys"], d["fkxs"])
    order = np.argsort(k2[1:, 0])
    assert np.all(np.diff(emk[1:, 0][order]) <= 0)
    assert emk[0, 1] < math.exp(-0.5 * 1e-2 * 0.2)


@pytest.mark.parametrize("n,count", [(256, 11524), (64, None)])
def test_trunc_mask_exact_rational(O, n, count):
    """The 2/3 circle of src/trunc.f90:13-15 evaluated with exact rationals
    agrees mode by mode (no mode of a power-of-two grid lies on it)."""
    o = mk(O, n)
    z = np.ones((n, n + 2))
    o.trunc(z)
    kept = np.abs(as_complex(z)) != 0
    nh = n // 2 + 1
    exact = np.zeros((n, nh), bool)
    for j in range(n):
        ky = j if j <= n // 2 else j - n
        for i in range(nh):
            exact[j, i] = Fraction(ky * ky + i * i, (n // 2) ** 2) <= Fraction(4, 9)
    exact[0, 0] = False
    assert np.array_equal(kept, exact)
    if count is not None:
        assert int(kept.sum()) + 1 == count  # the tabulated count includes the mean


@pytest.mark.parametrize("n,ring,nf", [(256, (15, 17), 96), (512, (31, 33), 204),
                                       (1024, (63, 65), 410)])
def test_iniforcing_modes(O, n, ring, nf):
    o = mk(O, n, ring=ring, alpv=1.0)
    assert o.nf == nf
    ikx, iky, ff = o.forcing_modes()
    # independent enumeration in the order of src/iniforcing.f90:15-68
    order = [(1, j) for j in range(2, n // 2 + 2)]
    order += [(i, j) for i in range(2, n // 2 + 1) for j in range(1, n + 1)]
    order += [(n // 2 + 1, j) for j in range(1, n // 2 + 2)]
    sel = []
    for i, j in order:
        kx, ky = i - 1, (j - 1 if j - 1 <= n // 2 else j - 1 - n)
        k2 = kx * kx + ky * ky
        inside = Fraction(k2, (n // 2) ** 2) <= Fraction(4, 9)
        if inside and ring[0] ** 2 < k2 <= ring[1] ** 2:
            sel.append((i, j))
    assert list(zip(ikx.tolist(), iky.tolist())) == sel
    assert np.all(ff == 1.0)
    assert o.inpz == pytest.approx(nf * 1.0)
    assert o.inpe == pytest.approx(sum(1.0 / math.sqrt((i - 1) ** 2 + ((j - 1) if j - 1 <= n // 2 else j - 1 - n) ** 2)
                                       for i, j in sel))


def test_nforc_overflow_is_an_error(O):
    with pytest.raises(RuntimeError, match="NFORC"):
        O.Oracle(params_for(1024, ring=(20, 200)))


# --------------------------------------------------------------------------
# nlt: src/nlt.f90 -- closed form and conservation
# --------------------------------------------------------------------------

@pytest.mark.parametrize("alpv", [0.5, 1.0, 2.0])
def test_nlt_three_mode_closed_form(O, alpv):
    """q = sum a_m cos(k_m.x + phi_m) with every sum and difference mode inside
    the truncation circle: N = -(u q_x + v q_y) with u = d_y psi,
    v = -d_x psi, psi = sum a_m |k_m|^-alpha cos(...), evaluated exactly in
    physical space."""
    n = 64
    o = mk(O, n, alpv=alpv)
    x = 2 * np.pi * np.arange(n) / n
    X, Y = np.meshgrid(x, x)
    modes = [(1.3, 2, 1, 0.3), (0.7, -3, 2, 1.1), (0.9, 4, -1, 2.0)]
    q = np.zeros((n, n))
    qx = np.zeros((n, n)); qy = np.zeros((n, n)); u = np.zeros((n, n)); v = np.zeros((n, n))
    for a, kx, ky, ph in modes:
        th = kx * X + ky * Y + ph
        kk = (kx * kx + ky * ky) ** (alpv / 2.0)
        q += a * np.cos(th)
        qx += -a * kx * np.sin(th)
        qy += -a * ky * np.sin(th)
        u += -a / kk * ky * np.sin(th)
        v += a / kk * kx * np.sin(th)
    expect = -(u * qx + v * qy)
    z = np.zeros((n, n + 2)); z[:, :n] = q
    o.fft_dir(z)
    o.nlt(z)
    o.fft_inv(z)
    assert np.abs(z[:, :n] - expect).max() < 1e-13 * np.abs(expect).max()


@pytest.mark.parametrize("alpv", [1.0, 2.0])
def test_nlt_conserves_enstrophy_and_energy(O, alpv):
    n = 128
    o = mk(O, n, alpv=alpv)
    z = random_spectral(n, 8)
    q = as_complex(z).copy()
    o.nlt(z)
    nh = n // 2 + 1
    N = as_complex(z)
    fac = np.full(q.shape, 2.0); fac[:, 0] = fac[:, -1] = 1.0
    kx = np.arange(nh); ky = np.fft.fftfreq(n, 1.0 / n)
    k2 = np.add.outer(ky ** 2, kx ** 2); k2[0, 0] = 1.0
    t = fac * (q.conj() * N).real
    scale = np.sum(fac * np.abs(q) * np.abs(N))
    assert abs(t.sum()) < 1e-13 * scale                        # enstrophy
    assert abs((t / k2 ** (alpv / 2)).sum()) < 1e-13 * scale   # energy


def test_nlt_single_mode_is_zero(O):
    n = 32
    o = mk(O, n)
    z = np.zeros((n, n + 2))
    as_complex(z)[3, 2] = 0.7 - 0.2j
    o.nlt(z)
    assert np.abs(z).max() < 1e-16


# --------------------------------------------------------------------------
# step / step4: src/rkstep.f90
# --------------------------------------------------------------------------

@pytest.mark.parametrize("rko", [2, 4])
def test_step_single_mode_is_integrating_factor(O, rko):
    n = 32
    o = mk(O, n, rko=rko, nu=1e-2, alpnu=1.0, mu=0.3, dt=0.05)
    E = o.defk()["emk"]
    z = np.zeros((n, n + 2))
    as_complex(z)[3, 2] = 0.7 - 0.2j
    z0 = as_complex(z).copy()
    (o.step4 if rko == 4 else o.step)(z)
    assert np.abs(as_complex(z) - E * E * z0).max() < 2e-16


def _smooth_state(n):
    z = np.zeros((n, n + 2))
    zc = as_complex(z)
    zc[1, 1] = 1.0 + 0.5j
    zc[2, 0] = 0.8
    zc[n - 2, 0] = 0.8
    zc[n - 1, 2] = 0.3 - 0.6j
    zc[3, 1] = -0.4 + 0.2j
    return z


@pytest.mark.parametrize("rko,order", [(2, 2), (4, 4)])
@pytest.mark.parametrize("dissipative", [False, True])
def test_rk_convergence_order(O, rko, order, dissipative):
    """Errors at fixed T shrink as dt^order; a wrong stage coefficient breaks
    the rate.  Lawson integrating factors keep the order (Hochbruck &
    Ostermann, Acta Numerica 2010, sec. 2.3)."""
    n, T = 32, 0.4
    kw = dict(nu=5e-3, alpnu=1.0, mu=0.5, alpmu=0.0) if dissipative else dict(nu=0.0, mu=0.0)
    if rko == 2 and dissipative:
        kw["alpnu"] = 2.0  # keep the unsplit factor (split needs alpnu = 1)

    def solve(steps):
        o = mk(O, n, rko=rko, dt=T / steps, famp=0.0, **kw)
        z = _smooth_state(n)
        for _ in range(steps):
            (o.step4 if rko == 4 else o.step)(z)
        return as_complex(z).copy()

    ref = solve(512)
    errs = [np.abs(solve(s) - ref).max() for s in (16, 32, 64)]
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(abs(r - order) < 0.35 for r in rates), rates


def test_split_step_uses_product_factor(O):
    n = 32
    o = mk(O, n, rko=2, nu=1e-2, alpnu=1.0, mu=0.0, alpmu=0.0, dt=0.05)
    assert o.split == 1
    d = o.defk()
    z = np.zeros((n, n + 2))
    as_complex(z)[3, 2] = 0.7 - 0.2j
    z0 = as_complex(z).copy()
    o.step(z)
    E = np.outer(d["emky"], d["emkx"])
    assert np.abs(as_complex(z) - E * E * z0).max() < 2e-16


# --------------------------------------------------------------------------
# The integrating factor from the physics alone: src/defk.f90:40-50 and the
# Lawson steps of src/rkstep.f90.  With N(z) = 0 the step must be the exact
# solution of dq/dt = -D(k) q over dt, D = nu k^(2 alpnu) + mu k^(2 alpmu):
# q(dt) = exp(-dt D) q(0).  Plane waves with parallel wave vectors have a
# vanishing Jacobian (u is normal to k, grad q along it), so a state made of
# them decays mode by mode.  The expected values below are computed from the
# params, never from the oracle's tables.
# --------------------------------------------------------------------------

PARALLEL_MODES = [(1, 2, 0.7 - 0.2j), (2, 4, -0.3 + 0.5j), (3, 6, 0.25 + 0.1j),
                  (5, 10, 0.05 - 0.02j)]  # (kx, ky, amplitude): k parallel to (1, 2)


def _parallel_state(n, xlx=2 * np.pi, xly=2 * np.pi):
    z = np.zeros((n, n + 2))
    zc = as_complex(z)
    for kx, ky, a in PARALLEL_MODES:
        zc[ky % n, kx] = a
    return z


@pytest.mark.parametrize("rko,alpnu,alpmu,nu,mu", [
    (4, 4.0, 0.0, 1e-5, 1.0),    # hyperviscosity k^8 and linear drag (the workloads)
    (4, 1.0, -1.0, 2e-2, 0.5),   # viscosity k^2 and large-scale drag k^-2
    (2, 2.0, 0.0, 1e-3, 0.2),    # unsplit RK2
])
def test_step_factor_is_exp_minus_dt_dissipation(O, rko, alpnu, alpmu, nu, mu):
    n, dt = 32, 0.05
    o = mk(O, n, rko=rko, nu=nu, alpnu=alpnu, mu=mu, alpmu=alpmu, dt=dt, famp=0.0)
    assert o.split == 0
    z = _parallel_state(n)
    (o.step4 if rko == 4 else o.step)(z)
    zc = as_complex(z)
    for kx, ky, a in PARALLEL_MODES:
        k = math.hypot(kx, ky)
        D = nu * k ** (2 * alpnu) + mu * k ** (2 * alpmu)
        want = math.exp(-dt * D) * a
        assert dt * D > 0.01  # the factor is far from 1, so a wrong exponent shows
        assert abs(zc[ky % n, kx] - want) <= 1e-14 * abs(a), (kx, ky, zc[ky % n, kx], want)
    # every other mode stays zero up to the rounding of the vanishing Jacobian
    for kx, ky, _ in PARALLEL_MODES:
        zc[ky % n, kx] = 0
    assert np.abs(zc).max() < 1e-15


def test_split_step_factor_is_exp_minus_dt_dissipation(O):
    """The split RK2 factor emkx(i)*emky(j) (src/defk.f90:58-68, used when
    alpnu = 1 and alpmu = 0) is exp(-dt/2 nu kx^2) * exp(-dt/2 (nu ky^2 + mu)),
    whose square is exp(-dt (nu k^2 + mu))."""
    n, dt, nu, mu = 32, 0.05, 2e-2, 0.7
    o = mk(O, n, rko=2, nu=nu, alpnu=1.0, mu=mu, alpmu=0.0, dt=dt, famp=0.0)
    assert o.split == 1
    z = _parallel_state(n)
    o.step(z)
    zc = as_complex(z)
    for kx, ky, a in PARALLEL_MODES:
        want = math.exp(-dt * (nu * (kx * kx + ky * ky) + mu)) * a
        assert abs(zc[ky % n, kx] - want) <= 1e-14 * abs(a)


def test_step_factor_in_a_larger_box(O):
    """xlx = 4 pi: wave numbers are multiples of dk = 1/2 (src/defk.f90:18-30),
    so index (i, j) is the wave vector (i/2, j/2)."""
    n, dt, nu, mu = 32, 0.05, 1e-2, 0.3
    o = mk(O, n, rko=4, nu=nu, alpnu=2.0, mu=mu, alpmu=0.0, dt=dt, famp=0.0,
           xlx=4 * np.pi)
    assert o.dkx == 0.5 and o.dky == 0.5
    z = _parallel_state(n)
    o.step4(z)
    zc = as_complex(z)
    for kx, ky, a in PARALLEL_MODES:
        k2 = (kx / 2) ** 2 + (ky / 2) ** 2
        want = math.exp(-dt * (nu * k2 ** 2 + mu)) * a
        assert abs(zc[ky % n, kx] - want) <= 1e-14 * abs(a)


# --------------------------------------------------------------------------
# diagnostics: src/diagnostics.f90
# --------------------------------------------------------------------------

@pytest.mark.parametrize("xlx", [2 * np.pi, 4 * np.pi])
@pytest.mark.parametrize("alpv", [1.0, 2.0])
def test_flux_is_the_derivative_of_the_spectrum(O, xlx, alpv):
    """fluxes (diagnostics.f90:84-109) against spectrum (:25-44): the
    per-shell transfer flzm(b) = sum fac Re(conj(z) N) is dk times the
    derivative of sz(z + eps N) at eps = 0, and flem likewise of sp; both
    spectra are quadratic in z, so a central difference is exact up to
    rounding for any eps.  This ties the flux weight (2 inside, 1 on the
    kx = 0 and kx = N/2 columns) and the missing /dk to the spectrum's
    weights, which the single-mode tests pin."""
    n = 64
    o = mk(O, n, alpv=alpv, xlx=xlx)
    dk = o.dkx
    z = random_spectral(n, 31)
    znl = z.copy()
    o.nlt(znl)
    eps = 0.5
    spp, szp = o.spectrum(z + eps * znl)
    spm, szm = o.spectrum(z - eps * znl)
    flzm = dk * (szp - szm) / (2 * eps)
    flem = dk * (spp - spm) / (2 * eps)
    fle, flz = o.fluxes(z, znl)
    want_z = np.cumsum(flzm[::-1])[::-1]
    want_e = np.cumsum(flem[::-1])[::-1]
    assert np.abs(flz - want_z).max() <= 1e-12 * np.abs(want_z).max()
    assert np.abs(fle - want_e).max() <= 1e-12 * np.abs(want_e).max()


def test_binning_in_a_larger_box(O):
    """xlx = 4 pi, dk = 1/2: a mode of wave vector (3, 4) (index (6, 8)) has
    |k| = 5, shell int(5/dk)+1 = 11 (diagnostics.f90:19,31-34), and the
    spectrum divides by dk (:42-44)."""
    n = 64
    o = mk(O, n, alpv=2.0, xlx=4 * np.pi)
    assert o.dkx == 0.5
    z = np.zeros((n, n + 2))
    as_complex(z)[8, 6] = 0.3 + 0.4j            # |q|^2 = 1/4, k^2 = 25
    sp, sz = o.spectrum(z)
    assert np.count_nonzero(sz) == 1 and np.count_nonzero(sp) == 1
    assert sz[10] == pytest.approx(0.25 / 0.5, rel=1e-15)
    assert sp[10] == pytest.approx(0.25 / 25.0 / 0.5, rel=1e-15)
    as_complex(z)[8, 6] = 0
    as_complex(z)[n - 3, 0] = 2.0               # kx = 0, ky = -3/2: weight 1/2, shell 4
    sp, sz = o.spectrum(z)
    assert np.count_nonzero(sz) == 1 and sz[3] == pytest.approx(0.5 * 4.0 / 0.5)
    # shell index of every mode: int(|k|/dk)+1 in grid-index units
    ib = o.shell_index()
    nh = n // 2 + 1
    for j in range(n):
        ky = j if j <= n // 2 else j - n
        for i in range(nh):
            if i or j:
                assert ib[j, i] == math.isqrt(i * i + ky * ky) + 1
    # a box of 8 pi (dk = 1/4) bins the same index pair into the same shell
    o4 = mk(O, n, alpv=2.0, xlx=8 * np.pi)
    assert np.array_equal(o4.shell_index(), ib)


def test_spectrum_single_mode(O):
    n = 64
    o = mk(O, n, alpv=2.0)
    z = np.zeros((n, n + 2))
    as_complex(z)[4, 3] = 0.5 + 0.5j          # |k| = 5 -> shell int(5)+1 = 6
    sp, sz = o.spectrum(z)
    assert np.count_nonzero(sp) == 1 and sp[5] == pytest.approx(0.5 / 25.0)
    assert sz[5] == pytest.approx(0.5)
    as_complex(z)[4, 3] = 0
    as_complex(z)[4, 0] = 2.0                  # kx = 0: weight 1/2
    sp, sz = o.spectrum(z)
    assert sz[4] == pytest.approx(0.5 * 4.0)


def test_spectrum_sums_to_glob_and_flux_closes(O):
    n = 128
    o = mk(O, n, alpv=1.0)
    z = random_spectral(n, 21)
    sp, sz = o.spectrum(z)
    g = o.glob(z)
    assert sp.sum() * o.dkx == pytest.approx(g[0], rel=1e-12)
    assert sz.sum() * o.dkx == pytest.approx(g[1], rel=1e-12)
    znl = z.copy()
    o.nlt(znl)
    fle, flz = o.fluxes(z, znl)
    # total transfer over all shells vanishes (conservation, flux(1) = 0)
    assert abs(flz[0]) < 1e-12 * np.abs(flz).max()
    assert abs(fle[0]) < 1e-12 * np.abs(fle).max()
    # flux is a suffix sum: consecutive differences are per-shell transfers
    assert flz[-1] == 0.0


def test_glob_dissipation_closed_form(O):
    n = 32
    o = mk(O, n, nu=2e-3, alpnu=2.0, mu=0.1, alpmu=0.0, alpv=2.0)
    z = np.zeros((n, n + 2))
    as_complex(z)[4, 3] = 0.3 + 0.4j          # |q|^2 = 0.25, k^2 = 25
    g = o.glob(z)
    assert g[0] == pytest.approx(0.25 / 25)                  # ee
    assert g[1] == pytest.approx(0.25)                       # zz
    assert g[2] == pytest.approx(2 * 2e-3 * 25 ** 2 * 0.25)  # disz_nu
    assert g[3] == pytest.approx(2 * 0.1 * 0.25)             # disz_mu
    assert g[4] == pytest.approx(2 * 2e-3 * 25 * 0.25)       # dise_nu
    assert g[5] == pytest.approx(2 * 0.1 / 25 * 0.25)        # dise_mu


def test_shell_index_is_integer_sqrt(O):
    n = 256
    o = mk(O, n)
    ib = o.shell_index()
    nh = n // 2 + 1
    for j in range(n):
        ky = j if j <= n // 2 else j - n
        for i in range(0, nh, 7):
            if i == 0 and j == 0:
                assert ib[0, 0] == 0
                continue
            assert ib[j, i] == math.isqrt(i * i + ky * ky) + 1


# --------------------------------------------------------------------------
# driver ordering: src/main.f90:89-143
# --------------------------------------------------------------------------

def test_run_is_the_driver_sequence(O):
    n = 32
    o = mk(O, n, rko=4, dt=1e-3)
    w = random_physical(n, 5)
    a = w.copy()
    seed_a, t, _ = o.run(a, 481929, 3)
    b = w.copy()
    o.fft_dir(b)
    o.trunc(b)
    s = 481929
    for _ in range(3):
        o.step4(b)
        s = o.forcing(b, s)
    o.fft_inv(b)
    assert seed_a == s and t == pytest.approx(3e-3)
    assert np.array_equal(a, b)


# --- tangent-linear model (checks the library's adjoint) -------------------

@pytest.mark.parametrize("alpv", [0.5, 1.0, 2.0])
def test_nlt_tangent_is_the_exact_derivative(oracle_mod, alpv):
    """N is quadratic in z, so a central difference of or_nlt equals
    N'(z) d up to rounding, for any step size."""
    n = 32
    o = oracle_mod.Oracle(params_for(n, alpv=alpv))
    z = random_spectral(n, 5)
    d = random_spectral(n, 6)
    tl = o.nlt_tl(z, d)
    eps = 0.25
    zp, zm = z + eps * d, z - eps * d
    o.nlt(zp)
    o.nlt(zm)
    fd = (zp - zm) / (2 * eps)
    assert rel_l2(tl, fd) < 1e-13
    # linear in the direction
    d2 = random_spectral(n, 7)
    assert rel_l2(o.nlt_tl(z, 2.0 * d - 3.0 * d2),
                  2.0 * tl - 3.0 * o.nlt_tl(z, d2)) < 1e-13


@pytest.mark.parametrize("rko", [4, 2])
def test_step_tangent_matches_finite_differences(oracle_mod, rko):
    n = 32
    p = params_for(n, alpv=1.0, rko=rko, famp=0.0)
    o = oracle_mod.Oracle(p)
    z = random_spectral(n, 8) * 4.0
    d = random_spectral(n, 9)
    zt, dt_ = z.copy(), d.copy()
    o.step_tl(zt, dt_)
    z4 = z.copy()
    o.step4(z4) if rko == 4 else o.step(z4)
    assert np.array_equal(zt, z4)  # the state advances as the plain step
    errs = []
    for eps in (1e-3, 1e-4):
        zp, zm = z + eps * d, z - eps * d
        for a in (zp, zm):
            o.step4(a) if rko == 4 else o.step(a)
        errs.append(rel_l2(dt_, (zp - zm) / (2 * eps)))
    assert errs[1] < 1e-7
    assert errs[1] < errs[0] * 0.05  # second-order convergence of the difference
