"""Batched ensembles (gt_ens_*): every member of a batched run must match the
same member run alone through the plain per-run loop (src/main.f90:89-136
with the single-run entry points), and the oracle on small cases.

With the default schedule (radix 8 first, one row per x-pass line) a member
goes through the same arithmetic as a solo run, so the match is exact; the
other schedules reorder FFT roundings and are held to 1e-12 relative L2.
"""
import numpy as np
import pytest

from conftest import rel_l2
from helpers import params_for
from gturbo_amd import ensemble as E
from gturbo_amd import ic
from gturbo_amd.params import InitialSpectrum, ensemble_members

pytestmark = pytest.mark.gpu


def _members():
    # uneven N and alpha, different forcing rings and dt
    ps = [params_for(128, alpv=0.5), params_for(256, alpv=1.0), params_for(512, alpv=2.0),
          params_for(128, alpv=2.0, ring=(5.0, 7.0)), params_for(1024, alpv=1.0),
          params_for(2048, alpv=2.0)]
    fields = [ic.physical_field(p, InitialSpectrum(seed=40 + m, p=4.0, k0=p.n / 32.0,
                                                   energy=0.5))
              for m, p in enumerate(ps)]
    seeds = [481929 + 7 * m for m in range(len(ps))]
    return ps, fields, seeds


def _batched(ps, fields, seeds, nsteps, cfg=None):
    ens = E.Ensemble(ps, seeds=seeds)
    for m, w in enumerate(fields):
        ens.set_field(m, w)
    if cfg is not None:
        ens.schedule = cfg
    ens.run(nsteps)
    ens.synchronize()
    out = [ens.get_field(m, E.SPECTRAL) for m in range(len(ps))]
    sd = [ens.get_seed(m) for m in range(len(ps))]
    ens.close()
    return out, sd


@pytest.fixture(scope="module")
def solo12():
    ps, fields, seeds = _members()
    print("solo runs of", [p.n for p in ps], flush=True)
    return [E.solo_run(p, w, s, 12) for p, w, s in zip(ps, fields, seeds)]


def test_default_schedule_matches_solo_exactly(solo12):
    ps, fields, seeds = _members()
    # 12 steps at 5 steps per graph: two graph replays and two eager steps
    cfg = [512, 512, 0, 0, 1 << 11, 5]
    out, sd = _batched(ps, fields, seeds, 12, cfg)
    for m, (z, s) in enumerate(zip(out, sd)):
        zs, ss = solo12[m]
        assert s == ss, m
        assert np.array_equal(z, zs), (m, rel_l2(z, zs))


@pytest.mark.parametrize("cfg", [
    [256, 256, 0, 0, 0, 12],
    [512, 256, 2, 0, 0, 3],
    [256, 512, 4, 1, 0, 7],
    [512, 512, 1, 1, (1 << 7) | (1 << 8) | (1 << 9) | (1 << 10) | (1 << 11), 1],
    [256, 256, 2, 0, (1 << 8) | (1 << 10), 256],
])
def test_every_schedule_matches_solo(solo12, cfg):
    ps, fields, seeds = _members()
    print("schedule", cfg, flush=True)
    out, sd = _batched(ps, fields, seeds, 12, cfg)
    for m, (z, s) in enumerate(zip(out, sd)):
        zs, ss = solo12[m]
        assert s == ss
        assert rel_l2(z, zs) < 1e-12, (m, rel_l2(z, zs))


def test_members_match_oracle(oracle_mod):
    ps = [params_for(128, alpv=a) for a in (0.5, 1.0, 2.0)] + [params_for(256, alpv=1.0)]
    fields = [ic.physical_field(p, InitialSpectrum(seed=7 + m, p=4.0, k0=8.0, energy=0.5))
              for m, p in enumerate(ps)]
    seeds = [481929, 1, 77, 123456789]
    out, sd = _batched(ps, fields, seeds, 10)
    for m, p in enumerate(ps):
        o = oracle_mod.Oracle(p)
        z = np.zeros((p.n, p.n + 2))
        z[:, : p.n] = fields[m][:, : p.n]
        o.fft_dir(z)
        o.trunc(z)
        seed = seeds[m]
        for _ in range(10):
            o.step4(z)
            seed = o.forcing(z, seed)
        assert sd[m] == seed
        assert rel_l2(out[m], z) < 1e-10, (m, rel_l2(out[m], z))


def test_rk2_ensemble_matches_solo():
    ps = [params_for(128, rko=2, alpv=1.0), params_for(256, rko=2)]
    fields = [ic.physical_field(p, InitialSpectrum(seed=3 + m, p=4.0, k0=8.0, energy=0.5))
              for m, p in enumerate(ps)]
    seeds = [5, 6]
    out, sd = _batched(ps, fields, seeds, 9, [512, 512, 0, 0, 0, 4])
    for m, p in enumerate(ps):
        zs, ss = E.solo_run(p, fields[m], seeds[m], 9)
        assert ss == sd[m]
        assert np.array_equal(out[m], zs)


def test_tuner_keeps_fields_and_picks_a_valid_schedule():
    ps, fields, seeds = _members()
    ens = E.Ensemble(ps, seeds=seeds)
    for m, w in enumerate(fields):
        ens.set_field(m, w)
    before = [ens.get_field(m, E.SPECTRAL) for m in range(len(ps))]
    cfg, ms = ens.tune(2)
    print("tuned", cfg, ms, flush=True)
    assert cfg[0] in (256, 512) and cfg[1] in (256, 512) and cfg[2] in (0, 1, 2, 4)
    assert cfg[3] in (0, 1) and ms > 0
    assert ens.schedule == cfg
    for m in range(len(ps)):
        assert np.array_equal(ens.get_field(m, E.SPECTRAL), before[m])
        assert ens.get_seed(m) == seeds[m]
    ens.close()


def test_field_round_trip_and_errors():
    p = params_for(256)
    ens = E.Ensemble([p])
    w = ic.physical_field(p, InitialSpectrum(seed=1, p=4.0, k0=8.0, energy=0.5))
    ens.set_field(0, w)
    back = ens.get_field(0, E.PHYSICAL)
    # fft_dir and trunc on entry (the float32 frame has content beyond 2/3)
    n = 256
    q = np.fft.rfft2(w.astype(np.float64))
    kx = np.arange(n // 2 + 1)
    ky = np.fft.fftfreq(n, 1.0 / n)
    KX, KY = np.meshgrid(kx, ky)
    q[(KY ** 2 + KX ** 2) / (n / 2) ** 2 > (2.0 / 3.0) ** 2] = 0.0
    q[0, 0] = 0.0
    assert rel_l2(back[:, :n], np.fft.irfft2(q, s=(n, n))) < 1e-13
    spec = ens.get_field(0, E.SPECTRAL)
    ens.set_field(0, spec, E.SPECTRAL)
    assert np.array_equal(ens.get_field(0, E.SPECTRAL), spec)
    bad = spec.copy()
    bad[0, 2 * 100] = 1.0  # kx = 100 lies outside the 2/3 rectangle (85)
    with pytest.raises(Exception, match="outside"):
        ens.set_field(0, bad, E.SPECTRAL)
    with pytest.raises(Exception):
        ens.schedule = [384, 512, 0, 0, 0, 1]
    ens.close()
    with pytest.raises(Exception, match="power of two"):
        E.Ensemble([params_for(64)])
    with pytest.raises(Exception, match="RK order"):
        E.Ensemble([params_for(128), params_for(128, rko=2)])
    with pytest.raises(Exception, match="2/3"):
        E.Ensemble([params_for(128, truncate=0)])


def test_ensemble_workload_every_member_matches_solo():
    """The 64-member workload (32 x 256^2, 24 x 512^2, 8 x 1024^2, alpha
    cycling through 0.5, 1, 2, dt = 0.25/N) over its 1000 steps, each member
    against its solo run."""
    wl = ensemble_members()
    steps = wl[0].params.npas
    ens = E.from_workloads(wl)
    print("ensemble of", len(wl), "members,", steps, "steps", flush=True)
    ens.run(steps)
    ens.synchronize()
    worst = 0.0
    for m, w in enumerate(wl):
        zs, ss = E.solo_run(w.params, ic.physical_field(w.params, w.ic), w.forcing_seed, steps)
        z = ens.get_field(m, E.SPECTRAL)
        assert ens.get_seed(m) == ss
        assert np.all(np.isfinite(z))
        worst = max(worst, rel_l2(z, zs))
        assert np.array_equal(z, zs), (m, rel_l2(z, zs))
    print("worst rel L2", worst, flush=True)
    ens.close()


@pytest.mark.parametrize("cfg", [None, [256, 512, 0, 0, 3072, 10]])
def test_workload_members_of_every_size_match_oracle(oracle_mod, cfg):
    """Members of the batched-ensemble workload at 256^2, 512^2 and 1024^2
    (alpha 1/2, 1 and 2, their own ring, nu, dt and seeds) advanced together
    for 20 steps, with the default schedule and with the schedule the
    autotuner picked for the whole workload on MI355X: each member within
    1e-10 relative L2 of the oracle, forcing seeds equal."""
    members = ensemble_members()
    pick = [0, 31, 33, 34, 56, 57, 58]  # 256/0.5, 256/1, 512/0.5, 512/1, 1024/2, 1024/0.5, 1024/1
    ws = [members[m] for m in pick]
    assert sorted({w.params.n for w in ws}) == [256, 512, 1024]
    assert sorted({w.params.alpv for w in ws}) == [0.5, 1.0, 2.0]
    ps = [w.params for w in ws]
    fields = [ic.physical_field(w.params, w.ic) for w in ws]
    seeds = [w.forcing_seed for w in ws]
    nsteps = 20
    out, sd = _batched(ps, fields, seeds, nsteps, cfg)
    for k, p in enumerate(ps):
        o = oracle_mod.Oracle(p)
        z = np.zeros((p.n, p.n + 2))
        z[:, : p.n] = fields[k][:, : p.n]
        o.fft_dir(z)
        o.trunc(z)
        seed = seeds[k]
        for _ in range(nsteps):
            o.step4(z)
            seed = o.forcing(z, seed)
        err = rel_l2(out[k], z)
        print(f"member {pick[k]} N={p.n} alpha={p.alpv}: rel L2 {err:.2e}", flush=True)
        assert sd[k] == seed
        assert err < 1e-10, (pick[k], err)
