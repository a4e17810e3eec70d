"""CPU checks of the index algebra the HIP transforms rely on (no GPU):

* the padded LDS index splits into a per-thread base plus compile-time
  offsets for every Stockham pass and for the mirror read of the
  half-spectrum split (`lpad_off`, `mirror_at` in the kernels);
* the in-place radix-8 transforms of a 4096-point line (`fft_dif`,
  `fft_dit`): a step-by-step simulation of their passes, with the same LDS
  positions, twiddles and thread/register ownership, reproduces the DFT,
  DIF leaving X[D(j) + 512 r] in thread j, register r, and DIT taking that
  order back to natural order; passes 2-4 touch only the wave's own run;
* the row-pair block layout of the inverse-y-pass output at N = 4096
  (`y1_off`).
"""
import numpy as np
import pytest


def lpad(e):
    return e + (e >> 4)


def lpad_off(c):
    return c + c // 16


def _passes(logn, sched):
    rl, p8 = 1 << (logn % 3), logn // 3
    if sched == 0 or rl == 1:
        return [(8, 8 ** i) for i in range(p8)] + ([(rl, 8 ** p8)] if rl > 1 else [])
    return [(rl, 1)] + [(8, rl * 8 ** i) for i in range(p8)]


@pytest.mark.parametrize("logn", range(4, 14))
def test_stockham_lds_offsets_split(logn):
    n, tpl = 1 << logn, (1 << logn) // 8
    for sched in (0, 1):
        for radix, ns in _passes(logn, sched):
            nb = 8 // radix
            for j in range(tpl):
                for b in range(nb):
                    jj = j + tpl * b
                    base = (jj // ns) * ns * radix + jj % ns
                    for r in range(radix):
                        assert lpad(base + r * ns) == lpad(base) + lpad_off(r * ns)
    for j in range(tpl):
        for r in range(8):
            assert lpad(j + tpl * r) == lpad(j) + lpad_off(tpl * r)
        if tpl % 16 == 0:  # mirror_at's split form
            pm = lpad(tpl - j)
            for r in range(5):
                want = lpad((n - j - tpl * r) & (n - 1))
                got = (0 if j == 0 else pm + lpad_off(7 * tpl)) if r == 0 else \
                    pm + lpad_off(tpl * (7 - r))
                assert got == want


N, T = 4096, 512


def ip_pad(e):
    return e + (e >> 3)


def D(j):
    return 64 * (j & 7) + 8 * ((j >> 3) & 7) + (j >> 6)


def _dft8(t, s):
    k = np.arange(8)
    return np.exp(s * 2j * np.pi * np.outer(k, k) / 8) @ t


def _tw(t, k, s):
    return t * np.exp(s * 2j * np.pi * k * np.arange(8) / N)


def _runs(j):
    """LDS positions of thread j's runs in passes 2, 3, 4 (element indices)."""
    b2 = (j >> 6) * 512 + (j & 63)
    b3 = (j >> 3) * 64 + (j & 7)
    return ([ip_pad(b2) + 72 * r for r in range(8)], [ip_pad(b3) + 9 * r for r in range(8)],
            [9 * j + r for r in range(8)])


def _dif(a, s):
    buf = np.zeros(N + N // 8, complex)
    a = a.copy()
    for j in range(T):
        a[j] = _tw(_dft8(a[j], s), j, s)
        for r in range(8):
            buf[ip_pad(j) + 576 * r] = a[j][r]
    for j in range(T):
        p2 = _runs(j)[0]
        buf[p2] = _tw(_dft8(buf[p2], s), 8 * (j & 63), s)
    for j in range(T):
        p3 = _runs(j)[1]
        buf[p3] = _tw(_dft8(buf[p3], s), 64 * (j & 7), s)
    for j in range(T):
        a[j] = _dft8(buf[_runs(j)[2]], s)
    return a


def _dit(a, s):
    buf = np.zeros(N + N // 8, complex)
    a = a.copy()
    for j in range(T):
        buf[_runs(j)[2]] = _dft8(a[j], s)
    for j in range(T):
        p3 = _runs(j)[1]
        buf[p3] = _dft8(_tw(buf[p3], 64 * (j & 7), s), s)
    for j in range(T):
        p2 = _runs(j)[0]
        buf[p2] = _dft8(_tw(buf[p2], 8 * (j & 63), s), s)
    for j in range(T):
        a[j] = _dft8(_tw(buf[[ip_pad(j) + 576 * r for r in range(8)]], j, s), s)
    return a


@pytest.mark.parametrize("s", [+1, -1])
def test_in_place_transforms_reproduce_the_dft(s):
    x = np.random.default_rng(3).standard_normal(N) + 1j * np.random.default_rng(4).standard_normal(N)
    ref = np.fft.ifft(x) * N if s > 0 else np.fft.fft(x)
    nat = np.array([[x[j + 512 * r] for r in range(8)] for j in range(T)])
    out = _dif(nat, s)
    want = np.array([[ref[D(j) + 512 * r] for r in range(8)] for j in range(T)])
    assert np.abs(out - want).max() < 1e-9 * np.abs(ref).max()
    perm = np.array([[x[D(j) + 512 * r] for r in range(8)] for j in range(T)])
    out2 = _dit(perm, s)
    want2 = np.array([[ref[j + 512 * r] for r in range(8)] for j in range(T)])
    assert np.abs(out2 - want2).max() < 1e-9 * np.abs(ref).max()


def test_in_place_layout_and_ownership():
    pads = [ip_pad(e) for e in range(N)]
    assert len(set(pads)) == N and max(pads) < N + N // 8
    assert all(D(D(j)) == j for j in range(T))  # digit reversal is an involution
    for j in range(T):
        w = j >> 6
        p2, p3, p4 = _runs(j)
        # pass 2: the wave's own 512 elements; passes 3, 4: one 64-element run
        # written and read by the same eight lanes
        lo = ip_pad(512 * w)
        assert all(lo <= p < lo + 576 for p in p2)
        run = ip_pad(64 * (j >> 3))
        assert all(run <= p < run + 72 for p in p3 + p4)


def _tw8192(m, s):
    return np.exp(s * 2j * np.pi * np.asarray(m) / 8192)


def _dif2(a, s):
    """fft_dif2: radix 2 on the thread's elements t + 1024 r, then 4096-point
    fft_dif on each half of the threads."""
    y = np.zeros((2, 4096), complex)
    for t in range(1024):
        for r in range(4):
            n = t + 1024 * r
            y[0][n] = a[t][r] + a[t][r + 4]
            y[1][n] = (a[t][r] - a[t][r + 4]) * _tw8192(n, s)
    out = np.zeros((1024, 8), complex)
    for h in range(2):
        nat = np.array([[y[h][j + 512 * r] for r in range(8)] for j in range(512)])
        out[512 * h:512 * h + 512] = _dif(nat, s)
    return out


def _dit2(a, s):
    y = np.zeros((2, 4096), complex)
    for h in range(2):
        nat = _dit(a[512 * h:512 * h + 512], s)
        for j in range(512):
            y[h][j + 512 * np.arange(8)] = nat[j]
    out = np.zeros((1024, 8), complex)
    for t in range(1024):
        for r in range(4):
            n = t + 1024 * r
            y1 = y[1][n] * _tw8192(n, s)
            out[t][r], out[t][r + 4] = y[0][n] + y1, y[0][n] - y1
    return out


@pytest.mark.parametrize("s", [+1, -1])
def test_split_8192_transforms_reproduce_the_dft(s):
    """fft_dif2 leaves thread (h, j), register r with X[2 (D(j) + 512 r) + h];
    fft_dit2 takes that order back to natural order."""
    n = 8192
    rng = np.random.default_rng(9)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = np.fft.ifft(x) * n if s > 0 else np.fft.fft(x)
    order = np.array([[2 * (D(t & 511) + 512 * r) + (t >> 9) for r in range(8)]
                      for t in range(1024)])
    nat = np.array([[x[t + 1024 * r] for r in range(8)] for t in range(1024)])
    assert np.abs(_dif2(nat, s) - ref[order]).max() < 1e-9 * np.abs(ref).max()
    natural = np.arange(n).reshape(8, 1024).T  # [t][r] = t + 1024 r
    assert np.abs(_dit2(x[order], s) - ref[natural]).max() < 1e-9 * np.abs(ref).max()


def y1_cols(kc):
    return (kc + 3) // 4 * 4


def y1_off(y, c, cols):
    """Element (row y, column c) of one y1 field at N = 4096 (blocks of 2
    rows x 4 columns, the two rows of a column adjacent)."""
    return (y >> 1) * (2 * cols) + (c // 4) * 8 + (c % 4) * 2 + (y & 1)


@pytest.mark.parametrize("kc", [1366, 1365, 1024, 683])
def test_row_pair_blocks_of_the_y1_fields(kc):
    """The y1 layout at 4096: a bijection onto one field; a row pair's four
    consecutive columns fill one aligned 128-byte line (8 elements), so the
    x pass reads whole lines; the y pass's lanes 2i, 2i+1 (rows s, s+1 of a
    column) write 32 contiguous bytes; the two columns of a tile fill half
    of each line they touch; and every column a tile writes fits the
    allocation (tiles of two columns, so up to kc + 1 columns)."""
    n, cols = 4096, y1_cols(kc)
    y, c = np.meshgrid(np.arange(n), np.arange(cols), indexing="ij")
    off = y1_off(y, c, cols)
    assert np.array_equal(np.sort(off.ravel()), np.arange(n * cols))
    for y0 in (0, 2, 2046, 4094):
        for c0 in (0, 4, cols - 4):
            line = {y1_off(y0 + dy, c0 + dc, cols) for dy in (0, 1) for dc in range(4)}
            lo = min(line)
            assert lo % 8 == 0 and line == set(range(lo, lo + 8))
    s = np.arange(0, n, 2)
    for col in (0, 1, 5, kc - 1):
        assert np.all(y1_off(s + 1, col, cols) == y1_off(s, col, cols) + 1)
    for t in (0, 1, (kc - 1) // 2):
        tile = {y1_off(y0, 2 * t + b, cols) // 8 * 8 for y0 in (0, 1) for b in (0, 1)}
        assert len(tile) == 1
        half = {y1_off(y0, 2 * t + b, cols) for y0 in (0, 1) for b in (0, 1)}
        assert len(half) == 4
    ntiles = (kc + 1) // 2
    assert 2 * ntiles <= cols  # the last tile's second column is inside the field
