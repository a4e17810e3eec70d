"""ctypes front end of the CPU oracle (oracle/gturbo_oracle.c).

TEST INFRASTRUCTURE ONLY: the oracle is the checker.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module;
the product package never does.

Arrays use the reference layout (src/modules.f90:4-9): a field is a float64
numpy array of shape (N, N+2) in C order, i.e. the Fortran array z(N+2, N)
with x (or the interleaved re/im of kx) fastest.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "libgturbo_oracle.so")
_lib = None

_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_I = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build(force=False):
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(_SO):
        subprocess.check_call(["make", "-s", "-B" if force else "all"], cwd=_HERE)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        L.or_create.restype = C.c_void_p
        L.or_create.argtypes = [C.c_int, _D]
        L.or_destroy.argtypes = [C.c_void_p]
        L.or_last_error.restype = C.c_char_p
        L.or_set_threads.argtypes = [C.c_int]
        L.or_get_scalars.argtypes = [C.c_void_p, _D, _I]
        L.or_get_defk.argtypes = [C.c_void_p, _D, _D, _D, _D, _I, _D, _D, _D]
        L.or_get_forcing_modes.argtypes = [C.c_void_p, _I, _I, _D]
        L.or_rann.restype = C.c_float
        L.or_rann.argtypes = [C.POINTER(C.c_int32)]
        for name in ("or_fft_dir", "or_fft_inv", "or_trunc"):
            getattr(L, name).argtypes = [C.c_void_p, _D]
        L.or_produ.argtypes = [C.c_void_p, _D, _D, _D]
        L.or_nlt.argtypes = [C.c_void_p, _D, _D, _D]
        L.or_step.argtypes = [C.c_void_p, _D, _D, _D, _D]
        L.or_step4.argtypes = [C.c_void_p, _D, _D, _D, _D, _D]
        L.or_forcing.argtypes = [C.c_void_p, _D, C.POINTER(C.c_int32)]
        L.or_nlt_tl.argtypes = [C.c_void_p, _D, _D, _D]
        L.or_step4_tl.argtypes = [C.c_void_p, _D, _D]
        L.or_step_tl.argtypes = [C.c_void_p, _D, _D]
        L.or_forcing_phases.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.c_int, _D]
        L.or_spectrum.argtypes = [C.c_void_p, _D, _D, _D]
        L.or_fluxes.argtypes = [C.c_void_p, _D, _D, _D, _D]
        L.or_glob.argtypes = [C.c_void_p, _D, _D]
        L.or_shell_index.argtypes = [C.c_void_p, _I]
        L.or_run.argtypes = [C.c_void_p, _D, C.POINTER(C.c_int32), C.c_int,
                             C.POINTER(C.c_double), C.c_void_p]
        _lib = L
    return _lib


def set_threads(n):
    lib().or_set_threads(int(n))


def rann(seed):
    """One draw of the reference LCG (src/forcing.f90:40-51); returns
    (value as float32, new seed)."""
    s = C.c_int32(int(seed))
    r = lib().or_rann(C.byref(s))
    return np.float32(r), s.value


class Oracle:
    """One simulation context: the state of the reference's modules after
    inieuler (src/inieuler.f90:13-88)."""

    def __init__(self, params):
        self.p = params
        self.n = int(params.n)
        arr = np.ascontiguousarray(params.as_array(), dtype=np.float64)
        h = lib().or_create(self.n, arr)
        if not h:
            raise RuntimeError(lib().or_last_error().decode())
        self.h = h
        sc = np.zeros(11)
        isc = np.zeros(4, np.int32)
        lib().or_get_scalars(h, sc, isc)
        (self.dkx, self.dky, self.dt2, self.dt3, self.dt6, self.sdt, self.rtrunc2,
         self.fkxsmax, self.fkysmax, self.inpz, self.inpe) = sc.tolist()
        self.nf, self.split, self.rko, self.truncate = (int(x) for x in isc)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().or_destroy(self.h)
                self.h = None
        except Exception:
            pass

    # shapes -----------------------------------------------------------
    @property
    def nh(self):
        return self.n // 2 + 1

    def field(self):
        return np.zeros((self.n, self.n + 2), np.float64)

    def defk(self):
        n, nh = self.n, self.nh
        out = dict(fkx=np.zeros(nh), fky=np.zeros(n), fkxs=np.zeros(nh),
                   fkys=np.zeros(n), jc=np.zeros(n, np.int32),
                   emk=np.zeros((n, nh)), emkx=np.zeros(nh), emky=np.zeros(n))
        lib().or_get_defk(self.h, out["fkx"], out["fky"], out["fkxs"], out["fkys"],
                          out["jc"], out["emk"], out["emkx"], out["emky"])
        return out

    def forcing_modes(self):
        ikx = np.zeros(self.nf, np.int32)
        iky = np.zeros(self.nf, np.int32)
        ff = np.zeros((self.nf, 2))
        if self.nf:
            lib().or_get_forcing_modes(self.h, ikx, iky, ff)
        return ikx, iky, ff

    # subroutines (in place on reference-layout arrays) ---------------
    def fft_dir(self, z):
        lib().or_fft_dir(self.h, z)

    def fft_inv(self, z):
        lib().or_fft_inv(self.h, z)

    def produ(self, z, u, v):
        lib().or_produ(self.h, z, u, v)

    def trunc(self, z):
        lib().or_trunc(self.h, z)

    def nlt(self, z, u=None, v=None):
        u = self.field() if u is None else u
        v = self.field() if v is None else v
        lib().or_nlt(self.h, z, u, v)

    def step(self, z, znl=None, u=None, v=None):
        f = self.field
        lib().or_step(self.h, z, f() if znl is None else znl,
                      f() if u is None else u, f() if v is None else v)

    def step4(self, z, znl=None, znew=None, u=None, v=None):
        f = self.field
        lib().or_step4(self.h, z, f() if znl is None else znl,
                       f() if znew is None else znew,
                       f() if u is None else u, f() if v is None else v)

    def nlt_tl(self, z, dz):
        """N'(z) dz (tangent of nlt at z)."""
        out = self.field()
        lib().or_nlt_tl(self.h, z, dz, out)
        return out

    def step_tl(self, z, dz):
        """One step of z (step or step4 by the RK order) and of its tangent
        dz, both in place."""
        if self.rko == 4:
            lib().or_step4_tl(self.h, z, dz)
        else:
            lib().or_step_tl(self.h, z, dz)

    def forcing(self, z, seed):
        s = C.c_int32(int(seed))
        lib().or_forcing(self.h, z, C.byref(s))
        return s.value

    def forcing_phases(self, seed, nsteps):
        s = C.c_int32(int(seed))
        cs = np.zeros((max(nsteps, 0), max(self.nf, 0), 2))
        if nsteps and self.nf:
            lib().or_forcing_phases(self.h, C.byref(s), int(nsteps), cs)
        return cs, s.value

    def spectrum(self, z):
        nb = self.n // 2
        sp, sz = np.zeros(nb), np.zeros(nb)
        lib().or_spectrum(self.h, z, sp, sz)
        return sp, sz

    def fluxes(self, z, znl):
        nb = self.n // 2
        fle, flz = np.zeros(nb), np.zeros(nb)
        lib().or_fluxes(self.h, z, znl, fle, flz)
        return fle, flz

    def glob(self, z):
        out = np.zeros(6)
        lib().or_glob(self.h, z, out)
        return out

    def shell_index(self):
        ib = np.zeros((self.n, self.nh), np.int32)
        lib().or_shell_index(self.h, ib)
        return ib

    def run(self, zphys, seed, npas, t=0.0, with_diag=False):
        """main.f90:89-143 without file output; zphys is overwritten with the
        final physical field.  Returns (seed, t, diag or None)."""
        s = C.c_int32(int(seed))
        tt = C.c_double(t)
        diag = None
        ptr = None
        if with_diag and self.p.imix > 0:
            nd = npas // self.p.imix
            diag = np.zeros((max(nd, 1), 4 * (self.n // 2) + 6))
            ptr = diag.ctypes.data
        lib().or_run(self.h, zphys, C.byref(s), int(npas), C.byref(tt), ptr)
        return s.value, tt.value, diag
