"""Python mirror of the reference's subroutine interface over the C ABI.

A :class:`Solver` is one simulation context (the module state of
src/modules.f90).  Its methods keep the reference names and argument meaning
(src/nlt.f90, src/rkstep.f90, src/fft_dir.f90, ...).  Arrays are numpy
float64 arrays of shape (N, N+2) (the Fortran array z(N+2,N)); they must be
made present with :meth:`enter_data` first, exactly as the reference's
OpenACC data region does (src/main.f90:67-83).  Device-resident arrays (torch
tensors) are attached with :meth:`map_data`.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check
from .params import Params

PLAN_D2Z = 1
PLAN_Z2D = 2

K_YSTART, K_YSTAGE, K_YFINAL, K_XPASS, K_FORCING, K_OTHER, K_YBOUND = range(7)
K_NAMES = ["y_start", "y_stage", "y_final", "x_pass", "forcing", "other", "y_bound"]


def _addr(a):
    if isinstance(a, int):
        return a
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    raise TypeError(f"cannot take the address of {type(a)}")


def _field_addr(a, n):
    """Address of a field array after checking that its memory really is the
    row-major (N, N+2) float64 array the library reads and writes: a
    Fortran-ordered array, a strided view or another dtype with enough bytes
    would otherwise be silently misread.  Raw integer addresses are taken as
    they are."""
    if isinstance(a, int):
        return a
    size = n * (n + 2)
    if isinstance(a, np.ndarray):
        if a.dtype != np.float64:
            raise TypeError(f"field arrays must be float64, got {a.dtype}")
        if not a.flags.c_contiguous:
            raise TypeError("field arrays must be C-contiguous (row-major (N, N+2))")
        if a.size != size:
            raise TypeError(f"field arrays hold N*(N+2) = {size} doubles, got shape {a.shape}")
        return a.ctypes.data
    if hasattr(a, "data_ptr"):
        import torch
        if a.dtype != torch.float64 or not a.is_contiguous() or a.numel() != size:
            raise TypeError(f"field tensors must be contiguous float64 with N*(N+2) = {size} "
                            f"elements, got {a.dtype} {tuple(a.shape)}")
        return a.data_ptr()
    raise TypeError(f"cannot take the address of {type(a)}")


class Solver:
    def __init__(self, params: Params):
        self.L = _lib.load()
        self.params = params
        self.n = int(params.n)
        h = C.c_void_p()
        arr = np.ascontiguousarray(params.as_array(), np.float64)
        check(self.L.gt_create(self.n, arr.ctypes.data, C.byref(h)))
        self.h = h
        self._keep = {}
        ii = np.zeros(7, np.int32)
        dd = np.zeros(11)
        check(self.L.gt_info(self.h, ii.ctypes.data, dd.ctypes.data))
        self.nf, self.split, self.rko, self.truncate = (int(x) for x in ii[:4])
        self.rect = tuple(int(x) for x in ii[4:7])
        (self.inpz, self.inpe, self.dt2, self.dt3, self.dt6, self.sdt, self.dkx,
         self.dky, self.rtrunc2, self.fkxsmax, self.fkysmax) = dd.tolist()
        self.plan_dir = self.plan_create(PLAN_D2Z)
        self.plan_inv = self.plan_create(PLAN_Z2D)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            check(self.L.gt_destroy(self.h))
            self.h = C.c_void_p()
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- tables -------------------------------------------------------------
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
        check(self.L.gt_get_defk(self.h, *(out[k].ctypes.data for k in
                                           ("fkx", "fky", "fkxs", "fkys", "jc", "emk",
                                            "emkx", "emky"))))
        return out

    def forcing_modes(self):
        ikx = np.zeros(max(self.nf, 1), np.int32)
        iky = np.zeros(max(self.nf, 1), np.int32)
        ff = np.zeros((max(self.nf, 1), 2))
        check(self.L.gt_get_forcing_modes(self.h, ikx.ctypes.data, iky.ctypes.data,
                                          ff.ctypes.data))
        return ikx[:self.nf], iky[:self.nf], ff[:self.nf]

    def plan_create(self, kind):
        p = C.c_int32()
        check(self.L.gt_plan_create(self.h, int(kind), C.byref(p)))
        return p.value

    # ---- data clauses -------------------------------------------------------
    def _fa(self, a):
        return _field_addr(a, self.n)

    def enter_data(self, a, copyin=True):
        check(self.L.gt_enter_data(self.h, self._fa(a), int(a.nbytes), 1 if copyin else 0))
        self._keep[self._fa(a)] = a

    def exit_data(self, a, copyout=True):
        check(self.L.gt_exit_data(self.h, self._fa(a), 1 if copyout else 0))
        self._keep.pop(self._fa(a), None)

    def update_host(self, a):
        check(self.L.gt_update_host(self.h, self._fa(a)))

    def update_device(self, a):
        check(self.L.gt_update_device(self.h, self._fa(a)))

    def map_data(self, t):
        """Attach a device tensor (torch, float64, N*(N+2) elements) as its own
        device copy; the tensor's address is the key."""
        nbytes = t.numel() * t.element_size()
        self._fa(t)
        check(self.L.gt_map_data(self.h, t.data_ptr(), t.data_ptr(), int(nbytes)))
        self._keep[t.data_ptr()] = t

    def unmap_data(self, t):
        check(self.L.gt_unmap_data(self.h, self._fa(t)))
        self._keep.pop(self._fa(t), None)

    def sync_ref(self, a):
        check(self.L.gt_sync_ref(self.h, self._fa(a)))

    def copy(self, dst, src):
        check(self.L.gt_copy(self.h, self._fa(dst), self._fa(src)))

    def synchronize(self):
        check(self.L.gt_synchronize(self.h))

    # ---- subroutines ----------------------------------------------------------
    def fft_dir(self, z, plan=None):
        check(self.L.gt_fft_dir(self.h, self._fa(z), self.plan_dir if plan is None else plan))

    def fft_inv(self, z, plan=None):
        check(self.L.gt_fft_inv(self.h, self._fa(z), self.plan_inv if plan is None else plan))

    def produ(self, z, u, v):
        check(self.L.gt_produ(self.h, self._fa(z), self._fa(u), self._fa(v)))

    def trunc(self, z):
        check(self.L.gt_trunc(self.h, self._fa(z)))

    def nlt(self, z, u, v):
        check(self.L.gt_nlt(self.h, self._fa(z), self._fa(u), self._fa(v)))

    def step(self, z, znl, u, v):
        check(self.L.gt_step(self.h, self._fa(z), self._fa(znl), self._fa(u), self._fa(v)))

    def step4(self, z, znl, znew, u, v):
        check(self.L.gt_step4(self.h, self._fa(z), self._fa(znl), self._fa(znew), self._fa(u),
                              self._fa(v)))

    def forcing(self, z, seed):
        s = C.c_int32(int(seed))
        check(self.L.gt_forcing(self.h, self._fa(z), C.byref(s)))
        return s.value

    def spectrum(self, z):
        nb = self.n // 2
        sp, sz = np.zeros(nb), np.zeros(nb)
        check(self.L.gt_spectrum(self.h, self._fa(z), sp.ctypes.data, sz.ctypes.data))
        return sp, sz

    def fluxes(self, z, znl):
        nb = self.n // 2
        fle, flz = np.zeros(nb), np.zeros(nb)
        check(self.L.gt_fluxes(self.h, self._fa(z), self._fa(znl), fle.ctypes.data,
                               flz.ctypes.data))
        return fle, flz

    def glob(self, z):
        out = np.zeros(6)
        check(self.L.gt_glob(self.h, self._fa(z), out.ctypes.data))
        return out

    def shell_index(self):
        ib = np.zeros((self.n, self.nh), np.int32)
        check(self.L.gt_shell_index(self.h, ib.ctypes.data))
        return ib

    # ---- timing ---------------------------------------------------------------
    def timing_enable(self, on=True, classes=None):
        """Event-time kernel launches on the stream; classes: names from
        K_NAMES to time (default: all)."""
        mask = -1 if classes is None else sum(1 << K_NAMES.index(k) for k in classes)
        check(self.L.gt_timing_classes(self.h, mask))
        check(self.L.gt_timing_enable(self.h, 1 if on else 0))

    def timing_read(self):
        ms = np.zeros(len(K_NAMES))
        cnt = np.zeros(len(K_NAMES), np.int32)
        check(self.L.gt_timing_read(self.h, ms.ctypes.data, cnt.ctypes.data))
        return {K_NAMES[i]: (float(ms[i]), int(cnt[i])) for i in range(len(K_NAMES))}

    # ---- the driver loop (src/main.f90:89-143) ---------------------------------
    def run(self, zphys, seed, npas, t=0.0, diag=None):
        """Run the reference's main loop on a host array holding the physical
        field; on return it holds the final physical field.  diag(ipas, sp, sz,
        fle, flz, glob) is called every imix steps.  Returns (seed, t)."""
        z, znl, znew, u, v = zphys, self.field(), self.field(), self.field(), self.field()
        self.enter_data(z)
        for a in (znl, znew, u, v):
            self.enter_data(a, copyin=False)
        try:
            self.fft_dir(z)
            if self.truncate == 1:
                self.trunc(z)
            p = self.params
            for ipas in range(1, int(npas) + 1):
                t = t + p.dt
                if self.rko == 4:
                    self.step4(z, znl, znew, u, v)
                else:
                    self.step(z, znl, u, v)
                if diag is not None and p.imix > 0 and ipas % p.imix == 0:
                    self.copy(znl, z)
                    self.nlt(znl, u, v)
                    sp, sz = self.spectrum(z)
                    fle, flz = self.fluxes(z, znl)
                    diag(ipas, sp, sz, fle, flz, self.glob(z))
                seed = self.forcing(z, seed)
            self.fft_inv(z)
            self.update_host(z)
        finally:
            for a in (znl, znew, u, v):
                self.exit_data(a, copyout=False)
            self.exit_data(z, copyout=True)
        return seed, t
