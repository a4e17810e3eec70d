"""Batched ensembles of forced runs over the C ABI (gt_ens_* of
include/gturbo.h).

The reference runs an ensemble as copies of one run directory submitted as
separate jobs (src/reset.sh:12-77, src/jobscript:14).  An :class:`Ensemble`
advances all its members -- each with its own params.dat values, N, initial
field and forcing seed -- together: one launch per pass for every member and
the RK step captured in a HIP graph.  :meth:`Ensemble.tune` runs the
library's autotuner over the pass schedules; :func:`autotune` keeps its
choice in a JSON file keyed by the ensemble's shape.
"""
from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

from . import _lib
from ._lib import check
from .params import Params

PHYSICAL = 0
SPECTRAL = 1
NCFG = 6
CFG_NAMES = ["y_threads", "x_threads", "waves_per_simd", "radix_small_first",
             "row_pair_mask", "steps_per_graph"]


class Ensemble:
    def __init__(self, members: list, seeds=None):
        self.L = _lib.load()
        self.params = list(members)
        self.n = [int(p.n) for p in self.params]
        nm = len(self.params)
        ns = np.asarray(self.n, np.int32)
        pa = np.ascontiguousarray(np.stack([p.as_array() for p in self.params]), np.float64)
        h = C.c_void_p()
        check(self.L.gt_ens_create(nm, ns.ctypes.data, pa.ctypes.data, C.byref(h)))
        self.h = h
        if seeds is not None:
            for m, s in enumerate(seeds):
                self.set_seed(m, int(s))

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            check(self.L.gt_ens_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __len__(self):
        return len(self.params)

    def set_field(self, m: int, w: np.ndarray, space: int = PHYSICAL):
        """Member m's field: a physical (N, N+2) or (N, N) array (fft_dir and
        trunc applied, as src/main.f90:89-92 does to the input frame), or a
        reference-layout spectrum with space=SPECTRAL."""
        n = self.n[m]
        z = np.zeros((n, n + 2))
        w = np.asarray(w, np.float64)
        z[:, : w.shape[1]] = w
        check(self.L.gt_ens_set_field(self.h, m, z.ctypes.data, space))

    def get_field(self, m: int, space: int = PHYSICAL) -> np.ndarray:
        n = self.n[m]
        z = np.zeros((n, n + 2))
        check(self.L.gt_ens_get_field(self.h, m, z.ctypes.data, space))
        return z

    def set_seed(self, m: int, seed: int):
        check(self.L.gt_ens_set_seed(self.h, m, int(seed)))

    def get_seed(self, m: int) -> int:
        s = C.c_int32()
        check(self.L.gt_ens_get_seed(self.h, m, C.byref(s)))
        return int(s.value)

    @property
    def schedule(self) -> list:
        c = np.zeros(NCFG, np.int32)
        check(self.L.gt_ens_schedule(self.h, c.ctypes.data, 0))
        return [int(x) for x in c]

    @schedule.setter
    def schedule(self, cfg):
        c = np.ascontiguousarray(cfg, np.int32)
        if c.shape != (NCFG,):
            raise ValueError(f"a schedule has {NCFG} entries")
        check(self.L.gt_ens_schedule(self.h, c.ctypes.data, 1))

    def tune(self, steps: int = 4):
        """Run the autotuner; returns (schedule, ms per RK step)."""
        c = np.zeros(NCFG, np.int32)
        ms = C.c_double()
        check(self.L.gt_ens_tune(self.h, int(steps), c.ctypes.data, C.byref(ms)))
        return [int(x) for x in c], float(ms.value)

    def run(self, nsteps: int):
        """nsteps x (RK step, then forcing) for every member; asynchronous."""
        check(self.L.gt_ens_run(self.h, int(nsteps)))

    def advance(self, nsteps: int, hold: bool = False):
        """As run; with hold the last step's forcing waits for the next
        advance/run (advance(0) applies it alone), so that the fields in
        between show the state the reference's driver takes diagnostics and
        snapshots of."""
        check(self.L.gt_ens_advance(self.h, int(nsteps), 1 if hold else 0))

    def synchronize(self):
        check(self.L.gt_ens_synchronize(self.h))

    def shape_key(self) -> str:
        p0 = self.params[0]
        sizes = {}
        for n in self.n:
            sizes[n] = sizes.get(n, 0) + 1
        return "rko%d:" % p0.rko + ",".join(f"{n}x{c}" for n, c in sorted(sizes.items()))


def autotune(ens: Ensemble, steps: int = 4, cache_path: str | None = None):
    """Schedule of ens from cache_path (a JSON file keyed by the ensemble's
    sizes and RK order) or from the autotuner, whose choice is then stored."""
    key = ens.shape_key()
    table = {}
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            table = json.load(f)
        if key in table:
            ens.schedule = table[key]["schedule"]
            return table[key]["schedule"], table[key]["ms_per_step"]
    cfg, ms = ens.tune(steps)
    if cache_path:
        table[key] = {"schedule": cfg, "ms_per_step": ms,
                      "fields": dict(zip(CFG_NAMES, cfg))}
        os.makedirs(os.path.dirname(os.path.abspath(cache_path)), exist_ok=True)
        with open(cache_path, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
    return cfg, ms


def from_workloads(workloads: list) -> Ensemble:
    """An ensemble of params.Workload members, each from its initial
    spectrum (gturbo_amd.ic) and forcing seed."""
    from . import ic
    ens = Ensemble([w.params for w in workloads], seeds=[w.forcing_seed for w in workloads])
    for m, w in enumerate(workloads):
        ens.set_field(m, ic.physical_field(w.params, w.ic))
    return ens


def solo_run(p: Params, w0: np.ndarray, seed: int, nsteps: int) -> tuple:
    """The plain per-run loop of src/main.f90:89-136 for one member through
    the single-run entry points: fft_dir + trunc of the input, then nsteps x
    (step or step4, forcing).  Returns the spectrum (reference layout) and
    the seed."""
    from .solver import Solver
    n = p.n
    with Solver(p) as s:
        z = np.zeros((n, n + 2))
        z[:, : w0.shape[1]] = w0
        znl = np.zeros_like(z)
        znew = np.zeros_like(z)
        u = np.zeros_like(z)
        v = np.zeros_like(z)
        for a in (z, znl, znew, u, v):
            s.enter_data(a)
        s.fft_dir(z)
        s.trunc(z)
        for _ in range(nsteps):
            if s.rko == 4:
                s.step4(z, znl, znew, u, v)
            else:
                s.step(z, znl, u, v)
            seed = s.forcing(z, seed)
        s.update_host(z)
        out = z.copy()
        for a in (z, znl, znew, u, v):
            s.exit_data(a, copyout=False)
    return out, seed
