"""Autotuner for the launch shape of the single-pass LSQR matvec kernel.

The kernel behind ``tomos_matvec_pair`` and the fused LSQR iteration keeps a
tile of rows x all n columns in registers.  Two shape parameters trade
registers (occupancy) against tile height and load width:

* ``row_groups`` G in {2, 4, 8}: lanes that share one column of a tile; a
  tile is G x 16 bytes tall per column, every lane holds ceil(n G / T)
  16-byte vectors;
* ``threads`` T in {256, 512} per block; (G, T) = (32, 1024) is the variant
  with 512-byte column runs that reads each tile twice (the second read
  from cache);
* ``blocks_per_cu``: grid size (0 = as many blocks as are resident at once).

``tune_fused`` times every candidate with HIP events on the real matrix,
installs the fastest through ``tomos_fused_config_set`` (a process-wide
table keyed by storage type and n) and returns all timings.  ``TuneCache``
keeps results in a JSON file keyed by the device, so a later process can
re-install them without timing again.  Scratch sizes never depend on the
chosen shape.
"""
import ctypes
import json
import os

import numpy as np

from . import _capi as C
from ._capi import DeviceBuffer

# largest n each (row groups, threads) shape has a kernel for
MAX_N = {(2, 256): 2048, (4, 256): 1024, (8, 256): 1024, (4, 512): 2048, (8, 512): 1024,
         (32, 1024): 2048}
BLOCKS_PER_CU = (0, 1, 2, 3, 4)


def candidates(n):
    """(row_groups, threads, blocks_per_cu) shapes with a kernel for this n."""
    out = []
    for (g, t), mx in MAX_N.items():
        if n <= mx:
            out.extend((g, t, b) for b in BLOCKS_PER_CU)
    return out


def device_key():
    cu, amaj, amin, mem = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_size_t()
    C.call("tomos_get_device_info", ctypes.byref(cu), ctypes.byref(amaj), ctypes.byref(amin),
           ctypes.byref(mem))
    return "gfx%d%d-%dcu" % (amaj.value, amin.value, cu.value)


def _time_pair(a_type, m, n, a_ptr, lda, reps, bufs, ev):
    pd, ud, td, sc, scb = bufs
    call = lambda: C.call("tomos_matvec_pair", a_type, m, n, a_ptr, lda, pd.ptr, -0.5, ud.ptr,
                          td.ptr, None, sc.ptr, scb, None)
    call()
    C.call("tomos_event_record", ev[0], None)
    for _ in range(reps):
        call()
    C.call("tomos_event_record", ev[1], None)
    f = ctypes.c_float()
    C.call("tomos_event_elapsed_ms", ctypes.byref(f), ev[0], ev[1])
    return f.value / reps


def installed(a_type, n):
    """The shape installed for (a_type, n) as a dict, or None."""
    g, b, t = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    C.call("tomos_fused_config_get", a_type, n, ctypes.byref(g), ctypes.byref(b), ctypes.byref(t))
    if g.value == 0 and b.value == 0 and t.value == 0:
        return None
    return {"row_groups": g.value, "blocks_per_cu": b.value, "threads": t.value}


def tune_fused(a_type, m, n, a_ptr, lda, reps=5, cands=None, install=True):
    """Time every candidate shape of the single-pass kernel on A (device
    pointer a_ptr: the hi plane for packed storage, lda) and install the
    fastest.  Returns (best, results); each entry is a dict with
    row_groups, threads, blocks_per_cu, ms and GBps (hi plane + vectors)."""
    rng = np.random.default_rng(0)
    pd = DeviceBuffer.from_numpy(rng.standard_normal(n))
    ud = DeviceBuffer.from_numpy(rng.standard_normal(m).astype(np.float32))
    td = DeviceBuffer(n * 8)
    scb = C.lib().tomos_matvec_pair_scratch_size(a_type, m, n)
    sc = DeviceBuffer(scb)
    ev = (ctypes.c_void_p(), ctypes.c_void_p())
    for e in ev:
        C.call("tomos_event_create", ctypes.byref(e))
    esz = 4 if a_type == C.F32 else 2
    byts = esz * m * n + 8.0 * m + 16.0 * n
    results = []
    try:
        for g, t, b in (cands or candidates(n)):
            C.call("tomos_fused_config_set", a_type, n, g, b, t)
            ms = _time_pair(a_type, m, n, a_ptr, lda, reps, (pd, ud, td, sc, scb), ev)
            results.append({"row_groups": g, "threads": t, "blocks_per_cu": b, "ms": ms,
                            "GBps": byts / (ms * 1e-3) / 1e9})
    finally:
        # drop only this (a_type, n) entry: shapes tuned for other storage
        # formats or widths stay installed
        C.call("tomos_fused_config_set", a_type, n, 0, 0, 0)
        for e in ev:
            C.lib().tomos_event_destroy(e)
    best = min(results, key=lambda r: r["ms"])
    if install:
        C.call("tomos_fused_config_set", a_type, n, best["row_groups"], best["blocks_per_cu"], best["threads"])
    return best, results


class TuneCache:
    """Tuned shapes in a JSON file: {device: {"fused/<a_type>/<n>": entry}}."""

    def __init__(self, path):
        self.path = path
        self.data = {}
        if os.path.exists(path):
            with open(path) as f:
                self.data = json.load(f)

    def put(self, a_type, n, entry):
        self.data.setdefault(device_key(), {})["fused/%d/%d" % (a_type, n)] = entry

    def get(self, a_type, n):
        return self.data.get(device_key(), {}).get("fused/%d/%d" % (a_type, n))

    def install(self):
        """Install every entry of this device; returns how many."""
        k = 0
        for key, e in self.data.get(device_key(), {}).items():
            _, a_type, n = key.split("/")
            C.call("tomos_fused_config_set", int(a_type), int(n), e["row_groups"], e["blocks_per_cu"],
                   e.get("threads", 256))
            k += 1
        return k

    def save(self):
        d = os.path.dirname(self.path)
        if d:
            os.makedirs(d, exist_ok=True)
        with open(self.path, "w") as f:
            json.dump(self.data, f, indent=1, sort_keys=True)
