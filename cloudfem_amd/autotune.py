"""Launch-shape tuner of the batched solver.

The batched theta-method solver (cf_batch_theta_run) runs its sparse
products with L lanes per row over tiles of R rows.  The best (L, R) depends
on the row lengths and on how the rows split into systems, so the first run
on a new batch shape times the nine candidates on the device
(cf_batch_autotune) and keeps the fastest, in memory and in a small JSON file
(CLOUDFEM_AMD_TUNE_CACHE, default ~/.cache/cloudfem_amd/tune.json) keyed by
device name and batch shape.  No reference counterpart.
"""

import ctypes
import json
import os
import threading

import numpy as np
import torch

from . import _native as N
from .device import ptr, stream_handle, workspace

LANES = (4, 8, 16)
TILE_MULT = (1, 2, 4)
_mem = {}
_lock = threading.Lock()


def cache_path():
    return os.environ.get("CLOUDFEM_AMD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "cloudfem_amd",
                                       "tune.json"))


def _load():
    try:
        with open(cache_path()) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _store(key, entry):
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        data = _load()
        data[key] = entry
        tmp = path + f".{os.getpid()}.tmp"
        with open(tmp, "w") as fh:
            json.dump(data, fh, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass   # a read-only home only costs a re-tune next time


def batch_key(dev, sys_rows, nnz):
    rows = np.diff(np.asarray(sys_rows))
    return (f"{torch.cuda.get_device_name(dev)}|batch_spmv|n_sys={len(rows)}|rows={int(rows.sum())}"
            f"|nnz={int(nnz)}|min={int(rows.min())}|max={int(rows.max())}")


def batch_spmv_config(sys_rows, indptr, indices, values, retune=False, reps=20):
    """(lanes, rows_per_tile, info) for the batch; times the candidates once
    per batch shape and device."""
    sys_rows = np.ascontiguousarray(sys_rows, dtype=np.int32)
    dev = values.device
    key = batch_key(dev, sys_rows, values.numel())
    with _lock:
        if not retune:
            if key in _mem:
                return _mem[key]["lanes"], _mem[key]["rows"], _mem[key]
            disk = _load().get(key)
            if disk:
                _mem[key] = disk
                return disk["lanes"], disk["rows"], disk
        lib = N.load()
        n_sys = len(sys_rows) - 1
        nb = ctypes.c_size_t()
        N.check(lib.cf_batch_workspace_bytes(n_sys, sys_rows.ctypes.data, values.numel(),
                                             N.PREC_FP64, ctypes.byref(nb)),
                "cf_batch_workspace_bytes")
        ws = workspace(nb.value, dev)
        lanes, rows = ctypes.c_int32(), ctypes.c_int32()
        us = (ctypes.c_float * 9)()
        N.check(lib.cf_batch_autotune(n_sys, sys_rows.ctypes.data, ptr(indptr), ptr(indices),
                                      ptr(values), values.numel(), int(reps),
                                      ctypes.byref(lanes), ctypes.byref(rows), us, ptr(ws),
                                      nb.value, stream_handle()), "cf_batch_autotune")
        entry = {"lanes": lanes.value, "rows": rows.value,
                 "timings_us": {f"L{LANES[i // 3]}xR{256 // LANES[i // 3] * TILE_MULT[i % 3]}":
                                round(float(us[i]), 2) for i in range(9)}}
        _mem[key] = entry
        _store(key, entry)
        return entry["lanes"], entry["rows"], entry
