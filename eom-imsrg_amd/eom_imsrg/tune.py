"""GEMM tile autotuner: per basis and device, time every tile shape of each
grouped MFMA GEMM family (64x64, 128x64, 128x128, 64x32, 32x64, 32x32,
each with and without XCD packing) with hipEvents on the launch stream, keep the fastest, and
persist the choice as JSON so later runs apply it without timing.

Families (include/imsrg_hip.h, IMSRG_GEMM_*): PPPP (222_pp_hh L1 R1), PPX
(pp x (ph|hh)), XX ((ph|hh)^T (ph|hh)), PH (222_ph W').  Each family is
timed through the whole commutator (IMSRG_K_COMMUTATOR), one family at a
time with the others held at their current choice: a tile shape changes
how the family's output meets the kernels that read it (the pppp epilogue
reads the GEMM's tiles), so the family alone is not the objective.

The cache file is EOM_IMSRG_TUNE_FILE if set, else
~/.cache/eom_imsrg/gemm_tiles.json; the entries measured on MI355X ship in
tuned_mi355x.json beside this module and are read when the cache has none.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os

import numpy as np
import torch

from . import _lib
from .basis import Basis

FAMILIES = {"PPPP": ("PPPP", "COMMUTATOR"), "PPX": ("PPX", "COMMUTATOR"), "XX": ("XX", "COMMUTATOR"),
            "PH": ("PH", "COMMUTATOR")}
XCD = 8
CANDIDATES = [t | x for x in (0, XCD) for t in range(8)]
SHIPPED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_mi355x.json")


def cache_path() -> str:
    return os.environ.get("EOM_IMSRG_TUNE_FILE",
                          os.path.join(os.path.expanduser("~"), ".cache", "eom_imsrg", "gemm_tiles.json"))


def device_tag() -> str:
    """the GPU's architecture (e.g. gfx950) and compute-unit count"""
    if not torch.cuda.is_available():
        return "cpu"
    p = torch.cuda.get_device_properties(0)
    arch = (getattr(p, "gcnArchName", "") or p.name or "gpu").split(":")[0]
    return f"{arch}x{p.multi_processor_count}"


def basis_key(b: Basis) -> str:
    """device tag + a digest of the orbit table and hole set"""
    h = hashlib.sha1()
    for a in (b.n, b.l, b.jj, b.tz, b.hole):
        h.update(np.ascontiguousarray(np.asarray(a, np.int64)).tobytes())
    return f"{device_tag()}|norb={b.norb}|{h.hexdigest()[:16]}"


def _load(path: str) -> dict:
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def lookup(b: Basis):
    """the stored choice for this basis and device, or None"""
    k = basis_key(b)
    for path in (cache_path(), SHIPPED):
        d = _load(path)
        if k in d:
            return d[k]
    return None


def apply(b: Basis, choice: dict) -> None:
    for fam, tile in choice.items():
        _lib.check(_lib.lib.imsrg_set_gemm_tile(b.handle, _lib.GEMM[FAMILIES[fam][0]], int(tile)))


def current(b: Basis) -> dict:
    return {fam: int(_lib.lib.imsrg_get_gemm_tile(b.handle, _lib.GEMM[FAMILIES[fam][0]])) for fam in FAMILIES}


def measure(b: Basis, iters: int = 10) -> tuple[dict, dict]:
    """time every candidate of every family; returns (best choice,
    {family: {tile: ms}}).  Leaves the best choice applied."""
    from . import synth
    from .operators import SqOp
    dev = torch.device("cuda:0")
    L = SqOp.from_numpy(b, synth.antihermitian(b, seed=1, sigma=0.05), -1)
    R = SqOp.from_numpy(b, synth.hermitian_random(b, seed=2, sigma=1.0), 1)
    RES, w1, w2 = SqOp.zeros(b, 1), SqOp.zeros(b), SqOp.zeros(b)
    LCC = torch.zeros(max(b.cclen, 1), dtype=torch.float64, device=dev)
    RCC = torch.zeros_like(LCC)
    WCC = torch.zeros(max(b.wcclen, 1), dtype=torch.float64, device=dev)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)

    def kernel_ms(which: str) -> float:
        ms = C.c_double()
        _lib.check(_lib.lib.imsrg_profile_kernel(b.handle, _lib.KERNEL[which], L.ptr, -1, R.ptr, 1, RES.ptr, 1,
                                                 w1.ptr, w2.ptr, LCC.data_ptr(), RCC.data_ptr(), WCC.data_ptr(),
                                                 int(iters), C.byref(ms), stream))
        return ms.value

    choice = current(b)
    table = {}
    for fam, (gname, kname) in FAMILIES.items():
        times = {}
        for tile in CANDIDATES:
            _lib.check(_lib.lib.imsrg_set_gemm_tile(b.handle, _lib.GEMM[gname], tile))
            kernel_ms(kname)                      # warm-up
            times[tile] = kernel_ms(kname)
        best = min(times, key=times.get)
        _lib.check(_lib.lib.imsrg_set_gemm_tile(b.handle, _lib.GEMM[gname], best))
        choice[fam] = best
        table[fam] = {str(t): round(v, 5) for t, v in times.items()}
    return choice, table


def tune(b: Basis, iters: int = 10, force: bool = False, persist: bool = True) -> dict:
    """apply the stored choice for this basis and device, or measure, apply
    and store one; returns {"choice": ..., "source": "cache" | "measured",
    "times_ms": ... (measured only)}"""
    if not force:
        c = lookup(b)
        if c is not None:
            apply(b, c)
            return {"choice": c, "source": "cache", "key": basis_key(b)}
    choice, table = measure(b, iters)
    if persist:
        path = cache_path()
        d = _load(path)
        d[basis_key(b)] = choice
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as fh:
            json.dump(d, fh, indent=1, sort_keys=True)
        os.replace(tmp, path)
    return {"choice": choice, "source": "measured", "times_ms": table, "key": basis_key(b)}
