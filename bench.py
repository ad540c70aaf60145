"""Benchmark: Magnus step of the IMSRG(2) flow at eMax 8 on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (eMax 8): HO orbits 2n+l <= 8 (90 orbits), 16O holes, hbar*omega =
20 MeV, synthetic seeded Hamiltonian (seed 0), normal ordered, white
generator.  A step is one iteration of magnus_decouple's loop
(IMSRG_MAGNUS.f90:100-125): MAGNUS_EXPAND, euler_step, BCH_EXPAND,
build_gs_white, mbpt2.  The timed steps are steps 1..K of the flow, on the
grouped + hipGraph path (`value`) and on the per-channel path; E0 after K
steps must agree between the two within 1e-8 MeV.  Inputs are resident in
HBM before timing starts.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "eom-imsrg_amd"))

METRIC = "Magnus step ms grouped-graph vs per-block GPU at dE0<1e-8 MeV; grad/fwd time"
FP64_MFMA_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def _parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--emax", type=int, default=8)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-grad", action="store_true")
    ap.add_argument("--grad-check", type=int, default=20,
                    help="central-difference check of dE0/dV on this many seeded two-body elements")
    ap.add_argument("--kernel-only", default=None,
                    help="launch one kernel class repeatedly and exit (for rocprofv3 counter runs)")
    return ap.parse_args()


def main() -> int:
    a = _parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the multi-GPU "
              "design is described in DESIGN.md and not built.", file=sys.stderr)
        return 2
    import numpy as np
    import torch
    from eom_imsrg import Basis, Flow, SqOp, synth, commutators as cm, _lib, workload as wl
    import ctypes as C

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 1
    dev = torch.device("cuda:0")
    b = Basis.oxygen16(a.emax)
    Hv = synth.hamiltonian_vacuum(b, seed=0, hw=20.0)
    H = SqOp.from_numpy(b, Hv, 1)
    cm.normal_order(H)
    Omega = SqOp.from_numpy(b, synth.antihermitian(b, seed=1, sigma=0.05), -1)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream().cuda_stream

    # ---- dominant kernels, timed with hipEvents on the launch stream
    RES = SqOp.zeros(b, 1)
    w1, w2 = SqOp.zeros(b), SqOp.zeros(b)
    LCC = torch.zeros(max(b.cclen, 1), dtype=torch.float64, device=dev)
    RCC = torch.zeros_like(LCC)
    WCC = torch.zeros(max(b.wcclen, 1), dtype=torch.float64, device=dev)

    def kernel_ms(which: str, path: int = 0, iters: int = 10) -> float:
        b.set_path(path)
        ms = C.c_double()
        _lib.check(_lib.lib.imsrg_profile_kernel(
            b.handle, _lib.KERNEL[which], Omega.ptr, -1, H.ptr, 1, RES.ptr, 1, w1.ptr, w2.ptr, LCC.data_ptr(),
            RCC.data_ptr(), WCC.data_ptr(), iters, C.byref(ms), C.c_void_p(stream)))
        b.set_path(0)
        return ms.value

    if a.kernel_only:
        t = kernel_ms(a.kernel_only, 0, 20)
        print(json.dumps({"kernel": a.kernel_only, "ms": t}))
        return 0

    kms = {k: kernel_ms(k) for k in ("PPPP", "PPHH_SMALL", "PH_GEMM", "PH_SCATTER", "K122", "PANDYA", "ONEBODY",
                                     "COMMUTATOR")}
    comm_plain_ms = kernel_ms("COMMUTATOR", 1, 3)

    # ---- the flow steps
    def run_steps(path: int, graphs: bool, nsteps: int, warm: int):
        b.set_path(path)
        f = Flow(b, graphs=graphs)

        def fresh():
            HS = H.clone()
            G = SqOp.zeros(b, -1)
            DG = SqOp.zeros(b, -1)
            cm.build_gs_white(HS, DG)
            return HS, G, DG

        if warm:
            HS, G, DG = fresh()
            for _ in range(warm):
                f.magnus_step(HS, G, DG, H)
        HS, G, DG = fresh()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        log = [f.magnus_step(HS, G, DG, H) for _ in range(nsteps)]
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        b.set_path(0)
        return dt / nsteps * 1e3, log

    g_ms, g_log = run_steps(0, True, a.steps, a.warmup)
    ncomm = [s[2] + s[3] - 1 for s in g_log]
    out_plain = {}
    if not a.no_plain:
        p_ms, p_log = run_steps(1, False, a.steps, min(a.warmup, 1))
        dE0 = abs(p_log[-1][0] - g_log[-1][0])
        out_plain = {"plain_ms_per_step": round(p_ms, 3), "speedup_vs_plain": round(p_ms / g_ms, 3),
                     "dE0_grouped_vs_plain_MeV": dE0, "dE0_ok": bool(dE0 < 1e-8)}

    # ---- gradient of the flowed E0 with respect to the vacuum two-body elements
    grad = None
    if not a.no_grad:
        grad = gradient_bench(b, Hv, a.grad_check)

    # ---- roofline of the dominant kernel
    cand = {
        "PPPP": ("mfma", wl.pppp_flops(b) / 1e12, "TFLOP/s", FP64_MFMA_PEAK_TFLOPS, wl.pppp_bytes(b)),
        "PH_GEMM": ("mfma", wl.ph_gemm_flops(b) / 1e12, "TFLOP/s", FP64_MFMA_PEAK_TFLOPS, wl.ph_gemm_bytes(b)),
        "PH_SCATTER": ("hbm", wl.scatter_bytes(b) / 1e9, "GB/s", HBM_PEAK_GBS, wl.scatter_bytes(b)),
    }
    dom = max(cand, key=lambda k: kms[k])
    bound, work, unit, peak, _ = cand[dom]
    achieved = work / (kms[dom] / 1e3)
    roof = {"kernel": dom, "bound": bound, "achieved": round(achieved, 3), "peak": peak, "unit": unit,
            "frac": round(achieved / peak, 4), "traffic": measured_traffic(dom, a.emax),
            "launch_ms": round(kms[dom], 4),
            "work_per_launch": work * (1e12 if unit == "TFLOP/s" else 1e9),
            "work_unit": "flop" if unit == "TFLOP/s" else "byte"}

    # ---- CPU baseline: the oracle (C++ restatement of the reference) on this host
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(b, a.emax, H, Omega, float(np.mean(ncomm)))

    F = wl.commutator_flops(b)
    res = {
        "metric": METRIC, "value": round(g_ms, 3), "unit": "ms", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(g_ms, 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"magnus_step_emax{a.emax}", "emax": a.emax, "orbits": b.norb, "holes": "16O",
                   "hw_MeV": 20.0, "seed": 0, "generator": "white", "path": "grouped+hipgraph",
                   "channels": b.nblocks, "E2": b.e2_count()},
        "commutators_per_step": [int(x) for x in ncomm],
        "E0_after_steps": g_log[-1][0],
        **out_plain,
        "commutator_ms": {"grouped": round(kms["COMMUTATOR"], 4), "plain": round(comm_plain_ms, 4)},
        "commutator_tflops": round(F / (kms["COMMUTATOR"] / 1e3) / 1e12, 3),
        "kernel_ms": {k: round(v, 4) for k, v in kms.items()},
        "grad_over_fwd": grad["grad_over_fwd"] if grad else None,
        "gradient": grad,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res))
    return 0


def gradient_bench(b, Hv, ncheck: int):
    """dE0/dV of the full flow (|mbpt2| < 1e-6, Omega checkpointed every 4
    steps): time of a plain forward flow, of the checkpointed forward plus the
    reverse sweep, and a central-difference check (h = 1e-5 MeV) on `ncheck`
    seeded hermitian two-body elements of the vacuum interaction"""
    import numpy as np
    import torch
    from eom_imsrg import Flow, SqOp, adjoint, commutators as cm

    T = torch.from_numpy(Hv).cuda()
    f = Flow(b, graphs=True)

    def plain_flow(Tv):
        H = SqOp(b, Tv.clone(), 1)
        cm.normal_order(H)
        G = SqOp.zeros(b, -1)
        log = f.magnus_decouple(H, G, ds=0.5)
        return H.E0, len(log)

    plain_flow(T)                                    # graph capture
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0, steps = plain_flow(T)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    Tb = torch.empty_like(T)
    adjoint.flow_forward(f, T, ckpt=4)
    adjoint.flow_backward(f, Tb)                     # warm-up of the reverse graphs / buffers
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0g, st = adjoint.flow_forward(f, T, ckpt=4)
    torch.cuda.synchronize()
    t_fw = time.perf_counter() - t0
    t0 = time.perf_counter()
    adjoint.flow_backward(f, Tb)
    torch.cuda.synchronize()
    t_bw = time.perf_counter() - t0
    g = Tb.cpu().numpy()
    gs = adjoint.symmetric_gradient(b, g, 1)
    out = {"steps": steps, "E0": e0, "forward_flow_s": round(t_fwd, 4),
           "checkpointed_forward_s": round(t_fw, 4), "reverse_sweep_s": round(t_bw, 4),
           "grad_over_fwd": round((t_fw + t_bw) / t_fwd, 3), "reverse_over_fwd": round(t_bw / t_fwd, 3),
           "checkpoint_every": 4, "E0_forward_vs_gradient_run": abs(e0 - e0g)}
    if ncheck > 0:
        # elements whose derivative stands clear of the difference quotient's
        # noise floor: the flowed E0 (~700 MeV) carries ~1e-12 MeV of
        # rounding, so (E+ - E-)/2h is good to about 1e-7 absolute
        rng = np.random.default_rng(2024)
        two0 = b.off_fpp + b.npart * b.npart
        cand = np.nonzero(np.abs(gs[two0:]) > 1e-2)[0] + two0
        pick = rng.choice(cand, size=min(ncheck, len(cand)), replace=False)
        mirror = mirror_index(b)
        h = 1e-5
        worst_rel, worst_abs, ok = 0.0, 0.0, True
        for x in pick:
            d = torch.zeros_like(T)
            d[int(x)] = 1.0
            if mirror[x] >= 0:
                d[int(mirror[x])] = 1.0
            ep, _ = plain_flow(T + h * d)
            em, _ = plain_flow(T - h * d)
            fd = (ep - em) / (2 * h)
            an = float(gs[x])
            ok = ok and abs(fd - an) <= 1e-6 * abs(an) + 2e-7
            worst_rel = max(worst_rel, abs(fd - an) / abs(an))
            worst_abs = max(worst_abs, abs(fd - an))
        out["fd_check"] = {"elements": int(len(pick)), "candidates_abs_grad_gt_1e-2": int(len(cand)),
                           "h_MeV": h, "max_rel_err": worst_rel, "max_abs_err": worst_abs,
                           "criterion": "|fd - grad| <= 1e-6 |grad| + 2e-7", "pass": bool(ok)}
    return out


def mirror_index(b):
    """flat index of the hermitian partner of every two-body element of a
    square block (-1 elsewhere or on the diagonal)"""
    import numpy as np
    m = -np.ones(b.oplen, np.int64)
    for q in range(b.nblocks):
        for g in (0, 3, 4):
            n1, _ = b.block_shape(q, g)
            if n1 == 0:
                continue
            o = int(b.goff[q, g])
            ix, jx = np.meshgrid(np.arange(n1), np.arange(n1), indexing="ij")
            src = o + ix + jx * n1
            dst = o + jx + ix * n1
            sel = ix != jx
            m[src[sel]] = dst[sel]
    return m


def measured_traffic(kernel: str, emax: int):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE passes scaled by a copy kernel of known
    size), or None when no summary covers this kernel and basis"""
    p = os.path.join(ROOT, "profiles", f"hbm_traffic_emax{emax}.md")
    if not os.path.exists(p):
        return None
    for line in open(p):
        f = line.split()
        if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel:
            return float(f[2])
    return None


def cpu_baseline(b, emax, H, Omega, comm_per_step):
    """the oracle's [Omega, H] at this eMax on the host cores, scaled to one
    Magnus step by the measured nested commutators per step"""
    import numpy as np
    sys.path.insert(0, ROOT)
    from oracle.oracle import OracleBasis, use_host_blas, lib
    nthreads = max(1, min(16, len(os.sched_getaffinity(0))))
    lib().orc_set_threads(nthreads)
    blas = use_host_blas()
    o = OracleBasis(b.n, b.l, b.jj, b.tz, b.hole)
    Hh, Oh = H.numpy(), Omega.numpy()
    t0 = time.perf_counter()
    o.commutator(Oh, -1, Hh, 1, 1)
    dt = time.perf_counter() - t0
    return {"value": round(dt * 1e3 * comm_per_step, 1), "unit": "ms", "cores": nthreads, "kind": "port",
            "sample": f"one [Omega,H] commutator at eMax {emax} on the CPU oracle "
                      f"({dt * 1e3:.0f} ms, OpenMP + {'OpenBLAS dgemm' if blas else 'own dgemm'}) x "
                      f"{comm_per_step:.2f} nested commutators per Magnus step"}


if __name__ == "__main__":
    sys.exit(main())
