"""Benchmark: Magnus step of the IMSRG(2) flow at eMax 8 on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (eMax 8): HO orbits 2n+l <= 8 (90 orbits), 16O holes, hbar*omega =
20 MeV, synthetic seeded Hamiltonian (seed 0), normal ordered, white
generator.  A step is one iteration of magnus_decouple's loop
(IMSRG_MAGNUS.f90:100-125): MAGNUS_EXPAND, euler_step, BCH_EXPAND,
build_gs_white, mbpt2.  The timed steps are steps of flows run until
|E0(s) - E0(s - ds)| < 1e-8 MeV: after a flow's last step the next step
starts the flow again from Omega = 0 (the reset is inside the timing), so
no timed step lies past convergence.  They run on the grouped + hipGraph
path (`value`) and on the per-channel path; E0 after K steps must agree
between the two within 1e-8 MeV.  Inputs are resident in HBM before timing
starts.  Whole converged flows (|mbpt2| < 1e-6, the reference's criterion,
and |dE0| < 1e-8 MeV) are timed as their own fields.
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
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--emax", type=int, default=8)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-grad", action="store_true")
    ap.add_argument("--grad-check", type=int, default=200,
                    help="central-difference check of dE0/dV on this many seeded two-body elements")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-instance eMax 6 batch")
    ap.add_argument("--batch-cpu-check", type=int, default=1,
                    help="instances of the batch whose flowed E0 is checked against the oracle on the host")
    ap.add_argument("--batch-seq", type=int, default=8,
                    help="single-instance flows timed one after another for the batch's comparison (0: none)")
    ap.add_argument("--no-emax10", action="store_true", help="skip the eMax 10 grouped + hipGraph steps")
    ap.add_argument("--no-setup", action="store_true",
                    help="skip the Hamiltonian set-up timings (me2j reader, HO one-body part, HO / HF normal ordering)")
    ap.add_argument("--no-tune", action="store_true", help="keep the default GEMM tiles (no autotuner)")
    ap.add_argument("--no-tensor", action="store_true", help="skip the rank-2 tensor BCH leg")
    ap.add_argument("--leg", choices=["gradient", "batch", "emax10", "setup", "tensor"], default=None,
                    help="run only this part and print its JSON object (for profiler runs)")
    ap.add_argument("--kernel-only", default=None,
                    help="launch one kernel class repeatedly and exit (for rocprofv3 counter runs)")
    ap.add_argument("--steps-only", action="store_true",
                    help="only the warm-up and timed grouped steps (for rocprofv3 runs of the Magnus step)")
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
    from eom_imsrg import tune
    if a.leg == "batch":
        print(json.dumps(batch_bench(64, 6, a.batch_seq, a.batch_cpu_check)))
        return 0
    if a.leg == "emax10":
        print(json.dumps(emax10_bench(3, 1, not a.no_tune)))
        return 0
    if a.leg == "setup":
        print(json.dumps(setup_bench(a.emax)))
        return 0
    if a.leg == "tensor":
        print(json.dumps(tensor_bench(a.emax, not a.no_cpu)))
        return 0
    b = Basis.oxygen16(a.emax)
    tuned = None if a.no_tune else tune.tune(b)
    if a.leg == "gradient":
        print(json.dumps(gradient_bench(b, synth.hamiltonian_vacuum(b, seed=0, hw=20.0), a.grad_check)))
        return 0
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

    kms, comm_plain_ms = {}, None
    if not a.steps_only:
        kernel_ms("COMMUTATOR", 0, 20)   # untimed: every class's first launches and the clocks settle
        kms = {k: kernel_ms(k) for k in ("PPPP", "PPHH_SMALL", "PH_GEMM", "PH_SCATTER", "K122", "PANDYA",
                                         "ONEBODY", "COMMUTATOR")}
        if not a.no_plain:
            comm_plain_ms = kernel_ms("COMMUTATOR", 1, 3)

    # ---- the flow steps: flows to |dE0| < 1e-8 MeV, restarted from Omega = 0 after convergence
    DE0_CRIT = 1e-8

    def run_steps(path: int, graphs: bool, nsteps: int, warm: int):
        b.set_path(path)
        f = Flow(b, graphs=graphs)
        HS, G, DG = SqOp.zeros(b, 1), SqOp.zeros(b, -1), SqOp.zeros(b, -1)
        st = {"Eold": 0.0, "flow_steps": 0, "flows": 0}

        def restart():
            HS.data.copy_(H.data)
            G.data.zero_()
            cm.build_gs_white(HS, DG)
            st["Eold"], st["flow_steps"] = E0_bare, 0

        def step():
            out = f.magnus_step(HS, G, DG, H)
            st["flow_steps"] += 1
            if abs(out[0] - st["Eold"]) < DE0_CRIT:
                st["flows"] += 1
                restart()
            else:
                st["Eold"] = out[0]
            return out

        restart()
        for _ in range(warm):
            step()
        restart()
        st["flows"] = 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        log = [step() for _ in range(nsteps)]
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        b.set_path(0)
        return dt / nsteps * 1e3, log, st["flows"]

    def converged_flow(path: int, graphs: bool, crit_mode: int, crit_tol: float):
        b.set_path(path)
        f = Flow(b, graphs=graphs)
        HS, G = H.clone(), SqOp.zeros(b, -1)
        if graphs:
            f.magnus_decouple(HS, G, crit_mode=crit_mode, crit_tol=crit_tol, maxsteps=100)   # graph capture
            HS, G = H.clone(), SqOp.zeros(b, -1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        log = f.magnus_decouple(HS, G, crit_mode=crit_mode, crit_tol=crit_tol, maxsteps=100)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        b.set_path(0)
        return {"s": round(dt, 4), "steps": int(len(log)), "E0": float(log[-1, 1]),
                "commutators": int(log[:, 3].sum()), "ms_per_step": round(dt / len(log) * 1e3, 3)}

    E0_bare = H.E0
    g_ms, g_log, g_flows = run_steps(0, True, a.steps, a.warmup)
    ncomm = [s[2] + s[3] - 1 for s in g_log]
    if a.steps_only:
        print(json.dumps({"ms_per_step": round(g_ms, 3), "steps": a.steps, "commutators_per_step": ncomm,
                          "E0_after_steps": g_log[-1][0]}))
        return 0
    flows = {"grouped_mbpt2_1e-6": converged_flow(0, True, 0, float(np.float32(1e-6))),
             "grouped_dE0_1e-8": converged_flow(0, True, 1, DE0_CRIT)}
    out_plain = {}
    if not a.no_plain:
        p_ms, p_log, _ = run_steps(1, False, a.steps, min(a.warmup, 1))
        dE0 = abs(p_log[-1][0] - g_log[-1][0])
        out_plain = {"plain_ms_per_step": round(p_ms, 3), "speedup_vs_plain": round(p_ms / g_ms, 3),
                     "dE0_grouped_vs_plain_MeV": dE0, "dE0_ok": bool(dE0 < 1e-8)}
        flows["plain_dE0_1e-8"] = converged_flow(1, False, 1, DE0_CRIT)
        flows["speedup_vs_plain_dE0_1e-8"] = round(flows["plain_dE0_1e-8"]["s"] / flows["grouped_dE0_1e-8"]["s"], 3)

    # ---- gradient of the flowed E0 with respect to the vacuum two-body elements
    grad = None
    if not a.no_grad:
        grad = gradient_bench(b, Hv, a.grad_check)

    # ---- many-instance batch (64 x eMax 6) and eMax 10 steps
    batch = None if a.no_batch else batch_bench(64, 6, a.batch_seq, a.batch_cpu_check)
    e10 = None if a.no_emax10 else emax10_bench(3, 1, not a.no_tune)
    setup = None if a.no_setup else setup_bench(a.emax)
    tens = None if a.no_tensor else tensor_bench(a.emax, not a.no_cpu)

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
            "work_unit": "flop" if unit == "TFLOP/s" else "byte",
            "mfma_busy": measured_mfma(dom, a.emax)}

    # ---- CPU baseline: the oracle (C++ restatement of the reference) on this host
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(b, a.emax, H, Omega)

    F = wl.commutator_flops(b)
    res = {
        "metric": METRIC, "value": round(g_ms, 3), "unit": "ms", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(g_ms, 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"magnus_step_emax{a.emax}", "emax": a.emax, "orbits": b.norb, "holes": "16O",
                   "hw_MeV": 20.0, "seed": 0, "generator": "white", "path": "grouped+hipgraph",
                   "channels": b.nblocks, "E2": b.e2_count(),
                   "gemm_tiles": tuned["choice"] if tuned else "default", "gemm_tiles_source": tuned["source"] if tuned else None},
        "gemm_tune": tuned,
        "commutators_per_step": [int(x) for x in ncomm],
        "E0_after_steps": g_log[-1][0],
        "flows_completed_in_timed_steps": g_flows,
        "converged_flows": flows,
        **out_plain,
        "commutator_ms": {"grouped": round(kms["COMMUTATOR"], 4),
                          "plain": round(comm_plain_ms, 4) if comm_plain_ms else None},
        "commutator_tflops": round(F / (kms["COMMUTATOR"] / 1e3) / 1e12, 3),
        "kernel_ms": {k: round(v, 4) for k, v in kms.items()},
        "grad_over_fwd": grad["grad_over_fwd"] if grad else None,
        "gradient": grad,
        "batch": batch,
        "emax10": e10,
        "setup": setup,
        "tensor": tens,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res))
    return 0


def gradient_bench(b, Hv, ncheck: int):
    """dE0/dV of the full flow (|mbpt2| < 1e-6, every step's states kept on
    the device; the Omega-every-4-steps setting alongside): time of a plain
    forward flow, of the recording forward plus the reverse sweep, and a
    central-difference check (h = 1e-5 MeV) on `ncheck` seeded hermitian
    two-body elements of the vacuum interaction"""
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

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, r

    plain_flow(T)                                    # graph capture
    fwd = [timed(lambda: plain_flow(T)) for _ in range(3)]
    t_fwd = float(np.median([x[0] for x in fwd]))
    e0, steps = fwd[0][1]
    Tb = torch.empty_like(T)
    # every step's states and nested-commutator terms kept on the device
    # (ckpt 0): the reverse sweep recomputes nothing
    free0 = torch.cuda.mem_get_info()[0]
    adjoint.flow_forward(f, T, ckpt=0)
    adjoint.flow_backward(f, Tb)                     # warm-up of the reverse graphs / buffers
    torch.cuda.synchronize()
    kept_gb = (free0 - torch.cuda.mem_get_info()[0]) / 1e9
    runs = []
    for _ in range(3):
        t_fw, (e0g, st) = timed(lambda: adjoint.flow_forward(f, T, ckpt=0))
        t_bw, _ = timed(lambda: adjoint.flow_backward(f, Tb))
        runs.append((t_fw + t_bw, t_fw, t_bw))
    _, t_fw, t_bw = sorted(runs)[1]                  # the median run
    # the memory-lean setting for reference: Omega and H(s) every 4 steps,
    # the segments recomputed in the reverse sweep
    Tb4 = torch.empty_like(T)
    adjoint.flow_forward(f, T, ckpt=4)
    adjoint.flow_backward(f, Tb4)
    runs4 = []
    for _ in range(3):
        t_fw4, _ = timed(lambda: adjoint.flow_forward(f, T, ckpt=4))
        t_bw4, _ = timed(lambda: adjoint.flow_backward(f, Tb4))
        runs4.append((t_fw4 + t_bw4, t_fw4, t_bw4))
    _, t_fw4, t_bw4 = sorted(runs4)[1]
    g = Tb.cpu().numpy()
    gs = adjoint.symmetric_gradient(b, g, 1)
    out = {"steps": steps, "E0": e0, "forward_flow_s": round(t_fwd, 4),
           "checkpointed_forward_s": round(t_fw, 4), "reverse_sweep_s": round(t_bw, 4),
           "grad_over_fwd": round((t_fw + t_bw) / t_fwd, 3), "reverse_over_fwd": round(t_bw / t_fwd, 3),
           "timing": "median of 3 plain flows; median of 3 forward + reverse sweeps (each setting)",
           "checkpoint_every": "every step kept (ckpt 0)", "device_memory_GB": round(kept_gb, 2),
           "ckpt4": {"checkpointed_forward_s": round(t_fw4, 4), "reverse_sweep_s": round(t_bw4, 4),
                     "grad_over_fwd": round((t_fw4 + t_bw4) / t_fwd, 3),
                     "max_abs_diff_vs_ckpt0": float((Tb4 - Tb).abs().max())},
           "E0_forward_vs_gradient_run": abs(e0 - e0g)}
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


def batch_bench(ninst: int, emax: int, nseq: int, ncpu: int):
    """ninst independent flows at eMax `emax` (instance k: seed k, two-body
    scale U(0.8, 1.2), hw in {16, 20, 24}) run to |mbpt2| < 1e-6 as one
    batch (grouped + hipGraph), against `nseq` of them flowed one after
    another by single-instance flows (scaled to ninst); the first `ncpu`
    instances' E0 checked against the oracle's flow on the host"""
    import numpy as np
    import torch
    from eom_imsrg import Basis, Batch, Flow, SqOp, synth, commutators as cm

    from eom_imsrg import tune
    b = Basis.oxygen16(emax)
    tuned = tune.tune(b)
    H = torch.empty(ninst, b.oplen, dtype=torch.float64, device="cuda")
    Hv = []
    for k in range(ninst):
        seed, scale, hw = synth.instance_parameters(k)
        v = synth.hamiltonian_vacuum(b, seed=seed, hw=hw, v_scale=scale)
        Hv.append(v)
        op = SqOp.from_numpy(b, v, 1)
        cm.normal_order(op)
        H[k].copy_(op.data)
    bt = Batch(b, ninst, graphs=True)

    def run_batch():
        HS, G = H.clone(), bt.zeros()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e0, st, nmax = bt.magnus_decouple(HS, G)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, e0, st, nmax

    run_batch()                                   # graph capture
    t_b, e0, st, nmax = run_batch()
    f = Flow(b, graphs=True)

    def run_one(k):
        HS, G = SqOp(b, H[k].clone(), 1), SqOp.zeros(b, -1)
        log = f.magnus_decouple(HS, G)
        return log[-1, 1], len(log)

    out = {"workload": f"magnus_decouple_emax{emax}_x{ninst}", "instances": ninst, "emax": emax,
           "batch_s": round(t_b, 4), "instance_steps_total": int(st.sum()), "batched_steps": int(nmax),
           "steps_per_s": round(float(st.sum()) / t_b, 2), "gemm_tune": tuned,
           "E0": [float(x) for x in e0], "steps": [int(x) for x in st]}
    if nseq > 0:
        run_one(0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        single = [run_one(k) for k in range(nseq)]
        torch.cuda.synchronize()
        t_s = (time.perf_counter() - t0) / nseq * ninst
        out.update({"sequential_single_s_est": round(t_s, 4), "sequential_sample": nseq,
                    "speedup_vs_sequential": round(t_s / t_b, 3),
                    "E0_batch_vs_single_max_MeV": max(abs(single[k][0] - e0[k]) for k in range(nseq)),
                    "steps_match_single": bool(all(single[k][1] == st[k] for k in range(nseq)))})
    if ncpu > 0:
        sys.path.insert(0, ROOT)
        from oracle.oracle import OracleBasis, lib
        lib().orc_set_threads(max(1, min(16, len(os.sched_getaffinity(0)))))
        o = OracleBasis(b.n, b.l, b.jj, b.tz, b.hole)
        tol = float(np.float32(1e-6))
        errs = []
        for k in range(ncpu):
            HSr, _, _, stc = o.magnus_decouple(o.normal_order(Hv[k]), 0.5, tol, tol, 0, tol)
            errs.append({"instance": k, "E0_cpu": float(HSr[0]), "E0_gpu": float(e0[k]),
                         "dE0_MeV": abs(float(HSr[0]) - float(e0[k])), "steps_cpu": int(stc), "steps_gpu": int(st[k])})
        out["cpu_check"] = errs
    return out


def emax10_bench(nsteps: int, warm: int, autotune: bool):
    """Magnus steps 1..nsteps of the eMax 10 flow (132 orbits, seed 0, hw 20)
    on the grouped + hipGraph path with autotuned GEMM tiles, BCH terms to
    1e-10"""
    import numpy as np
    import torch
    from eom_imsrg import Basis, Flow, SqOp, synth, commutators as cm, _lib, workload as wl, tune
    import ctypes as C

    b = Basis.oxygen16(10)
    tuned = tune.tune(b) if autotune else None
    H = SqOp.from_numpy(b, synth.hamiltonian_vacuum(b, seed=0, hw=20.0), 1)
    cm.normal_order(H)
    f = Flow(b, graphs=True)

    def fresh():
        HS, G, DG = H.clone(), SqOp.zeros(b, -1), SqOp.zeros(b, -1)
        cm.build_gs_white(HS, DG)
        return HS, G, DG

    HS, G, DG = fresh()
    for _ in range(warm):
        f.magnus_step(HS, G, DG, H, bch_conv=1e-10)
    HS, G, DG = fresh()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    log = [f.magnus_step(HS, G, DG, H, bch_conv=1e-10) for _ in range(nsteps)]
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / nsteps * 1e3
    # one commutator, hipEvent-timed on the launch stream
    Om = SqOp.from_numpy(b, synth.antihermitian(b, seed=1, sigma=0.05), -1)
    RES, w1, w2 = SqOp.zeros(b, 1), SqOp.zeros(b), SqOp.zeros(b)
    LCC = torch.zeros(max(b.cclen, 1), dtype=torch.float64, device="cuda")
    RCC = torch.zeros_like(LCC)
    WCC = torch.zeros(max(b.wcclen, 1), dtype=torch.float64, device="cuda")
    cms = C.c_double()
    _lib.check(_lib.lib.imsrg_profile_kernel(b.handle, _lib.KERNEL["COMMUTATOR"], Om.ptr, -1, H.ptr, 1, RES.ptr, 1,
                                             w1.ptr, w2.ptr, LCC.data_ptr(), RCC.data_ptr(), WCC.data_ptr(), 3,
                                             C.byref(cms), C.c_void_p(torch.cuda.current_stream().cuda_stream)))
    F = wl.commutator_flops(b)
    return {"workload": "magnus_step_emax10", "orbits": b.norb, "E2": b.e2_count(), "ms_per_step": round(ms, 3),
            "steps": nsteps, "bch_conv": 1e-10, "commutators_per_step": [int(x[2] + x[3] - 1) for x in log],
            "E0_after_steps": log[-1][0], "commutator_ms": round(cms.value, 3),
            "commutator_tflops": round(F / (cms.value / 1e3) / 1e12, 3), "gemm_tune": tuned}


def tensor_bench(emax: int, cpu: bool):
    """transform_sq_op_BCH of a rank-2 observable (BCH_TENSOR,
    IMSRG_MAGNUS.f90:389-464) at eMax `emax`: one [Omega, O] per term
    class and whole (hipEvents on the launch stream, mean of 5 after one
    warm-up call), and the whole BCH_TENSOR.  Omega: the commutator kernels'
    synthetic generator; O: synth.tensor_random (rank 4 = 2 lambda, even
    parity, hermitian).  CPU: the oracle's commutator at eMax 4 (1 thread,
    host BLAS) beside the GPU's at the same size."""
    import numpy as np
    import torch
    from eom_imsrg import Basis, SqOp, CCMat, synth, tensor, commutators as cm

    def run(emax_, reps, with_terms):
        b = Basis.oxygen16(emax_)
        tl = tensor.TensorLayout(b, 4, 0)
        G = SqOp.from_numpy(b, synth.antihermitian(b, seed=1, sigma=0.05), -1)
        O = tensor.TensorOp.from_numpy(tl, synth.tensor_random(tl, seed=5, herm=1), 1)
        RES = tensor.TensorOp.zeros(tl, 1)
        s = torch.cuda.current_stream()

        def timed(fn):
            fn()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(s)
            for _ in range(reps):
                fn()
            e1.record(s)
            torch.cuda.synchronize()
            return e0.elapsed_time(e1) / reps

        out = {"orbits": b.norb, "tensor_elements": tl.oplen,
               "commutator_ms": round(timed(lambda: tensor.TS_commutator(G, O, RES)), 3)}
        if with_terms:
            LCC = CCMat(b, -1)
            cm.calculate_cross_coupled(G, LCC)
            RCC = tensor.TensorCC(tl)
            w1, w2 = tensor.TensorOp.zeros(tl), tensor.TensorOp.zeros(tl)
            terms = {
                "111": lambda: tensor.TS_commutator_111(G, O, RES),
                "121": lambda: tensor.TS_commutator_121(G, O, RES),
                "211": lambda: tensor.TS_commutator_211(LCC, O, RES),
                "122": lambda: tensor.TS_commutator_122(G, O, RES),
                "212": lambda: tensor.TS_commutator_212(G, O, RES),
                "222_pp_hh": lambda: tensor.TS_commutator_222_pp_hh(G, O, RES, w1, w2),
                "221": lambda: tensor.TS_commutator_221(w1, w2, -1, RES),
                "single_pandya": lambda: tensor.calculate_single_pandya(O, RCC),
                "222_ph": lambda: tensor.TS_commutator_222_ph(LCC, RCC, O, RES),
            }
            out["term_ms"] = {k: round(timed(f), 3) for k, f in terms.items()}
            HS = O.clone()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            it, adv = tensor.BCH_TENSOR(G, HS)
            torch.cuda.synchronize()
            out["bch_tensor_s"] = round(time.perf_counter() - t0, 4)
            out["bch_terms"] = int(it)
            out["bch_last_term_norm"] = float(adv[it - 1])
        return out

    res = {"workload": f"bch_tensor_rank2_emax{emax}", "rank": 4, "dpar": 0, **run(emax, 5, True)}
    small = run(4, 5, False)
    res["emax4_commutator_ms"] = small["commutator_ms"]
    if cpu:
        sys.path.insert(0, ROOT)
        from oracle.oracle import OracleBasis, OracleTensor, use_host_blas
        from eom_imsrg import Basis as B_
        use_host_blas()
        b4 = B_.oxygen16(4)
        o4 = OracleBasis(b4.n, b4.l, b4.jj, b4.tz, b4.hole)
        ot = OracleTensor(o4, 4, 0)
        tl4 = tensor.TensorLayout(b4, 4, 0)
        G4 = synth.antihermitian(b4, seed=1, sigma=0.05)
        O4 = synth.tensor_random(tl4, seed=5, herm=1)
        ot.commutator(G4, -1, O4, 1, 1)
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            ot.commutator(G4, -1, O4, 1, 1)
            ts.append((time.perf_counter() - t0) * 1e3)
        res["cpu_oracle_emax4_commutator_ms"] = round(float(np.median(ts)), 1)
        res["gpu_over_cpu_emax4"] = round(res["cpu_oracle_emax4_commutator_ms"] / small["commutator_ms"], 1)
    return res


def setup_bench(emax: int):
    """The Hamiltonian set-up of a run from an interaction file at eMax:
    read_me2j_interaction (a synthetic eMax-8-style me2j.gz written first,
    untimed), calculate_h0_harm_osc, then normal ordering in the HO basis
    (normal_order) and in the Hartree-Fock basis (calc_HF).  Host-to-device
    transfers are inside the timings (the readers are synchronous)."""
    import tempfile
    import torch
    from eom_imsrg import Basis, SqOp, synth, interaction, hk_orbits, commutators as cm
    b = Basis.oxygen16(emax)
    orb = hk_orbits(emax)
    vals = synth.me2j_values(orb, seed=0)
    out = {"workload": f"hamiltonian_setup_emax{emax}", "me2j_values": int(len(vals))}
    H = SqOp.zeros(b, 1)
    with tempfile.TemporaryDirectory() as d:
        path = interaction.write_me2j(os.path.join(d, f"synthetic_eMax{emax:02d}_hw20.me2j.gz"), vals)
        out["me2j_gz_bytes"] = os.path.getsize(path)
        interaction.read_me2j(H, path, emax, 20.0, 1, 16)        # first read: device tables, file cache
        torch.cuda.synchronize()
        t = time.perf_counter()
        interaction.read_me2j(H, path, emax, 20.0, 1, 16)
        torch.cuda.synchronize()
        out["read_me2j_ms"] = round(1e3 * (time.perf_counter() - t), 2)
    t = time.perf_counter()
    interaction.h0_harm_osc(H, 20.0, 1, 16)
    torch.cuda.synchronize()
    out["h0_harm_osc_ms"] = round(1e3 * (time.perf_counter() - t), 3)
    Hho = H.clone()
    torch.cuda.synchronize()
    t = time.perf_counter()
    cm.normal_order(Hho)
    torch.cuda.synchronize()
    out["normal_order_ms"] = round(1e3 * (time.perf_counter() - t), 3)
    out["E0_ho"] = Hho.E0
    Hhf = H.clone()
    interaction.hartree_fock(Hhf)                                # first call builds the transform plan
    Hhf = H.clone()
    torch.cuda.synchronize()
    t = time.perf_counter()
    e, it, _ = interaction.hartree_fock(Hhf)
    torch.cuda.synchronize()
    out["hartree_fock_ms"] = round(1e3 * (time.perf_counter() - t), 2)
    out["hf_iterations"] = it
    out["e_hf"] = e
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


def measured_mfma(kernel: str, emax: int):
    """FP64-MFMA figures of `kernel` from the committed rocprofv3 counter
    summary (SQ_VALU_MFMA_BUSY_CYCLES, SQ_INSTS_VALU_MFMA_MOPS_F64,
    GRBM_GUI_ACTIVE passes), or None when no summary covers it"""
    p = os.path.join(ROOT, "profiles", f"mfma_util_emax{emax}.md")
    if not os.path.exists(p):
        return None
    out = {}
    for line in open(p):
        f = line.split()
        if len(f) == 4 and f[0] == "mfma" and f[1] == kernel:
            out[f[2]] = float(f[3])
    return out or None


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(b, emax, H, Omega):
    """The oracle (the C++/OpenMP restatement of the reference path, with the
    host BLAS dgemm at the reference's dgemm sites) on this host: two
    measured Magnus steps (steps 1 and 2 of the flow, one iteration of
    IMSRG_MAGNUS.f90:100-125 each), the median of 5 [Omega, H] commutators
    after one warm-up, and one commutator on a single thread.  Threads: the
    process's affinity cores, capped by OMP_NUM_THREADS when it is set (the
    GPU boxes grant each GPU a 16-core share and set it to 16)."""
    import numpy as np
    sys.path.insert(0, ROOT)
    from oracle.oracle import OracleBasis, use_host_blas, lib
    aff = len(os.sched_getaffinity(0))
    cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    nthreads = max(1, min(aff, cap) if cap > 0 else aff)
    lib().orc_set_threads(nthreads)
    blas = use_host_blas()
    o = OracleBasis(b.n, b.l, b.jj, b.tz, b.hole)
    Hh, Oh = H.numpy(), Omega.numpy()
    o.commutator(Oh, -1, Hh, 1, 1)                    # warm-up
    ts = []
    for _ in range(5):
        t0 = time.perf_counter()
        o.commutator(Oh, -1, Hh, 1, 1)
        ts.append(time.perf_counter() - t0)
    comm_ms = float(np.median(ts)) * 1e3
    lib().orc_set_threads(1)
    t0 = time.perf_counter()
    o.commutator(Oh, -1, Hh, 1, 1)
    comm1_ms = (time.perf_counter() - t0) * 1e3
    lib().orc_set_threads(nthreads)
    # Magnus steps 1 and 2: MAGNUS_EXPAND, euler_step, BCH_EXPAND (from the
    # bare H), build_gs_white, mbpt2, as the reference loop runs them
    tol = float(np.float32(1e-6))
    G = np.zeros(b.oplen)
    DG = o.white(Hh)
    steps = []
    for _ in range(2):
        t0 = time.perf_counter()
        dG, nm = o.magnus(DG, G, tol)
        G = o.add(G, 1.0, dG, 0.5)
        HS, nb, _ = o.bch(G, Hh, tol)
        DG = o.white(HS)
        o.mbpt2(HS)
        steps.append({"ms": round((time.perf_counter() - t0) * 1e3, 1), "commutators": int(nm + nb - 1)})
    step_ms = float(np.mean([x["ms"] for x in steps]))
    return {"value": round(step_ms, 1), "unit": "ms", "cores": nthreads, "kind": "port",
            "sample": f"Magnus steps 1 and 2 of the eMax {emax} flow on the CPU oracle (OpenMP, "
                      f"{'host OpenBLAS dgemm' if blas else 'own dgemm'}), mean ms per step; "
                      f"{steps[0]['commutators']} + {steps[1]['commutators']} nested commutators",
            "steps": steps,
            "commutator_ms_median_of_5": round(comm_ms, 1), "commutator_ms_1_thread": round(comm1_ms, 1),
            "nproc": os.cpu_count(), "affinity_cores": aff, "omp_num_threads_env": cap or None,
            "cpu_model": _cpu_model(), "blas": os.path.basename(blas) if blas else None}


if __name__ == "__main__":
    sys.exit(main())
