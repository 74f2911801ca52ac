#!/usr/bin/env python
"""Benchmark: HEA energy + gradient evals/s at 24 qubits on one MI355X.

Workload (one "step"): a batch of 32 parameter sets of a 24-qubit, 20-layer
hardware-efficient ansatz (ry, rz on every qubit, CNOT chain; 960 params),
complex64, TFIM energy (open chain, h = 1.0) and its full gradient by the
batched adjoint sweep with forward checkpoints every 5 layers.  Inputs
(parameters) are resident in HBM before timing; synthetic uniform[0, 2pi)
angles from numpy default_rng(1004).

Also reported: the plain per-circuit path (one kernel per gate, one
parameter set per call, same adjoint), the dominant kernel's roofline and a
CPU baseline (the reference's plain method, torch autograd, on host cores).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# HBM bytes per adjoint pass launch measured with rocprofv3 --pmc FETCH_SIZE /
# WRITE_SIZE on the default workload (24q, 20 layers, batch 32, checkpoints
# every 5 layers, default tiles) over whole energy+grad steps
# (--kernel-only eg, the timed work), FETCH scaled x2.0 and WRITE x1.0 by a
# 4 GiB copy profiled the same way.  Key: (qubits, layers, batch, ckpt, jit, first-fit
# scheduling).  Commands and per-counter figures:
# profiles/tcm_adj_passes_hea24_b32.md (compiled passes, searched schedule)
# and profiles/k_fused_adj_hea24_b32.md (generic kernels, first-fit schedule).
MEASURED_ADJ_TRAFFIC = {(24, 20, 32, 5, 0, 1): 17046845037, (24, 20, 32, 5, 1, 0): 16715319666}
# VALU wave-instructions per launch (rocprofv3 --pmc SQ_INSTS_VALU, mean over
# the adjoint / forward pass launches of one sweep), same workload and key;
# profiles/tcm_adj_passes_hea24_b32.md.
MEASURED_VALU_INSTS = {(24, 20, 32, 5, 1, 0): (3492724736, 1338048512)}
# MI355X VALU issue peak: 256 CUs x 4 SIMDs, one wave64 instruction per SIMD
# every 4 cycles at 2.4 GHz
VALU_PEAK_WAVE_INSTS = 256 * 4 * 2.4e9 / 4


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=24)
    ap.add_argument("--layers", type=int, default=20)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--ckpt", type=int, default=5)
    ap.add_argument("--tile", type=int, default=0, help="forward tile bits (0 = default)")
    ap.add_argument("--adj-tile", type=int, default=0, help="adjoint tile bits (0 = default)")
    ap.add_argument("--jit", type=int, default=1, help="1: compiled pass kernels (pass compiler), 0: generic")
    ap.add_argument("--reg-slots", type=int, default=0, help="register-resident qubits (0 = default)")
    ap.add_argument("--mfma", type=int, default=2,
                    help="c5: 2 = 5-qubit blocks grouped into LDS tile passes on the matrix cores, "
                         "1 = one streaming matrix-core pass per block")
    ap.add_argument("--greedy", type=int, default=0,
                    help="1: first-fit pass scheduling (no search for fewer passes)")
    ap.add_argument("--fusion", choices=["search", "window", "first_fit"], default="search",
                    help="c5: how gates are fused into 5-qubit blocks")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--config", choices=["c4", "c2", "c3", "c5"], default="c4",
                    help="c4 (default): 24q adjoint energy+grad; c2: 26q forward + energy; "
                         "c5: 30q fused-MFMA brickwork with a complex128 check")
    ap.add_argument("--kernel-only", choices=["eg", "fwd", "adj", "copy", "sweep"], default=None,
                    help="run only whole energy+grad steps (the timed work), the fused forward passes, the "
                         "adjoint sweep, a state copy of the same size, or the 1-qubit gate sweeps (for "
                         "counter profiles); prints no JSON")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: one GPU only; the batch of parameter sets would shard across GPUs with no "
              "exchange, so --gpus %d is not run" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import tensorcircuit_amd as tc
    from tensorcircuit_amd import engine
    from tensorcircuit_amd.templates.blocks import hea_layers

    if args.config == "c5":
        return run_c5(args)
    if args.config == "c2":
        return run_c2(args)
    if args.config == "c3":
        return run_c3(args)
    n, L, B = args.n, args.layers, args.batch
    dev = engine.device()
    c = tc.Circuit(n)
    hea_layers(c, np.zeros((2 * L, n)), L)
    specs, thetas = c._lower()
    P = len(thetas)
    codes, coeffs = [], []
    for i in range(n - 1):
        cc = [0] * n
        cc[i] = cc[i + 1] = 3
        codes.append(cc)
        coeffs.append(-1.0)
    for i in range(n):
        cc = [0] * n
        cc[i] = 1
        codes.append(cc)
        coeffs.append(-1.0)
    ham = engine.PauliSum(n, codes, coeffs)
    t_stage = time.perf_counter()
    prog = engine.Program(n, specs, P, fuse=1, ckpt_every=args.ckpt, tile_bits=args.tile,
                          adj_tile_bits=args.adj_tile, jit=args.jit, reg_slots=args.reg_slots,
                          greedy=args.greedy)
    jit_s = 0.0
    if args.jit:
        jit_s = prog.compile("complex64", 0) + prog.compile("complex64", 1)
    rng = np.random.default_rng(1004)
    params = torch.as_tensor(rng.uniform(0, 2 * np.pi, size=(B, P)).astype(np.float32)).to(dev)
    if args.kernel_only:
        st, ck = prog.forward(params, B, "complex64", want_ckpts=True)
        if args.kernel_only == "sweep":
            sprog, sparams = _sweep_program(n, B)
        for _ in range(max(args.steps, 1)):
            if args.kernel_only == "sweep":
                sprog.forward(sparams, B, "complex64", out=st)
            elif args.kernel_only == "fwd":
                prog.forward(params, B, "complex64", out=st)
            elif args.kernel_only == "adj":
                prog.vjp(params, ck, st, st, B, "complex64", False)
            elif args.kernel_only == "eg":
                prog.energy_grad(params, ham, B)
            else:
                torch.empty_like(st).copy_(st)
        torch.cuda.synchronize()
        print("kernel-only %s done" % args.kernel_only, file=sys.stderr)
        return
    e, g = prog.energy_grad(params, ham, B)
    torch.cuda.synchronize()
    staging_s = time.perf_counter() - t_stage
    for _ in range(args.warmup):
        prog.energy_grad(params, ham, B)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        e, g = prog.energy_grad(params, ham, B)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = B * args.steps / (t1 - t0)
    print("adjoint batched: %.3f ms/step (%d evals), %.2f evals/s" % (ms_per_step, B, value), file=sys.stderr)

    out = {
        "metric": "HEA energy+grad evals/s at 24q (adjoint vs plain); achieved HBM GB/s",
        "value": round(value, 3),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": "hea_tfim_energy_grad", "nqubits": n, "layers": L, "params_per_set": P,
                   "batch": B, "ckpt_every_layers": args.ckpt, "h": 1.0, "seed": 1004,
                   "fwd_passes": prog.npasses, "staging_s": round(staging_s, 3), "jit": args.jit,
                   "jit_compile_s": round(jit_s, 2), "reg_slots": prog.describe("complex64", 0)["sched"]["R"],
                   "sched": "first-fit" if args.greedy else "search"},
    }

    # ------------------------------------------------------------ plain path
    if not args.no_plain:
        pprog = engine.Program(n, specs, P, fuse=0, ckpt_every=args.ckpt)
        p1 = params[:1].contiguous()
        pprog.energy_grad(p1, ham, 1)
        torch.cuda.synchronize()
        nplain = 2
        t0 = time.perf_counter()
        for i in range(nplain):
            ep, gp = pprog.energy_grad(params[i:i + 1].contiguous(), ham, 1)
        torch.cuda.synchronize()
        tp = (time.perf_counter() - t0) / nplain
        # parity of the two paths on the same parameter sets
        dmax = float((gp[0] - g[nplain - 1]).abs().max())
        de = float((ep[0] - e[nplain - 1]).abs())
        out["plain_evals_per_s"] = round(1.0 / tp, 4)
        out["speedup_vs_plain"] = round(value * tp, 2)
        out["plain_vs_fused_max_abs_dgrad"] = dmax
        out["plain_vs_fused_abs_denergy"] = de
        print("plain per-circuit: %.1f ms/eval (%.2f evals/s), speedup %.1fx" % (tp * 1e3, 1 / tp, value * tp),
              file=sys.stderr)

    # -------------------------------------------------------------- roofline
    if not args.no_roofline:
        N = 1 << n
        bpa = 8  # bytes per complex64 amplitude
        st, ck = prog.forward(params, B, "complex64", want_ckpts=True)
        gs = st.clone()
        stream = torch.cuda.current_stream()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        reps = 3
        ev[0].record(stream)
        for _ in range(reps):
            prog.forward(params, B, "complex64", out=st)
        ev[1].record(stream)
        for _ in range(reps):
            prog.vjp(params, ck, st, gs, B, "complex64", False)
        ev[2].record(stream)
        torch.cuda.synchronize()
        fwd_ms = ev[0].elapsed_time(ev[1]) / reps
        adj_ms = ev[1].elapsed_time(ev[2]) / reps
        da = prog.describe("complex64", 1)
        n_adj = len(da["sched"]["passes"])
        n_fwd = prog.npasses
        fwd_launch_ms = fwd_ms / n_fwd
        adj_launch_ms = adj_ms / n_adj
        adj_bytes = 4 * bpa * N * B   # read + write psi and lambda per pass
        fwd_bytes = 2 * bpa * N * B   # read + write psi per pass
        adj_gbs = adj_bytes / (adj_launch_ms * 1e-3) / 1e9
        fwd_gbs = fwd_bytes / (fwd_launch_ms * 1e-3) / 1e9
        traffic = None
        valu = None
        if args.adj_tile == 0 and args.tile == 0 and args.reg_slots == 0:
            key = (n, L, B, args.ckpt, args.jit, args.greedy)
            traffic = MEASURED_ADJ_TRAFFIC.get(key)
            valu = MEASURED_VALU_INSTS.get(key)
        kname = "tcm_adj_k* (compiled adjoint passes)" if args.jit else "k_fused_adj"
        out["roofline"] = {"bound": "hbm", "kernel": kname, "achieved": round(adj_gbs, 1),
                           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(adj_gbs / HBM_PEAK_GBS, 4),
                           "traffic": traffic,
                           "traffic_source": None if traffic is None else
                           "recorded constant: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE over this workload's "
                           "adjoint pass dispatches (profiles/tcm_adj_passes_hea24_b32.md)",
                           "bytes_per_launch": adj_bytes,
                           "avg_launch_ms": round(adj_launch_ms, 4), "launches_per_step": n_adj}
        out["roofline_fwd"] = {"kernel": "tcm_fwd_k* (compiled forward passes)" if args.jit else "k_fused_fwd",
                               "achieved": round(fwd_gbs, 1), "unit": "GB/s",
                               "frac": round(fwd_gbs / HBM_PEAK_GBS, 4), "bytes_per_launch": fwd_bytes,
                               "avg_launch_ms": round(fwd_launch_ms, 4), "launches_per_step": n_fwd}
        if valu is not None:
            # the compiled passes issue more VALU work per byte than HBM can
            # feed when many gates share a pass: their binding roof is VALU issue
            va = valu[0] / (adj_launch_ms * 1e-3)
            vf = valu[1] / (fwd_launch_ms * 1e-3)
            out["roofline_valu"] = {"bound": "valu", "kernel": kname, "achieved": round(va / 1e9, 1),
                                    "peak": round(VALU_PEAK_WAVE_INSTS / 1e9, 1), "unit": "G wave-instr/s",
                                    "frac": round(va / VALU_PEAK_WAVE_INSTS, 4), "insts_per_launch": valu[0],
                                    "fwd_achieved": round(vf / 1e9, 1),
                                    "fwd_frac": round(vf / VALU_PEAK_WAVE_INSTS, 4)}
        print("fused fwd pass %.3f ms (%.0f GB/s), adjoint pass %.3f ms (%.0f GB/s)"
              % (fwd_launch_ms, fwd_gbs, adj_launch_ms, adj_gbs), file=sys.stderr)
        # gate sweeps: one 1-qubit gate per launch (ry then rz on every
        # qubit) over the same 4 GiB batch, next to a device copy of it
        sprog, sparams = _sweep_program(n, B)
        sprog.forward(sparams, B, "complex64", out=st)
        cp = torch.empty_like(st)
        ev[0].record(stream)
        for _ in range(reps):
            sprog.forward(sparams, B, "complex64", out=st)
        ev[1].record(stream)
        for _ in range(reps):
            cp.copy_(st)
        ev[2].record(stream)
        torch.cuda.synchronize()
        sweep_launch_ms = ev[0].elapsed_time(ev[1]) / reps / sprog.npasses
        copy_ms = ev[1].elapsed_time(ev[2]) / reps
        sweep_gbs = fwd_bytes / (sweep_launch_ms * 1e-3) / 1e9
        copy_gbs = fwd_bytes / (copy_ms * 1e-3) / 1e9
        del cp
        out["roofline_sweep"] = {"bound": "hbm", "kernel": "k_plain1 (one 1-qubit gate per launch)",
                                 "achieved": round(sweep_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": round(sweep_gbs / HBM_PEAK_GBS, 4), "bytes_per_launch": fwd_bytes,
                                 "avg_launch_ms": round(sweep_launch_ms, 4), "launches": sprog.npasses,
                                 "device_copy_gbs": round(copy_gbs, 1),
                                 "frac_of_copy": round(sweep_gbs / copy_gbs, 4)}
        print("gate sweep %.3f ms per 1-qubit gate (%.0f GB/s, %.3f of peak), device copy %.0f GB/s"
              % (sweep_launch_ms, sweep_gbs, sweep_gbs / HBM_PEAK_GBS, copy_gbs), file=sys.stderr)

    # ------------------------------------------- precision against complex128
    # the timed complex64 path against the complex128 run of the same program
    # on the first parameter sets (the same float32 angles, widened)
    nchk = min(2, B)
    prog128 = engine.Program(n, specs, P, fuse=1, ckpt_every=args.ckpt, tile_bits=args.tile,
                             adj_tile_bits=args.adj_tile, jit=args.jit, greedy=args.greedy)
    e128, g128 = prog128.energy_grad(params[:nchk].double().contiguous(), ham, nchk, "complex128")
    torch.cuda.synchronize()
    out["max_abs_dgrad_vs_c128"] = float((g[:nchk] - g128).abs().max())
    out["max_rel_denergy_vs_c128"] = float(((e[:nchk] - e128).abs() / e128.abs()).max())
    out["precision_check_sets"] = nchk
    print("vs complex128: max|dgrad| %.2e, max rel|dE| %.2e" % (out["max_abs_dgrad_vs_c128"],
                                                               out["max_rel_denergy_vs_c128"]), file=sys.stderr)
    del prog128, e128, g128

    # ---------------------------------------------------------- CPU baseline
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = _cpu_baseline(n, L)

    print(json.dumps(out))


def _cpu_baseline(n, L):
    """The reference's plain method (gate-by-gate tensordot, torch autograd,
    complex64) on the host cores.  The 12-qubit 4-layer case is timed in full
    (median of 5).  A 24-qubit 20-layer eval takes about two minutes, so it is
    sampled: energy + gradient at 1 and at 3 layers (each layer under
    per-layer rematerialisation), twice each after a discarded warm-up;
    20 layers = t(1) + 19 (t(3) - t(1)) / 2 per pair, and the two pairs give
    the spread."""
    import statistics

    from oracle import plain_torch

    # the box's CPU share for one GPU is OMP_NUM_THREADS; otherwise the affinity mask
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    if threads <= 0:
        threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    plain_torch.time_energy_grad(12, 1, seed=1001, threads=threads)  # warm-up
    c1 = [plain_torch.time_energy_grad(12, 4, seed=1001, threads=threads)[0] for _ in range(5)]
    c1_s = statistics.median(c1)
    plain_torch.time_energy_grad(n, 1, seed=1004, threads=threads, remat=True)  # warm-up, discarded
    samples, ests = [], []
    for _ in range(2):
        t1 = plain_torch.time_energy_grad(n, 1, seed=1004, threads=threads, remat=True)[0]
        t3 = plain_torch.time_energy_grad(n, 3, seed=1004, threads=threads, remat=True)[0]
        samples.append((round(t1, 3), round(t3, 3)))
        ests.append(t1 + (L - 1) * max(t3 - t1, 0.0) / 2)
    cpu_eval_s = sum(ests) / len(ests)
    print("cpu baseline: C1 %.4f s/eval; %dq (t1, t3) %s -> %.1f s per %d-layer eval (%d threads)"
          % (c1_s, n, samples, cpu_eval_s, L, threads), file=sys.stderr)
    return {"value": round(1.0 / cpu_eval_s, 6), "unit": "evals/s", "cores": threads, "kind": "port",
            "cpu_model": plain_torch.cpu_model(),
            "sample": "%dq HEA + TFIM energy and gradient by torch autograd (gate-by-gate tensordot, complex64, "
                      "each layer rematerialised), timed at 1 and 3 layers twice; %d layers = t(1) + %d (t(3) - "
                      "t(1)) / 2" % (n, L, L - 1),
            "samples_s_1_and_3_layers": samples,
            "eval_s_estimates": [round(x, 2) for x in ests],
            "c1_12q_4l_evals_per_s": round(1.0 / c1_s, 3),
            "c1_samples_s": [round(x, 5) for x in c1]}


def _sweep_program(n, B):
    """ry then rz on every qubit, one 1-qubit gate per launch (no fusion),
    for a batch of B states."""
    import torch

    import tensorcircuit_amd as tc
    from tensorcircuit_amd import engine

    sc = tc.Circuit(n)
    for i in range(n):
        sc.ry(i, theta=0.3 + 0.01 * i)
    for i in range(n):
        sc.rz(i, theta=0.7 - 0.01 * i)
    sspecs, sth = sc._lower()
    sprog = engine.Program(n, sspecs, len(sth), fuse=0)
    sparams = torch.tensor([float(x) for x in sth], dtype=torch.float32, device=engine.device()).repeat(B, 1)
    return sprog, sparams


def _timed(fn, steps, warmup):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, out


def run_c5(args):
    """30-qubit depth-20 brickwork (Haar SU(2) + CZ, seed 1005) fused into
    5-qubit 32x32 blocks applied on FP32 matrix cores; the same blocks in
    complex128 on the VALU serve as the on-GPU reference."""
    import torch

    from tensorcircuit_amd import engine, fusion, workloads

    n, depth = 30, 20
    t_stage = time.perf_counter()
    gates = workloads.brickwork(n, depth, 1005)
    blocks = fusion.fuse(gates, n, method=args.fusion)
    p32 = engine.Program(n, blocks, 0, fuse=1, mfma=args.mfma, tile_bits=args.tile)
    staging = time.perf_counter() - t_stage
    st32 = torch.empty(1, 1 << n, dtype=torch.complex64, device="cuda")
    dt32, _ = _timed(lambda: p32.forward(None, 1, "complex64", out=st32), args.steps, args.warmup)
    p64 = engine.Program(n, blocks, 0, fuse=1, mfma=0)
    st64 = torch.empty(1, 1 << n, dtype=torch.complex128, device="cuda")
    dt64, _ = _timed(lambda: p64.forward(None, 1, "complex128", out=st64), 1, 0)
    fid = abs(complex(engine.inner(st64, st32.to(torch.complex128), n)[0])) ** 2
    del st64
    nb = len(blocks)
    flops = nb * 8.0 * 32 * (1 << n)          # 32 complex MACs per amplitude per block
    bytes_ = p32.npasses * 2 * 8.0 * (1 << n)  # read + write of the c64 state per pass
    # matrix-core flops issued: four real products per block on the streaming
    # kernel, three (Gauss) in the tile passes
    issued = flops * (0.75 if args.mfma == 2 else 1.0)
    out = {"metric": "C5 fused-MFMA brickwork states/s (30q, depth 20)", "value": round(1.0 / dt32, 4),
           "unit": "states/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(dt32 * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "c64", "data": "synthetic",
           "config": {"workload": "brickwork_haar_su2_cz", "nqubits": n, "depth": depth, "gates": len(gates),
                      "fused_blocks": nb, "fusion": args.fusion, "seed": 1005, "staging_s": round(staging, 3),
                      "mfma_mode": args.mfma, "passes": p32.npasses},
           "fidelity_vs_c128": fid, "c128_ms": round(dt64 * 1e3, 3),
           "mfma": {"achieved": round(flops / dt32 / 1e12, 2), "peak": 157.3, "unit": "TFLOP/s",
                    "frac": round(flops / dt32 / 1e12 / 157.3, 4),
                    "issued_frac": round(issued / dt32 / 1e12 / 157.3, 4)},
           "hbm": {"achieved": round(bytes_ / dt32 / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                   "frac": round(bytes_ / dt32 / 1e9 / HBM_PEAK_GBS, 4)}}
    print("c5: %.2f ms/state (%d blocks), c128 %.1f ms, fidelity %.9f" % (dt32 * 1e3, nb, dt64 * 1e3, fid),
          file=sys.stderr)
    print(json.dumps(out))


def run_c3(args):
    """1024 parameter sets of a 16-qubit 8-layer HEA plus 256 circuits with
    qubit counts uniform in 10..20 (depth 6), complex64, then fp64 TFIM
    energies of every state; seed 1003 (qubit counts first, then angles),
    run through engine.UnevenBatch.  Staging: programs compiled into
    straight-line pass kernels; the larger
    programs' fusion width and tile autotuned (forward of the real batch);
    then a step-level choice among the programs' passes in sequence, the
    compiled multi-program launches (pass r of every program in one kernel)
    and a hybrid (the programs holding >= 5 % of the amplitudes on their own
    passes, the rest in one multi-program launch per pass; also with the two
    parts on two streams), each with default and with tuned programs, every
    candidate captured as a HIP graph and timed whole; the default (compiled
    passes in sequence, default programs) is replaced only by a candidate
    faster by 3 %."""
    import numpy as np
    import torch

    import tensorcircuit_amd as tc
    from tensorcircuit_amd import engine, workloads
    from tensorcircuit_amd.templates.blocks import hea_layers

    rng = np.random.default_rng(1003)
    ns = rng.integers(10, 21, 256)
    groups = {}
    for i, n in enumerate(ns):
        groups.setdefault(int(n), []).append(i)
    t_stage = time.perf_counter()

    sizes = [(16, 8, 1024)] + [(n, 6, len(groups[n])) for n in sorted(groups)]
    params, structs = [], []
    for n, L, B in sizes:
        P = 2 * L * n
        params.append(torch.as_tensor(rng.uniform(0, 2 * np.pi, size=(B, P)).astype(np.float32)).cuda())
        c = tc.Circuit(n)
        hea_layers(c, np.zeros((2 * L, n)), L, barrier=False)
        structs.append(c._lower())

    def hamiltonian(n):
        h = engine.PauliSum(n, *workloads.tfim_terms(n), jit=1)
        h.compile("complex64")
        return h

    hams = [hamiltonian(n) for n, _, _ in sizes]
    options = None
    if args.tile or args.reg_slots:
        options = [{"tile_bits": min(args.tile, n) if args.tile else 0, "reg_slots": args.reg_slots}
                   for n, _, _ in sizes]
    ub = engine.UnevenBatch([(n, specs, len(thetas)) for (n, _, _), (specs, thetas) in zip(sizes, structs)],
                            [B for _, _, B in sizes], hams=hams, params=params, options=options,
                            tune=not (args.tile or args.reg_slots))
    states = ub.states
    choice = ub.choice
    staging = time.perf_counter() - t_stage

    dt_eager, _ = _timed(ub.run_eager, args.steps, args.warmup)
    dt, es = _timed(ub.run, args.steps, args.warmup)
    # every state against a single-program run of its program (the first set)
    dmax = 0.0
    for prog, pa, st in zip(ub.programs, params, states):
        s1, _ = prog.forward(pa[:1], 1, "complex64")
        dmax = max(dmax, float((s1[0] - st[0]).abs().max()))
    nstates = sum(st.shape[0] for st in states)
    amps = sum(st.numel() for st in states)
    out = {"metric": "C3 batched uneven HEA states+energies/s", "value": round(nstates / dt, 2),
           "unit": "states/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "c64", "data": "synthetic",
           "config": {"workload": "hea_batched_uneven", "states": nstates, "amplitudes": amps,
                      "programs": len(sizes), "qubit_counts": {str(k): len(v) for k, v in sorted(groups.items())},
                      "seed": 1003, "staging_s": round(staging, 3), "eager_ms_per_step": round(dt_eager * 1e3, 3),
                      "hip_graph": True, "step_choice": choice,
                      "step_candidates_ms": {k: round(v, 4) for k, v in ub.candidate_ms.items()},
                      "program_tuning_ms": {k: {c: round(v, 4) for c, v in t.items()}
                                            for k, t in ub.tuning_ms.items()},
                      "tuned_options": {"%dq_x%d" % (n, B): engine.tune_key(o)
                                        for (n, _, B), o in zip(sizes, ub.tuned_options) if o},
                      "multi_plan_compile_s": round(ub.multi_plan.compile_seconds, 2) if ub.multi_plan else 0.0},
           "energy_16q_first": float(es[0][0].real), "max_abs_diff_vs_single_program": dmax}
    print("c3: %.3f ms for %d states (%d programs) as a HIP graph (%s), %.3f ms eager; candidates %s"
          % (dt * 1e3, nstates, len(sizes), choice, dt_eager * 1e3, out["config"]["step_candidates_ms"]),
          file=sys.stderr)
    print(json.dumps(out))


def run_c2(args):
    """26-qubit 10-layer HEA forward + TFIM energy (seed 1002) in complex64,
    checked against the complex128 run of the same program."""
    import numpy as np
    import torch

    import tensorcircuit_amd as tc
    from tensorcircuit_amd import engine, workloads
    from tensorcircuit_amd.templates.blocks import hea_layers

    n, L = 26, 10
    c = tc.Circuit(n)
    hea_layers(c, np.zeros((2 * L, n)), L)
    specs, thetas = c._lower()
    codes, w = workloads.tfim_terms(n)
    ham = engine.PauliSum(n, codes, w)
    p = torch.as_tensor(workloads.hea_params(1002, L, n).reshape(1, -1)).cuda()
    t_stage = time.perf_counter()
    prog = engine.Program(n, specs, len(thetas), fuse=1, jit=args.jit, reg_slots=args.reg_slots,
                          tile_bits=args.tile)
    if args.jit:
        prog.compile("complex64", 0)
        prog.compile("complex128", 0)
    staging = time.perf_counter() - t_stage
    st = torch.empty(1, 1 << n, dtype=torch.complex64, device="cuda")

    def step():
        prog.forward(p, 1, "complex64", out=st)
        return ham.expect(st)

    dt, e32 = _timed(step, args.steps, args.warmup)
    st64, _ = prog.forward(p.double(), 1, "complex128")
    e64 = ham.expect(st64)
    fid = abs(complex(engine.inner(st64, st.to(torch.complex128), n)[0])) ** 2
    out = {"metric": "C2 26q HEA forward + TFIM energy per s", "value": round(1.0 / dt, 3), "unit": "evals/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c64",
           "data": "synthetic", "config": {"workload": "hea_tfim_forward", "nqubits": n, "layers": L,
                                           "fwd_passes": prog.npasses, "tile_bits": prog.tile_bits,
                                           "seed": 1002, "jit": args.jit,
                                           "staging_s": round(staging, 2)},
           "energy_c64": float(e32[0].real), "energy_c128": float(e64[0].real),
           "abs_denergy": abs(float(e32[0].real) - float(e64[0].real)), "fidelity_vs_c128": fid}
    print("c2: %.2f ms/eval, E=%.8f (c128 %.8f)" % (dt * 1e3, out["energy_c64"], out["energy_c128"]),
          file=sys.stderr)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
