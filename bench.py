"""DC-PSE hot-path benchmark on one MI355X.

Workload (the configuration the headline metric is quoted on): 3D, 128^3
lattice of spacing h = 1/128 jittered by U(-0.1h, 0.1h), 10% of the particles
removed at random (seed 3), periodic unit cube, rCut = 2.4h -> about 1.89M
particles with 35..60 neighbours. Convergence order 2.

  setup     neighbour CSR on the GPU (support_ms), FP64 assembly of the
            Laplacian (one operator, 19 monomials) and Dx (10 monomials) in
            one pass sharing gathers, epsilon and moments (assembly_ms, HIP
            events; each operator alone in assembly_ms_by_operator; the pair
            with one step of FP64 iterative refinement per solve in
            assembly_ms_with_refinement).
  step      one fused FP32 pass applying Laplacian and Dx to
            f = sin(2 pi x) sin(2 pi y) sin(2 pi z) (inputs resident in HBM).
  value     M particle-op evals/s = particles * 2 operators / step time.
  fp64      the same fused pass in FP64 (fp64_apply_us; FP64 weights on the
            sliced-ELL layout), the reference's own apply precision.
  roofline  the apply kernel against HBM (8 TB/s): algorithmic bytes per
            launch / mean launch duration (HIP events on the launch stream).
  cpu       the CPU oracle (faithful restatement of the reference's CPU Dcpse:
            reference-form FP64 apply, Laplacian as 3 separate operators) timed
            on the host cores on a bounded sample, same process.

Prints one JSON line. --gpus above 1 exits: the benchmark runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "M particle-op evals/s + assembly ms, 3D 2.1M pts; FP32 vs FP64 at 1e-5 rel err"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def events_ms(fn, reps, torch):
    """Mean duration of fn() over reps, each bracketed by its own event pair."""
    s = [torch.cuda.Event(enable_timing=True) for _ in range(reps)]
    e = [torch.cuda.Event(enable_timing=True) for _ in range(reps)]
    for i in range(reps):
        s[i].record()
        fn()
        e[i].record()
    torch.cuda.synchronize()
    d = [s[i].elapsed_time(e[i]) for i in range(reps)]
    return sum(d) / reps, d


def read_profile_value(name, key):
    """A per-launch figure (``key = value`` line) of a committed rocprofv3 --pmc summary, if any."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            k, sep, v = line.strip().partition("=")
            if sep and k.strip() == key:
                try:
                    return float(v.split()[0])
                except (ValueError, IndexError):
                    return None
    return None


def read_traffic():
    """HBM bytes per fused apply launch (FETCH_SIZE + WRITE_SIZE, calibrated)."""
    return read_profile_value("apply_fused_lap_dx_3d128.pmc.txt", "traffic_bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=128, help="lattice edge (128 -> ~1.89M particles)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU baseline work")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline (profiling runs)")
    ap.add_argument("--tune", action="store_true",
                    help="autotune launch shapes (CSR shapes, SELL unroll, assembly particles/wave) and use the "
                         "fastest; the default is the fixed sliced-ELL kernel that the committed profiles cover")
    ap.add_argument("--layout", choices=["sell", "sell_bundle", "sell_bundle16", "csr"], default="sell_bundle16",
                    help="FP32 apply layout of the timed step")
    ap.add_argument("--order", default="lattice",
                    help="particle storage order (morton: ParticleSet.reorder, as openfpm's reorder)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run torch's copy kernel on 1 GiB (HBM counter calibration)")
    ap.add_argument("--asm-only", type=int, default=0, metavar="REPS",
                    help="only build the support and run REPS Laplacian + Dx assembly passes with the default "
                         "launch (assembly counter runs)")
    ap.add_argument("--no-uneven", action="store_true", help="skip the uneven-support assembly workload")
    ap.add_argument("--no-adaptive", action="store_true",
                    help="skip the adaptive-resolution workload (support, assembly, 100 RK4 steps)")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the 2D 1024^2 batched adjoint workload (16 fields, autograd)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "the DC-PSE benchmark runs on exactly one GPU; "
                                   f"--gpus {args.gpus} (WORLD_SIZE={world}) is not supported"}))
        sys.exit(2)

    sys.path.insert(0, ROOT)
    import numpy as np
    import torch

    from openfpm_numerics_amd import Derivative_x, Laplacian, ParticleSet, workloads as W
    from openfpm_numerics_amd.autotune import (ASSEMBLY_DEFAULT, set_assembly_launch, tune_apply, tune_assembly,
                                               tune_sell)
    from openfpm_numerics_amd.dcpse import Dcpse, PreparedApply, assemble_together

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.calibrate_copy:
        n = 1 << 28  # 2^28 floats = 1 GiB read + 1 GiB written per copy
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes_read_per_launch": 4 * n, "bytes_written_per_launch": 4 * n}))
        return

    log(f"[bench] generating workload n={args.n}")
    w = W.jittered_3d_holes(args.n, seed=3)
    N = w["pos"].shape[0]
    ps = ParticleSet(w["pos"], w["box_lo"], w["box_hi"], bc=[1, 1, 1], device=dev)
    rcut = w["rcut"]
    if args.order != "lattice":
        perm = ps.reorder(args.order)
        w["pos"] = np.ascontiguousarray(w["pos"][perm])

    # ---- support -------------------------------------------------------------
    ps.getVerlet(rcut)  # warm (code objects, workspace)
    torch.cuda.synchronize()
    sup = []
    for _ in range(3):  # host-inclusive (the count's size is read back before the fill): median of 3
        vl = None
        t0 = time.perf_counter()
        vl = ps.getVerlet(rcut)
        torch.cuda.synchronize()
        sup.append((time.perf_counter() - t0) * 1e3)
    support_ms = sorted(sup)[1]
    nnz = vl.nnz
    lens = vl.row_lengths()
    log(f"[bench] N={N} nnz={nnz} neighbours min/mean/max = {int(lens.min())}/{nnz / N:.1f}/{int(lens.max())} "
        f"support {support_ms:.2f} ms")

    # ---- assembly ---------------------------------------------------------------
    lap = Laplacian(ps, vl, 2, rcut)
    dx = Derivative_x(ps, vl, 2, rcut)
    torch.cuda.synchronize()

    def assemble_both():  # one pass for the pair (shared gathers, epsilon and moments)
        assemble_together([lap.dcpse, dx.dcpse], ps, ps, check=False)

    if args.asm_only:
        for _ in range(args.asm_only):
            assemble_both()
        torch.cuda.synchronize()
        print(json.dumps({"asm_only": args.asm_only, "particles": N, "nnz": nnz}))
        return
    asm_lap_ms, _ = events_ms(lambda: lap.dcpse.initializeStaticSize(ps, ps, check=False), 3, torch)
    asm_dx_ms, _ = events_ms(lambda: dx.dcpse.initializeStaticSize(ps, ps, check=False), 3, torch)
    # the pair with one step of FP64 iterative refinement per solve (an option,
    # off by default as in the reference), then the default again
    lap.dcpse.refine = dx.dcpse.refine = True
    assembly_refined_ms, _ = events_ms(assemble_both, 3, torch)
    lap.dcpse.refine = dx.dcpse.refine = False
    assembly_default_ms, _ = events_ms(assemble_both, 3, torch)
    # the launch shape chosen by the autotuner (static or persistent grid,
    # particles per batch, neighbour chunk, bucketing); the weights are the same
    (asm_cfg, _), asm_tune = tune_assembly(assemble_both)
    assembly_ms, _ = events_ms(assemble_both, 3, torch)
    lap.dcpse.check_info()
    dx.dcpse.check_info()
    log(f"[bench] assembly Lap {asm_lap_ms:.3f} ms, Dx {asm_dx_ms:.3f} ms, both {assembly_default_ms:.3f} ms "
        f"(static grid), {assembly_ms:.3f} ms with the tuned launch {asm_cfg} (with refinement "
        f"{assembly_refined_ms:.3f} ms)")
    set_assembly_launch(*ASSEMBLY_DEFAULT)

    # ---- batched uneven supports: about 20 to 200 neighbours in one launch ---------
    uneven = None
    if not args.no_uneven:
        wu = W.uneven_3d(1_000_000)
        psu = ParticleSet(wu["pos"], wu["box_lo"], wu["box_hi"], bc=[1, 1, 1], device=dev)
        vlu = psu.getVerlet(wu["rcut"])
        lu = vlu.row_lengths().cpu().numpy()
        lapu = Laplacian(psu, vlu, 2, wu["rcut"])
        dxu = Derivative_x(psu, vlu, 2, wu["rcut"])

        def assemble_uneven():
            assemble_together([lapu.dcpse, dxu.dcpse], psu, psu, check=False)

        uneven_static_ms, _ = events_ms(assemble_uneven, 3, torch)
        (ucfg, _), utune = tune_assembly(assemble_uneven)
        uneven_ms, _ = events_ms(assemble_uneven, 3, torch)
        lapu.dcpse.check_info()
        dxu.dcpse.check_info()
        set_assembly_launch(*ASSEMBLY_DEFAULT)
        uneven = {"workload": "3d_uneven_h0(1+1.1x)_periodic_one_rcut_order2_lap+dx", "particles": int(len(lu)),
                  "nnz": int(vlu.nnz),
                  "neighbours_min_p5_p50_p95_max": [int(lu.min()), float(np.percentile(lu, 5)),
                                                    float(np.median(lu)), float(np.percentile(lu, 95)),
                                                    int(lu.max())],
                  "assembly_ms": uneven_ms, "assembly_ms_static_grid": uneven_static_ms,
                  "launch": {"persistent": ucfg[0], "particles_per_batch": ucfg[1], "chunk": ucfg[2],
                             "bucket_width": ucfg[3]},
                  "autotune_ms": {"%d/%d/%d/%d" % k: v for k, v in utune.items()}}
        log(f"[bench] uneven supports N={len(lu)} n {lu.min()}..{lu.max()}: Lap+Dx assembly {uneven_ms:.3f} ms "
            f"(tuned {ucfg}), static grid {uneven_static_ms:.3f} ms")
        del lapu, dxu, vlu, psu

    # ---- adaptive resolution: 1M particles, h(x) = h0 (1 + 3x), per-particle
    # rCut = 2.4 h(x_p); 100 RK4 steps of df/dt = 0.01 Lap f with FP32 state
    # on the device against the same integration in FP64 ------------------------
    adaptive = None
    if not args.no_adaptive:
        from openfpm_numerics_amd import rk4

        wa = W.adaptive_3d(1_000_000, seed=5)
        na = wa["pos"].shape[0]
        psa = ParticleSet(wa["pos"], wa["box_lo"], wa["box_hi"], bc=[0, 1, 1], device=dev)
        psa.getVerlet(wa["rcut"])  # warm
        sup_a_ms, _ = events_ms(lambda: psa.getVerlet(wa["rcut"]), 1, torch)
        vla = psa.getVerlet(wa["rcut"])
        la = vla.row_lengths()
        lapa = Laplacian(psa, vla, 2, wa["rcut"])
        asm_a_ms, _ = events_ms(lambda: assemble_together([lapa.dcpse], psa, psa, check=False), 3, torch)
        lapa.dcpse.check_info()
        # x is not periodic: rows whose support crosses an x face are one-sided
        # and grow under explicit integration; they are held fixed (zero rows)
        xa, rca = wa["pos"][:, 0], wa["rcut"]
        held = torch.as_tensor((xa - rca < 0.0) | (xa + rca > 1.0), device=dev)
        nz_held = torch.repeat_interleave(held, la)
        for dt_ in (torch.float64, torch.float32):
            lapa.dcpse.weights(dt_)[nz_held] = 0
        rows_a = torch.repeat_interleave(torch.arange(na, device=dev), la)
        gersh = 2.0 * float(torch.zeros(na, dtype=torch.float64, device=dev)
                            .index_add_(0, rows_a, lapa.dcpse.weights(torch.float64).abs()).max())
        alpha, steps = 0.01, 100
        dt = min(1e-4, 0.5 * 2.78 / (alpha * gersh))
        fa0 = np.sin(2 * np.pi * wa["pos"][:, 1]) * np.sin(2 * np.pi * wa["pos"][:, 2])
        fa64_0 = torch.as_tensor(fa0, device=dev)
        fa32_0 = fa64_0.to(torch.float32)
        rk4(lapa, fa32_0.clone(), dt, 1, alpha=alpha)  # builds the apply layouts
        rk4(lapa, fa64_0.clone(), dt, 1, alpha=alpha)
        out = {}

        def run(f0, key):
            f = f0.clone()
            rk4(lapa, f, dt, steps, alpha=alpha)
            out[key] = f

        rk32_ms, _ = events_ms(lambda: run(fa32_0, "f32"), 1, torch)
        rk64_ms, _ = events_ms(lambda: run(fa64_0, "f64"), 1, torch)
        rel_a = float(torch.linalg.vector_norm(out["f32"].double() - out["f64"]) / torch.linalg.vector_norm(out["f64"]))
        adaptive = {"workload": "3d_adaptive_h0(1+3x)_x_open_yz_periodic_rcut2.4h(x)_order2_lap_rk4",
                    "particles": int(na), "nnz": int(vla.nnz),
                    "neighbours_min_max": [int(la.min()), int(la.max())], "support_ms": sup_a_ms,
                    "assembly_ms": asm_a_ms, "rk4_steps": steps, "dt": dt, "held_rows": int(held.sum()),
                    "rk4_fp32_ms": rk32_ms, "rk4_fp64_ms": rk64_ms,
                    "fp32_stage_us": rk32_ms / (4 * steps) * 1e3,
                    "fp32_vs_fp64_rel_l2": rel_a}
        log(f"[bench] adaptive N={na}: support {sup_a_ms:.2f} ms, Laplacian assembly {asm_a_ms:.2f} ms, "
            f"{steps} RK4 steps FP32 {rk32_ms:.2f} ms / FP64 {rk64_ms:.2f} ms, FP32 vs FP64 {rel_a:.2e}")
        del lapa, vla, psa, out

    # ---- 2D adjoint: 1024^2 jittered, Laplacian, a batch of 16 FP32 fields,
    # L = sum_b |Lap f_b - g_b|^2 and its gradient through the hand-written
    # transposed apply (torch.autograd.Function) ---------------------------------
    adjoint = None
    if not args.no_adjoint:
        from openfpm_numerics_amd import dcpse_apply

        wj = W.jittered_2d(1024, seed=4)
        psj = ParticleSet(wj["pos"], wj["box_lo"], wj["box_hi"], bc=[(wj["periodic"] >> d) & 1 for d in range(2)],
                          device=dev)
        vlj = psj.getVerlet(wj["rcut"])
        lapj = Laplacian(psj, vlj, 2, wj["rcut"])
        nj, B = psj.size_local(), 16
        rng = np.random.default_rng(4)
        fbj = torch.as_tensor(rng.standard_normal((B, nj)), dtype=torch.float32, device=dev)
        sj, _, _, _ = W.sin_field(wj["pos"])
        gbj = (lapj(torch.as_tensor(sj, device=dev))[None, :]
               + 0.01 * torch.as_tensor(rng.standard_normal((B, nj)), device=dev))

        def grad(dtype):
            f = fbj.detach().to(dtype).requires_grad_(True)
            ((dcpse_apply(lapj, f) - gbj.to(dtype)) ** 2).sum().backward()
            return f.grad

        torch.cuda.synchronize()
        t0 = time.perf_counter()
        lapj.dcpse.transposed(torch.float32)  # the transposed CSR, built once per assembly
        torch.cuda.synchronize()
        tbuild_ms = (time.perf_counter() - t0) * 1e3
        g32 = grad(torch.float32)
        step_ms, _ = events_ms(lambda: grad(torch.float32), 5, torch)
        fwd_ms, _ = events_ms(lambda: lapj.dcpse.computeDifferentialOperator(fbj), 5, torch)
        adj_ms, _ = events_ms(lambda: lapj.dcpse.applyTranspose(fbj), 5, torch)
        g64 = grad(torch.float64)
        rel_g = float(torch.linalg.vector_norm(g32.double() - g64) / torch.linalg.vector_norm(g64))
        adjoint = {"workload": "2d_jitter1024_rcut3.1h_order2_lap_batch16", "particles": int(nj), "nnz": int(vlj.nnz),
                   "fields": B, "transpose_build_ms": tbuild_ms, "forward_batch_ms": fwd_ms,
                   "transpose_batch_ms": adj_ms, "loss_grad_step_ms": step_ms,
                   "fp32_vs_fp64_grad_rel_l2": rel_g}
        log(f"[bench] adjoint 2D N={nj} B={B}: transpose build {tbuild_ms:.2f} ms, forward {fwd_ms:.3f} ms, "
            f"transposed {adj_ms:.3f} ms, loss + gradient {step_ms:.3f} ms, FP32 vs FP64 gradient {rel_g:.2e}")
        del lapj, vlj, psj

    # ---- apply --------------------------------------------------------------------
    f64, dx_exact, _, lap_exact = W.sin_field(w["pos"])
    f64_t = torch.as_tensor(f64, device=dev)
    f32_t = f64_t.to(torch.float32).contiguous()
    out_lap = torch.empty(N, dtype=torch.float32, device=dev)
    out_dx = torch.empty(N, dtype=torch.float32, device=dev)
    step = PreparedApply([lap.dcpse, dx.dcpse], f32_t, [out_lap, out_dx],
                         sell={"sell": True, "sell_bundle": "bundle", "sell_bundle16": "bundle16",
                               "csr": False}[args.layout])
    tuned = None
    sell_pipe = None
    if args.tune:
        # CSR kernels (several FP32 fields, FP64) and the sliced-ELL single-field kernel
        csr_step = PreparedApply([lap.dcpse, dx.dcpse], f32_t, [out_lap, out_dx], sell=False)
        tuned, app_tune = tune_apply(csr_step)
        log("[bench] CSR apply autotune (G,epl,vec,block) -> %s: %s" % (
            tuned[0], ", ".join(f"{k}:{v * 1e3:.1f}us" for k, v in app_tune.items())))
        (sell_pipe, sell_ms), sell_tune = tune_sell(step)
        log("[bench] SELL apply (groups, depth) -> %s: %s" % (
            sell_pipe, ", ".join(f"{k}:{v * 1e3:.1f}us" for k, v in sell_tune.items())))
        if tuned[1] < sell_ms:
            step = csr_step
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    launch_ms, per = events_ms(step, args.steps, torch)
    wall_s = time.perf_counter() - t0
    # whole-region timing (back-to-back launches, no per-launch events)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(args.steps):
        step()
    b.record()
    torch.cuda.synchronize()
    ms_per_step = a.elapsed_time(b) / args.steps
    evals = 2.0 * N
    value = evals / (ms_per_step * 1e-3) / 1e6

    # the same step when the caller hands over host buffers: pinned f uploaded,
    # both outputs downloaded (reported only; value keeps inputs resident)
    f_host = torch.empty(N, dtype=torch.float32, pin_memory=True)
    f_host.copy_(f32_t.cpu())
    o_host = [torch.empty(N, dtype=torch.float32, pin_memory=True) for _ in range(2)]

    def host_step():
        f32_t.copy_(f_host, non_blocking=True)
        step()
        o_host[0].copy_(out_lap, non_blocking=True)
        o_host[1].copy_(out_dx, non_blocking=True)

    for _ in range(3):
        host_step()
    torch.cuda.synchronize()
    reps_h = max(10, args.steps // 4)
    a.record()
    for _ in range(reps_h):
        host_step()
    b.record()
    torch.cuda.synchronize()
    pcie_ms = a.elapsed_time(b) / reps_h

    # algorithmic bytes of one fused launch (the per-neighbour unit of the
    # operator: a 4-byte column and one 4-byte weight per operator; row_ptr
    # (int64), f once, 2 outputs per particle). The 16-bit-offset layout moves
    # fewer bytes than this (2-byte columns): its measured traffic is below it.
    alg_bytes = nnz * (4 + 4 + 4) + (N + 1) * 8 + N * 4 + 2 * N * 4
    layout_bytes = nnz * ((2 if step.layout == "sell_bundle16" else 4) + 4 + 4) + (N + 1) * 8 + N * 4 + 2 * N * 4
    achieved = alg_bytes / (launch_ms * 1e-3) / 1e9
    traffic = read_traffic()

    # ---- accuracy -----------------------------------------------------------------
    lap64, dx64 = [t for t in __import__("openfpm_numerics_amd").apply_ops(
        [lap.dcpse, dx.dcpse], f32_t.to(torch.float64))]
    # ---- FP64 apply (the reference's precision): the fused pass on the FP64
    # sliced-ELL copy, which the SELL_AFTER-th FP64 apply builds (warm-up)
    from openfpm_numerics_amd import dcpse as _dcpse
    f64_d = f32_t.to(torch.float64)
    o64 = [torch.empty_like(f64_d) for _ in range(2)]

    def run64():
        _dcpse.apply_ops([lap.dcpse, dx.dcpse], f64_d, outs=o64)

    for _ in range(_dcpse.SELL_AFTER + 2):
        run64()
    torch.cuda.synchronize()
    reps64 = 20
    a.record()
    for _ in range(reps64):
        run64()
    b.record()
    torch.cuda.synchronize()
    fp64_apply_us = a.elapsed_time(b) / reps64 * 1e3
    if lap.dcpse._W.get("sell64") is None:
        raise RuntimeError("the FP64 apply did not switch to its sliced-ELL copy")
    log(f"[bench] FP64 fused apply {fp64_apply_us:.1f} us -> {2 * N / fp64_apply_us:.1f} M evals/s")
    # the same fused FP64 pass on the 16-bit-offset FP64 bundle (18 instead of
    # 20 bytes per neighbour and operator pair slot)
    o64b = [torch.empty_like(f64_d) for _ in range(2)]
    step64 = PreparedApply([lap.dcpse, dx.dcpse], f64_d, o64b, sell="bundle16")
    for _ in range(4):
        step64()
    fp64_b16_ms, _ = events_ms(step64, reps64, torch)
    fp64_b16_diff = max(float((x - y).abs().max() / y.abs().max()) for x, y in zip(o64b, o64))
    log(f"[bench] FP64 fused apply, 16-bit bundle: {fp64_b16_ms * 1e3:.1f} us (max rel. diff {fp64_b16_diff:.1e})")
    del step64, o64b
    err = {}
    for name, got, ref in (("laplacian", out_lap, lap64), ("dx", out_dx, dx64)):
        d = got.double() - ref
        err[name] = {"rel_l2": float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(ref)),
                     "max_rel": float(d.abs().max() / ref.abs().max())}
    # the same step with every operation of the row sums in FP32 (the
    # reference's float-property apply); FP64 sums stay the default
    from openfpm_numerics_amd import set_apply_accumulation
    acc32_out = [torch.empty_like(out_lap), torch.empty_like(out_dx)]
    step32 = PreparedApply([lap.dcpse, dx.dcpse], f32_t, acc32_out,
                           sell={"sell": True, "sell_bundle": "bundle", "sell_bundle16": "bundle16",
                                 "csr": False}[args.layout])
    acc32 = None
    if step32.sell:
        set_apply_accumulation(32)
        try:
            for _ in range(3):
                step32()
            acc32_ms, _ = events_ms(step32, max(20, args.steps // 2), torch)
        finally:
            set_apply_accumulation(64)
        err32 = {}
        for name, got, ref in (("laplacian", acc32_out[0], lap64), ("dx", acc32_out[1], dx64)):
            d = got.double() - ref
            err32[name] = {"rel_l2": float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(ref)),
                           "max_rel": float(d.abs().max() / ref.abs().max())}
        acc32 = {"launch_us": acc32_ms * 1e3, "evals_per_s_M": evals / (acc32_ms * 1e-3) / 1e6,
                 "fp32_error_vs_fp64": err32}
        log(f"[bench] FP32 accumulation: {acc32_ms * 1e3:.1f} us per launch, error {err32}")
    del step32
    ana = {
        "laplacian_max_rel": float((lap64.cpu().numpy() - lap_exact).__abs__().max() / np.abs(lap_exact).max()),
        "dx_max_rel": float((dx64.cpu().numpy() - dx_exact).__abs__().max() / np.abs(dx_exact).max()),
    }
    log(f"[bench] step {ms_per_step * 1e3:.1f} us ({launch_ms * 1e3:.1f} us per launch) -> {value:.1f} "
        f"M evals/s, {achieved:.0f} GB/s ({achieved / HBM_PEAK_GBS:.1%} of HBM); fp32 err {err}")

    # assembly roofline (FP64 flops of the reference's algorithm per particle)
    def asm_flops(n, m):
        return n * m * (m + 1) + m ** 3 / 3.0 + 2 * m * m + 3 * n * m
    nmean = nnz / N
    flops_lap = N * 3 * asm_flops(nmean, 19)
    flops_dx = N * asm_flops(nmean, 10)
    asm_tflops = (flops_lap + flops_dx) / (assembly_ms * 1e-3) / 1e12
    # FP64 flops the pair kernel executes (64 lanes x (2 FMA + ADD + MUL) wave
    # instructions, rocprofv3 counters of the default launch shape)
    asm_exec = read_profile_value("assemble_lap_dx_3d128.pmc.txt", "executed_fp64_flop_per_launch")

    # ---- CPU baseline -----------------------------------------------------------------
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        # every core this process may run on, capped by OMP_NUM_THREADS where the
        # host sets it (the GPU box gives one GPU's job a 16-core share that way)
        visible = len(os.sched_getaffinity(0))
        omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
        cores = min(visible, omp) if omp > 0 else visible
        rp = vl.row_ptr.cpu().numpy()
        col = vl.col.cpu().numpy()
        sigs = [(2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 0, 0)]
        asms = []
        for sgn in sigs:
            d = Dcpse(ps, vl, sgn, 2, rcut)
            asms.append(O.Assembly(d.localEps.cpu().numpy(), d.localEpsInvPow.cpu().numpy(),
                                   d.calcKernels.cpu().numpy(), d.info.cpu().numpy(), sgn, 2))
            del d
        f32h = f32_t.cpu().numpy().astype(np.float64)
        O.apply(asms[0], rp, col, f32h, nthreads=cores)  # warm
        reps = 0
        t0 = time.perf_counter()
        while True:
            for a_ in asms:
                O.apply(a_, rp, col, f32h, nthreads=cores)
            reps += 1
            el = time.perf_counter() - t0
            if el > 0.4 * args.cpu_seconds or reps >= 20:
                break
        cpu_step_s = el / reps
        # assembly on a particle sample, extrapolated to N
        S = 2000
        t0 = time.perf_counter()
        for sgn in sigs:
            O.assemble(w["pos"], rp, col, sgn, 2, w["box_lo"], w["box_hi"], w["periodic"],
                       p_range=(0, S), nthreads=cores)
        t_small = time.perf_counter() - t0
        S = int(min(N, max(S, S * 0.5 * args.cpu_seconds / max(t_small, 1e-3))))
        t0 = time.perf_counter()
        for sgn in sigs:
            O.assemble(w["pos"], rp, col, sgn, 2, w["box_lo"], w["box_hi"], w["periodic"],
                       p_range=(0, S), nthreads=cores)
        t_asm = time.perf_counter() - t0
        cpu_asm_ms = t_asm / S * N * 1e3
        cpu = {"value": evals / cpu_step_s / 1e6, "unit": "M particle-op evals/s", "cores": cores,
               "kind": "port",
               "sample": f"apply: full N={N}, Laplacian as 3 FP64 operators + Dx, reference form, {reps} reps; "
                         f"assembly: first {S} particles x 4 operators (dense E*V, pivoted QR), "
                         f"extrapolated to N",
               "assembly_ms_extrapolated": cpu_asm_ms,
               "step_ms": cpu_step_s * 1e3,
               "host_cores_visible": visible, "omp_num_threads": omp or None}
        log(f"[bench] CPU ({cores} threads of {visible} visible cores, OMP_NUM_THREADS={omp or 'unset'}): step {cpu_step_s * 1e3:.1f} ms -> {cpu['value']:.1f} M evals/s; "
            f"assembly {cpu_asm_ms:.0f} ms (extrapolated from {S})")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "M particle-op evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "3d_lattice128_jitter0.1h_minus10pct_periodic_rcut2.4h_order2_lap+dx",
                   "particles": N, "nnz": nnz, "particle_order": args.order, "operators": ["Laplacian", "Derivative_x"],
                   "apply_precision": "fp32", "assembly_precision": "fp64",
                   "apply_layout": step.layout,
                   "apply_launch": ({"sell_groups": (sell_pipe or (2, 2))[0], "sell_depth": (sell_pipe or (2, 2))[1]}
                                    if step.sell else
                                    {"group_lanes": tuned[0][0], "elems_per_lane": tuned[0][1],
                                     "vector_width": tuned[0][2], "block": tuned[0][3]})},
        "assembly_ms": assembly_ms,
        "assembly_launch": {"persistent": asm_cfg[0], "particles_per_batch": asm_cfg[1], "chunk": asm_cfg[2],
                            "bucket_width": asm_cfg[3]},
        "assembly_ms_static_grid": assembly_default_ms,
        "assembly_ms_by_operator": {"laplacian": asm_lap_ms, "dx": asm_dx_ms},
        "assembly_uneven": uneven,
        "adaptive_rk4": adaptive,
        "adjoint_2d": adjoint,
        "assembly_ms_with_refinement": assembly_refined_ms,
        "support_ms": support_ms,
        "fp64_apply_us": fp64_apply_us,
        "fp64_apply_evals_per_s_M": 2 * N / fp64_apply_us,
        "fp64_apply_bundle16_us": fp64_b16_ms * 1e3,
        "fp32_error_vs_fp64": err,
        "fp32_accumulation": acc32,
        "fp64_error_vs_analytic": ana,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": ("k_apply_sell%s<2,%d,0,%s,false>%s" % ("", (sell_pipe or (2, 2))[0],
                                                                "true" if step.layout == "sell_bundle16" else "false",
                                                          {"sell_bundle": " (interleaved bundle)",
                                                           "sell_bundle16": " (interleaved bundle, 16-bit columns)"}
                                                          .get(step.layout, "")))
                     if step.sell else "k_apply_v4<G,NCH,2,1,diff>", "alg_bytes_per_launch": alg_bytes, "layout_min_bytes_per_launch": layout_bytes,
                     "launch_ms": launch_ms},
        "roofline_assembly": {"bound": "fp64", "achieved": asm_tflops, "peak": FP64_PEAK_TFLOPS,
                              "unit": "TFLOP/s", "frac": asm_tflops / FP64_PEAK_TFLOPS,
                              "flops": flops_lap + flops_dx,
                              "flops_basis": "reference algorithm (Laplacian as three operators)",
                              "executed_flops": asm_exec,
                              "executed_achieved": (asm_exec / (assembly_ms * 1e-3) / 1e12) if asm_exec else None,
                              "executed_frac": ((asm_exec / (assembly_ms * 1e-3) / 1e12) / FP64_PEAK_TFLOPS
                                                if asm_exec else None)},
        "cpu_baseline": cpu,
        "wall_s_timed_region": wall_s,
        "host_buffers": {"ms_per_step": pcie_ms, "value": evals / (pcie_ms * 1e-3) / 1e6,
                         "note": "pinned host f uploaded and both outputs downloaded around each step"},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
