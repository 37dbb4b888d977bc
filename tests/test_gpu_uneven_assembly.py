"""Batched assembly of uneven supports: one launch for particles with 20 to
200 neighbours (uneven_3d: spacing h0 (1 + 1.1 x), one cut-off), in every
launch shape the autotuner chooses from (static or persistent grid, particles
per batch, neighbour chunk, bucketing by row length). Replaces the
grid-stride assembleLocalMatrices_gpu<<<numSMs, 256>>> of
src/DCPSE/Dcpse.cuh:668-669 for the reference's CPU algorithm,
src/DCPSE/Dcpse.hpp:690-809.

Tolerances: every launch shape gives the default shape's weights bit for bit
(each particle's arithmetic does not depend on which wavefront runs it); the
default against the oracle per row <= 1e-10 (or 16 cond(A) u, _dcpse_tol).
"""
import numpy as np
import pytest
import torch
from _dcpse_tol import check_weights

pytestmark = pytest.mark.gpu


def _uneven(cuda, n=30_000):
    from openfpm_numerics_amd import ParticleSet, workloads as W

    w = W.uneven_3d(n)
    ps = ParticleSet(w["pos"], w["box_lo"], w["box_hi"], bc=[1, 1, 1], device=cuda)
    vl = ps.getVerlet(w["rcut"])
    return w, ps, vl


def test_uneven_supports_span_20_to_200(oracle, cuda):
    w, ps, vl = _uneven(cuda)
    rp, col = oracle.support(w["pos"], w["rcut"], w["box_lo"], w["box_hi"], w["periodic"])
    assert np.array_equal(vl.row_ptr.cpu().numpy(), rp) and np.array_equal(vl.col.cpu().numpy(), col)
    lens = np.diff(rp)
    print(f"[uneven] N={len(lens)} n min/p5/p50/p95/max = {lens.min()}/{np.percentile(lens, 5):.0f}/"
          f"{np.median(lens):.0f}/{np.percentile(lens, 95):.0f}/{lens.max()}", flush=True)
    assert lens.min() >= 19 and np.percentile(lens, 5) <= 35 and np.percentile(lens, 95) >= 150


def test_launch_shapes_agree_and_match_oracle(oracle, cuda):
    from openfpm_numerics_amd import Derivative_x, Laplacian
    from openfpm_numerics_amd.autotune import ASSEMBLY_DEFAULT, ASSEMBLY_SPACE, set_assembly_launch
    from openfpm_numerics_amd.dcpse import assemble_together

    w, ps, vl = _uneven(cuda)
    lap = Laplacian(ps, vl, 2, w["rcut"])
    dx = Derivative_x(ps, vl, 2, w["rcut"])
    # the pair pass (shared moments) with the default launch is the reference
    # shape; the operators assembled alone round differently
    set_assembly_launch(*ASSEMBLY_DEFAULT)
    assemble_together([lap.dcpse, dx.dcpse], ps, ps)
    ref_K = [lap.dcpse.calcKernels.clone(), dx.dcpse.calcKernels.clone()]
    ref_W = [lap.dcpse.weights(torch.float32).clone(), dx.dcpse.weights(torch.float32).clone()]
    ref_eps = lap.dcpse.localEps.clone()
    try:
        # beyond the tuning space: one particle per batch, the largest batches
        # (a static batch of 64 runs as 63), and batch sizes that leave the
        # last group of a three-particle factorisation part empty
        extra = ((1, 1, 32, 4), (1, 64, 64, 1), (0, 64, 64, 0), (0, 16, 64, 0), (0, 3, 32, 32), (0, 1, 64, 0))
        for cfg in ASSEMBLY_SPACE + extra:
            print(f"[uneven] launch {cfg}", flush=True)
            set_assembly_launch(*cfg)
            assemble_together([lap.dcpse, dx.dcpse], ps, ps)
            for op, K, W in zip((lap, dx), ref_K, ref_W):
                assert torch.equal(op.dcpse.calcKernels, K), cfg
                assert torch.equal(op.dcpse.weights(torch.float32), W), cfg
            assert torch.equal(lap.dcpse.localEps, ref_eps), cfg
    finally:
        set_assembly_launch(*ASSEMBLY_DEFAULT)
    rp, col = vl.row_ptr.cpu().numpy(), vl.col.cpu().numpy()
    lo, hi, per = w["box_lo"], w["box_hi"], w["periodic"]
    refs = [oracle.assemble(w["pos"], rp, col, b, 2, lo, hi, per, dense_ev=False)
            for b in ((2, 0, 0), (0, 2, 0), (0, 0, 2))]
    rel, n_cond, worst = check_weights(ref_K[0].cpu().numpy(), sum(r.K for r in refs), rp, w["pos"], w["pos"], col,
                                       oracle.basis((2, 0, 0), 2), period=[1.0, 1.0, 1.0], label="lap")
    print(f"[uneven] Laplacian weights vs oracle {rel:.2e} ({n_cond} rows by cond(A))", flush=True)
    ref_dx = oracle.assemble(w["pos"], rp, col, (1, 0, 0), 2, lo, hi, per, dense_ev=False)
    rel, n_cond, worst = check_weights(ref_K[1].cpu().numpy(), ref_dx.K, rp, w["pos"], w["pos"], col,
                                       oracle.basis((1, 0, 0), 2), period=[1.0, 1.0, 1.0], label="dx")
    print(f"[uneven] Dx weights vs oracle {rel:.2e} ({n_cond} rows by cond(A))", flush=True)


def test_autotuner_keeps_the_fastest_shape(cuda):
    from openfpm_numerics_amd import Laplacian
    from openfpm_numerics_amd.autotune import ASSEMBLY_DEFAULT, set_assembly_launch, tune_assembly

    w, ps, vl = _uneven(cuda)
    lap = Laplacian(ps, vl, 2, w["rcut"])
    before = lap.dcpse.calcKernels.clone()
    space = ((0, 8, 32, 0), (1, 8, 32, 0), (1, 8, 32, 16), (1, 16, 64, 16))
    try:
        (cfg, ms), results = tune_assembly(lambda: lap.dcpse.initializeStaticSize(ps, ps, check=False), space=space)
        print(f"[uneven] autotuner: {cfg} {ms:.3f} ms of {results}", flush=True)
        assert cfg in space and ms == min(results.values())
        lap.dcpse.initializeStaticSize(ps, ps)
        assert torch.equal(lap.dcpse.calcKernels, before)
    finally:
        set_assembly_launch(*ASSEMBLY_DEFAULT)
