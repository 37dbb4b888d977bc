"""Runs the C++ drop-in program (tests/cpp/dropin_test.cpp): the operator
classes of include/openfpm_dcpse used as the reference's own tests use theirs
(Dcpse_2D/3D_test, dcpse_op_tests, dcpse_op_save_load, dcpse_op_test_lap),
plus FP32 fields, in-place expressions, the row accessors, checkMomenta, FP64 applies on the sliced-ELL copy
and the DCPSE_VERBOSE diagnostics."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "dropin_test")


@pytest.mark.gpu
def test_cpp_dropin_program(cuda, tmp_path):
    assert os.path.exists(BIN), "tests/cpp/dropin_test is not built (run __graft_entry__.build())"
    proc = subprocess.run([BIN], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(proc.stdout)
    print(proc.stderr)
    assert proc.returncode == 0, proc.stdout[-4000:] + proc.stderr[-2000:]
    assert "0 failure(s)" in proc.stdout
    assert proc.stdout.count("[  OK  ]") == 11


def _read(d):
    import numpy as np

    out = {}
    for line in open(os.path.join(d, "manifest.txt")):
        name, n = line.split()
        stem, ext = name.rsplit(".", 1)
        dt = {"f64": np.float64, "i64": np.int64, "i32": np.int32}[ext]
        a = np.fromfile(os.path.join(d, name), dtype=dt)
        assert a.size == int(n), name
        out[stem] = a
    return out


def _rel(got, want):
    import numpy as np

    return float(np.abs(got - want).max() / np.abs(want).max())


@pytest.mark.gpu
def test_cpp_parity_against_oracle(cuda, oracle, tmp_path):
    """The C++ drop-in's outputs against the oracle (FP64, max-abs / max|ref|
    <= 1e-10): expressions v = 2 Dx(P) + Dy(P), Lap(P), a mixed sum and an
    in-place update (vector_dist_operators.hpp:1113-1138 semantics); the
    per-particle computeDifferentialOperator(key, o1) and (key, o1, i) driven
    by a VECT_DCPSE-shaped node and its value_nz rows (DCPSE_op.hpp:29-106);
    the operator classes over a Dcpse_type with only the reference contract;
    reassembly after particles move (update); a two-set operator on a subset
    domain; and kernel files passed between the C++ and Python layers."""
    import numpy as np
    import scipy.sparse as sp
    import torch

    from openfpm_numerics_amd import Derivative_x, LOAD, ParticleSet

    assert os.path.exists(BIN), "tests/cpp/dropin_test is not built (run __graft_entry__.build())"
    d = str(tmp_path)
    proc = subprocess.run([BIN, d], capture_output=True, text=True, timeout=300)
    print(proc.stdout)
    assert proc.returncode == 0, proc.stdout[-4000:] + proc.stderr[-2000:]
    a = _read(d)
    lo, hi, rc = [0.0, 0.0], [1.0, 1.0], 3.1 / 48
    pos = a["expr_pos"].reshape(-1, 2)
    P = a["expr_P"]
    rp, col = oracle.support(pos, rc, lo, hi, 3)
    ops = {b: oracle.assemble(pos, rp, col, b, 2, lo, hi, 3) for b in ((1, 0), (0, 1), (2, 0), (0, 2), (1, 1))}
    D = {b: oracle.apply(asm, rp, col, P) for b, asm in ops.items()}
    lap = D[(2, 0)] + D[(0, 2)]
    checks = {
        "v = 2 Dx(P) + Dy(P)": (a["expr_v"], 2 * D[(1, 0)] + D[(0, 1)]),
        "Lap(P)": (a["expr_lap"], lap),
        "P - 0.5 Lap(P) + Dx(P)": (a["expr_w"], P - 0.5 * lap + D[(1, 0)]),
        "q = q + 0.25 Dy(q)": (a["expr_inplace"], P + 0.25 * D[(0, 1)]),
        "node Dx": (a["keyed_dx"], D[(1, 0)]),
        "node Dxy": (a["keyed_dxy"], D[(1, 1)]),
        "(key, o1, 0)": (a["keyed_comp0"], D[(1, 0)]),
        "(key, o1, 1)": (a["keyed_comp1"], oracle.apply(ops[(1, 0)], rp, col, a["keyed_f5"])),
        "reference-shaped 2 Dx + Dy": (a["ref_v"], 2 * D[(1, 0)] + D[(0, 1)]),
        "reference-shaped Lap": (a["ref_lap"], lap),
    }
    for name, (got, want) in checks.items():
        err = _rel(got, want)
        print(f"[cpp parity] {name}: {err:.2e}")
        assert err <= 1e-10, (name, err)
    nz = sp.csr_matrix((a["nz_val"], a["nz_col"], a["nz_rowptr"]), shape=(len(P), len(P)))
    ref_nz = oracle.value_nz(ops[(1, 1)], rp, col)
    assert np.abs((nz - ref_nz).toarray()).max() <= 1e-10 * np.abs(ref_nz.data).max()

    # reassembly after motion: the operators at the final positions
    h = 2.0 / 40
    upos = a["upd_pos"].reshape(-1, 2)
    urp, ucol = oracle.support(upos, 3.1 * h, [-1, -1], [1, 1], 0)
    C = a["upd_C"]
    for b, key in (((2, 0), "upd_dxx"), ((0, 2), "upd_dyy")):
        asm = oracle.assemble(upos, urp, ucol, b, 2, [-1, -1], [1, 1], 0)
        err = _rel(a[key], oracle.apply(asm, urp, ucol, C))
        print(f"[cpp parity] after motion {key}: {err:.2e}")
        assert err <= 1e-10, (key, err)
    k = np.arange(41 * 41)
    lattice = np.stack([-1 + (k % 41) * h, -1 + (k // 41) * h], 1)  # grid order, dimension 0 fastest
    assert 1e-4 < np.abs(upos - lattice).max() < 0.1  # the particles moved, by a few steps' worth

    # two sets: subset domain
    ids = a["two_ids"]
    trp, tcol = oracle.support(pos, rc, lo, hi, 3, dom_ids=ids)
    asm = oracle.assemble(pos, trp, tcol, (1, 0), 2, lo, hi, 3, pos_dom=pos[ids])
    want = oracle.apply(asm, trp, tcol, P, f_dom=P[ids])
    for key in ("two_keyed", "two_bulk"):
        err = _rel(a[key], want)
        print(f"[cpp parity] two sets {key}: {err:.2e}")
        assert err <= 1e-10, (key, err)

    # C++ -> Python: the C++ kernel file loads into the Python operator
    ps = ParticleSet(pos, lo, hi, bc=[1, 1], device=cuda)
    vl = ps.getVerlet(rc)
    dx_py = Derivative_x(ps, vl, 2, rc)
    dx_ld = Derivative_x(ps, vl, 2, rc, LOAD)
    dx_ld.load(ps, os.path.join(d, "cpp_dx"))
    assert torch.equal(dx_ld.dcpse.calcKernels, dx_py.dcpse.calcKernels)
    Pt = torch.as_tensor(P, device=cuda)
    assert _rel(dx_ld(Pt).cpu().numpy(), D[(1, 0)]) <= 1e-10

    # Python -> C++: a Python kernel file loads into the C++ operator; a file of
    # another support is refused by both layers
    dx_py.dcpse.save(os.path.join(d, "py_dx"))
    vl2 = ps.getVerlet(3.4 / 48)
    Derivative_x(ps, vl2, 2, 3.4 / 48).dcpse.save(os.path.join(d, "py_dx_other"))
    with pytest.raises(ValueError):
        dx_ld.load(ps, os.path.join(d, "py_dx_other"))
    proc = subprocess.run([BIN, d, "load-py"], capture_output=True, text=True, timeout=300)
    print(proc.stdout)
    assert proc.returncode == 0, proc.stdout[-4000:] + proc.stderr[-2000:]
    b = _read(d)
    assert _rel(b["pyload_dx"], dx_py(Pt).cpu().numpy()) <= 1e-12
