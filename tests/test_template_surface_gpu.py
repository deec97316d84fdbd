"""The C++ template surface <simd2/srgemm.hpp>, compiled and run.

simd2_amd/simd2_template_driver (tests/cpp/srgemm_template_driver.cpp, built
by simd2_amd/csrc/Makefile) includes the header the way code written against
the reference's templates does and calls simd2::srgemm<AddOp, MulOp, T> for
all nine semirings on an off-grid column-major shape (kernel/host_srgemm.h
:16-30, kernel/srgemm.cuh column-major convention), the comp_update(int*)
overload (kernel/converge.cuh:71-85) and the __half overloads of
tensor_srgemm, f2h_device and h2f_device (kernel/tensor_srgemm.cuh:197-218,
kernel/precision.cuh:21-33).  It writes its inputs and outputs to files;
this test checks them against the oracle: bit for bit for the idempotent
semirings, or_and and the integer-valued minus_square, plus_mul within 1e-5
relative of FP64, the FP16 emulation against its numpy restatement."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "simd2_amd", "simd2_template_driver")


def bits(x):
    return np.ascontiguousarray(x).view(np.uint32)


@pytest.fixture(scope="module")
def outputs(tmp_path_factory, gpu):
    assert os.path.exists(DRIVER), f"{DRIVER} is not built (make -C simd2_amd/csrc)"
    d = tmp_path_factory.mktemp("template")
    print(f"running {DRIVER} {d}", flush=True)
    r = subprocess.run([DRIVER, str(d)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    arrays = {}
    with open(d / "manifest.txt") as f:
        for line in f:
            name, dtype, rows, cols = line.split()
            dt = np.float32 if dtype == "f32" else np.int32
            arrays[name] = np.fromfile(d / f"{name}.bin", dtype=dt).reshape(int(rows), int(cols))
    return arrays


@pytest.mark.parametrize("op", ["min_plus", "plus_mul", "max_plus", "max_min", "max_mul",
                                "min_mul", "min_max", "minus_square"])
def test_float_semirings_through_template(outputs, oracle_mod, op):
    # files hold the column-major matrices as row-major transposes
    A, B, C, D = (outputs[f"{op}_{x}"].T for x in "ABCD")
    if op == "plus_mul":
        # within 1e-5 of the sum of the terms' magnitudes (MFMA accumulation order)
        A64, B64, C64 = (x.astype(np.float64) for x in (A, B, C))
        err = np.abs(D - (A64 @ B64 + C64)) / (np.abs(A64) @ np.abs(B64) + np.abs(C64))
        assert err.max() <= 1e-5, err.max()
        return
    want = oracle_mod.srgemm_f32(op, A, B, C, True, threads=8)
    assert np.array_equal(bits(D), bits(want)), op


def test_or_and_through_template(outputs, oracle_mod):
    A, B, C, D = (outputs[f"or_and_{x}"].T for x in "ABCD")
    want = oracle_mod.srgemm_i32(A, B, C, True)
    assert np.array_equal(D, want)


def test_comp_update_int_overload(outputs):
    same, changed = outputs["comp_update_flags"][0]
    assert (same, changed) == (0, 1)  # unchanged matrix, then one changed entry
    assert np.array_equal(outputs["comp_update_prev_after"], outputs["comp_update_cur"])


def test_half_overloads(outputs):
    from oracle import fp16_emulation as emu
    a = outputs["tensor_A"]
    assert np.array_equal(outputs["tensor_A_roundtrip"], a.astype(np.float16).astype(np.float32))
    want = emu.tensor_srgemm(a.astype(np.float16), outputs["tensor_B"].astype(np.float16),
                             outputs["tensor_C"], "min_plus", True)
    assert np.array_equal(bits(outputs["tensor_D"]), bits(want))
