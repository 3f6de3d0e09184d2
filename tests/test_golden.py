"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the pinned FP64 oracle): the oracle must keep reproducing them (CPU), and
libraft must match them through the C ABI (GPU)."""
import glob
import os

import numpy as np
import pytest

import oracle

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return dict(np.load(os.path.join(HERE, name), allow_pickle=False))


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, float) - np.asarray(b, float)) / max(np.linalg.norm(b), 1e-300))


FBP = sorted(os.path.basename(p) for p in glob.glob(os.path.join(HERE, "fbp_*.npz")))
RADON = sorted(os.path.basename(p) for p in glob.glob(os.path.join(HERE, "radon_*.npz")))


@pytest.mark.parametrize("name", FBP)
def test_oracle_reproduces_fbp_fixture(name):
    f = load(name)
    out = oracle.fbp(f["tomo"], f["angles"], f["paramf"], f["parami"])
    np.testing.assert_allclose(out, f["obj"], rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("name", RADON)
def test_oracle_reproduces_radon_fixture(name):
    f = load(name)
    np.testing.assert_allclose(oracle.radon(f["phantom"], f["angles"]), f["proj"], rtol=1e-13, atol=1e-14)


def test_oracle_reproduces_eem_fixture():
    f = load("eem_n32_a32.npz")
    for k in (1, 5, 20):
        x, kl, ex = oracle.eem(f["tomo"], f["angles"], k)
        np.testing.assert_allclose(x, f[f"obj_k{k}"], rtol=1e-11, atol=1e-13)
        np.testing.assert_allclose(kl, f[f"kl_k{k}"], rtol=1e-11)
        np.testing.assert_array_equal(ex, f[f"ex_k{k}"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", FBP)
def test_gpu_matches_fbp_fixture(gpu, name):
    import sscRaft

    f = load(name)
    pf, pi = f["paramf"], f["parami"]
    obj = np.zeros((int(pi[2]), int(pi[0]), int(pi[0])), np.float32)
    sscRaft.libraft.getFBPMultiGPU(np.array([0], np.intc).ctypes.data_as(sscRaft.ctypes.c_void_p), 1,
                                   obj.ctypes.data_as(sscRaft.ctypes.c_void_p),
                                   np.ascontiguousarray(f["tomo"], np.float32).ctypes.data_as(sscRaft.ctypes.c_void_p),
                                   np.ascontiguousarray(f["angles"], np.float32).ctypes.data_as(
                                       sscRaft.ctypes.c_void_p),
                                   np.ascontiguousarray(pf, np.float32).ctypes.data_as(sscRaft.ctypes.c_void_p),
                                   np.ascontiguousarray(pi, np.int32).ctypes.data_as(sscRaft.ctypes.c_void_p))
    sscRaft.raft_check("getFBPMultiGPU")
    assert rel(obj, f["obj"]) <= 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("name", RADON)
def test_gpu_matches_radon_fixture(gpu, name):
    import sscRaft

    f = load(name)
    nx = f["phantom"].shape[-1]
    out = sscRaft.radon_RT(f["phantom"], f["angles"], [0], pixel=2.0 / nx)
    assert rel(out, f["proj"]) <= 1e-5


@pytest.mark.gpu
def test_gpu_matches_eem_fixture(gpu):
    import sscRaft

    f = load("eem_n32_a32.npz")
    for k in (1, 5, 20):
        x, kl, ex = sscRaft.eEMRT_GPU_KL(f["tomo"], f["angles"], k, [0], 0)
        assert rel(x, f[f"obj_k{k}"]) <= (1e-5 if k <= 5 else 1e-4)
        np.testing.assert_allclose(kl, f[f"kl_k{k}"], rtol=1e-4)
        np.testing.assert_array_equal(ex, f[f"ex_k{k}"])
