"""The learned correction gate (care_amd/csrc/mlp.hip, mlp_device.h).

CPU: the numpy oracle (oracle/mlp_oracle.py) is pinned by finite differences
of its float64 network and closed forms of its bfloat16 rounding; the
device kernels, run through the CPU emulation of the HIP API, must match the
oracle's bf16-mirroring gradient and logits.  GPU: the same comparisons on
the MI355X (BF16 matrix instructions), training that learns a known rule,
and the recorded decision columns of the gated correction mode.
"""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

import care_amd
from tests import helpers

sys.path.insert(0, os.path.join(helpers.ROOT, "oracle"))
import mlp_oracle  # noqa: E402

EMU_DIR = os.path.join(helpers.ROOT, "tests", "emu")
EMU_LIB = os.path.join(EMU_DIR, "_build", "libcare_amd_emu.so")


def _data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, 1.5, size=(n, 32)).astype(np.float32)
    y = (x[:, 0] + 0.5 * x[:, 3] > 0.9).astype(np.uint8)
    return x, y


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


# ---------------------------------------------------------------- oracle pins

def test_oracle_bf16_rounding_closed_forms():
    vals = np.array([1.0, -2.5, 0.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -9, 3.0e-3],
                    dtype=np.float32)
    got = mlp_oracle.bf16(vals)
    # halfway cases round to the even mantissa: 1+2^-8 -> 1, 1+3*2^-8 -> 1+2^-6
    want = [1.0, -2.5, 0.0, 1.0, 1.0 + 2.0 ** -6, 1.0, None]
    for g, w in zip(got[:6], want[:6]):
        assert g == w
    assert abs(got[6] - 3.0e-3) / 3.0e-3 < 2.0 ** -8
    # idempotent
    assert np.array_equal(mlp_oracle.bf16(got.astype(np.float32)), got)


def test_oracle_gradient_matches_finite_differences():
    x, y = _data(40, 1)
    p = mlp_oracle.init_params(3).astype(np.float64)
    p[4096:4160] = 0.05  # non-zero b1 so that ReLU kinks are away from zero
    loss, g = mlp_oracle.loss_and_grad(x, y, p, bf16_round=False)
    rng = np.random.default_rng(4)
    idx = np.concatenate([rng.choice(2048, 12, replace=False), 2048 + rng.choice(64, 4, replace=False),
                          2112 + rng.choice(2048, 12, replace=False), 4160 + rng.choice(32, 4, replace=False),
                          4192 + rng.choice(512, 8, replace=False), 4704 + np.arange(16), [4720, 4728, 4736]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp, _ = mlp_oracle.loss_and_grad(x, y, pp, bf16_round=False)
        lm, _ = mlp_oracle.loss_and_grad(x, y, pm, bf16_round=False)
        num = (lp - lm) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 + 1e-4 * abs(num), (i, num, g[i])


def test_oracle_bf16_gradient_close_to_exact():
    x, y = _data(256, 5)
    p = mlp_oracle.init_params(6)
    _, g_exact = mlp_oracle.loss_and_grad(x, y, p, bf16_round=False)
    _, g_bf = mlp_oracle.loss_and_grad(x, y, p, bf16_round=True)
    assert _rel(g_bf, g_exact) < 0.05


# ------------------------------------------------------- emulated device kernels

@pytest.fixture(scope="module")
def emu_engine():
    subprocess.run(["sh", os.path.join(EMU_DIR, "build_emu.sh")], check=True)
    lib = care_amd._bind(ctypes.CDLL(EMU_LIB), care_amd.ENGINE_SYMBOLS)
    eng = care_amd.Engine(care_amd.default_options(), device=0, lib=lib)
    yield eng
    eng.close()


def test_emulated_gradient_matches_oracle(emu_engine):
    x, y = _data(100, 7)
    p = mlp_oracle.init_params(8)
    g, loss = emu_engine.mlp_gradient(x, y, p)
    lo, go = mlp_oracle.loss_and_grad(x, y, p)
    assert abs(loss - lo) < 1e-4 * max(1.0, abs(lo))
    assert _rel(g, go) < 2e-3


def test_emulated_logits_match_oracle(emu_engine):
    x, _ = _data(70, 9)
    p = mlp_oracle.init_params(10)
    z = emu_engine.mlp_infer(x, p)
    zo, _ = mlp_oracle.forward(x, p)
    assert np.max(np.abs(z - zo)) < 1e-3 * max(1.0, np.max(np.abs(zo)))


def test_emulated_training_learns_rule(emu_engine):
    x, y = _data(512, 11)
    p0 = mlp_oracle.init_params(12)
    p, losses = emu_engine.mlp_train(x, y, p0, epochs=15, batch=64, lr=3e-3, seed=2)
    assert losses[-1] < 0.5 * losses[0]
    z, _ = mlp_oracle.forward(x, p)
    assert np.mean((z > 0) == (y == 1)) > 0.9


# ------------------------------------------------------------------- GPU

@pytest.mark.gpu
def test_gpu_gradient_matches_oracle():
    x, y = _data(4096, 21)
    p = mlp_oracle.init_params(22)
    with care_amd.Engine(care_amd.default_options(), device=0) as eng:
        g, loss = eng.mlp_gradient(x, y, p)
    lo, go = mlp_oracle.loss_and_grad(x, y, p)
    assert abs(loss - lo) < 1e-4 * max(1.0, abs(lo))
    assert _rel(g, go) < 2e-3


@pytest.mark.gpu
def test_gpu_logits_match_oracle():
    x, _ = _data(1000, 23)
    p = mlp_oracle.init_params(24)
    with care_amd.Engine(care_amd.default_options(), device=0) as eng:
        z = eng.mlp_infer(x, p)
    zo, _ = mlp_oracle.forward(x, p)
    assert np.max(np.abs(z - zo)) < 1e-3 * max(1.0, np.max(np.abs(zo)))


@pytest.mark.gpu
def test_gpu_training_learns_rule():
    x, y = _data(20000, 25)
    p0 = mlp_oracle.init_params(26)
    with care_amd.Engine(care_amd.default_options(), device=0) as eng:
        p, losses = eng.mlp_train(x, y, p0, epochs=5, batch=256, lr=3e-3, seed=3)
    assert losses[-1] < 0.5 * losses[0]
    xt, yt = _data(5000, 27)
    z, _ = mlp_oracle.forward(xt, p)
    assert np.mean((z > 0) == (yt == 1)) > 0.95
