"""care_correct_batch_graph: the device work of a correction call captured
into a HIP graph once and replayed while calls repeat it.

Every replay must give care_correct_batch's results (status, corrected rows
and candidate corrections), also with other work enqueued between replays,
and a call with another range or other options must capture anew rather
than replay a stale graph.  Calls span several internal device batches, so
the graph holds several query/correction stages and the overlapped result
copies.  Result arrays that are not page-locked are refused.  The CPU test
runs the kernels through the HIP emulation (its capture records launches
and copies and the graph launch runs them in order); the GPU test runs the
same checks on the device.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import care_amd
from tests import helpers

EMU_DIR = os.path.join(helpers.ROOT, "tests", "emu")
EMU_LIB = os.path.join(EMU_DIR, "_build", "libcare_amd_emu.so")

SMALL = {"profile": "repeat-pe", "seed": 241, "genome": 2000}
GPU_CASE = {"profile": "repeat-pe", "seed": 242, "genome": 60000}


def _pinned(lib, n, pitch, cap):
    return (care_amd.pinned_empty((n,), np.uint8, lib=lib), care_amd.pinned_empty((n, pitch), np.uint8, lib=lib),
            care_amd.pinned_empty((max(cap, 1),), np.uint32, lib=lib),
            care_amd.pinned_empty((max(cap, 1), pitch), np.uint8, lib=lib))


def _snapshot(res, n, ml):
    status, seqs, cids, cseqs = res
    st = status[:n].copy()
    rows = seqs[:n, :ml][(st & 1) != 0].copy()
    cands = sorted(zip(cids.tolist(), [bytes(r[:ml]) for r in cseqs]))
    return st, rows, cands


def _same(a, b):
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1], b[1])
    assert a[2] == b[2]


def _check(lib, case, d, batch):
    reads, _ = helpers.simulate(case, d)
    rs = care_amd.ReadSet([reads], paired_end=True, qual_bits=2)
    opt = care_amd.default_options(kmer_length=20, num_hash_maps=32, coverage=30.0, paired_end=1, use_quality=1,
                                   correct_candidates=1)
    eng = care_amd.Engine(opt, device=0, lib=lib)
    eng.set_launch_config(batch, 0)
    eng.upload_readset(rs)
    eng.build()
    n = rs.n & ~1
    ml = rs.max_length
    pitch = (ml + 15) // 16 * 16
    cap = n * 64
    lib = eng.lib
    eager_out = _pinned(lib, n, pitch, cap)
    graph_out = _pinned(lib, n, pitch, cap)
    want = _snapshot(eng.correct(0, n, candidate_capacity=cap, out=eager_out), n, ml)
    assert (want[0] & 1).sum() > 0 and len(want[2]) > 0
    launches = eng.stats().launches
    assert launches >= 3  # several device batches in one graph
    for rep in range(3):
        graph_out[0][:] = 0xEE
        graph_out[1][:] = 0
        got = _snapshot(eng.correct(0, n, candidate_capacity=cap, out=graph_out, graph=True), n, ml)
        _same(want, got)
        assert eng.stats().launches == launches
        # other work on the device between replays
        eng.query(0, min(n, 64))
        _same(want, _snapshot(eng.correct(0, n, candidate_capacity=cap, out=eager_out), n, ml))
    # another range: captured anew
    m = (n // 3) & ~1
    want2 = _snapshot(eng.correct(2, m, candidate_capacity=cap, out=eager_out), m, ml)
    for rep in range(2):
        _same(want2, _snapshot(eng.correct(2, m, candidate_capacity=cap, out=graph_out, graph=True), m, ml))
    # and back: the first range again
    _same(want, _snapshot(eng.correct(0, n, candidate_capacity=cap, out=graph_out, graph=True), n, ml))
    # pageable result arrays are refused
    with pytest.raises(care_amd.CareError, match="page-locked"):
        eng.correct(0, n, candidate_capacity=cap, graph=True)
    eng.close()
    rs.close()


def test_emulated_graph_replays_match_batch(tmp_path):
    subprocess.run(["sh", os.path.join(EMU_DIR, "build_emu.sh")], check=True)
    lib = care_amd._bind(ctypes.CDLL(EMU_LIB), care_amd.ENGINE_SYMBOLS)
    _check(lib, SMALL, str(tmp_path), 128)


@pytest.mark.gpu
def test_gpu_graph_replays_match_batch(tmp_path):
    print("graph-captured correction calls vs care_correct_batch (repeat-pe seed 242)", flush=True)
    _check(None, GPU_CASE, str(tmp_path), 2048)
