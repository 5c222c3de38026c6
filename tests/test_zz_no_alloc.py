"""care_correct_batch_noalloc (include/care_amd.h): the correction entry
point that allocates nothing, its batch scratch in a fixed device array.

Two calls in a row, the second over fewer anchors, must give exactly the
results of care_correct_batch over the same anchors: anchor status and
corrected rows, and the candidate corrections (compared as a set: their
order depends on the device's work queue).  The CPU test runs the kernels
through the HIP emulation; the GPU test runs on the device.
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

# paired-end with candidate correction: h = 48 and coverage 40 give about
# 183 KB of scratch per anchor, so a call runs in several device batches
SMALL = {"profile": "large-pe", "seed": 241, "genome": 7000}  # two device batches
GPU_CASE = {"profile": "large-pe", "seed": 242, "genome": 120000}
OPTIONS = dict(kmer_length=20, num_hash_maps=48, coverage=40.0, paired_end=1, use_quality=1, correct_candidates=1)


def _results(eng, first, n, noalloc):
    status, seqs, cids, cseqs = eng.correct(first, n, noalloc=noalloc)
    width = eng.max_length
    rows = seqs[(status & 1) != 0, :width].copy()
    cands = sorted((int(i), bytes(cseqs[k, :width])) for k, i in enumerate(cids))
    return status.copy(), rows, cands


def _check(lib, case, d, calls):
    reads, _ = helpers.simulate(case, d)
    rs = care_amd.ReadSet([reads], paired_end=True, qual_bits=8)
    eng = care_amd.Engine(care_amd.default_options(**OPTIONS), device=0, lib=lib)
    eng.upload_readset(rs)
    eng.build()
    for first, n in calls(rs.n):
        print("calls of %d anchors from %d" % (n, first), flush=True)
        want = _results(eng, first, n, False)
        got = _results(eng, first, n, True)
        assert np.array_equal(want[0], got[0])
        assert np.array_equal(want[1], got[1])
        assert want[2] == got[2] and len(got[2]) > 0
    eng.close()
    rs.close()


def test_emulated_noalloc_matches_batch(tmp_path):
    subprocess.run(["sh", os.path.join(EMU_DIR, "build_emu.sh")], check=True)
    lib = care_amd._bind(ctypes.CDLL(EMU_LIB), care_amd.ENGINE_SYMBOLS)
    _check(lib, SMALL, str(tmp_path), lambda n: [(0, n & ~1), (2, (n // 3) & ~1)])


@pytest.mark.gpu
def test_gpu_noalloc_matches_batch(tmp_path):
    print("allocation-free correction on the device", flush=True)
    _check(None, GPU_CASE, str(tmp_path), lambda n: [(0, n & ~1), (1000, (n // 3) & ~1)])
