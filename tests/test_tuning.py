"""Launch shapes (care_tuning) never change a result.

The query and correction kernels are run under the default shapes, under
shapes that keep every hit list in global scratch instead of LDS (LDS
capacity 16 hits) with two wavefronts per anchor, the smallest MSA chunk and
512-anchor device batches, under shapes that merge hit lists of more than 64
entries list by list instead of sorting them in registers (in LDS and in
global scratch), with page-locked result arrays (per-batch copies
overlapping later batches),
and under the shapes the autotuner keeps; candidate lists, anchor status
and corrected anchors must be identical, and the candidate lists equal the
oracle's.  The CPU tests drive the kernels through the HIP
emulation (tests/emu); the GPU tests run the same checks on the device.
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

SMALL = {
    "pe_q2": {"profile": "repeat-pe", "seed": 211, "genome": 1500,
              "options": dict(kmer_length=20, num_hash_maps=32, coverage=30.0, paired_end=1, use_quality=1),
              "qual_bits": 2},
    "se_mixed": {"profile": "mixed-se", "seed": 212, "genome": 1500,
                 "options": dict(kmer_length=20, num_hash_maps=48, coverage=50.0, paired_end=0, use_quality=1,
                                 exclude_ambiguous=1, correct_candidates=1),
                 "qual_bits": 8},
}
GPU = {"profile": "repeat-pe", "seed": 213, "genome": 40000,
       "options": dict(kmer_length=20, num_hash_maps=32, coverage=30.0, paired_end=1, use_quality=1),
       "qual_bits": 2}

# 512-anchor batches: many device batches enqueued back to back, candidate
# results appended across batches
SPILL = dict(query_lds_hits=16, correct_threads=128, msa_chunk=16, batch_anchors=512)
# lists longer than 64 hits merged list by list instead of sorted in registers,
# three wavefronts per anchor
MERGE = dict(query_register_hits=64, correct_threads=192)


def _readset(case, d):
    reads, _ = helpers.simulate(case, d)
    return care_amd.ReadSet([reads], paired_end=bool(case["options"]["paired_end"]), qual_bits=case["qual_bits"])


def _run(eng, rs, tuning=None, autotune=False):
    if tuning is not None:
        eng.set_tuning(**tuning)
    chosen = eng.autotune(0, rs.n) if autotune else None
    ids, offs = eng.query(0, rs.n)
    status, seqs, cids, cseqs = eng.correct(0, rs.n)
    raw = eng.stats().raw_hits
    corrected = (status & 1) != 0
    rows = seqs[corrected, :rs.max_length].copy()
    order = np.lexsort((np.arange(len(cids)), cids))  # candidate results in id order
    return {"ids": ids.copy(), "offs": offs.copy(), "status": status.copy(), "rows": rows,
            "cids": cids[order].copy(), "cseqs": cseqs[order, :rs.max_length].copy(), "raw": np.array([raw])}, chosen


def _same(a, b):
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def _oracle_lists(rs, opt):
    orc = helpers.oracle()
    seqs = np.ascontiguousarray(rs.seqs)
    lens = np.ascontiguousarray(rs.lengths)
    t = orc.orc_tables_build(seqs.ctypes.data, rs.pitch, lens.ctypes.data, rs.n, opt["kmer_length"],
                             opt["num_hash_maps"], opt["coverage"])
    buf = np.zeros(opt["num_hash_maps"] * 400, dtype=np.uint32)
    out = []
    for i in range(rs.n):
        c = orc.orc_tables_query(t, seqs[i].ctypes.data, int(lens[i]), i, buf.ctypes.data, len(buf))
        out.append(buf[:c].copy())
    orc.orc_tables_free(t)
    return out


def _launches(n, batch):
    """Device batches of a correction call over n anchors: batches of
    `batch`, the last batch + batch / 8 anchors split so that the call ends
    in a batch of batch / 8 (a call of more than one batch)."""
    tail = (batch // 8) & ~1
    done = k = 0
    while done < n:
        rem = n - done
        done += rem - tail if (n > batch and tail >= 2 and tail < rem <= batch + tail) else min(batch, rem)
        k += 1
    return k


def _check(lib, case, d, device_autotune):
    rs = _readset(case, d)
    opt = care_amd.default_options(**case["options"])
    results = []
    for tuning, auto in ((None, False), (SPILL, False), (None, device_autotune), (MERGE, False),
                         (dict(SPILL, **MERGE), False)):
        eng = care_amd.Engine(opt, device=0, lib=lib)
        eng.upload_readset(rs)
        eng.build()
        res, chosen = _run(eng, rs, tuning, auto)
        if tuning == SPILL:
            got = eng.tuning()
            assert got["query_lds_hits"] == 16 and got["correct_threads"] == 128 and got["msa_chunk"] == 16
            assert got["batch_anchors"] == 512 and eng.stats().launches == _launches(rs.n, 512)
            # page-locked result arrays: each batch's results are copied while
            # the next batches run; they must equal the pageable call's
            pitch = (rs.max_length + 15) // 16 * 16
            ps = care_amd.pinned_empty((rs.n,), np.uint8, lib=eng.lib)
            pq = care_amd.pinned_empty((rs.n, pitch), np.uint8, lib=eng.lib)
            status, seqs, _, _ = eng.correct(0, rs.n, out=(ps, pq))
            assert np.array_equal(status, res["status"])
            assert np.array_equal(seqs[(status & 1) != 0, :rs.max_length], res["rows"])
        if chosen is not None:
            assert chosen["correct_threads"] in (128, 192, 256)
            assert chosen["query_lds_hits"] % 16 == 0 and chosen["correct_blocks_per_cu"] >= 1
        eng.close()
        results.append(res)
    for r in results[1:]:
        _same(results[0], r)
    if not case["options"].get("exclude_ambiguous"):
        want = _oracle_lists(rs, case["options"])
        ids, offs = results[1]["ids"], results[1]["offs"]
        assert all(np.array_equal(ids[offs[i]:offs[i + 1]], want[i]) for i in range(rs.n))
    # the forced shapes take the other paths: on average an anchor has more
    # than 64 table hits (merged) and more than 16 unique candidates (scratch)
    assert results[0]["raw"][0] > 64 * rs.n
    assert results[1]["offs"][-1] > 16 * rs.n // 4
    rs.close()


@pytest.fixture(scope="module")
def emu():
    subprocess.run(["sh", os.path.join(EMU_DIR, "build_emu.sh")], check=True)
    return care_amd._bind(ctypes.CDLL(EMU_LIB), care_amd.ENGINE_SYMBOLS)


@pytest.mark.parametrize("name", sorted(SMALL))
def test_emulated_shapes_give_identical_results(emu, name, tmp_path):
    _check(emu, SMALL[name], str(tmp_path), device_autotune=(name == "pe_q2"))


def test_tuning_rejects_invalid_shapes(emu, tmp_path):
    rs = _readset(SMALL["pe_q2"], str(tmp_path))
    eng = care_amd.Engine(care_amd.default_options(**SMALL["pe_q2"]["options"]), device=0, lib=emu)
    eng.upload_readset(rs)
    for bad in (dict(correct_threads=64), dict(correct_threads=512), dict(msa_chunk=8), dict(query_lds_hits=24),
                dict(query_lds_hits=8192), dict(batch_anchors=3), dict(query_register_hits=100),
                dict(query_register_hits=4096)):
        with pytest.raises(care_amd.CareError):
            eng.set_tuning(**bad)
    with pytest.raises(care_amd.CareError):
        eng.set_launch_config(0, 512)
    eng.close()
    rs.close()


@pytest.mark.gpu
def test_gpu_shapes_give_identical_results(tmp_path):
    print("tuning shapes on the device", flush=True)
    _check(None, GPU, str(tmp_path), device_autotune=True)
