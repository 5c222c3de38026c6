"""The stream-ordered interface on device buffers (include/care_amd.h):
read-store gathers by id and candidate queries of sequences given on the
device, enqueued on a caller stream.

Gathered rows, lengths, qualities and ambiguity flags must equal the host
read set's; the candidate lists of the store's own reads, queried from
device copies of their rows, must equal the synchronous query's (and the
oracle's); without read ids a read's list keeps the read itself.  The CPU
tests run the kernels through the HIP emulation (device buffers are host
memory there); the GPU tests run the same checks on the device.
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

SMALL = {"profile": "mixed-se", "seed": 221, "genome": 1500}
GPU_CASE = {"profile": "mixed-se", "seed": 222, "genome": 30000}


def _check(lib, case, d):
    reads, _ = helpers.simulate(case, d)
    rs = care_amd.ReadSet([reads], paired_end=False, qual_bits=8)
    opt = care_amd.default_options(kmer_length=20, num_hash_maps=24, coverage=50.0, use_quality=1, exclude_ambiguous=0)
    eng = care_amd.Engine(opt, device=0, lib=lib)
    eng.upload_readset(rs)
    eng.build()
    seqs, lens, amb, quals = rs.seqs, rs.lengths, rs.ambiguous, rs.quals
    n = rs.n
    stream = eng.stream_create()
    rng = np.random.default_rng(3)
    ids = rng.integers(0, n, size=min(n, 700)).astype(np.uint32)
    ids[5] = n + 3  # outside the store: zero row
    m = len(ids)
    pitch = rs.pitch + 2
    d_ids = eng.to_device(ids)
    d_rows = eng.device_array(m * pitch * 4)
    d_lens = eng.device_array(m * 4)
    d_amb = eng.device_array(m)
    qp = rs.max_length + 3
    d_q = eng.device_array(m * qp)
    L = eng.lib
    eng._ck(L.care_readstore_gather_sequences_async(eng.ctx, d_ids, m, d_rows, pitch, stream))
    eng._ck(L.care_readstore_gather_lengths_async(eng.ctx, d_ids, m, d_lens, stream))
    eng._ck(L.care_readstore_ambiguous_async(eng.ctx, d_ids, m, d_amb, stream))
    eng._ck(L.care_readstore_gather_qualities_async(eng.ctx, d_ids, m, d_q, qp, stream))
    eng.stream_sync(stream)
    rows = eng.from_device(d_rows, (m, pitch), np.uint32)
    glens = eng.from_device(d_lens, (m,), np.int32)
    gamb = eng.from_device(d_amb, (m,), np.uint8)
    gq = eng.from_device(d_q, (m, qp), np.uint8)
    valid = ids < n
    assert np.array_equal(rows[valid, :rs.pitch], seqs[ids[valid]])
    assert not rows[valid, rs.pitch:].any() and not rows[~valid].any()
    assert np.array_equal(glens[valid], lens[ids[valid]]) and not glens[~valid].any()
    assert np.array_equal(gamb[valid], amb[ids[valid]])
    qbytes = quals.view(np.uint8).reshape(n, -1)[:, :rs.max_length]  # 8-bit qualities, bytes in read order
    for k in np.nonzero(valid)[0][:50]:
        assert np.array_equal(gq[k, :lens[ids[k]]], qbytes[ids[k], :lens[ids[k]]])

    # candidate queries of device copies of the store's rows
    want_ids, want_offs = eng.query(0, n)
    maxc = L.care_minhash_max_candidates(eng.ctx)
    ws = L.care_minhash_query_workspace_bytes(eng.ctx, n)
    d_seqs = eng.to_device(seqs)
    d_all_lens = eng.to_device(lens)
    d_self = eng.to_device(np.arange(n, dtype=np.uint32))
    d_cand = eng.device_array(n * maxc * 4)
    d_cnt = eng.device_array(n * 4)
    d_ws = eng.device_array(ws)
    eng._ck(L.care_minhash_query_async(eng.ctx, d_seqs, rs.pitch, d_all_lens, d_self, n, d_cand, d_cnt, d_ws, ws,
                                       stream))
    eng.stream_sync(stream)
    cand = eng.from_device(d_cand, (n, maxc), np.uint32)
    cnt = eng.from_device(d_cnt, (n,), np.int32)
    for i in range(n):
        assert np.array_equal(cand[i, :cnt[i]], want_ids[want_offs[i]:want_offs[i + 1]])
    # without read ids nothing is dropped as the anchor itself
    eng._ck(L.care_minhash_query_async(eng.ctx, d_seqs, rs.pitch, d_all_lens, None, n, d_cand, d_cnt, d_ws, ws,
                                       stream))
    eng.stream_sync(stream)
    cand2 = eng.from_device(d_cand, (n, maxc), np.uint32)
    cnt2 = eng.from_device(d_cnt, (n,), np.int32)
    for i in range(n):
        got = cand2[i, :cnt2[i]]
        if cnt[i] or cnt2[i]:
            assert np.array_equal(np.setdiff1d(got, [i]), cand[i, :cnt[i]])
    # too small a workspace is refused
    assert L.care_minhash_query_async(eng.ctx, d_seqs, rs.pitch, d_all_lens, d_self, n, d_cand, d_cnt, d_ws, ws - 1,
                                      stream) == care_amd.CARE_OK + 4
    for p in (d_ids, d_rows, d_lens, d_amb, d_q, d_seqs, d_all_lens, d_self, d_cand, d_cnt, d_ws):
        eng.device_free(p)
    eng.stream_destroy(stream)
    eng.close()
    rs.close()
    return want_ids, want_offs


@pytest.fixture(scope="module")
def emu():
    subprocess.run(["sh", os.path.join(EMU_DIR, "build_emu.sh")], check=True)
    return care_amd._bind(ctypes.CDLL(EMU_LIB), care_amd.ENGINE_SYMBOLS)


def test_emulated_device_interface(emu, tmp_path):
    _check(emu, SMALL, str(tmp_path))


@pytest.mark.gpu
def test_gpu_device_interface(tmp_path):
    print("stream-ordered gathers and queries on the device", flush=True)
    _check(None, GPU_CASE, str(tmp_path))
