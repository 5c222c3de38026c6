"""The gated correction mode (correction_type 3): candidate limit by map
hits, no refinement, every non-HQ decision column left to the learned gate.

The oracle restates the mode with a constant gate (accept or reject every
column) and dumps the 32 features of every decision column; the engine must
reproduce the corrected reads byte for byte under the matching constant
gates (gate parameters all zero but the output bias) and record the same
decision columns with bit-identical features.  The candidate limit itself is
pinned against a brute-force count of map hits.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import care_amd
from tests import helpers

EMU_DIR = os.path.join(helpers.ROOT, "tests", "emu")
EMU_EXE = os.path.join(EMU_DIR, "_build", "care-gpu-emu")
EMU_LIB = os.path.join(EMU_DIR, "_build", "libcare_amd_emu.so")

SMALL = {
    "se": {"profile": "uniform-se", "seed": 301, "genome": 1500,
           "args": ["-c", "30", "--pairmode", "SE", "-k", "20", "-h", "16"]},
    "pe_q": {"profile": "repeat-pe", "seed": 302, "genome": 2000,
             "args": ["-c", "30", "--pairmode", "PE", "-q", "--qualityScoreBits", "2", "-k", "20", "-h", "32"]},
}
GPU_CASE = {"profile": "repeat-pe", "seed": 303, "genome": 40000,
            "args": ["-c", "30", "--pairmode", "PE", "-q", "--qualityScoreBits", "2", "-k", "20", "-h", "32"]}
# the gate on candidate corrections (correctionTypeCands 3)
CAND_SMALL = {"profile": "repeat-pe", "seed": 304, "genome": 2000,
              "args": ["-c", "30", "--pairmode", "PE", "-q", "--qualityScoreBits", "2", "-k", "20", "-h", "32",
                       "--candidateCorrection", "--correctionTypeCands", "3"]}
CAND_GPU = {"profile": "large-pe", "seed": 305, "genome": 40000,
            "args": ["-c", "40", "--pairmode", "PE", "-q", "-k", "20", "-h", "48",
                     "--candidateCorrection", "--correctionTypeCands", "3"]}
LIMIT = 12

REC = np.dtype([("id", "<u4"), ("pos", "<u2"), ("orig", "u1"), ("cons", "u1"), ("f", "<f4", (32,))])


def constant_gate(path, accept):
    p = np.zeros(care_amd.MLP_PARAMS, dtype=np.float32)
    p[-1] = 100.0 if accept else -100.0
    p.tofile(path)
    return p


def _oracle(case, d, gate, dump=None, gate_cands=1, dump_cands=None):
    reads, _ = helpers.simulate(case, d)
    out = os.path.join(d, "oracle%d_%d" % (gate, gate_cands))
    args = [helpers.ORACLE_EXE, "-i", reads, "-o", "r.fastq", "-d", out, "-t", "4", "--correctionType", "3",
            "--candidateLimit", str(LIMIT), "--gateConstant", str(gate), "--gateConstantCands", str(gate_cands)] + case["args"]
    if dump:
        args += ["--gateDump", dump]
    if dump_cands:
        args += ["--gateDumpCands", dump_cands]
    p = subprocess.run(args, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert p.returncode == 0, p.stdout[-2000:]
    return helpers.md5(os.path.join(out, "r.fastq"))


def _engine_run(exe, case, d, gate, gate_cands=None):
    reads, _ = helpers.simulate(case, d)
    gf = os.path.join(d, "gate%d.bin" % gate)
    constant_gate(gf, gate > 0)
    extra = []
    if gate_cands is not None:
        gfc = os.path.join(d, "gatec%d.bin" % gate_cands)
        constant_gate(gfc, gate_cands > 0)
        extra = ["--ml-cands-gatefile", gfc]
    out = os.path.join(d, "eng%d_%s" % (gate, gate_cands))
    p = subprocess.run([exe, "-i", reads, "-o", "r.fastq", "-d", out, "-g", "0", "-t", "4", "--correctionType", "3",
                        "--candidateLimit", str(LIMIT), "--ml-gatefile", gf] + extra + case["args"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=900)
    assert p.returncode == 0, p.stdout[-2000:]
    return helpers.md5(os.path.join(out, "r.fastq"))


def _records(eng, rs, candidates=False):
    f, ids, pos, orig, cons = eng.gate_columns(0, rs.n, candidates=candidates)
    rec = np.zeros(len(ids), dtype=REC)
    rec["id"], rec["pos"], rec["orig"], rec["cons"], rec["f"] = ids, pos, orig, cons, f
    return np.sort(rec, order=["id", "pos"])


def _options(case):
    a = case["args"]
    kw = dict(kmer_length=int(a[a.index("-k") + 1]), num_hash_maps=int(a[a.index("-h") + 1]),
              coverage=float(a[a.index("-c") + 1]), paired_end=int("PE" in a), use_quality=int("-q" in a),
              correction_type=3, candidate_limit=LIMIT)
    if "--candidateCorrection" in a:
        kw["correct_candidates"] = 1
    return kw


def _readset(case, d):
    reads, _ = helpers.simulate(case, d)
    a = case["args"]
    qb = int(a[a.index("--qualityScoreBits") + 1]) if "--qualityScoreBits" in a else (8 if "-q" in a else 0)
    return care_amd.ReadSet([reads], paired_end="PE" in a, qual_bits=qb)


def _cand_records(lib, case, d):
    """Candidate decision columns of the engine (with a constant accepting
    gate) and of the oracle, sorted by read, position and features."""
    dump = os.path.join(d, "gate_cands.bin")
    _oracle(case, d, 1, gate_cands=1, dump_cands=dump)
    want = np.fromfile(dump, dtype=REC)
    rs = _readset(case, d)
    eng = care_amd.Engine(care_amd.default_options(**_options(case)), device=0, lib=lib)
    eng.upload_readset(rs)
    eng.build()
    eng.mlp_load(constant_gate(os.path.join(d, "g1.bin"), True))
    f, ids, pos, orig, cons = eng.gate_columns(0, rs.n, candidates=True)
    eng.close()
    got = np.zeros(len(ids), dtype=REC)
    got["id"], got["pos"], got["orig"], got["cons"], got["f"] = ids, pos, orig, cons, f
    # a read can be corrected as the candidate of several anchors: compare
    # the multisets of records
    key = lambda r: np.lexsort([r["f"][:, k] for k in range(31, -1, -1)] + [r["cons"], r["orig"], r["pos"], r["id"]])
    return got[key(got)], want[key(want)]


def test_limited_query_oracle_bruteforce():
    orc = helpers.oracle()
    rng = np.random.default_rng(9)
    g = "".join(rng.choice(list("ACGT"), 600))
    reads = []
    for _ in range(150):
        p = int(rng.integers(0, 540))
        reads.append(g[p:p + 60])
    k, h, cov = 12, 8, 4.0
    seqs = np.zeros((len(reads), 4), dtype=np.uint32)
    for i, r in enumerate(reads):
        seqs[i, :] = helpers.encode(orc, r)
    lens = np.full(len(reads), 60, dtype=np.int32)
    sig = np.zeros((len(reads), h), dtype=np.uint64)
    for i in range(len(reads)):
        orc.orc_signature(seqs[i].ctypes.data, 60, k, h, sig[i].ctypes.data)
    cap = max(10, int(cov * 2.5))
    t = orc.orc_tables_build(seqs.ctypes.data, 4, lens.ctypes.data, len(reads), k, h, cov)
    buf = np.zeros(2000, dtype=np.uint32)
    for i in range(len(reads)):
        hits = {}
        for j in range(h):
            same = [r for r in range(len(reads)) if sig[r, j] == sig[i, j]]
            if 2 <= len(same) <= cap:
                for r in same:
                    if r != i:
                        hits[r] = hits.get(r, 0) + 1
        for limit in (3, 7, 1000):
            order = sorted(hits, key=lambda r: (-hits[r], r))[:limit]
            c = orc.orc_tables_query_limited(t, seqs[i].ctypes.data, 60, i, limit, buf.ctypes.data, len(buf))
            assert buf[:c].tolist() == sorted(order)
    orc.orc_tables_free(t)


# ------------------------------------------------------------ emulated kernels

@pytest.fixture(scope="module")
def emu():
    subprocess.run(["sh", os.path.join(EMU_DIR, "build_emu.sh")], check=True)
    # device memory and LDS filled with a non-zero byte: features read from
    # MSA columns nothing computed would differ from the oracle's
    os.environ.setdefault("EMU_POISON", "0xA5")
    return care_amd._bind(ctypes.CDLL(EMU_LIB), care_amd.ENGINE_SYMBOLS)


@pytest.mark.parametrize("name", sorted(SMALL))
@pytest.mark.parametrize("gate", [1, -1])
def test_emulated_gated_mode_matches_oracle(emu, name, gate, tmp_path):
    case = SMALL[name]
    assert _engine_run(EMU_EXE, case, str(tmp_path), gate) == _oracle(case, str(tmp_path), gate)


@pytest.mark.parametrize("name", sorted(SMALL))
def test_emulated_gate_columns_match_oracle(emu, name, tmp_path):
    case = SMALL[name]
    dump = str(tmp_path / "gate.bin")
    _oracle(case, str(tmp_path), 1, dump)
    want = np.sort(np.fromfile(dump, dtype=REC), order=["id", "pos"])
    rs = _readset(case, str(tmp_path))
    eng = care_amd.Engine(care_amd.default_options(**_options(case)), device=0, lib=emu)
    eng.upload_readset(rs)
    eng.build()
    got = _records(eng, rs)
    eng.close()
    if name == "pe_q":
        assert len(want) > 20
    assert np.array_equal(got, want)


def _plain_columns(lib, case, d):
    """Decision columns over the plain method's MSAs (every candidate,
    refinement first): the engine's care_gate_columns_plain and the oracle's
    dump of the same mode."""
    reads, _ = helpers.simulate(case, d)
    dump = os.path.join(d, "gate_plain.bin")
    p = subprocess.run([helpers.ORACLE_EXE, "-i", reads, "-o", "r.fastq", "-d", os.path.join(d, "oracle_plain"), "-t",
                        "4", "--correctionType", "3", "--candidateLimit", "0", "--gateRefine", "--gateConstant", "1",
                        "--gateDump", dump] + case["args"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert p.returncode == 0, p.stdout[-2000:]
    want = np.sort(np.fromfile(dump, dtype=REC), order=["id", "pos"])
    rs = _readset(case, d)
    eng = care_amd.Engine(care_amd.default_options(**_options(case)), device=0, lib=lib)
    eng.upload_readset(rs)
    eng.build()
    f, ids, pos, orig, cons = eng.gate_columns(0, rs.n, plain=True)
    reduced = _records(eng, rs)
    eng.close()
    got = np.zeros(len(ids), dtype=REC)
    got["id"], got["pos"], got["orig"], got["cons"], got["f"] = ids, pos, orig, cons, f
    return np.sort(got, order=["id", "pos"]), want, reduced


def test_emulated_plain_method_columns_match_oracle(emu, tmp_path):
    got, want, reduced = _plain_columns(emu, SMALL["pe_q"], str(tmp_path))
    assert len(want) > 20
    _assert_same_records(got, want)
    # every candidate and refinement: other MSAs than the reduced pipeline's
    assert len(got) != len(reduced) or not np.array_equal(got, reduced)


def _assert_same_records(got, want):
    assert len(got) == len(want), (len(got), len(want))
    bad = np.nonzero(got != want)[0]
    if len(bad):
        i = bad[0]
        fd = np.nonzero(got[i]["f"] != want[i]["f"])[0]
        raise AssertionError("%d differing records; first %d: engine %s / oracle %s, features %s: %s vs %s" % (
            len(bad), i, got[i][["id", "pos", "orig", "cons"]], want[i][["id", "pos", "orig", "cons"]], fd.tolist(),
            got[i]["f"][fd].tolist(), want[i]["f"][fd].tolist()))


@pytest.mark.parametrize("gates", [(1, 1), (1, -1), (-1, -1)])
def test_emulated_candidate_gate_matches_oracle(emu, gates, tmp_path):
    """Candidate corrections through the candidate gate (accept / reject every
    candidate decision column) equal the oracle's restatement."""
    ga, gc = gates
    assert _engine_run(EMU_EXE, CAND_SMALL, str(tmp_path), ga, gc) == _oracle(CAND_SMALL, str(tmp_path), ga, gate_cands=gc)


def test_emulated_candidate_gate_columns_match_oracle(emu, tmp_path):
    got, want = _cand_records(emu, CAND_SMALL, str(tmp_path))
    assert len(want) > 20
    _assert_same_records(got, want)


# ------------------------------------------------------------------- GPU

@pytest.mark.gpu
@pytest.mark.parametrize("gate", [1, -1])
def test_gpu_gated_mode_matches_oracle(gate, tmp_path):
    print("gated mode, constant gate %d" % gate, flush=True)
    assert _engine_run(care_amd.CARE_GPU, GPU_CASE, str(tmp_path), gate) == _oracle(GPU_CASE, str(tmp_path), gate)


@pytest.mark.gpu
def test_gpu_plain_method_columns_match_oracle(tmp_path):
    print("plain-method gate columns", flush=True)
    got, want, _ = _plain_columns(None, GPU_CASE, str(tmp_path))
    assert len(want) > 100
    _assert_same_records(got, want)


@pytest.mark.gpu
def test_gpu_gate_columns_match_oracle(tmp_path):
    print("gate columns", flush=True)
    dump = str(tmp_path / "gate.bin")
    _oracle(GPU_CASE, str(tmp_path), 1, dump)
    want = np.sort(np.fromfile(dump, dtype=REC), order=["id", "pos"])
    rs = _readset(GPU_CASE, str(tmp_path))
    with care_amd.Engine(care_amd.default_options(**_options(GPU_CASE)), device=0) as eng:
        eng.upload_readset(rs)
        eng.build()
        got = _records(eng, rs)
    assert len(want) > 100
    assert np.array_equal(got, want)


@pytest.mark.gpu
def test_gpu_limited_candidate_lists_match_oracle(tmp_path):
    print("limited candidate lists", flush=True)
    rs = _readset(GPU_CASE, str(tmp_path))
    kw = _options(GPU_CASE)
    with care_amd.Engine(care_amd.default_options(**kw), device=0) as eng:
        eng.upload_readset(rs)
        eng.build()
        ids, offs = eng.query(0, rs.n)
    orc = helpers.oracle()
    seqs = np.ascontiguousarray(rs.seqs)
    lens = np.ascontiguousarray(rs.lengths)
    t = orc.orc_tables_build(seqs.ctypes.data, rs.pitch, lens.ctypes.data, rs.n, kw["kmer_length"], kw["num_hash_maps"],
                             kw["coverage"])
    buf = np.zeros(4096, dtype=np.uint32)
    bad = 0
    for i in range(rs.n):
        c = orc.orc_tables_query_limited(t, seqs[i].ctypes.data, int(lens[i]), i, LIMIT, buf.ctypes.data, len(buf))
        if not np.array_equal(ids[offs[i]:offs[i + 1]], buf[:c]):
            bad += 1
    orc.orc_tables_free(t)
    assert bad == 0


@pytest.mark.gpu
@pytest.mark.parametrize("gates", [(1, 1), (1, -1)])
def test_gpu_candidate_gate_matches_oracle(gates, tmp_path):
    print("candidate gate, constant gates %s" % (gates,), flush=True)
    ga, gc = gates
    assert _engine_run(care_amd.CARE_GPU, CAND_GPU, str(tmp_path), ga, gc) == _oracle(CAND_GPU, str(tmp_path), ga,
                                                                                       gate_cands=gc)


@pytest.mark.gpu
def test_gpu_candidate_gate_columns_match_oracle(tmp_path):
    print("candidate gate columns", flush=True)
    got, want = _cand_records(None, CAND_GPU, str(tmp_path))
    assert len(want) > 100
    _assert_same_records(got, want)


# What the trained network gains over the method without it, on a held-out
# read set of the generator it was trained on (a seed used neither for
# training nor for the threshold): the same reduced pipeline (top 20
# candidates, no refinement) deciding by the classic thresholds loses true
# positives, accepting every consensus multiplies false positives; the gate
# keeps the plain path's true positives within 2 % at far fewer false
# positives than accepting everything.
GAIN_CASE = {"profile": "repeat-pe", "seed": 20240777, "genome": 2000000}


@pytest.mark.gpu
def test_gpu_gate_gain_over_reduced_pipeline_without_it(tmp_path):
    from care_amd import gate as gate_mod
    print("gate ablation on a held-out repeat-pe set (seed %d)" % GAIN_CASE["seed"], flush=True)
    reads, truth = helpers.simulate(GAIN_CASE, str(tmp_path))
    res = gate_mod.ablation(reads, truth, str(tmp_path))
    print(res, flush=True)
    plain, classic, gated, accept = (res[k] for k in ("plain", "reduced_classic", "reduced_mlp_gate",
                                                      "reduced_accept_all"))
    assert gated["TP"] >= 1.1 * classic["TP"]
    assert gated["FP"] * 4 <= accept["FP"]
    assert gated["TP"] >= 0.98 * plain["TP"]
