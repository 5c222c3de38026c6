"""The engine's HIP kernel sources run on the CPU through the HIP emulation
in tests/emu (fibers for threads, wave and block barriers), driven by the
unchanged care-gpu host, and must reproduce the oracle's corrected output
byte for byte.  This checks the kernels' semantics (alignment, filters,
MSA, refinement, classic and forest correction, candidate correction) on
small seeded inputs without a GPU; it cannot see GPU timing or hardware
behaviour.  EMU_POISON fills device memory and LDS with a non-zero byte, so
reads of memory nothing wrote change the output; EMU_SCHED varies the order
in which waves run between barriers.
"""
import os
import subprocess

import pytest

from tests import helpers
from tests.forest_files import ANCHOR_DESC, CANDS_DESC, write_forest

EMU_DIR = os.path.join(helpers.ROOT, "tests", "emu")
EMU_EXE = os.path.join(EMU_DIR, "_build", "care-gpu-emu")

SMALL = {
    "se_small": {"profile": "uniform-se", "seed": 201, "genome": 1500,
                 "args": ["-c", "30", "--pairmode", "SE", "-k", "20", "-h", "16"]},
    "pe_small_q2": {"profile": "repeat-pe", "seed": 202, "genome": 2000,
                    "args": ["-c", "30", "--pairmode", "PE", "-q", "--qualityScoreBits", "2", "-k", "20", "-h", "32"]},
    "se_mixed_small": {"profile": "mixed-se", "seed": 203, "genome": 2000,
                       "args": ["-c", "50", "--pairmode", "SE", "--excludeAmbiguous", "-q", "-k", "20", "-h", "48",
                                "--candidateCorrection"]},
    # paired-end with Ns (N rule, excluded ambiguous reads) and 1-bit qualities
    "pe_ambig_small": {"profile": "repeat-pe", "seed": 206, "genome": 2000, "sim_args": ["--nfrac", "0.05"],
                       "args": ["-c", "30", "--pairmode", "PE", "-q", "--qualityScoreBits", "1", "--excludeAmbiguous",
                                "-k", "20", "-h", "32", "--candidateCorrection"]},
    # reduced pipeline (top 20 candidates by map hits, no refinement) with
    # the classic decision
    "pe_reduced_small": {"profile": "repeat-pe", "seed": 205, "genome": 2000,
                         "args": ["-c", "30", "--pairmode", "PE", "-q", "--qualityScoreBits", "2", "-k", "20", "-h",
                                  "32", "--candidateLimit", "20", "--refinements", "0"]},
    # 400 bp reads: alignment windows longer than the unrolled word steps
    "se_long_small": {"profile": "uniform-se", "seed": 207, "genome": 3000, "sim_args": ["--length", "400"],
                      "args": ["-c", "30", "--pairmode", "SE", "-k", "20", "-h", "16", "--candidateCorrection"]},
    "pe_forest_small": {"profile": "repeat-pe", "seed": 204, "genome": 1500,
                        "args": ["-c", "30", "--pairmode", "PE", "-q", "-k", "20", "-h", "32",
                                 "--candidateCorrection", "--correctionType", "1", "--correctionTypeCands", "1"],
                        "forest": True},
}


@pytest.fixture(scope="module")
def emu_exe():
    subprocess.run(["sh", os.path.join(EMU_DIR, "build_emu.sh")], check=True)
    return EMU_EXE


def _args(case, d):
    args = list(case["args"])
    if case.get("forest"):
        write_forest(os.path.join(d, "a.rf"), ANCHOR_DESC, 11)
        write_forest(os.path.join(d, "c.rf"), CANDS_DESC, 12)
        args += ["--ml-forestfile", os.path.join(d, "a.rf"), "--ml-cands-forestfile", os.path.join(d, "c.rf")]
    return args


def _run(exe, case, d, env_extra=None):
    reads, _ = helpers.simulate(case, d)
    env = dict(os.environ)
    env.update(env_extra or {})
    out = os.path.join(d, "out")
    p = subprocess.run([exe, "-i", reads, "-o", "r.fastq", "-d", out, "-g", "0", "-t", "4"] + _args(case, d),
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env, timeout=600)
    assert p.returncode == 0, p.stdout[-3000:]
    return helpers.md5(os.path.join(out, "r.fastq"))


def _oracle(case, d):
    reads, _ = helpers.simulate(case, d)
    out = os.path.join(d, "oracle")
    p = subprocess.run([helpers.ORACLE_EXE, "-i", reads, "-o", "r.fastq", "-d", out, "-t", "4"] + _args(case, d),
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert p.returncode == 0, p.stdout[-3000:]
    return helpers.md5(os.path.join(out, "r.fastq"))


@pytest.mark.parametrize("name", sorted(SMALL))
def test_emulated_kernels_match_oracle(emu_exe, name, tmp_path):
    case = SMALL[name]
    want = _oracle(case, str(tmp_path))
    assert _run(emu_exe, case, str(tmp_path)) == want


@pytest.mark.parametrize("env", [{"EMU_POISON": "0xA5", "EMU_SCHED": "random"}])
def test_emulated_kernels_poisoned_memory_and_wave_order(emu_exe, env, tmp_path):
    case = SMALL["se_mixed_small"]
    want = _oracle(case, str(tmp_path))
    assert _run(emu_exe, case, str(tmp_path), env) == want


# Kernel index checks (CARE_CHECK, every LDS region and global row index the
# kernels write or gather) and the address / undefined-behaviour sanitizers,
# on the launch grids of an MI355X (256 compute units) with small device
# batches: the output must still be the oracle's.
CHECKED_EXE = os.path.join(EMU_DIR, "_build", "checked", "care-gpu-emu")


@pytest.fixture(scope="module")
def checked_exe():
    subprocess.run(["sh", os.path.join(EMU_DIR, "build_emu.sh"), "checked"], check=True)
    return CHECKED_EXE


@pytest.mark.slow
@pytest.mark.parametrize("name", ["se_small", "se_mixed_small", "pe_small_q2"])
def test_checked_kernels_match_oracle(checked_exe, name, tmp_path):
    case = dict(SMALL[name])
    case["args"] = case["args"] + ["--batchsize", "512"]
    want = _oracle(dict(SMALL[name]), str(tmp_path))
    env = {"EMU_CUS": "256", "ASAN_OPTIONS": "detect_leaks=0:detect_stack_use_after_return=0"}
    assert _run(checked_exe, case, str(tmp_path), env) == want


def test_padded_pitch_upload_matches_minimal_pitch(tmp_path):
    """Rows wider than the minimal pitch are repacked on upload: results are
    those of the minimal-pitch upload."""
    import ctypes

    import numpy as np

    import care_amd
    subprocess.run(["sh", os.path.join(EMU_DIR, "build_emu.sh")], check=True)
    lib = care_amd._bind(ctypes.CDLL(os.path.join(EMU_DIR, "_build", "libcare_amd_emu.so")), care_amd.ENGINE_SYMBOLS)
    case = SMALL["se_mixed_small"]
    reads, _ = helpers.simulate(case, str(tmp_path))
    rs = care_amd.ReadSet([reads], paired_end=False, qual_bits=8)
    opt = care_amd.default_options(kmer_length=20, num_hash_maps=48, coverage=50.0, use_quality=1,
                                   exclude_ambiguous=1, correct_candidates=1)
    seqs, quals = rs.seqs, rs.quals
    runs = []
    for pad in (0, 5):
        s = np.zeros((rs.n, rs.pitch + pad), dtype=np.uint32)
        s[:, :rs.pitch] = seqs
        s[:, rs.pitch:] = 0xDEADBEEF
        q = np.zeros((rs.n, rs.qpitch + pad), dtype=np.uint32)
        q[:, :rs.qpitch] = quals
        q[:, rs.qpitch:] = 0xDEADBEEF
        eng = care_amd.Engine(opt, device=0, lib=lib)
        eng.upload(s, rs.lengths, rs.ambiguous, q, 8, rs.max_length)
        eng.build()
        ids, offs = eng.query(0, rs.n)
        status, out, cids, cseqs = eng.correct(0, rs.n)
        order = np.lexsort((np.arange(len(cids)), cids))
        runs.append((ids.copy(), status.copy(), out.copy(), cids[order].copy(), cseqs[order].copy()))
        eng.close()
    for a, b in zip(runs[0], runs[1]):
        assert np.array_equal(a, b)
    rs.close()


def test_candidate_result_overflow_recorrects_in_halves(emu_exe, tmp_path):
    """With room for one candidate result per anchor, chunks whose candidate
    corrections do not fit are corrected again in halves: the output is the
    oracle's."""
    case = SMALL["se_mixed_small"]
    want = _oracle(case, str(tmp_path))
    reads, _ = helpers.simulate(case, str(tmp_path))
    out = os.path.join(str(tmp_path), "out")
    p = subprocess.run([emu_exe, "-i", reads, "-o", "r.fastq", "-d", out, "-g", "0", "-t", "4",
                        "--candidateResultsPerAnchor", "1"] + _args(case, str(tmp_path)),
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    assert p.returncode == 0, p.stdout[-3000:]
    assert "corrected in halves" in p.stdout
    assert helpers.md5(os.path.join(out, "r.fastq")) == want


def test_emulated_two_gzip_mate_files_and_gzip_output(emu_exe, tmp_path):
    """The mates of a paired-end case as two gzip'ed files, gzip output:
    the decompressed output equals the interleaved single-file run's."""
    import gzip
    case = SMALL["pe_small_q2"]
    d = str(tmp_path)
    want = _run(emu_exe, case, d)
    reads, _ = helpers.simulate(case, d)
    with open(reads) as f:
        lines = f.readlines()
    m1, m2 = os.path.join(d, "m1.fq.gz"), os.path.join(d, "m2.fq.gz")
    with gzip.open(m1, "wt") as a, gzip.open(m2, "wt") as b:
        for r in range(len(lines) // 4):
            (a if r % 2 == 0 else b).writelines(lines[4 * r:4 * r + 4])
    out = os.path.join(d, "out2")
    p = subprocess.run([emu_exe, "-i", m1, "-i", m2, "-o", "r.fq.gz", "-d", out, "-g", "0", "-t", "4", "--gzoutput"]
                       + _args(case, d), stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    assert p.returncode == 0, p.stdout[-3000:]
    plain = os.path.join(d, "r2.fastq")
    with gzip.open(os.path.join(out, "r.fq.gz"), "rb") as src, open(plain, "wb") as dst:
        dst.write(src.read())
    assert helpers.md5(plain) == want
