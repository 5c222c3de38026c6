"""Known-answer pins of the oracle's single rules (oracle/care_oracle.cpp).

Every expected value here is fixed by the rule as the reference states it,
worked out by hand on a small constructed input (a pileup, a support
profile, a list of mismatch ratios, a forest with two trees), not by
re-running the oracle's formula: a wrong sign, index, tie rule or threshold
in the oracle changes at least one of them (tests/oracle_mutations.py
applies such one-token mutations and checks that this file then fails).

  murmur64           MurmurHash3 x64_128 built on the oracle's finalizer
                     reproduces the published SMHasher verification value
                     0x6384BA69 (hashers.cuh:129-137 is MurmurHash3's fmix64)
  overlap weight     include/hostdevicefunctions.cuh:93-97 at exact points
  MSA build          src/msa.cpp:9-173: column frame, counts, weights,
                     consensus ties (A < C < G < T), support, orig coverage
  properties / HQ    src/msa.cpp:304-376 at the fgeq epsilon and the HQ
                     percent loop boundaries
  anchor correction  src/msa.cpp:379-453 (consensus vs per-column rule)
  different region   src/msa.cpp:532-734 (column, found base, keep/discard,
                     last significant base, 0.9-weight veto)
  SE filter          include/corrector.hpp:1053-1149 (2/3/4 x error rate)
  candidate window   src/msa.cpp:457-522, include/corrector.hpp:1539-1623
  edits              include/corrector.hpp:1639-1761 (<= len/7, ambiguous)
  forest vote        include/forest.hpp:217-226 on hand-built trees
  combine            src/correctionresultoutput.cpp:730-1023
"""
import ctypes
import os
import struct

import numpy as np
import pytest

from tests import helpers

F32 = np.float32
M64 = (1 << 64) - 1


def _lib():
    lib = ctypes.CDLL(os.environ.get("CARE_ORACLE_LIB", helpers.ORACLE_LIB))
    c_p, c_i, c_f = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
    lib.orc_murmur64.argtypes = [ctypes.c_uint64]
    lib.orc_murmur64.restype = ctypes.c_uint64
    lib.orc_msa_build.argtypes = [c_i, c_i, c_p, c_p, c_i, c_p, c_i, c_p, c_i, c_p, c_p, c_p, c_i, c_p, c_p, c_p, c_p,
                                  c_p, c_p, c_p]
    lib.orc_msa_properties.argtypes = [c_p, c_p, c_i, c_i, c_i, c_f, c_f, c_f, c_p]
    lib.orc_msa_corrected_anchor.argtypes = [c_i, c_i, c_p, c_p, c_i, c_p, c_i, c_p, c_i, c_p, c_p, c_p, c_f, c_f,
                                             c_f, c_p]
    lib.orc_msa_different_region.argtypes = [c_i, c_p, c_i, c_p, c_i, c_p, c_p, c_p, c_i, c_p]
    lib.orc_se_filter_threshold.argtypes = [c_p, c_i, c_f, c_f, c_f]
    lib.orc_se_filter_threshold.restype = c_f
    lib.orc_candidate_in_range.argtypes = [c_i] * 6
    lib.orc_overlap_weight.argtypes = [c_i, c_i, c_f]
    lib.orc_overlap_weight.restype = c_f
    lib.orc_make_edits.argtypes = [c_p, c_p, c_i, c_i, c_p, c_p]
    lib.orc_forest_decide.argtypes = [ctypes.c_char_p, c_i, ctypes.c_uint32, c_f, c_p]
    lib.orc_combine.argtypes = [c_p, c_i, c_i, c_i, c_p, c_p, c_p]
    return lib


@pytest.fixture(scope="module")
def orc():
    return _lib()


# ---------------------------------------------------------------------------
# murmur finalizer
# ---------------------------------------------------------------------------
def _rotl(x, r):
    return ((x << r) | (x >> (64 - r))) & M64


def _murmur3_x64_128(key, seed, fmix64):
    """MurmurHash3_x64_128 (Appleby, public domain), with the 64-bit
    finalizer supplied by the caller."""
    c1, c2 = 0x87C37B91114253D5, 0x4CF5AD432745937F
    h1 = h2 = seed
    nb = len(key) // 16
    for i in range(nb):
        k1 = int.from_bytes(key[16 * i:16 * i + 8], "little")
        k2 = int.from_bytes(key[16 * i + 8:16 * i + 16], "little")
        k1 = (_rotl((k1 * c1) & M64, 31) * c2) & M64
        h1 ^= k1
        h1 = (((_rotl(h1, 27) + h2) & M64) * 5 + 0x52DCE729) & M64
        k2 = (_rotl((k2 * c2) & M64, 33) * c1) & M64
        h2 ^= k2
        h2 = (((_rotl(h2, 31) + h1) & M64) * 5 + 0x38495AB5) & M64
    tail = key[16 * nb:]
    k1 = k2 = 0
    for i in range(len(tail) - 1, 7, -1):
        k2 ^= tail[i] << (8 * (i - 8))
    if len(tail) > 8:
        h2 ^= (_rotl((k2 * c2) & M64, 33) * c1) & M64
    for i in range(min(len(tail), 8) - 1, -1, -1):
        k1 ^= tail[i] << (8 * i)
    if tail:
        h1 ^= (_rotl((k1 * c1) & M64, 31) * c2) & M64
    h1 ^= len(key)
    h2 ^= len(key)
    h1 = (h1 + h2) & M64
    h2 = (h2 + h1) & M64
    h1, h2 = fmix64(h1), fmix64(h2)
    h1 = (h1 + h2) & M64
    h2 = (h2 + h1) & M64
    return h1.to_bytes(8, "little") + h2.to_bytes(8, "little")


def test_murmur_finalizer_gives_published_verification_value(orc):
    # SMHasher's verification: hash keys {}, {0}, {0,1}, ... {0..254} with
    # seed 256 - i, hash the concatenated digests with seed 0, first four
    # bytes little-endian.  MurmurHash3_x64_128 is published as 0x6384BA69.
    fmix = lambda x: int(orc.orc_murmur64(x))  # noqa: E731
    key = bytes(range(256))
    digests = b"".join(_murmur3_x64_128(key[:i], 256 - i, fmix) for i in range(256))
    final = _murmur3_x64_128(digests, 0, fmix)
    assert int.from_bytes(final[:4], "little") == 0x6384BA69
    assert orc.orc_murmur64(0) == 0


# ---------------------------------------------------------------------------
# overlap weight
# ---------------------------------------------------------------------------
def test_overlap_weight_points(orc):
    # w = 1 - sqrt(nOps / (overlap * ratio)) with ratio 0.2
    assert orc.orc_overlap_weight(0, 100, 0.2) == 1.0
    assert orc.orc_overlap_weight(5, 100, 0.2) == 0.5  # sqrt(1/4)
    assert orc.orc_overlap_weight(1, 20, 0.2) == 0.5
    assert orc.orc_overlap_weight(20, 100, 0.2) == 0.0
    assert abs(orc.orc_overlap_weight(9, 100, 0.25) - 0.4) < 1e-6  # 1 - sqrt(9 / 25)


# ---------------------------------------------------------------------------
# MSA
# ---------------------------------------------------------------------------
def _cands(rows):
    pitch = max([len(r) for r in rows] + [1])
    buf = np.zeros((max(len(rows), 1), pitch), dtype=np.uint8)
    for i, r in enumerate(rows):
        buf[i, :len(r)] = np.frombuffer(r.encode(), dtype=np.uint8)
    lens = np.array([len(r) for r in rows] + [0], dtype=np.int32)
    return buf, pitch, lens


def _msa(orc, anchor, rows, shifts, weights, anchorQ=None, candQ=None):
    buf, pitch, lens = _cands(rows)
    sh = np.array(list(shifts) + [0], dtype=np.int32)
    w = np.array(list(weights) + [0], dtype=np.float32)
    cap = 64
    dims = np.zeros(2, dtype=np.int32)
    cons = np.zeros(cap, dtype=np.uint8)
    sup = np.zeros(cap, dtype=np.float32)
    cov = np.zeros(cap, dtype=np.int32)
    cnt = np.zeros((4, cap), dtype=np.int32)
    wt = np.zeros((4, cap), dtype=np.float32)
    oc = np.zeros(cap, dtype=np.int32)
    a = np.frombuffer(anchor.encode(), dtype=np.uint8).copy()
    useQ = anchorQ is not None
    aq = np.frombuffer(anchorQ.encode(), dtype=np.uint8).copy() if useQ else None
    if useQ:
        qbuf, qpitch, _ = _cands(candQ)
    rc = orc.orc_msa_build(int(useQ), len(anchor), a.ctypes.data, aq.ctypes.data if useQ else None, len(rows),
                           buf.ctypes.data, pitch, qbuf.ctypes.data if useQ else None, qpitch if useQ else 0,
                           lens.ctypes.data, sh.ctypes.data, w.ctypes.data, cap, dims.ctypes.data, cons.ctypes.data,
                           sup.ctypes.data, cov.ctypes.data, cnt.ctypes.data, wt.ctypes.data, oc.ctypes.data)
    assert rc == 0
    n = int(dims[0])
    return {"cols": n, "ab": int(dims[1]), "cons": cons[:n].tobytes().decode(), "support": sup[:n], "coverage": cov[:n],
            "counts": cnt[:, :n], "weights": wt[:, :n], "origCov": oc[:n]}


def test_msa_column_frame_and_coverage(orc):
    # anchor at columns 3..6; one candidate starting 2 after the anchor,
    # one 3 before it: 9 columns
    m = _msa(orc, "ACGT", ["GTAC", "CCAC"], [2, -3], [0.5, 0.5])
    assert (m["cols"], m["ab"]) == (9, 3)
    assert list(m["coverage"]) == [1, 1, 1, 2, 1, 2, 2, 1, 1]
    # column 3: anchor A + candidate 1's last base C
    assert list(m["counts"][:, 3]) == [1, 1, 0, 0]
    assert m["weights"][0, 3] == 1.0 and m["weights"][1, 3] == 0.5
    assert m["cons"] == "CCAACGTAC"
    assert m["support"][3] == F32(1.0) / F32(1.5)
    # a candidate entirely before the anchor frame start is impossible; the
    # candidate wholly inside adds coverage only where it lies
    m2 = _msa(orc, "ACGTAC", ["GT"], [2], [1.0])
    assert (m2["cols"], m2["ab"]) == (6, 0)
    assert list(m2["coverage"]) == [1, 1, 2, 2, 1, 1]


@pytest.mark.parametrize("anchor,cand,want", [("A", "C", "A"), ("C", "G", "C"), ("G", "T", "G"), ("T", "G", "G"),
                                              ("C", "A", "A"), ("T", "A", "A")])
def test_msa_consensus_ties_go_to_the_earlier_base(orc, anchor, cand, want):
    # equal weights of two bases: the consensus is the earlier of A, C, G, T
    m = _msa(orc, anchor, [cand], [0], [1.0])
    assert m["cons"] == want
    assert m["support"][0] == 0.5
    assert m["coverage"][0] == 2
    assert m["origCov"][0] == 1


def test_msa_weighted_consensus_support_and_orig_coverage(orc):
    # column 0: anchor A (1.0) vs C 0.75 + C 0.5 = 1.25 -> C, support 1.25 / 2.25
    # column 1: anchor C (1.0) + C 0.75 vs G 0.5 -> C, support 1.75 / 2.25
    m = _msa(orc, "AC", ["CC", "CG"], [0, 0], [0.75, 0.5])
    assert m["cons"] == "CC"
    assert m["support"][0] == F32(1.25) / F32(2.25)
    assert m["support"][1] == F32(1.75) / F32(2.25)
    assert list(m["counts"][:, 0]) == [1, 2, 0, 0]
    assert list(m["counts"][:, 1]) == [0, 2, 1, 0]
    assert list(m["origCov"]) == [1, 2]  # count of the anchor's own base


def test_msa_quality_weights(orc):
    # '+' is phred 10: weight 1 - 10^-1 = 0.9; '5' is phred 20: 0.99; a
    # candidate's base weight is its overlap weight times its quality weight
    m = _msa(orc, "A", ["C"], [0], [0.5], anchorQ="+", candQ=["5"])
    wq10 = F32(1.0) - F32(0.1)
    wq20 = F32(1.0) - F32(0.01)
    assert abs(float(m["weights"][0, 0]) - float(wq10)) < 1e-7
    assert abs(float(m["weights"][1, 0]) - float(F32(0.5) * wq20)) < 1e-7
    assert m["cons"] == "A"  # 0.9 > 0.495


# ---------------------------------------------------------------------------
# properties and HQ
# ---------------------------------------------------------------------------
def _props(orc, support, coverage, err=0.06, cov=30.0, mcov=0.6):
    s = np.array(support, dtype=np.float32)
    c = np.array(coverage, dtype=np.int32)
    out = np.zeros(4, dtype=np.float32)
    hq = orc.orc_msa_properties(s.ctypes.data, c.ctypes.data, len(s), 0, len(s), err, cov, mcov, out.ctypes.data)
    return hq, out


def test_properties_values(orc):
    hq, out = _props(orc, [1.0, 0.5, 0.75], [31, 7, 12])
    assert out[0] == F32(2.25) / F32(3) and out[1] == 0.5 and out[2] == 7 and out[3] == 31
    assert hq == 0


def test_hq_rule_boundaries(orc):
    # err 0.06, coverage 30, m_coverage 0.6: avg >= 0.94, min >= 0.82,
    # min coverage >= 3; HQ iff the smallest percent p <= 6 with
    # avg >= 1 - p/100 and min >= 1 - 3p/100 is at most 3
    assert _props(orc, [1.0] * 4, [30] * 4)[0] == 1
    assert _props(orc, [0.97], [30])[0] == 1          # p = 3
    assert _props(orc, [0.9699], [30])[0] == 0        # p = 4
    assert _props(orc, [1.0] * 9 + [0.93], [30] * 10)[0] == 1   # min needs p >= 2.33 -> 3
    assert _props(orc, [1.0] * 9 + [0.90], [30] * 10)[0] == 0   # min needs p >= 3.33 -> 4
    assert _props(orc, [1.0] * 4, [30, 30, 3, 30])[0] == 1      # coverage 3 = 0.6 / 6 * 30
    assert _props(orc, [1.0] * 4, [30, 30, 2, 30])[0] == 0


def test_fgeq_epsilon(orc):
    # thresholds compare with a 1e-5 tolerance: 0.97 - 9e-6 still makes p = 3,
    # 0.97 - 2e-5 does not
    assert _props(orc, [0.97 - 9e-6], [30])[0] == 1
    assert _props(orc, [0.97 - 2e-5], [30])[0] == 0


# ---------------------------------------------------------------------------
# classic anchor correction
# ---------------------------------------------------------------------------
def _corrected(orc, anchor, rows, err=0.06, cov=30.0, mcov=0.6):
    buf, pitch, lens = _cands(rows)
    sh = np.zeros(len(rows) + 1, dtype=np.int32)
    w = np.ones(len(rows) + 1, dtype=np.float32)
    a = np.frombuffer(anchor.encode(), dtype=np.uint8).copy()
    out = np.zeros(len(anchor), dtype=np.uint8)
    hq = orc.orc_msa_corrected_anchor(0, len(anchor), a.ctypes.data, None, len(rows), buf.ctypes.data, pitch, None, 0,
                                      lens.ctypes.data, sh.ctypes.data, w.ctypes.data, err, cov, mcov, out.ctypes.data)
    return out.tobytes().decode(), hq


def _columns(spec, n=30):
    """Candidate rows from per-column base lists (spec[col] = list of n bases)."""
    return ["".join(spec[c][i] for c in range(len(spec))) for i in range(n)]


def test_anchor_correction_per_column_rule(orc):
    # 30 candidates of weight 1 against anchor AAAAA:
    #   col 0: 30 C        support 30/31 > 0.9, orig count 1  -> C
    #   col 1: 2 A, 28 C   support 28/31 > 0.9, orig count 3  -> A
    #   col 2: 1 A, 29 C   support 29/31 > 0.9, orig count 2  -> C
    #   col 3: 30 A        -> A
    #   col 4: 15 A, 15 C  support 16/31 (min support fails: per-column rule)
    spec = [["C"] * 30, ["A"] * 2 + ["C"] * 28, ["A"] + ["C"] * 29, ["A"] * 30, ["A"] * 15 + ["C"] * 15]
    got, hq = _corrected(orc, "AAAAA", _columns(spec))
    assert got == "CACAA" and hq == 0
    # without column 4 every threshold holds: the consensus is taken
    got, hq = _corrected(orc, "AAAA", _columns(spec[:4]))
    assert got == "CCCA" and hq == 0


# ---------------------------------------------------------------------------
# different region
# ---------------------------------------------------------------------------
def _region(orc, anchor, rows, weights, coverage=30):
    buf, pitch, lens = _cands(rows)
    sh = np.zeros(len(rows) + 1, dtype=np.int32)
    w = np.array(list(weights) + [0], dtype=np.float32)
    a = np.frombuffer(anchor.encode(), dtype=np.uint8).copy()
    rem = np.zeros(len(rows) + 1, dtype=np.uint8)
    found = orc.orc_msa_different_region(len(anchor), a.ctypes.data, len(rows), buf.ctypes.data, pitch,
                                         lens.ctypes.data, sh.ctypes.data, w.ctypes.data, coverage, rem.ctypes.data)
    return found, list(rem[:len(rows)])


def test_different_region_discards_rows_matching_a_foreign_base(orc):
    # coverage 30: significant count int(30 * 0.3f) = 9.  Column 1: anchor A,
    # 12 A and 10 C candidates of weight 0.5 -> consensus A, C significant;
    # the anchor does not carry C, so the C rows go.  Column 3 makes the
    # anchor differ from the consensus (all candidates G there).
    rows = ["AAAG"] * 12 + ["ACAG"] * 10
    found, rem = _region(orc, "AAAA", rows, [0.5] * 22)
    assert found == 1 and rem == [0] * 12 + [1] * 10


def test_different_region_keeps_rows_matching_the_anchors_base(orc):
    # column 1: anchor C, 12 A and 10 C rows: A 6.0 vs C 1 + 5.0 tie -> A
    # consensus; C (11 counts) is significant and is the anchor's base, so
    # rows that do not carry C go
    rows = ["AAAG"] * 12 + ["ACAG"] * 10
    found, rem = _region(orc, "ACAA", rows, [0.5] * 22)
    assert found == 1 and rem == [1] * 12 + [0] * 10


def test_different_region_takes_the_last_significant_base(orc):
    # column 1: anchor A; C, G and T rows 10 each (weight 0.5): consensus C
    # (G and T only tie it); G and T are both significant, the last (T) is
    # taken, so T rows go
    rows = ["ACAG"] * 10 + ["AGAG"] * 10 + ["ATAG"] * 10
    found, rem = _region(orc, "AAAA", rows, [0.5] * 30)
    assert found == 1 and rem == [0] * 20 + [1] * 10


def test_different_region_veto_and_no_mismatch(orc):
    rows = ["AAAG"] * 12 + ["ACAG"] * 10
    # one discarded row aligned with weight >= 0.9: nothing is removed
    found, rem = _region(orc, "AAAA", rows, [0.5] * 21 + [0.9])
    assert found == 1 and rem == [0] * 22
    # anchor equal to the consensus everywhere: no search at all
    rows2 = ["AAAA"] * 12 + ["ACAA"] * 10
    found, rem = _region(orc, "AAAA", rows2, [0.5] * 22)
    assert found == 0 and rem == [0] * 22
    # a foreign base below the significant count (8 < 9) is not a column,
    # one at it (9) is
    rows3 = ["AAAG"] * 12 + ["ACAG"] * 8
    found, _ = _region(orc, "AAAA", rows3, [0.5] * 20)
    assert found == 0
    rows4 = ["AAAG"] * 12 + ["ACAG"] * 9
    found, rem = _region(orc, "AAAA", rows4, [0.5] * 21)
    assert found == 1 and rem == [0] * 12 + [1] * 9


# ---------------------------------------------------------------------------
# shifted Hamming: ties and the last shift
# ---------------------------------------------------------------------------
def _align(seq_a, seq_c, min_overlap=30, max_err=0.2, min_ratio=0.3):
    o = helpers.oracle()
    a = helpers.encode(o, seq_a)
    c = helpers.encode(o, seq_c)
    out = np.zeros(4, dtype=np.int32)
    valid = o.orc_align(a.ctypes.data, len(seq_a), c.ctypes.data, len(seq_c), min_overlap, max_err, min_ratio,
                        out.ctypes.data)
    return valid, list(out)  # score, overlap, shift, nOps


def test_alignment_tie_goes_to_the_first_shift():
    # ACAC.. against CACA..: shift 0 mismatches everywhere; shifts +1 and -1
    # both match 99 bases exactly (score 0 + 200 - 2 * 99 = 2).  Positive
    # shifts are tried first, so +1 is kept.
    valid, (score, ov, shift, nops) = _align("AC" * 50, "CA" * 50)
    assert valid == 1 and (score, ov, shift, nops) == (2, 99, 1, 0)


def test_alignment_mismatch_limit():
    # overlap 100 at shift 0: mismatches must stay below int(100 * 0.2) = 20
    rng = np.random.default_rng(6)
    a = list(rng.choice(list("ACGT"), 100))
    for k, want in ((19, 1), (20, 0)):
        c = list(a)
        for p in range(0, 5 * k, 5):
            c[p] = "ACGT"[("ACGT".index(c[p]) + 1) % 4]
        valid, (score, ov, shift, nops) = _align("".join(a), "".join(c))
        assert valid == want
        if want:
            assert (score, ov, shift, nops) == (19, 100, 0, 19)


def test_alignment_reaches_the_last_negative_shift():
    # the candidate overlaps the anchor only by its last 30 bases: the
    # last negative shift, -(100 - 30), is the only valid one
    rng = np.random.default_rng(5)
    a = "".join(rng.choice(list("ACGT"), 100))
    c = "".join(rng.choice(list("ACGT"), 70)) + a[:30]
    valid, (score, ov, shift, nops) = _align(a, c)
    assert valid == 1 and (ov, shift, nops) == (30, -70, 0) and score == 0 + 200 - 60


# ---------------------------------------------------------------------------
# single-end mismatch-ratio filter
# ---------------------------------------------------------------------------
def _thr(orc, ratios, err=0.06, cov=30.0, mcov=0.6):
    r = np.array(ratios, dtype=np.float32)
    return orc.orc_se_filter_threshold(r.ctypes.data, len(r), err, cov, mcov)


def test_se_filter_threshold_levels(orc):
    # error rate 0.06: levels 0.12, 0.18, 0.24; enough = 30 * 0.6 = 18
    two, three, four = F32(2) * F32(0.06), F32(3) * F32(0.06), F32(4) * F32(0.06)
    assert _thr(orc, [0.1] * 18) == two
    assert _thr(orc, [0.1] * 17 + [0.15]) == three
    assert _thr(orc, [0.1] * 17 + [0.2]) == four
    assert _thr(orc, [0.1] * 17 + [0.25]) == np.finfo(np.float32).min  # 17 at every level
    assert _thr(orc, [0.3] * 40) == np.finfo(np.float32).min           # none below 4x
    # a ratio equal to a level is not below it
    assert _thr(orc, [0.1] * 17 + [float(two)]) == three


# ---------------------------------------------------------------------------
# candidate correction window
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("qb,qe,classic,forest", [(5, 105, 1, 1), (4, 104, 0, 0), (35, 135, 1, 1), (36, 135, 0, 1),
                                                  (30, 136, 0, 0), (-1, 99, 0, 0), (20, 120, 1, 1)])
def test_candidate_window(orc, qb, qe, classic, forest):
    # anchor columns [20, 120), 15 new columns
    assert orc.orc_candidate_in_range(qb, qe, 20, 120, 15, 0) == classic
    assert orc.orc_candidate_in_range(qb, qe, 20, 120, 15, 1) == forest


# ---------------------------------------------------------------------------
# edits
# ---------------------------------------------------------------------------
def _edits(orc, corrected, original, ambiguous=0):
    pos = np.zeros(len(original), dtype=np.int32)
    bases = np.zeros(len(original), dtype=np.uint8)
    n = orc.orc_make_edits(corrected.encode(), original.encode(), len(original), ambiguous, pos.ctypes.data,
                           bases.ctypes.data)
    return n, list(pos[:max(n, 0)]), bases[:max(n, 0)].tobytes().decode()


def test_edits_limit_len_over_7(orc):
    orig = "ACGTACGTACGTAC"  # 14 bases: at most 2 edits
    assert _edits(orc, "TCGTACGTACGTAA", orig) == (2, [0, 13], "TA")
    assert _edits(orc, "TCGTACGTACGTAC", orig) == (1, [0], "T")
    assert _edits(orc, orig, orig) == (0, [], "")
    assert _edits(orc, "TCGTTCGTACGTAA", orig)[0] == -1  # 3 edits: whole sequence
    assert _edits(orc, "TCGTACGTACGTAC", orig, ambiguous=1)[0] == -1
    # 13 bases: at most 1 edit
    assert _edits(orc, "TCGTACGTACGTA", orig[:13])[0] == 1
    assert _edits(orc, "TCGTACGTACGTT", orig[:13])[0] == -1


# ---------------------------------------------------------------------------
# forest vote
# ---------------------------------------------------------------------------
def _node(att, split, left, right):
    """Preorder bytes of a node; left/right: float leaf probability or a
    nested (bytes, internal count) subtree."""
    flag = (2 if isinstance(left, float) else 0) | (1 if isinstance(right, float) else 0)
    out = struct.pack("<B", att) + struct.pack("<d", split) + struct.pack("<B", flag)
    n = 1
    for child in (left, right):
        if isinstance(child, float):
            out += struct.pack("<f", child)
        else:
            out += child[0]
            n += child[1]
    return out, n


def _write_forest(path, desc, trees):
    body = b""
    for t, n in trees:
        body += struct.pack("<I", n) + t
    d = desc.encode()
    with open(path, "wb") as f:
        f.write(struct.pack("<Q", len(d)) + d + struct.pack("<I", len(trees)) + body)


def test_forest_votes(orc, tmp_path):
    # tree 1: f0 <= 0.5 ? 0.9 : 0.1
    # tree 2: f1 <= 0.25 ? (f2 <= 0.75 ? 0.2 : 0.6) : 0.8
    t1 = _node(0, 0.5, 0.9, 0.1)
    t2 = _node(1, 0.25, _node(2, 0.75, 0.2, 0.6), 0.8)
    path = str(tmp_path / "a.rf")
    _write_forest(path, "13 extract_anchor_v3", [t1, t2])

    def vote(f, thr=0.5, trees=0xFFFFFFFF):
        x = np.zeros(14, dtype=np.float32)
        x[:len(f)] = f
        return orc.orc_forest_decide(path.encode(), 0, trees, thr, x.ctypes.data)

    assert vote([0.4, 0.1, 0.9]) == 1   # (0.9 + 0.6) / 2 = 0.75
    assert vote([0.6, 0.1, 0.5]) == 0   # (0.1 + 0.2) / 2 = 0.15
    assert vote([0.6, 0.3, 0.0]) == 0   # (0.1 + 0.8) / 2 = 0.45
    assert vote([0.6, 0.3, 0.0], thr=0.45) == 1
    # features equal to the splits go left (<=): (0.9 + 0.2) / 2 = 0.55; with
    # strict < they would go right: (0.1 + 0.8) / 2 = 0.45
    assert vote([0.5, 0.25, 0.75]) == 1
    assert vote([0.5, 0.25, 0.75], thr=0.54) == 1
    assert vote([0.5, 0.25, 0.75], thr=0.56) == 0
    assert vote([0.6, 0.1, 0.9], trees=1) == 0   # tree 1 alone: 0.1
    assert vote([0.4, 0.3, 0.9], trees=1) == 1   # tree 1 alone: 0.9
    # the candidate descriptor does not match an anchor forest
    x = np.zeros(14, dtype=np.float32)
    assert orc.orc_forest_decide(path.encode(), 1, 0xFFFFFFFF, 0.5, x.ctypes.data) == -1


# ---------------------------------------------------------------------------
# combine
# ---------------------------------------------------------------------------
def _combine(orc, original, results, ambiguous=0):
    """results: list of (flags, corrected sequence)."""
    n = len(results)
    flags = np.array([f for f, _ in results] + [0], dtype=np.int32)
    seqs = np.frombuffer("".join(s for _, s in results).encode() or b"\0", dtype=np.uint8).copy()
    out = np.zeros(len(original), dtype=np.uint8)
    label = orc.orc_combine(original.encode(), len(original), ambiguous, n, flags.ctypes.data, seqs.ctypes.data,
                            out.ctypes.data)
    return label, out.tobytes().decode()


def test_combine_rules(orc):
    o = "ACGTACGTACGTAC"
    a, b = "TCGTACGTACGTAC", "ACGTACGTACGTAA"
    many = "TTTTACGTACGTAC"  # more than len/7 edits: stored whole
    assert _combine(orc, o, []) == (0, o)
    assert _combine(orc, o, [(3, a)]) == (3, a)                       # HQ anchor
    assert _combine(orc, o, [(0, b), (3, a), (0, b)]) == (3, a)       # HQ anchor wins wherever it is
    assert _combine(orc, o, [(1, a)]) == (1, a)                       # LQ anchor alone
    assert _combine(orc, o, [(1, a), (0, b)]) == (1, a)               # two results: anchor alone
    assert _combine(orc, o, [(1, a), (0, a), (0, a)]) == (2, a)       # three agreeing results
    assert _combine(orc, o, [(0, b), (1, b), (0, b)]) == (2, b)
    assert _combine(orc, o, [(1, a), (0, a), (0, b)]) == (0, o)       # disagreement: uncorrected
    assert _combine(orc, o, [(0, a), (0, a), (0, a)]) == (0, o)       # no anchor result
    assert _combine(orc, o, [(1, many), (0, many), (0, many)]) == (2, many)
    assert _combine(orc, o, [(1, many), (0, a), (0, a)]) == (0, o)
    assert _combine(orc, o, [(1, o), (0, o), (0, o)]) == (2, o)       # empty edit lists agree
