"""Pins of the CPU oracle (oracle/care_oracle.cpp).

1. End to end: the oracle's corrected output equals the reference care-cpu
   output (golden digests, tests/golden/golden.json) on every seeded case.
2. Closed forms and brute force for the building blocks: 2-bit codec and
   reverse complement, HiLo planes, quality binning, quality weights,
   canonical k-mer minhash signatures, table contents, shifted-Hamming
   alignment and the orientation rule.
"""
import os
import random
import subprocess

import numpy as np
import pytest

from tests import helpers
from tests.cases import CASES, case_args

BASES = "ACGT"


@pytest.fixture(scope="module")
def orc():
    return helpers.oracle()


def _rand_seq(rng, n):
    return "".join(rng.choice(BASES) for _ in range(n))


def _revcomp(s):
    return s[::-1].translate(str.maketrans("ACGT", "TGCA"))


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_output_equals_reference(name, tmp_path):
    case = CASES[name]
    reads, truth = helpers.simulate(case, str(tmp_path))
    g = helpers.golden()[name]
    assert helpers.md5(reads) == g["input_md5"]
    p = subprocess.run([helpers.ORACLE_EXE, "-i", reads, "-o", "o.fastq", "-d", str(tmp_path / "out"), "-t", "8"]
                       + case_args(case, str(tmp_path)), stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert p.returncode == 0, p.stdout
    assert helpers.md5(str(tmp_path / "out" / "o.fastq")) == g["output_md5"]


def test_codec_roundtrip_and_layout(orc):
    rng = random.Random(1)
    for n in [1, 15, 16, 17, 31, 32, 33, 100, 150, 250]:
        s = _rand_seq(rng, n)
        enc = helpers.encode(orc, s)
        # layout: base i in bits 31-2(i%16) .. of word i//16, A0 C1 G2 T3
        for i, c in enumerate(s):
            assert (int(enc[i // 16]) >> (30 - 2 * (i % 16))) & 3 == BASES.index(c)
        out = (ctypes_buf := bytearray(n))
        import ctypes
        buf = ctypes.create_string_buffer(n)
        orc.orc_decode_2bit(enc.ctypes.data, n, buf)
        assert buf.raw.decode() == s
        rc = np.zeros_like(enc)
        orc.orc_revcomp_2bit(enc.ctypes.data, n, rc.ctypes.data)
        orc.orc_decode_2bit(rc.ctypes.data, n, buf)
        assert buf.raw.decode() == _revcomp(s)
        del out, ctypes_buf


def test_hilo_planes(orc):
    rng = random.Random(2)
    for n in [5, 32, 33, 64, 100, 250]:
        s = _rand_seq(rng, n)
        enc = helpers.encode(orc, s)
        words = (n + 31) // 32
        out = np.zeros(2 * words, dtype=np.uint32)
        orc.orc_to_hilo(enc.ctypes.data, n, out.ctypes.data)
        for i, c in enumerate(s):
            code = BASES.index(c)
            assert (int(out[i // 32]) >> (31 - i % 32)) & 1 == code >> 1
            assert (int(out[words + i // 32]) >> (31 - i % 32)) & 1 == code & 1


def test_quality_binning(orc):
    import ctypes
    # bins of include/qualityscorecompression.hpp:28-68
    qs = "".join(chr(33 + q) for q in range(0, 42))
    for bits, bins in [(2, [(4, "$"), (13, "("), (20, "1"), (999, "9")]), (1, [(13, "("), (999, "5")])]:
        enc = np.zeros(8, dtype=np.uint32)
        orc.orc_encode_quality(qs.encode(), len(qs), bits, enc.ctypes.data)
        buf = ctypes.create_string_buffer(len(qs))
        orc.orc_decode_quality(enc.ctypes.data, len(qs), bits, buf)
        for q, got in zip(range(0, 42), buf.raw.decode()):
            want = next(ch for bound, ch in bins if q <= bound)
            assert got == want, (bits, q)
    enc = np.zeros(16, dtype=np.uint32)
    orc.orc_encode_quality(qs.encode(), len(qs), 8, enc.ctypes.data)
    buf = ctypes.create_string_buffer(len(qs))
    orc.orc_decode_quality(enc.ctypes.data, len(qs), 8, buf)
    assert buf.raw.decode() == qs


def test_quality_weights(orc):
    # 1 - 10^(-(q-33)/10), floor 0.001, characters below '!' weigh 0.001
    for c in range(256):
        w = orc.orc_quality_weight(c)
        if c <= 33:
            assert w == np.float32(0.001)
        else:
            ref = max(0.001, 1.0 - 10.0 ** (-(c - 33) / 10.0))
            assert abs(w - ref) < 2e-7


def test_murmur_is_bijective_on_sample(orc):
    xs = [random.Random(3).getrandbits(64) for _ in range(20000)]
    hs = {orc.orc_murmur64(x) for x in xs}
    assert len(hs) == len(set(xs))
    assert orc.orc_murmur64(0) == 0


def _bruteforce_signature(orc, s, k, h):
    mask = (1 << (2 * k)) - 1
    out = []
    kmers = []
    for p in range(len(s) - k + 1):
        f = s[p:p + k]
        r = _revcomp(f)
        val = lambda t: int("".join(str(BASES.index(c)) for c in t), 4)
        kmers.append(min(val(f), val(r)))
    for j in range(h):
        if not kmers:
            out.append(((1 << 64) - 1) & mask)
            continue
        out.append(min(orc.orc_murmur64((x + j) & ((1 << 64) - 1)) for x in kmers) & mask)
    return out


def test_signature_bruteforce(orc):
    rng = random.Random(4)
    for n, k, h in [(40, 20, 4), (33, 16, 3), (25, 20, 2), (15, 20, 2), (60, 31, 2)]:
        s = _rand_seq(rng, n)
        enc = helpers.encode(orc, s)
        out = np.zeros(h, dtype=np.uint64)
        orc.orc_signature(enc.ctypes.data, n, k, h, out.ctypes.data)
        assert [int(x) for x in out] == _bruteforce_signature(orc, s, k, h)


def test_tables_bruteforce(orc):
    rng = random.Random(5)
    genome = _rand_seq(rng, 400)
    reads = []
    for _ in range(120):
        p = rng.randrange(0, 400 - 60)
        r = genome[p:p + 60]
        reads.append(_revcomp(r) if rng.random() < 0.5 else r)
    k, h, cov = 12, 6, 4.0
    pitch = 4
    seqs = np.zeros((len(reads), pitch), dtype=np.uint32)
    for i, r in enumerate(reads):
        seqs[i, :] = helpers.encode(orc, r)
    lens = np.full(len(reads), 60, dtype=np.int32)
    sig = np.zeros((len(reads), h), dtype=np.uint64)
    for i in range(len(reads)):
        orc.orc_signature(seqs[i].ctypes.data, 60, k, h, sig[i].ctypes.data)
    cap = max(10, int(cov * 2.5))
    t = orc.orc_tables_build(seqs.ctypes.data, pitch, lens.ctypes.data, len(reads), k, h, cov)
    buf = np.zeros(1000, dtype=np.uint32)
    for i in range(len(reads)):
        want = set()
        for j in range(h):
            same = [r for r in range(len(reads)) if sig[r, j] == sig[i, j]]
            if 2 <= len(same) <= cap:
                want.update(same)
            got_bucket = orc.orc_tables_bucket(t, j, int(sig[i, j]), buf.ctypes.data, len(buf))
            assert sorted(buf[:got_bucket].tolist()) == (sorted(same) if 2 <= len(same) <= cap else [])
        want.discard(i)
        c = orc.orc_tables_query(t, seqs[i].ctypes.data, 60, i, buf.ctypes.data, len(buf))
        assert buf[:c].tolist() == sorted(want)
    orc.orc_tables_free(t)


def _bruteforce_align(a, c, min_overlap, max_err, ratio):
    la, lc = len(a), len(c)
    total = la + lc
    minov = max(min_overlap, int(np.float32(la) * np.float32(ratio)))
    order = list(range(0, la - minov + 1)) + [-x for x in range(1, lc - minov + 1)]
    best = None
    for s in order:
        ov = min(la - s, lc) if s >= 0 else min(la, lc + s)
        lim = int(np.float32(ov) * np.float32(max_err))
        sa, sc = max(s, 0), max(-s, 0)
        m = sum(1 for i in range(ov) if a[sa + i] != c[sc + i])
        if m < lim:
            score = m + total - 2 * ov
            if best is None or score < best[0]:
                best = (score, s)
    if best is None:
        return (total, 0, -lc, 0, 0)
    score, s = best
    ov = min(lc, la - s) - max(-s, 0)
    return (score, ov, s, score - total + 2 * ov, 1)


def test_alignment_bruteforce(orc):
    rng = random.Random(6)
    out = np.zeros(4, dtype=np.int32)
    for trial in range(300):
        la = rng.choice([40, 64, 100, 150, 97])
        lc = rng.choice([40, 64, 100, 150, 120])
        g = _rand_seq(rng, la + lc + 50)
        pa = rng.randrange(0, 40)
        a = g[pa:pa + la]
        pc = max(0, pa + rng.randrange(-lc + 20, la - 10))
        c = list(g[pc:pc + lc].ljust(lc, "A"))
        for _ in range(rng.randrange(0, 6)):
            i = rng.randrange(lc)
            c[i] = rng.choice(BASES)
        c = "".join(c)
        ea, ec = helpers.encode(orc, a), helpers.encode(orc, c)
        valid = orc.orc_align(ea.ctypes.data, la, ec.ctypes.data, lc, 30, 0.2, 0.3, out.ctypes.data)
        got = (int(out[0]), int(out[1]), int(out[2]), int(out[3]), int(valid))
        want = _bruteforce_align(a, c, 30, 0.2, 0.3)
        if want[4] == 0:
            assert valid == 0
        else:
            assert got == want, (trial, a, c)


def test_alignment_closed_forms(orc):
    rng = random.Random(7)
    out = np.zeros(4, dtype=np.int32)
    s = _rand_seq(rng, 100)
    e = helpers.encode(orc, s)
    assert orc.orc_align(e.ctypes.data, 100, e.ctypes.data, 100, 30, 0.2, 0.3, out.ctypes.data) == 1
    assert list(out) == [0, 100, 0, 0]
    g = _rand_seq(rng, 200)
    for shift in [-60, -1, 1, 17, 70]:
        a = g[60:160]
        c = g[60 + shift:160 + shift]
        ea, ec = helpers.encode(orc, a), helpers.encode(orc, c)
        assert orc.orc_align(ea.ctypes.data, 100, ec.ctypes.data, 100, 30, 0.2, 0.3, out.ctypes.data) == 1
        assert int(out[2]) == shift and int(out[3]) == 0 and int(out[1]) == 100 - abs(shift)


def test_orientation_rule(orc):
    def call(f, fv, r, rv):
        fa = np.array(f, dtype=np.int32)
        ra = np.array(r, dtype=np.int32)
        return orc.orc_choose_orientation(fa.ctypes.data, fv, ra.ctypes.data, rv, 100, 0.3, 30, 0.06)
    # lower mismatch ratio wins; equal ratio -> longer overlap, tie -> reverse
    assert call([0, 80, 0, 2], 1, [0, 80, 0, 4], 1) == 1
    assert call([0, 80, 0, 4], 1, [0, 80, 0, 2], 1) == 2
    assert call([0, 90, 0, 0], 1, [0, 80, 0, 0], 1) == 1
    assert call([0, 80, 0, 0], 1, [0, 80, 0, 0], 1) == 2
    # overlap below max(30, 0.3*100) or ratio >= 0.24 -> none
    assert call([0, 29, 0, 0], 1, [0, 0, 0, 0], 0) == 3
    assert call([0, 50, 0, 12], 1, [0, 0, 0, 0], 0) == 3
    assert call([0, 50, 0, 11], 1, [0, 0, 0, 0], 0) == 1
