"""Input parsing of the host (care_amd/host/care_host.cpp): four-line FASTQ
files are parsed in parallel blocks, everything else (FASTA, multi-line
records) sequentially like kseq (include/readlibraryio.hpp).  Both paths
must give the same read store: the same records written as four-line
FASTQ, as multi-line FASTQ, as FASTA, gzip-compressed, with CRLF line ends,
split over several files or over two mate files must pack to the same
sequences, lengths and ambiguity flags (and qualities, where the format has
them).  Blocks are 32 MiB, so a 70 MiB file crosses two block boundaries.
"""
import gzip
import os

import numpy as np
import pytest

import care_amd


def _records(n, seed, lmin=60, lmax=180):
    rng = np.random.default_rng(seed)
    recs = []
    for i in range(n):
        L = int(rng.integers(lmin, lmax + 1))
        s = "".join(rng.choice(list("ACGT"), L))
        if i % 37 == 0:
            s = s[:5] + "N" + s[6:]
        q = "".join(chr(33 + int(v)) for v in rng.integers(2, 42, L))
        recs.append(("r%d extra words" % i, s, q))
    return recs


def _write_fastq4(path, recs, eol="\n", opener=open):
    with opener(path, "wt", newline="") as f:
        for h, s, q in recs:
            f.write("@%s%s%s%s+%s%s%s" % (h, eol, s, eol, eol, q, eol))


def _write_fastq_multiline(path, recs):
    with open(path, "w") as f:
        for h, s, q in recs:
            f.write("@%s\n%s\n%s\n+%s\n%s\n%s\n" % (h, s[:50], s[50:], h, q[:50], q[50:]))


def _write_fasta(path, recs):
    with open(path, "w") as f:
        for h, s, _ in recs:
            f.write(">%s\n%s\n%s\n" % (h, s[:40], s[40:]))


def _load(files, paired=False, qbits=2):
    rs = care_amd.ReadSet(files, paired_end=paired, qual_bits=qbits)
    out = (rs.n, rs.lengths, rs.seqs, rs.ambiguous, rs.quals if qbits else None)
    rs.close()
    return out


def _same(a, b, quals=True):
    assert a[0] == b[0]
    for x, y in zip(a[1:4], b[1:4]):
        assert np.array_equal(x, y)
    if quals:
        assert np.array_equal(a[4], b[4])


def test_layouts_pack_alike(tmp_path):
    recs = _records(3000, 1)
    d = str(tmp_path)
    p4, pm, pf, pgz, pcr = (os.path.join(d, n) for n in ("a.fq", "m.fq", "a.fa", "a.fq.gz", "crlf.fq"))
    _write_fastq4(p4, recs)
    _write_fastq_multiline(pm, recs)
    _write_fasta(pf, recs)
    _write_fastq4(pgz, recs, opener=gzip.open)
    _write_fastq4(pcr, recs, eol="\r\n")
    want = _load([p4])
    assert want[0] == 3000
    assert np.array_equal(want[1], [len(s) for _, s, _ in recs])
    assert want[3].sum() == len(range(0, 3000, 37))
    _same(_load([pm]), want)
    _same(_load([pgz]), want)
    _same(_load([pcr]), want)
    _same(_load([pf], qbits=0), _load([p4], qbits=0), quals=False)
    # several single-end files: the N rule's counter restarts per file
    pa, pb = os.path.join(d, "part1.fq"), os.path.join(d, "part2.fq")
    _write_fastq4(pa, recs[:1000])
    _write_fastq4(pb, recs[1000:])
    both = _load([pa, pb])
    assert both[0] == 3000 and np.array_equal(both[1], want[1])
    pma, pmb = os.path.join(d, "mpart1.fq"), os.path.join(d, "mpart2.fq")
    _write_fastq_multiline(pma, recs[:1000])
    _write_fastq_multiline(pmb, recs[1000:])
    _same(_load([pma, pmb]), both)


def test_two_mate_files_interleave(tmp_path):
    recs = _records(2001, 2)
    d = str(tmp_path)
    m1, m2, inter = (os.path.join(d, n) for n in ("m1.fq", "m2.fq", "i.fq"))
    _write_fastq4(m1, recs[0:2000:2] + [recs[2000]])  # one unmatched mate at the end
    _write_fastq4(m2, recs[1:2000:2])
    _write_fastq4(inter, recs[:2000])
    got = _load([m1, m2], paired=True)
    _same(got, _load([inter], paired=True))
    # the same mates in multi-line records (sequential parser)
    mm1, mm2 = os.path.join(d, "mm1.fq"), os.path.join(d, "mm2.fq")
    _write_fastq_multiline(mm1, recs[0:2000:2])
    _write_fastq_multiline(mm2, recs[1:2000:2])
    _same(_load([mm1, mm2], paired=True), got)


def test_blocks_and_missing_final_newline(tmp_path):
    """About 70 MiB: records cut across 32 MiB block ends; the last record
    has no trailing newline."""
    recs = _records(1000, 3, 150, 150)
    d = str(tmp_path)
    big, big_m = os.path.join(d, "big.fq"), os.path.join(d, "big_m.fq")
    reps = 230
    with open(big, "w") as f:
        for r in range(reps):
            for h, s, q in recs:
                f.write("@%s\n%s\n+\n%s\n" % (h, s, q))
    with open(big, "rb+") as f:
        f.seek(-1, os.SEEK_END)
        f.truncate()
    assert os.path.getsize(big) > 64 << 20
    with open(big_m, "w") as f:
        for r in range(reps):
            for h, s, q in recs:
                f.write("@%s\n%s\n%s\n+\n%s\n%s\n" % (h, s[:70], s[70:], q[:70], q[70:]))
    got = _load([big])
    assert got[0] == reps * 1000
    _same(got, _load([big_m]))


def test_malformed_quality_is_an_error(tmp_path):
    p = os.path.join(str(tmp_path), "bad.fq")
    with open(p, "w") as f:
        f.write("@a\nACGT\n+\nIII\n@b\nACGT\n+\nIIII\n")
    with pytest.raises(care_amd.CareError, match="quality"):
        care_amd.ReadSet([p], qual_bits=8)
