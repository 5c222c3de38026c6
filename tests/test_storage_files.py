"""Reference file formats of the read store and the minhash tables
(care_amd/host/storage_files.cpp):

  preprocessed reads  include/chunkedreadstorage.hpp:158-400 (saveToFile,
                      loadFromFile), lengths as a LengthStore
                      (include/lengthstorage.hpp:113-204)
  hash tables         include/gpu/fakegpuminhasher.cuh:491-525 (writeToStream,
                      loadFromStream), per map a CpuReadOnlyMultiValueHashTable
                      (include/cpuhashtable.hpp:627-649) with an
                      AoSCpuSingleValueHashTable lookup (:63-121, :221-248)

The files are read here by a parser written from the reference's field
order, independent of the C++ writer: the LengthStore by the reference's
64-bit window extraction (getBits), the lookup by replaying the reference's
insertion (ascending keys at murmur(key) % capacity, linear probing) and
its query with the stored maxProbes.  Table contents are checked against
the oracle's tables; sequences, qualities and ambiguity flags against the
host read set.  A save -> load round trip must reproduce the corrected
output byte for byte, and a loaded file must be what the correction uses.
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

SE_CASE = {"profile": "mixed-se", "seed": 231, "genome": 2000}
SE_ARGS = ["-c", "50", "--pairmode", "SE", "-q", "--qualityScoreBits", "2", "-k", "20", "-h", "24"]
PE_CASE = {"profile": "repeat-pe", "seed": 232, "genome": 30000}
PE_ARGS = ["-c", "30", "--pairmode", "PE", "-q", "--qualityScoreBits", "2", "-k", "20", "-h", "32"]


class Reader:
    def __init__(self, path):
        with open(path, "rb") as f:
            self.b = f.read()
        self.p = 0

    def take(self, dtype, count=1):
        dt = np.dtype(dtype)
        n = dt.itemsize * count
        assert self.p + n <= len(self.b), "file ends early"
        a = np.frombuffer(self.b, dtype=dt, count=count, offset=self.p)
        self.p += n
        return a

    def one(self, dtype):
        return self.take(dtype)[0].item()


def length_at(data, bits, min_len, max_len, i):
    """LengthStore::getLength (include/lengthstorage.hpp:132-150, getBits
    :213-240): the bits of element i taken from the 64-bit window of the
    word holding its first bit and the next word."""
    if min_len == max_len:
        return min_len
    first = bits * i
    idx = first // 32
    begin = first - idx * 32
    end = begin + bits
    left = int(data[idx])
    right = int(data[idx]) if idx == len(data) - 1 else int(data[idx + 1])
    window = (left << 32) | right
    return ((window >> (64 - end)) & ((1 << bits) - 1)) + min_len


def parse_reads(path):
    r = Reader(path)
    f = {"n": r.one("<u8"), "upper": r.one("<i4"), "lower": r.one("<i4"), "has_qual": r.one("u1"),
         "qbits": r.one("<i4")}
    sizes = r.take("<u8", 4)
    start = r.p
    ls = {"datatbits": r.one("<i4"), "bits": r.one("<i4"), "min": r.one("<i4"), "max": r.one("<i4"),
          "mask": r.one("<u4"), "num": r.one("<i8")}
    raw_elems, raw_bytes = r.one("<u8"), r.one("<u8")
    assert raw_bytes == 4 * raw_elems
    ls["data"] = r.take("<u4", raw_elems)
    assert r.p - start == sizes[0]
    start = r.p
    pitch, nseq = r.one("<u8"), r.one("<u8")
    f["pitch"] = pitch
    f["seqs"] = r.take("<u4", nseq).reshape(f["n"], pitch) if f["n"] else np.zeros((0, pitch), np.uint32)
    assert r.p - start == sizes[1]
    start = r.p
    qpitch, nq = r.one("<u8"), r.one("<u8")
    f["qpitch"] = qpitch
    f["quals"] = r.take("<u4", nq).reshape(f["n"], qpitch) if qpitch else None
    assert r.p - start == sizes[2]
    start = r.p
    namb = r.one("<u8")
    f["ambig_ids"] = r.take("<u4", namb)
    assert r.p - start == sizes[3]
    assert r.p == len(r.b)
    f["ls"] = ls
    f["lengths"] = np.array([length_at(ls["data"], ls["bits"], ls["min"], ls["max"], i) for i in range(f["n"])],
                            dtype=np.int32)
    return f


SLOT = np.dtype([("key", "<u8"), ("offset", "<u4"), ("count", "<u2"), ("pad", "<u2")])


def parse_tables(path):
    r = Reader(path)
    t = {"k": r.one("<i4"), "threshold": r.one("<i4"), "load": r.take("<f4")[0], "num_maps": r.one("<i4"), "maps": []}
    for _ in range(t["num_maps"]):
        nv = r.one("<u8")
        m = {"values": r.take("<u4", nv), "load": r.take("<f4")[0], "num_keys": r.one("<u8"),
             "max_probes": r.one("<u8"), "size": r.one("<u8"), "capacity": r.one("<u8")}
        elems = r.one("<u8")
        assert elems == m["capacity"]
        m["slots"] = np.frombuffer(r.take("u1", elems * 16).tobytes(), dtype=SLOT)
        t["maps"].append(m)
    assert r.p == len(r.b)
    return t


def check_lookup(orc, m):
    """Replays AoSCpuSingleValueHashTable construction and query on a
    parsed map: slots must sit exactly where ascending-key insertion puts
    them, and every key must be found within maxProbes."""
    cap = m["capacity"]
    assert cap == int(np.float32(m["size"]) / np.float32(m["load"]))
    slots = m["slots"]
    empty = (slots["key"] == np.uint64(2**64 - 1)) & (slots["offset"] == 2**32 - 1) & (slots["count"] == 2**16 - 1)
    keys = np.sort(slots["key"][~empty])
    assert len(keys) == m["num_keys"] == m["size"]
    want = np.full(cap, np.uint64(2**64 - 1), dtype=np.uint64)
    max_probes = 0
    for key in keys.tolist():
        pos = orc.orc_murmur64(key) % cap
        probes = 0
        while want[pos] != np.uint64(2**64 - 1):
            pos = (pos + 1) % cap
            probes += 1
        want[pos] = key
        max_probes = max(max_probes, probes)
    assert np.array_equal(want, slots["key"])
    assert max_probes == m["max_probes"]
    # query: value ranges tile the values array in key order
    order = np.argsort(slots["key"][~empty], kind="stable")
    used = slots[~empty][order]
    assert np.array_equal(used["offset"], np.concatenate([[0], np.cumsum(used["count"])[:-1]]).astype(np.uint32))
    assert int(used["count"].sum()) == len(m["values"])
    return used


@pytest.fixture(scope="module")
def emu_exe():
    subprocess.run(["sh", os.path.join(EMU_DIR, "build_emu.sh")], check=True)
    return EMU_EXE


def _run(exe, reads, d, args, extra, out_name="r.fastq"):
    out = os.path.join(d, "out")
    p = subprocess.run([exe, "-i", reads, "-o", out_name, "-d", out, "-g", "0", "-t", "4"] + args + extra,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    return p, os.path.join(out, out_name)


def _check_files(rs, reads_file, tables_file, k, h, coverage, qbits):
    orc = helpers.oracle()
    f = parse_reads(reads_file)
    lens = rs.lengths
    assert f["n"] == rs.n and f["has_qual"] == 1 and f["qbits"] == qbits
    assert (f["upper"], f["lower"]) == (int(lens.max()), int(lens.min()))
    assert f["ls"]["datatbits"] == 32 and f["ls"]["num"] == rs.n
    assert (f["ls"]["min"], f["ls"]["max"]) == (int(lens.min()), int(lens.max()))
    assert f["ls"]["bits"] == int(lens.max() - lens.min()).bit_length()
    assert np.array_equal(f["lengths"], lens)
    assert np.array_equal(f["seqs"], rs.seqs)
    assert np.array_equal(f["quals"], rs.quals)
    assert np.array_equal(f["ambig_ids"], np.nonzero(rs.ambiguous)[0].astype(np.uint32))
    assert len(f["ambig_ids"]) > 0

    t = parse_tables(tables_file)
    cap = max(10, min(int(np.float32(coverage) * np.float32(2.5)), 65535))
    assert (t["k"], t["threshold"], t["num_maps"]) == (k, cap, h)
    assert abs(t["load"] - 0.8) < 1e-7
    seqs = np.ascontiguousarray(rs.seqs)
    tab = orc.orc_tables_build(seqs.ctypes.data, rs.pitch, lens.ctypes.data, rs.n, k, h, coverage)
    sigs = np.zeros((rs.n, h), dtype=np.uint64)
    for i in range(rs.n):
        orc.orc_signature(seqs[i].ctypes.data, int(lens[i]), k, h, sigs[i].ctypes.data)
    buf = np.zeros(70000, dtype=np.uint32)
    try:
        for mi, m in enumerate(t["maps"]):
            used = check_lookup(orc, m)
            file_keys = set(used["key"].tolist())
            for s in used:
                c = orc.orc_tables_bucket(tab, mi, int(s["key"]), buf.ctypes.data, len(buf))
                assert c == s["count"]
                got = np.sort(m["values"][s["offset"]:s["offset"] + s["count"]])
                assert np.array_equal(got, np.sort(buf[:c]))
            # every key the oracle keeps is in the file
            for key in set(sigs[:, mi].tolist()) - file_keys:
                assert orc.orc_tables_bucket(tab, mi, key, buf.ctypes.data, len(buf)) == 0
    finally:
        orc.orc_tables_free(tab)


def test_length_store_known_answer(tmp_path):
    """Lengths 100, 101, 103, 100: two bits each (3 = 0b11), values
    00 01 11 00 from the top of word 0 -> 0x1C000000; an N makes its read
    ambiguous."""
    d = str(tmp_path)
    fq = os.path.join(d, "t.fastq")
    rng = np.random.default_rng(5)
    with open(fq, "w") as f:
        for i, L in enumerate([100, 101, 103, 100]):
            s = "".join(rng.choice(list("ACGT"), L))
            if i == 2:
                s = s[:10] + "N" + s[11:]
            f.write("@r%d\n%s\n+\n%s\n" % (i, s, "I" * L))
    rs = care_amd.ReadSet([fq], qual_bits=2)
    path = os.path.join(d, "r.bin")
    rs.save(path, qual_bits=2)
    p = parse_reads(path)
    assert p["ls"]["bits"] == 2 and p["ls"]["mask"] == 3
    assert p["ls"]["data"].tolist() == [0x1C000000]
    assert p["lengths"].tolist() == [100, 101, 103, 100]
    assert p["ambig_ids"].tolist() == [2]
    assert p["pitch"] == 7 and p["qpitch"] == 7  # ceil(103 / 16), ceil(2 * 103 / 32)
    # all 'I' (Q40) -> bin 3 everywhere, last word of a 100-base read padded
    assert p["quals"][0, 0] == 0xFFFFFFFF and p["quals"][0, 6] == 0xFFFFFFFF << 24 & 0xFFFFFFFF
    back = care_amd.ReadSet.load_preprocessed(path, True, 2)
    for a in ("seqs", "lengths", "ambiguous", "quals"):
        assert np.array_equal(getattr(back, a), getattr(rs, a)), a
    assert back.max_length == 103
    # without qualities the section is skipped
    noq = care_amd.ReadSet.load_preprocessed(path, False, 8)
    assert noq.qpitch == 0 and np.array_equal(noq.seqs, rs.seqs)
    with pytest.raises(care_amd.CareError, match="bits per quality score"):
        care_amd.ReadSet.load_preprocessed(path, True, 8)
    with open(path, "rb") as f:
        data = f.read()
    cut = os.path.join(d, "cut.bin")
    for n in (5, 60, len(data) - 3):
        with open(cut, "wb") as f:
            f.write(data[:n])
        with pytest.raises(care_amd.CareError):
            care_amd.ReadSet.load_preprocessed(cut, True, 2)


def test_length_store_wide_values(tmp_path):
    """Lengths 1..140 in a scrambled order: 8-bit fields (139 < 256) that
    straddle word boundaries, decoded by the reference's window rule."""
    d = str(tmp_path)
    fq = os.path.join(d, "t.fasta")
    lens = np.random.default_rng(6).permutation(np.arange(1, 141))
    with open(fq, "w") as f:
        for i, L in enumerate(lens):
            f.write(">r%d\n%s\n" % (i, "ACGT" * (L // 4) + "A" * (L % 4)))
    rs = care_amd.ReadSet([fq])
    path = os.path.join(d, "r.bin")
    rs.save(path)
    p = parse_reads(path)
    assert p["ls"]["bits"] == 8 and p["has_qual"] == 0 and p["qpitch"] == 0
    assert np.array_equal(p["lengths"], lens)
    back = care_amd.ReadSet.load_preprocessed(path, False, 8)
    assert np.array_equal(back.lengths, lens) and np.array_equal(back.seqs, rs.seqs)
    with pytest.raises(care_amd.CareError, match="none are present"):
        care_amd.ReadSet.load_preprocessed(path, True, 8)


def test_emulated_save_load_roundtrip(emu_exe, tmp_path):
    d = str(tmp_path)
    reads, _ = helpers.simulate(SE_CASE, d)
    rfile, tfile = os.path.join(d, "reads.bin"), os.path.join(d, "tables.bin")
    p, out1 = _run(emu_exe, reads, d, SE_ARGS, ["--save-preprocessedreads-to", rfile, "--save-hashtables-to", tfile],
                   "a.fastq")
    assert p.returncode == 0, p.stdout[-3000:]
    assert "Saved reads" in p.stdout and "Saved minhasher" in p.stdout
    rs = care_amd.ReadSet([reads], qual_bits=2)
    _check_files(rs, rfile, tfile, 20, 24, 50.0, 2)
    want = helpers.md5(out1)
    for extra, name in ((["--load-preprocessedreads-from", rfile, "--load-hashtables-from", tfile], "b.fastq"),
                        (["--load-preprocessedreads-from", rfile], "c.fastq"),
                        (["--load-hashtables-from", tfile], "d.fastq")):
        p, out = _run(emu_exe, reads, d, SE_ARGS, extra, name)
        assert p.returncode == 0, p.stdout[-3000:]
        assert helpers.md5(out) == want, extra
    assert "Loaded 24 hash tables" in p.stdout
    # the loaded tables are the ones used: a file of fewer maps changes the result
    p, out = _run(emu_exe, reads, d, SE_ARGS[:-1] + ["12"], ["--load-hashtables-from", tfile], "e.fastq")
    assert p.returncode == 0 and "Loaded 12 hash tables" in p.stdout
    assert helpers.md5(out) != want
    # and so are the loaded reads: a file holding every read reverse-complemented
    f = parse_reads(rfile)
    with open(rfile, "rb") as fh:
        data = bytearray(fh.read())
    seq_at = 21 + 32 + 44 + 4 * len(f["ls"]["data"]) + 16
    orc = helpers.oracle()
    rc = np.zeros_like(f["seqs"])
    for i in range(f["n"]):
        row = np.ascontiguousarray(f["seqs"][i])
        orc.orc_revcomp_2bit(row.ctypes.data, int(f["lengths"][i]), rc[i].ctypes.data)
    data[seq_at:seq_at + rc.nbytes] = rc.tobytes()
    changed = os.path.join(d, "changed.bin")
    with open(changed, "wb") as fh:
        fh.write(bytes(data))
    p, out = _run(emu_exe, reads, d, SE_ARGS, ["--load-preprocessedreads-from", changed], "f.fastq")
    assert p.returncode == 0 and helpers.md5(out) != want


def test_emulated_load_errors(emu_exe, tmp_path):
    d = str(tmp_path)
    reads, _ = helpers.simulate(SE_CASE, d)
    rfile, tfile = os.path.join(d, "reads.bin"), os.path.join(d, "tables.bin")
    p, _ = _run(emu_exe, reads, d, SE_ARGS, ["--save-preprocessedreads-to", rfile, "--save-hashtables-to", tfile])
    assert p.returncode == 0, p.stdout[-3000:]
    with open(tfile, "rb") as fh:
        data = fh.read()
    cut = os.path.join(d, "cut.bin")
    with open(cut, "wb") as fh:
        fh.write(data[:len(data) // 2])
    p, _ = _run(emu_exe, reads, d, SE_ARGS, ["--load-hashtables-from", cut], "x.fastq")
    assert p.returncode != 0 and ("ends early" in p.stdout or "beyond the end" in p.stdout)
    p, _ = _run(emu_exe, reads, d, SE_ARGS, ["--load-hashtables-from", os.path.join(d, "missing.bin")], "y.fastq")
    assert p.returncode != 0 and "cannot open" in p.stdout
    # the read file must describe the inputs
    other = os.path.join(d, "other.fastq")
    with open(reads) as src, open(other, "w") as dst:
        dst.writelines(src.readlines()[:400])
    p, _ = _run(emu_exe, other, d, SE_ARGS, ["--load-preprocessedreads-from", rfile], "z.fastq")
    assert p.returncode != 0 and "preprocessed reads file holds" in p.stdout


def test_emulated_spilled_results(emu_exe, tmp_path):
    """-m below what the reads already take: every correction result goes
    to a temporary file in --tempdir, which is gone after the run, and the
    output is unchanged."""
    d = str(tmp_path)
    reads, _ = helpers.simulate(SE_CASE, d)
    p, out1 = _run(emu_exe, reads, d, SE_ARGS + ["--candidateCorrection"], [], "a.fastq")
    assert p.returncode == 0, p.stdout[-3000:]
    assert "temporary files" not in p.stdout
    tmp = os.path.join(d, "spill")
    for mem in ("1K", "300K"):
        p, out2 = _run(emu_exe, reads, d, SE_ARGS + ["--candidateCorrection"], ["-m", mem, "--tempdir", tmp],
                       "b%s.fastq" % mem)
        assert p.returncode == 0, p.stdout[-3000:]
        assert "kept in temporary files in " + tmp in p.stdout
        assert helpers.md5(out2) == helpers.md5(out1)
        assert os.listdir(tmp) == []


def test_engine_map_export_import_emulated(tmp_path):
    """care_minhash_export_map gives each map in key order; importing the
    exported maps into a fresh context reproduces every candidate list."""
    subprocess.run(["sh", os.path.join(EMU_DIR, "build_emu.sh")], check=True)
    lib = care_amd._bind(ctypes.CDLL(os.path.join(EMU_DIR, "_build", "libcare_amd_emu.so")),
                         care_amd.ENGINE_SYMBOLS)
    reads, _ = helpers.simulate({"profile": "mixed-se", "seed": 233, "genome": 1500}, str(tmp_path))
    rs = care_amd.ReadSet([reads], qual_bits=0)
    opt = care_amd.default_options(kmer_length=20, num_hash_maps=16, coverage=30.0)
    a = care_amd.Engine(opt, device=0, lib=lib)
    a.upload_readset(rs)
    a.build()
    maps = [a.export_map(m) for m in range(16)]
    want = a.query(0, rs.n)
    sigs = a.signatures(0, rs.n)
    for m, (keys, offs, cnts, vals) in enumerate(maps):
        assert np.all(keys[1:] > keys[:-1])
        assert np.array_equal(offs, np.concatenate([[0], np.cumsum(cnts)[:-1]]).astype(np.uint32))
        assert np.all((cnts >= 2) & (cnts <= 75))
        for key, o, c in zip(keys[:200], offs[:200], cnts[:200]):
            ids = vals[o:o + c]
            assert np.all(np.diff(ids.astype(np.int64)) > 0)
            assert np.all(sigs[ids, m] == key)
            assert np.array_equal(np.sort(ids), np.nonzero(sigs[:, m] == key)[0])
    a.close()
    opt2 = care_amd.default_options(kmer_length=15, num_hash_maps=40, coverage=30.0)
    b = care_amd.Engine(opt2, device=0, lib=lib)
    b.upload_readset(rs)
    # shuffled value order within keys: the import sorts them
    shuffled = []
    rng = np.random.default_rng(1)
    for keys, offs, cnts, vals in maps:
        v = vals.copy()
        for o, c in zip(offs, cnts):
            v[o:o + c] = rng.permutation(v[o:o + c])
        shuffled.append((keys, offs, cnts, v))
    b.import_maps(shuffled, 20, 75)
    got = b.query(0, rs.n)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    kk, hh, cc = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    assert lib.care_minhash_geometry(b.ctx, ctypes.byref(kk), ctypes.byref(hh), ctypes.byref(cc)) == 0
    assert (kk.value, hh.value, cc.value) == (20, 16, 75)
    # bad inputs are refused
    keys, offs, cnts, vals = maps[0]
    with pytest.raises(care_amd.CareError, match="repeated key"):
        b.import_maps([(np.concatenate([keys, keys[:1]]), np.concatenate([offs, offs[:1]]),
                        np.concatenate([cnts, cnts[:1]]), vals)], 20, 75)
    with pytest.raises(care_amd.CareError, match="outside the store"):
        bad = vals.copy()
        bad[0] = rs.n
        b.import_maps([(keys, offs, cnts, bad)], 20, 75)
    with pytest.raises(care_amd.CareError, match="outside 1..10"):
        b.import_maps([(keys, offs, cnts, vals)], 20, 10)
    b.close()
    rs.close()


@pytest.mark.gpu
def test_gpu_save_load_roundtrip(tmp_path):
    """care-gpu on the device: save both files, check their layout against
    the oracle, load them back and reproduce the output byte for byte."""
    print("save/load round trip on the device (repeat-pe seed 232)", flush=True)
    d = str(tmp_path)
    reads, _ = helpers.simulate(PE_CASE, d)
    exe = os.path.join(helpers.ROOT, "care_amd", "_build", "care-gpu")
    rfile, tfile = os.path.join(d, "reads.bin"), os.path.join(d, "tables.bin")
    p, out1 = _run(exe, reads, d, PE_ARGS, ["--save-preprocessedreads-to", rfile, "--save-hashtables-to", tfile],
                   "a.fastq")
    assert p.returncode == 0, p.stdout[-3000:]
    rs = care_amd.ReadSet([reads], paired_end=True, qual_bits=2)
    f = parse_reads(rfile)
    assert np.array_equal(f["seqs"], rs.seqs) and np.array_equal(f["lengths"], rs.lengths)
    t = parse_tables(tfile)
    orc = helpers.oracle()
    for m in t["maps"][:4]:
        check_lookup(orc, m)
    p, out2 = _run(exe, reads, d, PE_ARGS,
                   ["--load-preprocessedreads-from", rfile, "--load-hashtables-from", tfile], "b.fastq")
    assert p.returncode == 0, p.stdout[-3000:]
    assert helpers.md5(out2) == helpers.md5(out1)
