"""Shared helpers of the test suite: simulator inputs, oracle access."""
import ctypes
import hashlib
import json
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
READSIM = os.path.join(ROOT, "care_amd", "_build", "readsim")
ORACLE_LIB = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
ORACLE_EXE = os.path.join(ROOT, "oracle", "_build", "care-oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden", "golden.json")


def md5(path):
    h = hashlib.md5()
    with open(path, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 20), b""):
            h.update(chunk)
    return h.hexdigest()


def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def simulate(case, directory):
    reads = os.path.join(directory, "reads.fastq")
    truth = os.path.join(directory, "truth.fastq")
    if not os.path.exists(reads):
        subprocess.run([READSIM, "gen", "--profile", case["profile"], "--seed", str(case["seed"]),
                        "--genome", str(case["genome"]), "--out", reads, "--truth", truth]
                       + [str(a) for a in case.get("sim_args", [])], check=True)
    return reads, truth


def oracle():
    # CARE_ORACLE_LIB: another build of the oracle (tests/oracle_mutations.py)
    lib = ctypes.CDLL(os.environ.get("CARE_ORACLE_LIB", ORACLE_LIB))
    u32p = ctypes.POINTER(ctypes.c_uint32)
    lib.orc_encode_2bit.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p]
    lib.orc_decode_2bit.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p]
    lib.orc_revcomp_2bit.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    lib.orc_to_hilo.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    lib.orc_encode_quality.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    lib.orc_decode_quality.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p]
    lib.orc_murmur64.argtypes = [ctypes.c_uint64]
    lib.orc_murmur64.restype = ctypes.c_uint64
    lib.orc_signature.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    lib.orc_quality_weight.argtypes = [ctypes.c_int]
    lib.orc_quality_weight.restype = ctypes.c_float
    lib.orc_align.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                              ctypes.c_float, ctypes.c_float, ctypes.c_void_p]
    lib.orc_choose_orientation.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_float, ctypes.c_int, ctypes.c_float]
    lib.orc_tables_build.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_float]
    lib.orc_tables_build.restype = ctypes.c_void_p
    lib.orc_tables_free.argtypes = [ctypes.c_void_p]
    lib.orc_tables_query.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_void_p,
                                     ctypes.c_int]
    lib.orc_tables_bucket.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_int]
    lib.orc_tables_query_limited.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32,
                                             ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
    return lib


def encode(lib, seq):
    n = (len(seq) + 15) // 16
    out = np.zeros(max(n, 1), dtype=np.uint32)
    lib.orc_encode_2bit(seq.encode(), len(seq), out.ctypes.data)
    return out
