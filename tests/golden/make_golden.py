"""Writes tests/golden/golden.json: input and output digests of the seeded
end-to-end cases (tests/cases.py), with the base-level evaluation of the
output against the simulator's truth.

The output digests are regression vectors of the corrected FASTQ.  They were
first recorded from care-cpu compiled from the reference sources with
warpcore's vendored copy of hpc_helpers standing in for the un-vendored
include/hpc_helpers submodule; such a build does not count as the reference,
so the digests are not an independent pin of the oracle (the oracle's pins are
tests/test_oracle.py and tests/test_oracle_pins.py).  This script regenerates
them from the oracle (oracle/_build/care-oracle) and checks that the oracle
still reproduces the recorded values; run with --write to replace the file.
Inputs are regenerated from their seeds by care_amd/_build/readsim; only the
parameters and digests are stored.
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from tests.cases import CASES, case_args  # noqa: E402

ORACLE = os.path.join(ROOT, "oracle", "_build", "care-oracle")
READSIM = os.path.join(ROOT, "care_amd", "_build", "readsim")
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.json")


def md5(path):
    h = hashlib.md5()
    with open(path, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 20), b""):
            h.update(chunk)
    return h.hexdigest()


def main(write):
    out = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, case in CASES.items():
            reads = os.path.join(tmp, name + ".fastq")
            truth = os.path.join(tmp, name + ".truth.fastq")
            subprocess.run([READSIM, "gen", "--profile", case["profile"], "--seed", str(case["seed"]),
                            "--genome", str(case["genome"]), "--out", reads, "--truth", truth]
                           + [str(a) for a in case.get("sim_args", [])], check=True)
            odir = os.path.join(tmp, name + "_out")
            subprocess.run([ORACLE, "-i", reads, "-o", "out.fastq", "-d", odir, "-t", "8"] + case_args(case, tmp),
                           check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
            ev = subprocess.run([READSIM, "eval", "--orig", reads, "--corr", os.path.join(odir, "out.fastq"),
                                 "--truth", truth], check=True, stdout=subprocess.PIPE, text=True)
            out[name] = {"input_md5": md5(reads), "output_md5": md5(os.path.join(odir, "out.fastq")),
                         "eval": json.loads(ev.stdout)}
            print(name, out[name])
    if write:
        with open(GOLDEN, "w") as f:
            json.dump(out, f, indent=1, sort_keys=True)
    else:
        with open(GOLDEN) as f:
            old = json.load(f)
        bad = [n for n in out if old.get(n, {}).get("output_md5") != out[n]["output_md5"]]
        print("differs from the recorded digests:" if bad else "all digests as recorded", *bad)
        return 1 if bad else 0
    return 0


if __name__ == "__main__":
    sys.exit(main("--write" in sys.argv[1:]))
