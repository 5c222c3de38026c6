"""Seeded end-to-end cases shared by the golden-fixture script and the tests.

Each case: simulator profile, seed and genome size (care_amd/sim/readsim.cpp)
and the corrector options (same flags for care-cpu, care-oracle, care-gpu).
Read counts stay below 65536 per file so that the reference's stale quality
padding (reused encode buffers, include/chunkedreadstorageconstruction.hpp:
216-237) cannot occur.
"""

CASES = {
    # single-end, no qualities (uniform 100 bp, 1% errors)
    "se_plain": {"profile": "uniform-se", "seed": 101, "genome": 60000,
                 "args": ["-c", "30", "--pairmode", "SE", "-k", "20", "-h", "16"]},
    # paired-end with repeats, 2-bit qualities
    "pe_quality2": {"profile": "repeat-pe", "seed": 102, "genome": 120000,
                    "args": ["-c", "30", "--pairmode", "PE", "-q", "--qualityScoreBits", "2", "-k", "20", "-h", "32"]},
    # variable length 100-250, Ns, exclude ambiguous, 8-bit qualities
    "se_mixed_ambig": {"profile": "mixed-se", "seed": 103, "genome": 120000,
                       "args": ["-c", "50", "--pairmode", "SE", "--excludeAmbiguous", "-q", "-k", "20", "-h", "48"]},
    # paired-end with candidate correction and quality labels
    "pe_candidates": {"profile": "large-pe", "seed": 104, "genome": 120000,
                      "args": ["-c", "40", "--pairmode", "PE", "-q", "--candidateCorrection", "-k", "20", "-h", "48",
                               "--correctionQualityLabels"]},
    # forest classifier for anchors and candidates (random forests written by
    # tests/forest_files.py from the given seeds)
    "pe_forest": {"profile": "repeat-pe", "seed": 106, "genome": 60000,
                  "args": ["-c", "30", "--pairmode", "PE", "-q", "-k", "20", "-h", "32", "--candidateCorrection",
                           "--correctionType", "1", "--correctionTypeCands", "1"],
                  "forest": {"anchor": 1, "cands": 2}},
    "se_forest_anchor": {"profile": "uniform-se", "seed": 107, "genome": 40000,
                         "args": ["-c", "30", "--pairmode", "SE", "-q", "-k", "20", "-h", "16", "--correctionType", "1"],
                         "forest": {"anchor": 3}},
    # paired-end, fixed length, 2% of the reads with 1-3 Ns: the N rule,
    # excluded ambiguous reads and 1-bit qualities on a paired-end input
    "pe_ambig_q1": {"profile": "repeat-pe", "seed": 108, "genome": 60000, "sim_args": ["--nfrac", "0.02"],
                    "args": ["-c", "30", "--pairmode", "PE", "-q", "--qualityScoreBits", "1", "--excludeAmbiguous",
                             "-k", "20", "-h", "32", "--candidateCorrection"]},
    # variable length without excluding ambiguous reads, 1-bit qualities
    "se_mixed_q1": {"profile": "mixed-se", "seed": 105, "genome": 80000,
                    "args": ["-c", "50", "--pairmode", "SE", "-q", "--qualityScoreBits", "1", "-k", "18", "-h", "24",
                             "--candidateCorrection"]},
}


def case_args(case, directory):
    """Command-line options of a case; writes its forest files into directory."""
    import os
    from tests.forest_files import ANCHOR_DESC, CANDS_DESC, write_forest
    args = list(case["args"])
    forest = case.get("forest", {})
    if "anchor" in forest:
        path = os.path.join(directory, "anchor.rf")
        write_forest(path, ANCHOR_DESC, forest["anchor"])
        args += ["--ml-forestfile", path]
    if "cands" in forest:
        path = os.path.join(directory, "cands.rf")
        write_forest(path, CANDS_DESC, forest["cands"])
        args += ["--ml-cands-forestfile", path]
    return args
