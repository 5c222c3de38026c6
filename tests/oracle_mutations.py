"""Checks that the oracle's pins catch one-token mistakes.

Each mutation below changes one token of oracle/care_oracle.cpp (a constant,
a comparison, an index, a sign; one changes two tokens that enforce the same
rule twice), builds that variant of the oracle library
into a temporary directory, and runs the oracle's pin tests
(tests/test_oracle_pins.py and the closed-form tests of tests/test_oracle.py)
against it through CARE_ORACLE_LIB.  Every mutation must make at least one
test fail; the script reports each and exits non-zero if any survives.

    python tests/oracle_mutations.py            all mutations
    python tests/oracle_mutations.py -k murmur  those whose name contains it
"""
import os
import shutil
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "care_oracle.cpp")

# (name, text, replacement, occurrence index when the text is not unique)
MUTATIONS = [
    ("murmur multiplier", "x *= 0xff51afd7ed558ccdull;", "x *= 0xff51afd7ed558ccfull;", 0),
    ("murmur shift", "x ^= x >> 33;", "x ^= x >> 31;", 1),
    ("overlap weight without sqrt", "return 1.0f - std::sqrt(float(nOps)", "return 1.0f - (float(nOps)", 0),
    ("consensus tie to the later base", "if (wC[col] > cw) { cons = 'C'", "if (wC[col] >= cw) { cons = 'C'", 0),
    ("support over another weight", "support[col] = cw / colw;", "support[col] = wA[col] / colw;", 0),
    ("column frame sign", "anchorBegin = std::max(-startindex, 0);", "anchorBegin = std::max(startindex, 0);", 0),
    ("orig coverage of C", "case 'C': origWeights[col] = wC[col]; origCoverages[col] = cC[col];",
     "case 'C': origWeights[col] = wC[col]; origCoverages[col] = cA[col];", 0),
    ("HQ min-support factor", "fgeq(p.min_support, std::fma(-3.0f, factor, 1.0f))",
     "fgeq(p.min_support, std::fma(-2.0f, factor, 1.0f))", 0),
    ("HQ half percent", "fleq(float(smallest), pct * 0.5f)", "fleq(float(smallest), pct * 1.0f)", 0),
    ("fgeq epsilon", "return d < 1e-5f;", "return d < 1e-4f;", 0),
    ("min coverage threshold", "const float covT = mcov / 6.0f * cov;", "const float covT = mcov / 5.0f * cov;", 0),
    ("anchor support bound", "(support[c] > 0.90f && origCoverages[c] <= 2)",
     "(support[c] > 0.95f && origCoverages[c] <= 2)", 0),
    ("anchor orig coverage bound", "(support[c] > 0.90f && origCoverages[c] <= 2)",
     "(support[c] > 0.90f && origCoverages[c] < 2)", 0),
    ("significant count factor", "const int sig = int(datasetCoverage * 0.3f);",
     "const int sig = int(datasetCoverage * 0.4f);", 0),
    ("significant count comparison", "if (i != ci && sig <= counts[i]) s = i;",
     "if (i != ci && sig < counts[i]) s = i;", 0),
    ("keep matching rows", "const bool keepMatching = orig == found;", "const bool keepMatching = orig != found;", 0),
    ("very good alignment veto", "if (remove[c] && in.weights[c] >= 0.90f)", "if (remove[c] && in.weights[c] > 0.90f)",
     0),
    ("SE filter middle level", "if (ratios[i] < 3 * base) counts[1]++;", "if (ratios[i] < 4 * base) counts[1]++;", 0),
    ("SE filter enough count", "const float goodT = estimatedCoverage * mCoverage;",
     "const float goodT = estimatedCoverage / mCoverage;", 0),
    ("SE filter first level", "if (counts[0] >= goodT) thr = 2 * base;", "if (counts[0] > goodT) thr = 2 * base;", 0),
    ("candidate window start", "return ab - ncc <= qb && qb <= ab + ncc", "return ab - ncc < qb && qb <= ab + ncc", 0),
    ("forest candidate window end", "if (forestOrGate) return qb >= ab - ncc && qe <= ae + ncc;",
     "if (forestOrGate) return qb >= ab - ncc && qe < ae + ncc;", 0),
    ("edit limit", "const int maxEdits = len / 7;", "const int maxEdits = len / 6;", 0),
    ("forest split comparison", "if (double(f[n.att]) <= n.split) {", "if (double(f[n.att]) < n.split) {", 0),
    ("forest left leaf flag", "if (n.flag & 2) return n.lhs.prob;", "if (n.flag & 1) return n.lhs.prob;", 0),
    ("combine minimum results", "if (res.size() >= 3) {", "if (res.size() > 3) {", 0),
    ("combine HQ anchor", "    if (anchor->hq) {\n        apply(anchor);", "    if (!anchor->hq) {\n        apply(anchor);", 0),
    ("orientation tie", "if (ratio < rratio) return ratio < maxErr ? 1 : 3;",
     "if (ratio <= rratio) return ratio < maxErr ? 1 : 3;", 0),
    # ties keep the first shift twice over (the pruning bound m < best -
    # total + 2 ov and score < best): this one changes both tokens
    ("alignment first minimum", ["best - total + 2 * ov);", "            if (score < best) {"],
     ["best - total + 2 * ov + 1);", "            if (score <= best) {"], 0),
    ("alignment mismatch limit", "std::min(int(float(ov) * maxErrorRate),", "std::min(int(float(ov) * maxErrorRate) + 1,", 0),
    ("alignment shift order", "for (int s = -1; s >= -lc + minov; s--)", "for (int s = -1; s > -lc + minov; s--)", 0),
    ("quality bin comparison", "if (Q <= bound) {", "if (Q < bound) {", 0),
]

TESTS = ["tests/test_oracle_pins.py", "tests/test_oracle.py", "-k", "not equals_reference"]


def mutate(text, old, new, nth):
    if isinstance(old, list):
        for o, n in zip(old, new):
            text = mutate(text, o, n, nth)
        return text
    pos = -1
    for _ in range(nth + 1):
        pos = text.find(old, pos + 1)
        if pos < 0:
            raise ValueError("mutation text not found: %r" % old)
    return text[:pos] + new + text[pos + len(old):]


def main(argv):
    sel = argv[argv.index("-k") + 1] if "-k" in argv else ""
    with open(SRC) as f:
        base = f.read()
    survivors = []
    tmp = tempfile.mkdtemp(prefix="oracle_mut_")
    try:
        for i, (name, old, new, nth) in enumerate(MUTATIONS):
            if sel and sel not in name:
                continue
            d = os.path.join(tmp, str(i))
            os.makedirs(d)
            shutil.copy(os.path.join(ROOT, "oracle", "care_oracle.h"), d)
            src = os.path.join(d, "care_oracle.cpp")
            with open(src, "w") as f:
                f.write(mutate(base, old, new, nth))
            lib = os.path.join(d, "liboracle.so")
            subprocess.run(["g++", "-O1", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared", "-w",
                            "-o", lib, src, "-lz"], check=True)
            env = dict(os.environ, CARE_ORACLE_LIB=lib)
            p = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider"] + TESTS, cwd=ROOT,
                               env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
            caught = p.returncode != 0
            failed = [l.split("::")[-1].split(" ")[0] for l in p.stdout.splitlines() if l.startswith("FAILED")]
            print("%-32s %s %s" % (name, "caught" if caught else "SURVIVED", ",".join(failed)), flush=True)
            if not caught:
                survivors.append(name)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    print("%d mutations, %d survived%s" % (len([m for m in MUTATIONS if sel in m[0]]), len(survivors),
                                           (": " + ", ".join(survivors)) if survivors else ""))
    return 1 if survivors else 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
