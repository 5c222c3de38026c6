"""Training data, training and evaluation of the learned correction gate.

The gate (correction_type 3, see include/care_amd.h) decides, for every
anchor column whose MSA consensus differs from the anchor base in a non-HQ
MSA, whether the correction is applied.  Its training columns come from the
gated pipeline itself run on simulated reads with known truth: a column is
labelled 1 when the consensus base equals the true base.  Training runs on
the device (care_mlp_train: BF16 MFMA forward/backward, Adam).

    python -m care_amd.gate train --seeds 8 --out care_amd/gate_repeat_pe.bin
    python -m care_amd.gate evaluate --gate care_amd/gate_repeat_pe.bin
    python -m care_amd.gate train-forest --out care_amd/forest_repeat_pe.rf
    python -m care_amd.gate ablation
"""
import argparse
import re
import json
import os
import subprocess
import sys
import tempfile
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

import care_amd

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_GATE = os.path.join(HERE, "gate_repeat_pe.bin")
# the same network trained on the plain method's decision columns (every
# candidate, refinement) instead of the reduced pipeline's
PLAIN_COLUMNS_GATE = os.path.join(HERE, "gate_plain_columns.bin")

# the paired-end repeat workload's generator and options
PROFILE = "repeat-pe"
GENOME = 5_000_000
OPTIONS = dict(coverage=30.0, paired_end=1, use_quality=1, kmer_length=20, num_hash_maps=32)
QUAL_BITS = 2
CLI_ARGS = ["-c", "30", "--pairmode", "PE", "-q", "--qualityScoreBits", "2", "-k", "20", "-h", "32"]
CANDIDATE_LIMIT = 20
TRAIN_SEED_BASE = 7_000_000
VALIDATION_SEED = 6_999_999  # threshold selection
EVAL_SEED = 20240601  # the benchmark's read set, never used for training or selection
SWEEP = [0.95, 0.96, 0.97, 0.975, 0.98, 0.985, 0.99, 0.995]


def he_init(seed):
    """He-uniform initial gate parameters (zero biases), FP32."""
    rng = np.random.default_rng(seed)
    parts = []
    for n_out, n_in in ((64, 32), (32, 64), (16, 32)):
        lim = np.sqrt(6.0 / n_in)
        parts.append(rng.uniform(-lim, lim, size=n_out * n_in))
        parts.append(np.zeros(n_out))
    parts.append(rng.uniform(-np.sqrt(6.0 / 16), np.sqrt(6.0 / 16), size=16))
    parts.append(np.zeros(1))
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == care_amd.MLP_PARAMS
    return p


def save(params, path):
    np.asarray(params, dtype="<f4").tofile(path)


def load(path):
    p = np.fromfile(path, dtype="<f4")
    if p.size != care_amd.MLP_PARAMS:
        raise ValueError("%s: expected %d float32 values, found %d" % (path, care_amd.MLP_PARAMS, p.size))
    return p


def truth_codes(path, n, length):
    """Base codes (A0 C1 G2 T3, 4 for anything else) of the error-free reads
    of a simulator truth FASTQ, [n][length]."""
    lut = np.full(256, 4, dtype=np.uint8)
    for i, c in enumerate(b"ACGT"):
        lut[c] = i
    out = np.full((n, length), 4, dtype=np.uint8)
    with open(path, "rb") as f:
        for i, line in enumerate(f):
            if i % 4 == 1:
                r = i // 4
                s = np.frombuffer(line.rstrip(b"\n"), dtype=np.uint8)
                out[r, :len(s)] = lut[s]
    return out


def options(**kw):
    o = dict(OPTIONS)
    o.update(correction_type=3, candidate_limit=CANDIDATE_LIMIT)
    o.update(kw)
    return care_amd.default_options(**o)


def collect(reads, truth, device=0, plain=False):
    """Gate decision columns of one read set and their labels: over the
    reduced pipeline's MSAs (the ones the gate decides on), or with plain
    over the plain method's (every candidate, refinement)."""
    rs = care_amd.ReadSet([reads], paired_end=True, qual_bits=QUAL_BITS)
    with care_amd.Engine(options(), device=device) as eng:
        eng.upload_readset(rs)
        eng.build()
        f, ids, pos, orig, cons = eng.gate_columns(0, rs.n, plain=plain)
    t = truth_codes(truth, rs.n, rs.max_length)
    tb = t[ids, pos]
    keep = tb < 4
    labels = (tb == cons).astype(np.uint8)
    rs.close()
    return f[keep], labels[keep]


def simulate_many(seeds, directory, workers=8):
    def one(seed):
        reads = os.path.join(directory, "r%d.fastq" % seed)
        truth = reads + ".truth"
        if not os.path.exists(reads):
            care_amd.simulate(PROFILE, seed, reads, truth, genome=GENOME)
        return reads, truth
    with ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(one, seeds))


def train(args):
    work = args.work or tempfile.mkdtemp(prefix="care_gate_")
    os.makedirs(work, exist_ok=True)
    seeds = [TRAIN_SEED_BASE + i for i in range(args.seeds)]
    xs, ys = [], []
    for w0 in range(0, len(seeds), args.workers):
        t0 = time.perf_counter()
        sets = simulate_many(seeds[w0:w0 + args.workers], work, workers=args.workers)
        print("gate: simulated %d read sets in %.1f s" % (len(sets), time.perf_counter() - t0), flush=True)
        for reads, truth in sets:
            x, y = collect(reads, truth, plain=args.columns == "plain")
            xs.append(x)
            ys.append(y)
            os.remove(reads)
            os.remove(truth)
            print("gate: %s: %d columns, %.3f positive" % (os.path.basename(reads), len(y), y.mean()), flush=True)
    x = np.concatenate(xs)
    y = np.concatenate(ys)
    print("gate: %d training columns, %.3f positive" % (len(y), y.mean()), flush=True)
    p0 = he_init(args.init_seed)
    with care_amd.Engine(options(), device=0) as eng:
        t1 = time.perf_counter()
        p, losses = eng.mlp_train(x, y, p0, epochs=args.epochs, batch=args.batch, lr=args.lr, seed=args.init_seed)
        t2 = time.perf_counter()
        z = eng.mlp_infer(x, p)
    acc = float(np.mean((z >= 0) == (y == 1)))
    steps = args.epochs * ((len(y) + args.batch - 1) // args.batch)
    flops = 3 * 9248.0 * len(y) * args.epochs
    info = {"columns": int(len(y)), "column_source": args.columns, "positive": float(y.mean()),
            "epochs": args.epochs, "batch": args.batch,
            "lr": args.lr, "epoch_loss": [float(v) for v in losses], "train_accuracy": acc,
            "train_seconds": t2 - t1, "adam_steps": steps, "train_tflops": flops / (t2 - t1) / 1e12}
    print("gate: " + json.dumps(info), flush=True)
    save(p, args.out)
    with open(args.out + ".json", "w") as f:
        json.dump(info, f, indent=1)
    select(args.out, work)


def select(gate_path, work):
    """Acceptance threshold on a validation read set (never the benchmark's):
    the fewest false positives among the thresholds that keep at least
    99.5% of the plain path's true positives (half of the allowed 1% loss is
    kept as margin for other read sets).  Stored beside the gate."""
    os.makedirs(work, exist_ok=True)
    (vreads, vtruth), = simulate_many([VALIDATION_SEED], work, workers=1)
    plain = care_amd.evaluate(vreads, _run_cli(vreads, os.path.join(work, "vplain"), []), vtruth)
    sweep = {}
    best = None
    for t in SWEEP:
        r = care_amd.evaluate(vreads, _run_cli(vreads, os.path.join(work, "v%g" % t), _gated_args(gate_path, t)),
                              vtruth)
        sweep[repr(t)] = {"TP": r["TP"], "FP": r["FP"], "FN": r["FN"]}
        print("gate: validation threshold %g: TP %d FP %d (plain TP %d FP %d)" % (t, r["TP"], r["FP"], plain["TP"],
                                                                                plain["FP"]), flush=True)
        if r["TP"] >= 0.995 * plain["TP"] and r["FP"] <= plain["FP"] and (best is None or r["FP"] < sweep[repr(best)]["FP"]):
            best = t
    try:
        with open(gate_path + ".json") as f:
            info = json.load(f)
    except (OSError, ValueError):
        info = {}
    info["validation"] = {"seed": VALIDATION_SEED, "plain": {"TP": plain["TP"], "FP": plain["FP"], "FN": plain["FN"]},
                          "sweep": sweep, "rule": "fewest FP with TP >= 0.995 plain TP and FP <= plain FP"}
    info["threshold"] = best if best is not None else 0.5
    print("gate: chosen threshold %g" % info["threshold"], flush=True)
    with open(gate_path + ".json", "w") as f:
        json.dump(info, f, indent=1)


def _gated_args(gate_path, threshold):
    return ["--correctionType", "3", "--candidateLimit", str(CANDIDATE_LIMIT), "--ml-gatefile",
            os.path.abspath(gate_path), "--thresholdAnchor", repr(threshold)]


def chosen_threshold(gate_path):
    """Acceptance threshold stored beside a trained gate (0.5 if none)."""
    try:
        with open(gate_path + ".json") as f:
            return float(json.load(f).get("threshold", 0.5))
    except (OSError, ValueError):
        return 0.5


def _run_cli(reads, outdir, extra):
    cmd = [care_amd.CARE_GPU, "-i", reads, "-o", "c.fastq", "-d", outdir, "-g", "0", "-t", "16"] + CLI_ARGS + extra
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError(p.stdout[-2000:])
    return os.path.join(outdir, "c.fastq")


def _step_ms(rs, opt, gate, steps=3):
    with care_amd.Engine(opt, device=0) as eng:
        eng.upload_readset(rs)
        eng.build()
        if gate is not None:
            eng.mlp_load(gate)
        eng.correct(0, rs.n)
        ts = []
        for _ in range(steps):
            t = time.perf_counter()
            eng.correct(0, rs.n)
            ts.append((time.perf_counter() - t) * 1e3)
    return float(np.mean(ts))


def evaluate(args):
    """Plain vs gated on a held-out read set: base-level TP/FP/FN of the
    corrected output (care-gpu command line) at each acceptance threshold,
    and reads/s of both device paths (reads resident, tables built)."""
    work = args.work or tempfile.mkdtemp(prefix="care_gate_eval_")
    os.makedirs(work, exist_ok=True)
    (reads, truth), = simulate_many([args.seed], work, workers=1)
    gate_path = os.path.abspath(args.gate)
    res = {"plain": care_amd.evaluate(reads, _run_cli(reads, os.path.join(work, "plain"), []), truth), "gated": {}}
    for t in (args.thresholds or [chosen_threshold(gate_path)]):
        r = care_amd.evaluate(reads, _run_cli(reads, os.path.join(work, "gated_%g" % t), _gated_args(gate_path, t)),
                              truth)
        r["tp_ratio"] = r["TP"] / max(1, res["plain"]["TP"])
        r["fp_le_plain"] = r["FP"] <= res["plain"]["FP"]
        res["gated"][repr(t)] = r
        print("gate: threshold %g: %s" % (t, json.dumps(r)), flush=True)
    rs = care_amd.ReadSet([reads], paired_end=True, qual_bits=QUAL_BITS)
    gate = load(gate_path)
    plain_ms = _step_ms(rs, care_amd.default_options(**OPTIONS), None)
    gated_ms = _step_ms(rs, options(), gate)
    res["timing"] = {"reads": rs.n, "plain_ms_per_step": plain_ms, "gated_ms_per_step": gated_ms,
                     "plain_reads_per_s": rs.n / (plain_ms / 1e3), "gated_reads_per_s": rs.n / (gated_ms / 1e3),
                     "speedup": plain_ms / gated_ms}
    rs.close()
    print(json.dumps(res), flush=True)
    return res


# ---------------------------------------------------------------------------
# The method without the gate, on the same reduced pipeline (the top
# CANDIDATE_LIMIT candidates, no refinement): the classic threshold decision
# and a random forest in the reference's .rf format over the reference's
# anchor features (extract_anchor_v3, include/classification.hpp:748-800),
# trained like the reference trains its forests (sklearn's
# RandomForestClassifier, ml/care.py:231-282) on the same decision columns
# as the gate.
# ---------------------------------------------------------------------------
DEFAULT_FOREST = os.path.join(HERE, "forest_repeat_pe.rf")
REDUCED_ARGS = ["--candidateLimit", str(CANDIDATE_LIMIT), "--refinements", "0"]
# the 13 anchor features of extract_anchor_v3 are 13 of the gate's 32
# (same operands, same float divisions): orig count / coverage, orig weight
# / weight sum, orig weight / orig count, consensus count / coverage,
# support, consensus weight / consensus count, weight sum / coverage,
# coverage / estimated coverage, weight sum / estimated coverage, average
# and minimum support, maximum and minimum coverage / estimated coverage
V3_FROM_GATE = [0, 1, 4, 2, 3, 5, 30, 6, 7, 12, 13, 15, 14]


def write_forest(clf, path, desc="13 extract_anchor_v3"):
    """A fitted sklearn forest in CARE's .rf format (include/forest.hpp:38-55):
    u64 descriptor length and descriptor, u32 tree count, per tree the u32
    internal-node count and the nodes in preorder (feature u8, split f64,
    flag u8: bit 1 left child is a leaf, bit 0 right child is a leaf), a
    leaf being the f32 fraction of positive samples."""
    import struct
    out = bytearray()
    d = desc.encode()
    out += struct.pack("<Q", len(d)) + d + struct.pack("<I", len(clf.estimators_))
    for est in clf.estimators_:
        t = est.tree_
        left, right = t.children_left, t.children_right

        def leaf(i):
            return left[i] == right[i]

        body = bytearray()
        stack = [0]
        while stack:
            i = stack.pop()
            if leaf(i):
                v = t.value[i][0]
                body += struct.pack("<f", float(v[1] / (v[0] + v[1])))
                continue
            flag = (2 if leaf(left[i]) else 0) | (1 if leaf(right[i]) else 0)
            body += struct.pack("<B", int(t.feature[i])) + struct.pack("<d", float(t.threshold[i])) + struct.pack(
                "<B", flag)
            stack.append(right[i])
            stack.append(left[i])
        internal = int(np.sum(left != right))
        out += struct.pack("<I", internal) + body
    with open(path, "wb") as f:
        f.write(bytes(out))


def train_forest(args):
    from sklearn.ensemble import RandomForestClassifier
    work = args.work or tempfile.mkdtemp(prefix="care_forest_")
    os.makedirs(work, exist_ok=True)
    seeds = [TRAIN_SEED_BASE + 1000 + i for i in range(args.seeds)]
    xs, ys = [], []
    for reads, truth in simulate_many(seeds, work, workers=min(8, len(seeds))):
        x, y = collect(reads, truth)
        xs.append(x[:, V3_FROM_GATE])
        ys.append(y)
        os.remove(reads)
        os.remove(truth)
    x = np.concatenate(xs)
    y = np.concatenate(ys)
    rng = np.random.default_rng(args.seed)
    if len(y) > args.rows:
        pick = rng.choice(len(y), size=args.rows, replace=False)
        x, y = x[pick], y[pick]
    t0 = time.perf_counter()
    clf = RandomForestClassifier(n_estimators=args.trees, max_leaf_nodes=args.leaves, n_jobs=args.jobs,
                                 random_state=args.seed).fit(x.astype(np.float32), y)
    info = {"columns": int(len(y)), "positive": float(y.mean()), "trees": args.trees, "max_leaf_nodes": args.leaves,
            "fit_seconds": time.perf_counter() - t0, "train_accuracy": float(clf.score(x, y))}
    print("forest: " + json.dumps(info), flush=True)
    write_forest(clf, args.out)
    with open(args.out + ".json", "w") as f:
        json.dump(info, f, indent=1)
    select_forest(args.out, work)


def _forest_args(path, threshold):
    return ["--correctionType", "1", "--ml-forestfile", os.path.abspath(path), "--thresholdAnchor",
            repr(threshold)] + REDUCED_ARGS


FOREST_SWEEP = [0.5, 0.7, 0.9, 0.95, 0.97, 0.98, 0.99, 0.995, 0.999]


def select_forest(path, work):
    """Forest acceptance threshold on the validation read set, by the gate's
    rule (fewest FP with TP >= 0.995 plain TP and FP <= plain FP; else the
    threshold with the most TP among those with FP <= plain FP)."""
    os.makedirs(work, exist_ok=True)
    (vreads, vtruth), = simulate_many([VALIDATION_SEED], work, workers=1)
    plain = care_amd.evaluate(vreads, _run_cli(vreads, os.path.join(work, "vplain"), []), vtruth)
    sweep, best = {}, None
    for t in FOREST_SWEEP:
        r = care_amd.evaluate(vreads, _run_cli(vreads, os.path.join(work, "f%g" % t), _forest_args(path, t)), vtruth)
        sweep[repr(t)] = {"TP": r["TP"], "FP": r["FP"], "FN": r["FN"]}
        print("forest: validation threshold %g: TP %d FP %d (plain TP %d FP %d)" % (t, r["TP"], r["FP"], plain["TP"],
                                                                                  plain["FP"]), flush=True)
    ok = [t for t in FOREST_SWEEP if sweep[repr(t)]["TP"] >= 0.995 * plain["TP"] and sweep[repr(t)]["FP"] <= plain["FP"]]
    if ok:
        best = min(ok, key=lambda t: sweep[repr(t)]["FP"])
    else:
        le = [t for t in FOREST_SWEEP if sweep[repr(t)]["FP"] <= plain["FP"]] or FOREST_SWEEP
        best = max(le, key=lambda t: sweep[repr(t)]["TP"])
    try:
        with open(path + ".json") as f:
            info = json.load(f)
    except (OSError, ValueError):
        info = {}
    info["validation"] = {"seed": VALIDATION_SEED, "plain": {k: plain[k] for k in ("TP", "FP", "FN")}, "sweep": sweep}
    info["threshold"] = best
    print("forest: chosen threshold %g" % best, flush=True)
    with open(path + ".json", "w") as f:
        json.dump(info, f, indent=1)


def ablation(reads, truth, work, gate_path=DEFAULT_GATE, forest_path=DEFAULT_FOREST):
    """What the gate gains over the method without it: the full plain path,
    and the reduced pipeline (top CANDIDATE_LIMIT candidates by map hits, no
    refinement) deciding its non-consensus anchor columns by the classic
    thresholds, by the forest, by the gate, and by accepting every
    consensus.  Base-level TP/FP/FN and STEP2 seconds of the command line."""
    rows = {"plain": [], "reduced_classic": REDUCED_ARGS}
    if os.path.exists(forest_path):
        rows["reduced_forest"] = _forest_args(forest_path, chosen_threshold(forest_path))
    if os.path.exists(gate_path):
        rows["reduced_mlp_gate"] = _gated_args(gate_path, chosen_threshold(gate_path))
    if os.path.exists(PLAIN_COLUMNS_GATE):
        rows["reduced_mlp_gate_plain_columns"] = _gated_args(PLAIN_COLUMNS_GATE, chosen_threshold(PLAIN_COLUMNS_GATE))
    if os.path.exists(gate_path):
        # a gate that accepts every column: logit threshold far below any output
        rows["reduced_accept_all"] = _gated_args(gate_path, 1e-6)
    res = {}
    for name, extra in rows.items():
        outdir = os.path.join(work, "ablation_" + name)
        cmd = [care_amd.CARE_GPU, "-i", reads, "-o", "c.fastq", "-d", outdir, "-g", "0", "-t", "16"] + CLI_ARGS + extra
        p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if p.returncode != 0:
            raise RuntimeError(p.stdout[-2000:])
        r = care_amd.evaluate(reads, os.path.join(outdir, "c.fastq"), truth)
        m = re.search(r"elapsed time \(STEP2\): ([0-9.]+)s", p.stdout)
        res[name] = {"TP": r["TP"], "FP": r["FP"], "FN": r["FN"], "step2_s": float(m.group(1)) if m else None}
    return res


LARGE_ARGS = ["-c", "40", "--pairmode", "PE", "-q", "-k", "20", "-h", "48", "--candidateCorrection"]


def evaluate_large(args):
    """The largest configuration's generator (large-pe: 5 % repeats at 2 %
    divergence, 2x150 bp PE at 40x, 8-bit qualities, h = 48, candidate
    correction) at a reduced genome size: base-level TP/FP/FN and STEP 2
    time of the command line for the plain path, the gate on anchors and the
    gate on anchors and candidate corrections (same network and threshold,
    candidate thresholds as given)."""
    work = args.work or tempfile.mkdtemp(prefix="care_gate_large_")
    os.makedirs(work, exist_ok=True)
    reads = os.path.join(work, "large.fastq")
    truth = reads + ".truth"
    if not os.path.exists(reads):
        care_amd.simulate("large-pe", args.seed, reads, truth, genome=args.genome)
    gate_path = os.path.abspath(args.gate)
    t = chosen_threshold(gate_path)

    def run(name, extra):
        outdir = os.path.join(work, name)
        cmd = [care_amd.CARE_GPU, "-i", reads, "-o", "c.fastq", "-d", outdir, "-g", "0", "-t", "16"] + LARGE_ARGS + extra
        p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if p.returncode != 0:
            raise RuntimeError(p.stdout[-2000:])
        m = re.search(r"elapsed time \(STEP2\): ([0-9.]+)s", p.stdout)
        r = care_amd.evaluate(reads, os.path.join(outdir, "c.fastq"), truth)
        r = {k: r[k] for k in ("TP", "FP", "FN")}
        r["step2_s"] = float(m.group(1)) if m else None
        print("gate large: %s %s" % (name, json.dumps(r)), flush=True)
        return r

    res = {"reads": None, "plain": run("plain", []), "gated_anchors": run("gated_anchors", _gated_args(gate_path, t)),
           "gated_both": {}}
    for tc in args.thresholds_cands:
        res["gated_both"][repr(tc)] = run("gated_both_%g" % tc, _gated_args(gate_path, t) + [
            "--correctionTypeCands", "3", "--thresholdCands", repr(tc)])
    print(json.dumps(res), flush=True)
    return res


def timing(rows, epochs=1):
    """Device training and inference throughput of the gate on synthetic
    columns (a fixed linear rule), for kernel profiling."""
    rng = np.random.default_rng(1)
    x = rng.uniform(0.0, 1.5, size=(rows, care_amd.MLP_FEATURES)).astype(np.float32)
    y = (x[:, 0] + 0.5 * x[:, 3] > 0.9).astype(np.uint8)
    with care_amd.Engine(care_amd.default_options(), device=0) as eng:
        t = time.perf_counter()
        p, losses = eng.mlp_train(x, y, he_init(1), epochs=epochs, batch=4096, lr=1e-3, seed=1)
        dt = time.perf_counter() - t
        t = time.perf_counter()
        eng.mlp_infer(x, p)
        di = time.perf_counter() - t
    res = {"rows": rows, "epochs": epochs, "train_s": dt, "train_tflops": 3 * 9248.0 * rows * epochs / dt / 1e12,
           "infer_s": di, "infer_tflops": 9248.0 * rows / di / 1e12, "loss": [float(v) for v in losses]}
    print(json.dumps(res), flush=True)
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--seeds", type=int, default=8)
    t.add_argument("--epochs", type=int, default=5)
    t.add_argument("--batch", type=int, default=4096)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--init-seed", type=int, default=1)
    t.add_argument("--workers", type=int, default=8)
    t.add_argument("--work", default=None)
    t.add_argument("--out", default=DEFAULT_GATE)
    t.add_argument("--columns", choices=["reduced", "plain"], default="reduced",
                   help="training columns from the reduced pipeline the gate runs in, or from the plain method")
    tm = sub.add_parser("timing")
    tm.add_argument("--rows", type=int, default=4_000_000)
    tm.add_argument("--epochs", type=int, default=1)
    v = sub.add_parser("select")
    v.add_argument("--gate", default=DEFAULT_GATE)
    v.add_argument("--work", default=None)
    e = sub.add_parser("evaluate")
    e.add_argument("--gate", default=DEFAULT_GATE)
    e.add_argument("--seed", type=int, default=EVAL_SEED)
    e.add_argument("--work", default=None)
    e.add_argument("--thresholds", type=lambda v: [float(x) for x in v.split(",")], default=None,
                   help="acceptance thresholds to evaluate (default: the one chosen at training)")
    el = sub.add_parser("evaluate-large")
    el.add_argument("--gate", default=DEFAULT_GATE)
    el.add_argument("--seed", type=int, default=20240605)
    el.add_argument("--genome", type=int, default=10_000_000)
    el.add_argument("--work", default=None)
    el.add_argument("--thresholds-cands", type=lambda v: [float(x) for x in v.split(",")], default=[0.5, 0.9, 0.99])
    tf = sub.add_parser("train-forest")
    tf.add_argument("--seeds", type=int, default=4)
    tf.add_argument("--rows", type=int, default=2_000_000)
    tf.add_argument("--trees", type=int, default=64)
    tf.add_argument("--leaves", type=int, default=512)
    tf.add_argument("--jobs", type=int, default=16)
    tf.add_argument("--seed", type=int, default=1)
    tf.add_argument("--work", default=None)
    tf.add_argument("--out", default=DEFAULT_FOREST)
    sf = sub.add_parser("select-forest")
    sf.add_argument("--forest", default=DEFAULT_FOREST)
    sf.add_argument("--work", default=None)
    ab = sub.add_parser("ablation")
    ab.add_argument("--seed", type=int, default=EVAL_SEED)
    ab.add_argument("--work", default=None)
    a = ap.parse_args(argv)
    if a.cmd == "train":
        train(a)
    elif a.cmd == "train-forest":
        train_forest(a)
    elif a.cmd == "select-forest":
        select_forest(a.forest, a.work or tempfile.mkdtemp(prefix="care_forest_sel_"))
    elif a.cmd == "ablation":
        work = a.work or tempfile.mkdtemp(prefix="care_ablation_")
        os.makedirs(work, exist_ok=True)
        (reads, truth), = simulate_many([a.seed], work, workers=1)
        print(json.dumps(ablation(reads, truth, work)), flush=True)
    elif a.cmd == "timing":
        timing(a.rows, a.epochs)
    elif a.cmd == "evaluate-large":
        evaluate_large(a)
    elif a.cmd == "select":
        select(a.gate, a.work or tempfile.mkdtemp(prefix="care_gate_sel_"))
    else:
        evaluate(a)


if __name__ == "__main__":
    sys.exit(main())
