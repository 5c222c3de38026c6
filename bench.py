"""Benchmark: reads corrected per second on one MI355X.

Workload (default): the paired-end repeat genome of the reference's
headline configuration -- 5 Mbp uniform genome with 10 repeat families,
2x150 bp pairs at 30x (~1M reads), substitution rate rising 0.1%->2% along
the read, phred qualities; options -q --qualityScoreBits 2 -k 20 -h 32
-c 30 --pairmode PE.  Reads are simulated (seeded) and packed once; the
read store and the minhash tables are resident in HBM before timing.

One step = correcting every read of the set once (the reference's STEP2):
minhash query, alignment, filters, MSA, refinement and anchor correction on
the device, results copied back to host memory.  Two paths are timed on the
same resident read set: the plain path (bit-identical to care-cpu) and the
MLP-gated path (top candidates by map hits, no refinement, a learned gate
on every non-HQ decision column; weights care_amd/gate_repeat_pe.bin,
trained on other seeds of the same generator).  `value` is the gated
path's reads/s; both paths' reads/s and their base-level TP/FP/FN on the
read set's truth (care-gpu command line output) are in `config`, with the
ablation of the gate: the same reduced pipeline deciding by the classic
thresholds, by a random forest and by accepting every consensus.

Prints one JSON line (see the fields below).  --gpus above 1 is refused: this
build runs on one device.
"""

import argparse
import json

import numpy as np
import os
import re
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOADS = {
    "repeat-pe": {
        "profile": "repeat-pe", "seed": 20240601, "genome": 5000000,
        "args": ["-c", "30", "--pairmode", "PE", "-q", "--qualityScoreBits", "2", "-k", "20", "-h", "32"],
        "options": dict(coverage=30.0, paired_end=1, use_quality=1, kmer_length=20, num_hash_maps=32),
        "qual_bits": 2,
        "max_length": 150,
        "label": "repeat-pe: 5 Mbp genome, 10 repeat families, 2x150 bp PE at 30x, -q 2-bit, k=20, h=32",
    },
    "uniform-se": {
        "profile": "uniform-se", "seed": 12345, "genome": 1000000,
        "args": ["-c", "30", "--pairmode", "SE", "-k", "20", "-h", "16"],
        "options": dict(coverage=30.0, paired_end=0, use_quality=0, kmer_length=20, num_hash_maps=16),
        "qual_bits": 0,
        "max_length": 100,
        "label": "uniform-se: 1 Mbp genome, 100 bp SE at 30x (300k reads), k=20, h=16",
    },
    "large-pe": {
        "profile": "large-pe", "seed": 20240605, "genome": 100000000,
        "args": ["-c", "40", "--pairmode", "PE", "-q", "-k", "20", "-h", "48", "--candidateCorrection"],
        "options": dict(coverage=40.0, paired_end=1, use_quality=1, kmer_length=20, num_hash_maps=48,
                        correct_candidates=1),
        "qual_bits": 8,
        "max_length": 150,
        "truth": False,  # 8.5 GB of reads; base-level quality is measured on the smaller workloads
        "label": "large-pe: 100 Mbp genome, 5% repeats at 2% divergence, 2x150 bp PE at 40x (26.7M reads), "
                 "-q 8-bit, k=20, h=48, --candidateCorrection",
    },
}

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE (copy-kernel calibrated), kept
# with the commands that produced them in profiles/
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_k_correct_k_query_repeat-pe.txt")


SQ_FILE = os.path.join(ROOT, "profiles", "sq_counters_k_correct_k_query_repeat-pe.txt")
# VALU issue peak: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz, 32-bit lane operations per second
VALU_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12


def measured_valu_per_anchor(kernel, workload):
    """VALU instructions per anchor of `kernel` from the committed counter summary, or None."""
    try:
        with open(SQ_FILE) as f:
            for line in f:
                part = line.split()
                if len(part) == 4 and part[0] == "valu_per_anchor" and part[1] == kernel and part[2] == workload:
                    return float(part[3])
    except OSError:
        pass
    return None


def measured_traffic(kernel, workload):
    """Bytes per launch of `kernel` on `workload` from the committed counter summary, or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                part = line.split()
                if len(part) == 4 and part[0] == "traffic" and part[1] == kernel and part[2] == workload:
                    return int(part[3])
    except OSError:
        pass
    return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def prepare(workload, cache, genome=None):
    import care_amd
    w = dict(WORKLOADS[workload])
    if genome:
        w["genome"] = genome
    os.makedirs(cache, exist_ok=True)
    reads = os.path.join(cache, "%s_%d_%d.fastq" % (w["profile"], w["seed"], w["genome"]))
    truth = reads + ".truth" if w.get("truth", True) else os.devnull
    if not os.path.exists(reads):
        log("bench: simulating %s" % w["label"])
        t0 = time.perf_counter()
        care_amd.simulate(w["profile"], w["seed"], reads + ".tmp", truth, genome=w["genome"])
        os.replace(reads + ".tmp", reads)
        log("bench: simulated in %.1f s" % (time.perf_counter() - t0))
    return reads, truth


def cpu_baseline(reads, w, sample, threads):
    """The CPU restatement of care-cpu (oracle/care_oracle.cpp), timed on a
    bounded sample of anchors (--fixedNumberOfReads) of the same read set."""
    exe = os.path.join(ROOT, "oracle", "_build", "care-oracle")
    if not os.path.exists(exe):
        return None
    outdir = os.path.join(os.path.dirname(reads), "oracle_out")
    cmd = [exe, "-i", reads, "-o", "oracle.fastq", "-d", outdir, "-t", str(threads),
           "--fixedNumberOfReads", str(sample)] + w["args"]
    os.makedirs(outdir, exist_ok=True)
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    m = re.search(r"step2 ([0-9.]+) s step3 [0-9.]+ s -> ([0-9.]+) reads/s", p.stdout)
    if p.returncode != 0 or not m:
        log("bench: cpu baseline failed: %s" % p.stdout[-500:])
        return None
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    cpu_baseline.output = os.path.join(outdir, "oracle.fastq")
    return {"value": float(m.group(2)), "unit": "reads/s", "cores": threads, "kind": "port",
            "sample": "first %d anchors of the same read set (tables built over all reads, not timed), "
                      "query+correction %.2f s on %d threads, %s" % (sample, float(m.group(1)), threads, model)}


def time_path(care_amd, rs, opt, steps, warmup, gate=None, batch=0, autotune=True, graph=True):
    """Average wall time of one correction step over all reads (reads and
    tables resident in HBM), the correction kernel's HIP-event time and
    launch count, and the candidates of the last step."""
    eng = care_amd.Engine(opt, device=0)
    if batch:
        eng.set_launch_config(batch, 0)
    eng.upload_readset(rs)
    tb = time.perf_counter()
    eng.build()
    build_ms = (time.perf_counter() - tb) * 1e3
    log("bench: tables built in %.1f ms (%d maps)" % (build_ms, opt.num_hash_maps))
    if gate is not None:
        eng.mlp_load(gate)
    n = rs.n
    # launch shapes timed on the first anchors (set-up, like the table build)
    tuning = eng.autotune(0, min(n, 65536)) if autotune else eng.tuning()
    log("bench: %s launch shapes %s" % ("gated" if gate is not None else "plain", json.dumps(tuning)))
    # result arrays allocated once and reused by every step, as a caller
    # correcting batch after batch would
    pitch = (rs.max_length + 15) // 16 * 16
    # with candidate correction the reads go through in chunks of 2^18
    # anchors (as the command line does), each with room for 64 candidate
    # results per anchor; otherwise in one call
    chunk = 1 << 18 if opt.correct_candidates else n
    cap = chunk * 64 if opt.correct_candidates else 0
    # result arrays in page-locked host memory (copied by DMA)
    out = (care_amd.pinned_empty((chunk,), np.uint8), care_amd.pinned_empty((chunk, pitch), np.uint8),
           care_amd.pinned_empty((max(cap, 1),), np.uint32), care_amd.pinned_empty((max(cap, 1), pitch), np.uint8))

    class Totals:
        pass

    def step():
        t = Totals()
        t.query_ms = t.correct_ms = 0.0
        t.launches = t.candidates = t.raw_hits = t.corrected = 0
        for first in range(0, n, chunk):
            nb = min(chunk, n - first)
            status, _, _, _ = eng.correct(first, nb, candidate_capacity=cap, out=out)
            s = eng.stats()
            t.query_ms += s.query_ms
            t.correct_ms += s.correct_ms
            t.launches += s.launches
            t.candidates += s.candidates
            t.raw_hits += s.raw_hits
            t.corrected += int((status[:nb] & 1).sum())
        return t

    for i in range(warmup):
        step()
    step_ms, corr_ms, query_ms = [], [], []
    st = None
    for i in range(steps):
        ts = time.perf_counter()
        st = step()
        te = time.perf_counter()
        step_ms.append((te - ts) * 1e3)
        corr_ms.append(st.correct_ms)
        query_ms.append(st.query_ms)
        log("bench: %s step %d %.1f ms (query %.1f ms, correct kernel %.1f ms)" % (
            "gated" if gate is not None else "plain", i + 1, step_ms[-1], st.query_ms, st.correct_ms))
    # the same step as a HIP graph (care_correct_batch_graph): captured by the
    # first call, replayed by the timed ones; results checked against the
    # last eager step
    graph_ms = None
    if graph and chunk == n:
        ref_status = out[0][:n].copy()
        eng.correct(0, n, candidate_capacity=cap, out=out, graph=True)
        gms = []
        for i in range(steps):
            ts = time.perf_counter()
            eng.correct(0, n, candidate_capacity=cap, out=out, graph=True)
            gms.append((time.perf_counter() - ts) * 1e3)
        if not np.array_equal(out[0][:n], ref_status):
            raise RuntimeError("graph replay differs from the eager step")
        graph_ms = sum(gms) / len(gms)
        log("bench: %s graph step %.1f ms" % ("gated" if gate is not None else "plain", graph_ms))
    eng.close()
    return {"ms": sum(step_ms) / len(step_ms), "correct_ms": sum(corr_ms) / len(corr_ms),
            "query_ms": sum(query_ms) / len(query_ms), "launches": max(1, st.launches), "candidates": int(st.candidates),
            "raw_hits": int(st.raw_hits), "h": int(opt.num_hash_maps), "tuning": tuning, "build_ms": build_ms,
            "corrected": int(st.corrected), "graph_ms": graph_ms}


def step_times(stdout):
    """STEP1 / STEP2 wall seconds printed by care-gpu."""
    t = {}
    for k in ("STEP1", "STEP2"):
        m = re.search(r"# elapsed time \(%s\): ([0-9.]+)s" % k, stdout)
        if m:
            t[k.lower() + "_s"] = float(m.group(1))
    return t


def quality(care_amd, reads, truth, w, extra, outdir):
    """Base-level TP/FP/FN of the care-gpu command line's output, with its
    STEP1 / STEP2 wall times (STEP2: the correction of every read including
    the host's edit lists)."""
    os.makedirs(outdir, exist_ok=True)
    cmd = [care_amd.CARE_GPU, "-i", reads, "-o", "c.fastq", "-d", outdir, "-g", "0", "-t", "16"] + w["args"] + extra
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError(p.stdout[-2000:])
    r = care_amd.evaluate(reads, os.path.join(outdir, "c.fastq"), truth)
    out = {k: r[k] for k in ("TP", "FP", "FN")}
    out.update(step_times(p.stdout))
    return out


def fastq_records(path):
    with open(path, "rb") as f:
        while True:
            h = f.readline()
            if not h:
                return
            s = f.readline()
            f.readline()
            f.readline()
            yield s


def parity(care_amd, reads, w, sample, oracle_out, outdir):
    """Reads whose corrected sequence differs between care-gpu (plain path)
    and the oracle's output on the same first `sample` anchors."""
    os.makedirs(outdir, exist_ok=True)
    cmd = [care_amd.CARE_GPU, "-i", reads, "-o", "p.fastq", "-d", outdir, "-g", "0", "-t", "16",
           "--fixedNumberOfReads", str(sample)] + w["args"]
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError(p.stdout[-2000:])
    diff = total = 0
    for a, b in zip(fastq_records(os.path.join(outdir, "p.fastq")), fastq_records(oracle_out)):
        total += 1
        diff += a != b
    return {"anchors": sample, "reads_compared": total, "differing_reads": diff, "against": "oracle (care-oracle)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="repeat-pe", choices=sorted(WORKLOADS))
    ap.add_argument("--cache", default=os.path.join("/tmp", "care_amd_bench"))
    ap.add_argument("--cpu-sample", type=int, default=400000)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-quality", action="store_true")
    ap.add_argument("--batch", type=int, default=0)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="skip the HIP-graph replays of the step")
    ap.add_argument("--genome", type=int, default=0,
                    help="genome size of the workload's generator (default: the workload's own)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (requested %d)" % max(args.gpus, world),
              file=sys.stderr)
        sys.exit(2)

    import care_amd
    from care_amd import gate as gate_mod
    w = dict(WORKLOADS[args.workload])
    if args.genome:
        w["genome"] = args.genome
        w["label"] = re.sub(r" \([0-9.]+[kM] reads\)", "",
                            re.sub(r"\d+ Mbp genome", "%g Mbp genome" % (args.genome / 1e6), w["label"]))
    reads, truth = prepare(args.workload, args.cache, args.genome)
    t0 = time.perf_counter()
    rs = care_amd.ReadSet([reads], paired_end=bool(w["options"]["paired_end"]), qual_bits=w["qual_bits"])
    n = rs.n
    log("bench: %d reads packed in %.2f s" % (n, time.perf_counter() - t0))

    plain = time_path(care_amd, rs, care_amd.default_options(**w["options"]), args.steps, args.warmup,
                      batch=args.batch, autotune=not args.no_autotune, graph=not args.no_graph)
    gated = None
    gate_path = gate_mod.DEFAULT_GATE
    gated_ok = args.workload == "repeat-pe" and os.path.exists(gate_path)
    if gated_ok:
        threshold = gate_mod.chosen_threshold(gate_path)
        gopt = care_amd.default_options(**dict(w["options"], correction_type=3,
                                                candidate_limit=gate_mod.CANDIDATE_LIMIT,
                                                threshold_anchor=threshold))
        gated = time_path(care_amd, rs, gopt, args.steps, args.warmup, gate=gate_mod.load(gate_path),
                          batch=args.batch, autotune=not args.no_autotune, graph=not args.no_graph)
    gated_large = {}
    if args.workload == "large-pe" and os.path.exists(gate_path):
        # the gate on anchors, and on anchors and candidate corrections
        threshold = gate_mod.chosen_threshold(gate_path)
        for name, extra in (("mlp_gated_anchors", {}),
                            ("mlp_gated_anchors_and_candidates", dict(correction_type_cands=3, threshold_cands=0.99))):
            gopt = care_amd.default_options(**dict(w["options"], correction_type=3,
                                                    candidate_limit=gate_mod.CANDIDATE_LIMIT,
                                                    threshold_anchor=threshold, **extra))
            gated_large[name] = time_path(care_amd, rs, gopt, args.steps, args.warmup, gate=gate_mod.load(gate_path),
                                          batch=args.batch, autotune=not args.no_autotune)
    rs.close()

    qual = None
    if not args.no_quality and w.get("truth", True) and os.path.exists(truth):
        work = os.path.join(os.path.dirname(reads), "quality")
        qual = {"plain": quality(care_amd, reads, truth, w, [], os.path.join(work, "plain"))}
        if gated_ok:
            qual["gated"] = quality(care_amd, reads, truth, w, gate_mod._gated_args(gate_path, threshold),
                                    os.path.join(work, "gated"))
            # the method without the gate on the same reduced pipeline: the
            # classic decision, the forest, and accepting every consensus
            for name, extra in (("reduced_classic", gate_mod.REDUCED_ARGS),
                                ("reduced_forest", gate_mod._forest_args(gate_mod.DEFAULT_FOREST,
                                                                         gate_mod.chosen_threshold(gate_mod.DEFAULT_FOREST))),
                                ("reduced_accept_all", gate_mod._gated_args(gate_path, 1e-6)),
                                ("reduced_mlp_plain_columns",
                                 gate_mod._gated_args(gate_mod.PLAIN_COLUMNS_GATE,
                                                      gate_mod.chosen_threshold(gate_mod.PLAIN_COLUMNS_GATE)))):
                if name == "reduced_mlp_plain_columns" and not os.path.exists(gate_mod.PLAIN_COLUMNS_GATE):
                    continue
                if name == "reduced_forest" and not os.path.exists(gate_mod.DEFAULT_FOREST):
                    continue
                qual[name] = quality(care_amd, reads, truth, w, extra, os.path.join(work, name))
        log("bench: quality %s" % json.dumps(qual))

    # the gated path is the headline only where it meets the target's
    # quality rule: false positives no higher, true positives within 1 %
    rule = None
    if gated is not None and qual is not None and "gated" in qual:
        gq, pq = qual["gated"], qual["plain"]
        rule = {"fp_not_higher": gq["FP"] <= pq["FP"], "tp_within_1pct": gq["TP"] >= 0.99 * pq["TP"],
                "tp_change_pct": round(100.0 * (gq["TP"] - pq["TP"]) / max(1, pq["TP"]), 3),
                "fp_change_pct": round(100.0 * (gq["FP"] - pq["FP"]) / max(1, pq["FP"]), 3)}
        rule["met"] = rule["fp_not_higher"] and rule["tp_within_1pct"]
    head = gated if gated is not None and (rule is None or rule["met"]) else plain
    value = n / (head["ms"] / 1e3)

    # algorithmic bytes of the fused correction kernel per step: anchor and
    # candidate 2-bit rows, candidate lengths, candidate qualities, the
    # 16-byte per-candidate record written once, the corrected anchor out
    pitch_bytes = (w["max_length"] + 15) // 16 * 4
    qbytes = ((w["qual_bits"] * w["max_length"] + 31) // 32) * 4 if w["qual_bits"] else 0

    def roofline(r):
        alg = n * (pitch_bytes + w["max_length"] + 1) + r["candidates"] * (pitch_bytes + 4 + 16 + qbytes)
        launch_ms = r["correct_ms"] / r["launches"]
        achieved = (alg / r["launches"]) / (launch_ms / 1e3) / 1e9
        kernel = "k_correct<%s,false>" % ("true" if r is gated else "false")  # noqa
        return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic(kernel, args.workload),
                "kernel": kernel, "launches_per_step": r["launches"],
                "bytes_per_launch": int(alg / r["launches"]), "avg_launch_ms": round(launch_ms, 4)}

    def roofline_valu(r):
        # issued VALU lane operations per second against the VALU issue peak:
        # the correction kernels are VALU-issue bound (profiles/sq_counters_*)
        kernel = "k_correct<%s,false>" % ("true" if r is gated else "false")  # noqa
        per = measured_valu_per_anchor(kernel, args.workload)
        if per is None:
            return None
        launch_ms = r["correct_ms"] / r["launches"]
        achieved = per * 64 * (n / r["launches"]) / (launch_ms / 1e3) / 1e12
        return {"bound": "valu", "achieved": round(achieved, 2), "peak": round(VALU_PEAK_TOPS, 2), "unit": "Tops/s",
                "frac": round(achieved / VALU_PEAK_TOPS, 4), "kernel": kernel,
                "valu_instructions_per_anchor": per, "source": "profiles/sq_counters_k_correct_k_query_repeat-pe.txt"}

    def roofline_query(r):
        # per anchor: h stored signatures, h probes of (key, offset/count),
        # the hit values, candidate ids and the count written
        alg = n * (r["h"] * (8 + 8 + 8) + 4) + 4 * r["raw_hits"] + 4 * r["candidates"]
        launch_ms = r["query_ms"] / r["launches"]
        achieved = (alg / r["launches"]) / (launch_ms / 1e3) / 1e9
        return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic("k_query", args.workload),
                "kernel": "k_query", "raw_hits_per_step": r["raw_hits"],
                "bytes_per_launch": int(alg / r["launches"]), "avg_launch_ms": round(launch_ms, 4)}

    cpu = None
    par = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        sample = min(args.cpu_sample, n)
        cpu = cpu_baseline(reads, w, sample, threads)
        if cpu is not None and w.get("parity", True):
            par = parity(care_amd, reads, w, sample, cpu_baseline.output,
                         os.path.join(os.path.dirname(reads), "parity"))
            log("bench: parity %s" % json.dumps(par))

    def summary(r):
        return {"reads_per_s": round(n / (r["ms"] / 1e3), 1), "ms_per_step": round(r["ms"], 3),
                "query_kernel_ms": round(r["query_ms"], 3), "correct_kernel_ms": round(r["correct_ms"], 3),
                "candidates_per_step": r["candidates"], "corrected_anchors": r["corrected"], "launch_shapes": r["tuning"],
                "table_build_ms": round(r["build_ms"], 1),
                "graph_ms_per_step": None if r["graph_ms"] is None else round(r["graph_ms"], 3)}

    config = {"workload": w["label"], "reads": n, "plain": summary(plain)}
    if gated is not None:
        config["mlp_gated"] = dict(summary(gated), threshold=threshold, candidate_limit=gate_mod.CANDIDATE_LIMIT,
                                   method="top-%d candidates by map hits, no refinement, MLP gate on non-HQ "
                                          "decision columns" % gate_mod.CANDIDATE_LIMIT)
        config["speedup_gated_vs_plain"] = round(plain["ms"] / gated["ms"], 3)
    for name, r in gated_large.items():
        config[name] = summary(r)
    if qual is not None:
        config["quality"] = qual
    if rule is not None:
        config["gated_quality_rule"] = rule
    config["headline_path"] = "mlp_gated" if head is gated else "plain"
    if par is not None:
        config["parity"] = par
    line = {
        "metric": "reads corrected/s on 1 MI355X: plain vs MLP-gated at equal base-level TP/FP"
                  if gated is not None else "reads corrected/s on 1 MI355X (plain classic path)",
        "value": round(value, 1),
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "i32+f32+bf16" if gated is not None else "i32+f32",
        "data": "synthetic",
        "config": config,
        "roofline": roofline(head),
        "cpu_baseline": cpu,
    }
    line["roofline_valu"] = roofline_valu(head)
    if gated is not None:
        line["roofline_plain"] = roofline(plain)
        line["roofline_valu_plain"] = roofline_valu(plain)
    line["roofline_query"] = roofline_query(plain)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
