"""Benchmark: reads corrected per second on one MI355X.

Workload (default): the paired-end repeat genome of the reference's
headline configuration -- 5 Mbp uniform genome with 10 repeat families,
2x150 bp pairs at 30x (~1M reads), substitution rate rising 0.1%->2% along
the read, phred qualities; options -q --qualityScoreBits 2 -k 20 -h 32
-c 30 --pairmode PE.  Reads are simulated (seeded) and packed once; the
read store and the minhash tables are resident in HBM before timing.

One step = correcting every read of the set once (the reference's STEP2):
minhash query, alignment, filters, MSA, refinement and anchor correction on
the device, results copied back to host memory.

Prints one JSON line (see the fields below).  --gpus above 1 is refused: this
build runs on one device.
"""

import argparse
import json
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
        "label": "repeat-pe: 5 Mbp genome, 10 repeat families, 2x150 bp PE at 30x, -q 2-bit, k=20, h=32",
    },
    "uniform-se": {
        "profile": "uniform-se", "seed": 12345, "genome": 1000000,
        "args": ["-c", "30", "--pairmode", "SE", "-k", "20", "-h", "16"],
        "options": dict(coverage=30.0, paired_end=0, use_quality=0, kmer_length=20, num_hash_maps=16),
        "qual_bits": 0,
        "label": "uniform-se: 1 Mbp genome, 100 bp SE at 30x (300k reads), k=20, h=16",
    },
}

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def prepare(workload, cache):
    import care_amd
    w = WORKLOADS[workload]
    os.makedirs(cache, exist_ok=True)
    reads = os.path.join(cache, "%s_%d_%d.fastq" % (w["profile"], w["seed"], w["genome"]))
    truth = reads + ".truth"
    if not os.path.exists(reads):
        log("bench: simulating %s" % w["label"])
        care_amd.simulate(w["profile"], w["seed"], reads + ".tmp", truth, genome=w["genome"])
        os.replace(reads + ".tmp", reads)
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
    return {"value": float(m.group(2)), "unit": "reads/s", "cores": threads, "kind": "port",
            "sample": "first %d anchors of the same read set (tables built over all reads, not timed), "
                      "query+correction %.2f s on %d threads, %s" % (sample, float(m.group(1)), threads, model)}


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
    ap.add_argument("--batch", type=int, default=0)
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (requested %d)" % max(args.gpus, world),
              file=sys.stderr)
        sys.exit(2)

    import care_amd
    w = WORKLOADS[args.workload]
    reads, truth = prepare(args.workload, args.cache)
    t0 = time.perf_counter()
    rs = care_amd.ReadSet([reads], paired_end=bool(w["options"]["paired_end"]), qual_bits=w["qual_bits"])
    opt = care_amd.default_options(**w["options"])
    eng = care_amd.Engine(opt, device=0)
    if args.batch:
        eng.set_launch_config(args.batch, 0)
    eng.upload_readset(rs)
    eng.build()
    t1 = time.perf_counter()
    n = rs.n
    log("bench: %d reads packed, uploaded and indexed in %.2f s" % (n, t1 - t0))

    for i in range(args.warmup):
        eng.correct(0, n)
        log("bench: warmup step %d done" % (i + 1))
    step_ms = []
    corr_ms = []
    query_ms = []
    launches = 1
    cands = 0
    status = None
    for i in range(args.steps):
        ts = time.perf_counter()
        status, _, _, _ = eng.correct(0, n)
        te = time.perf_counter()
        st = eng.stats()
        step_ms.append((te - ts) * 1e3)
        corr_ms.append(st.correct_ms)
        query_ms.append(st.query_ms)
        cands = st.candidates
        launches = max(1, st.launches)
        log("bench: step %d %.1f ms (query %.1f ms, correct kernel %.1f ms)" % (i + 1, step_ms[-1], st.query_ms,
                                                                             st.correct_ms))
    ms = sum(step_ms) / len(step_ms)
    value = n / (ms / 1e3)

    # algorithmic bytes of the fused correction kernel per step: anchor and
    # candidate 2-bit rows, candidate lengths, candidate qualities, the
    # 16-byte per-candidate record written once, the corrected anchor out
    pitch_bytes = rs.pitch * 4
    qbytes = rs.qpitch * 4 if w["qual_bits"] else 0
    alg_bytes = n * (pitch_bytes + rs.max_length + 1) + cands * (pitch_bytes + 4 + 16 + qbytes)
    kern_ms = sum(corr_ms) / len(corr_ms)
    launch_ms = kern_ms / launches
    achieved = (alg_bytes / launches) / (launch_ms / 1e3) / 1e9

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        cpu = cpu_baseline(reads, w, min(args.cpu_sample, n), threads)

    line = {
        "metric": "reads corrected/s on 1 MI355X (plain classic path)",
        "value": round(value, 1),
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "i32+f32",
        "data": "synthetic",
        "config": {"workload": w["label"], "reads": n, "candidates_per_step": int(cands),
                   "query_kernel_ms": round(sum(query_ms) / len(query_ms), 3),
                   "correct_kernel_ms": round(kern_ms, 3),
                   "corrected_anchors": int((status & 1).sum()) if status is not None else None},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                     "kernel": "k_correct", "launches_per_step": launches,
                     "bytes_per_launch": int(alg_bytes / launches), "avg_launch_ms": round(launch_ms, 4)},
        "cpu_baseline": cpu,
    }
    eng.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
