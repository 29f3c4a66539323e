"""CPU baselines of the m-discriminator configurations, timed on this host.

    python -m oracle.cpu_baseline [--threads T] [--quick]

Test infrastructure (the oracle is the checker, never the product): run by
bench.py in a child process, so that OMP_PROC_BIND / OMP_NUM_THREADS are set
before the OpenMP runtime starts.  Prints one JSON line.

Method of the reference's sweep (cuda_code/run.py:49-63): 5 runs, the fastest
and slowest dropped, the mean of the other 3 reported.  Inputs come from the
same seeded SplitMix64 streams as the GPU workloads (seed 0x5EED0000 + config).

  config 1  the reference's own SeqMdiscr<Mod4> (SeqMdiscr.h:10-57, compiled
            from source into oracle/_ref) on 2^20 uniform values
  config 2  SeqMdiscr<Mod8> on all 2^28 uniform values, and on the Zipf(1.5)
            variant; the OpenMP restatement of OpenmpMdiscr.h:7-83 on the same
            2^28 values with OMP_PROC_BIND=close
  config 3  the SoA restatement of the reference's tuple seq
            (MdiscrHost_tuples.cu.h:56-120) on 2^26 (key, f32 bits, index)
            records, m = 16, 32, 64
  config 4  the per-segment seq restatement (MdiscrSegmTest.cu:8-41) over the
            10,000 lognormal segments (median 4096, sigma 1.5), m = 4
The reference's tuple and segment seqs sit in CUDA sources that need nvcc, so
configs 3 and 4 time the restatement ("port").
"""
import argparse
import json
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def log(msg):
    print(f"[cpu_baseline] {msg}", file=sys.stderr, flush=True)


def host_info():
    info = {"nproc": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity")
            else None}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        kv = {}
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        info["model"] = kv.get("Model name")
        cores = kv.get("Core(s) per socket")
        sockets = kv.get("Socket(s)")
        if cores and sockets:
            info["physical_cores"] = int(cores) * int(sockets)
        info["threads_per_core"] = int(kv["Thread(s) per core"]) if "Thread(s) per core" in kv \
            else None
    except (OSError, subprocess.SubprocessError, ValueError):
        pass
    try:
        with open("/proc/meminfo") as f:
            for line in f:
                if line.startswith("MemAvailable:"):
                    info["mem_available_gib"] = round(int(line.split()[1]) / (1 << 20), 1)
    except OSError:
        pass
    return info


def timed(fn, runs):
    """run.py:49-55: `runs` runs, min and max dropped (when runs >= 3), mean."""
    import oracle
    ts = []
    for _ in range(runs):
        t0 = oracle.now()
        fn()
        ts.append(oracle.now() - t0)
    ts.sort()
    mid = ts[1:-1] if len(ts) >= 3 else ts
    return sum(mid) / len(mid), ts


def line(units, t, unit_name, cores, kind, sample, ts):
    return {"value": round(units / t / 1e9, 4), "unit": unit_name, "cores": cores, "kind": kind,
            "mean_ms": round(t * 1e3, 2), "runs_ms": [round(x * 1e3, 2) for x in ts],
            "sample": sample}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=0,
                    help="OpenMP threads (default: OMP_NUM_THREADS, else the affinity mask)")
    ap.add_argument("--runs", type=int, default=5)
    ap.add_argument("--quick", action="store_true", help="small sizes (a smoke run)")
    args = ap.parse_args()

    import numpy as np
    import oracle
    from mdiscriminator_amd import workloads as W

    base = W.SEED_BASE
    R = args.runs
    res = {"host": host_info(), "method": f"{R} runs, fastest and slowest dropped, mean of the "
                                         f"rest (cuda_code/run.py:49-63)"}
    have_ref = oracle.ref_available()
    n1 = 1 << (12 if args.quick else 20)
    n2 = 1 << (16 if args.quick else 28)
    n3 = 1 << (14 if args.quick else 26)
    nseg = 100 if args.quick else 10000

    # config 1: 2^20, m = 4
    x = W.uniform_u31(n1, base + 1)
    if have_ref:
        t, ts = timed(lambda: oracle.ref_seq_mdiscr(x, 4, specialized=True), R)
        res["config1_seq"] = line(n1, t, "Gelem/s", 1, "reference", f"SeqMdiscr<Mod4> on {n1} "
                                  "uniform uint32 (cuda_code/SeqMdiscr.h, compiled from source)", ts)
    else:
        t, ts = timed(lambda: oracle.seq_mdiscr(x, 4, oracle.MASK), R)
        res["config1_seq"] = line(n1, t, "Gelem/s", 1, "port", f"seq_mdiscr m=4 on {n1}", ts)
    log(f"config 1: {res['config1_seq']['value']} Gelem/s")

    # config 2: 2^28, m = 8 (uniform, Zipf) and the OpenMP restatement
    log(f"config 2: generating {n2} values")
    x = W.uniform_u31(n2, base + 2)
    if have_ref:
        t, ts = timed(lambda: oracle.ref_seq_mdiscr(x, 8, specialized=True), R)
        res["config2_seq"] = line(n2, t, "Gelem/s", 1, "reference", f"SeqMdiscr<Mod8> on all {n2} "
                                  "uniform uint32 (compiled from source)", ts)
    else:
        t, ts = timed(lambda: oracle.seq_mdiscr(x, 8, oracle.MASK), R)
        res["config2_seq"] = line(n2, t, "Gelem/s", 1, "port", f"seq_mdiscr m=8 on {n2}", ts)
    log(f"config 2 seq: {res['config2_seq']['value']} Gelem/s")
    threads = args.threads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or \
        res["host"].get("affinity_cpus") or 1
    need_gib = n2 * (4 + 4 + 4 + 8 * 8) / (1 << 30)   # in, out, sizes, classes + 2 tuple arrays
    avail = res["host"].get("mem_available_gib")
    if avail is not None and avail < need_gib * 1.25:
        res["config2_openmp"] = {"skipped": f"RAM: needs {need_gib:.1f} GiB, "
                                            f"{avail} GiB available"}
    else:
        t, ts = timed(lambda: oracle.openmp_mdiscr(x, 8, oracle.MASK, nthreads=threads), R)
        res["config2_openmp"] = line(n2, t, "Gelem/s", threads, "port",
                                     f"OpenmpMdiscr restatement (serial tuple scan, "
                                     f"OpenmpMdiscr.h:39-42) on all {n2} uniform uint32, "
                                     f"OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')}", ts)
        log(f"config 2 openmp: {res['config2_openmp']['value']} Gelem/s on {threads} threads")
    del x
    z = W.zipf8_values(n2, base + 2)
    rz = max(3, R - 2)
    if have_ref:
        t, ts = timed(lambda: oracle.ref_seq_mdiscr(z, 8, specialized=True), rz)
        res["config2_seq_zipf"] = line(n2, t, "Gelem/s", 1, "reference",
                                       f"SeqMdiscr<Mod8> on {n2} Zipf(1.5) values", ts)
        log(f"config 2 zipf seq: {res['config2_seq_zipf']['value']} Gelem/s")
    del z

    # config 3: SoA records, m = 16, 32, 64
    cols = [W.uniform_u31(n3, base + 3), W.normal_f32_bits(n3, base + 3 + 100),
            np.arange(n3, dtype=np.uint32)]
    res["config3_seq"] = {}
    for m in (16, 32, 64):
        t, ts = timed(lambda: oracle.seq_mdiscr_soa(cols, m, 0, oracle.MASK), R)
        res["config3_seq"][f"m{m}"] = line(n3, t, "Grec/s", 1, "port",
                                           f"SoA seq restatement on {n3} (key, f32 bits, index) "
                                           f"records, class = key & {m - 1}", ts)
        log(f"config 3 m={m}: {res['config3_seq'][f'm{m}']['value']} Grec/s")
    del cols

    # config 4: per-segment seq over lognormal segments
    lengths = W.lognormal_lengths(nseg, base + 4)
    off = W.offsets_from_lengths(lengths)
    n4 = int(off[-1])
    x = W.uniform_u31(n4, base + 4 + 100)
    t, ts = timed(lambda: oracle.seq_mdiscr_segm(x, off, 4, oracle.MASK), R)
    res["config4_seq"] = line(n4, t, "Gelem/s", 1, "port",
                              f"per-segment seq restatement over {nseg} lognormal segments "
                              f"({n4} values), m=4", ts)
    log(f"config 4: {res['config4_seq']['value']} Gelem/s")
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
