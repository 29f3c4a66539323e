"""Benchmark of the plain m-discriminator on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

A step is one stable partition of N = 2^28 uint32 values uniform in [0,2^31)
into m = 8 classes (class = x & 7), inputs resident in HBM, through the C ABI
(mdiscr_partition_u32).  Prints ONE JSON line:
  value         whole-job Gelem/s over the K timed steps
  roofline      the partition kernel's algorithmic 8 B/elem over its average
                launch time (HIP events on the launch stream) vs 8 TB/s
  cpu_baseline  the reference's own SeqMdiscr (compiled from source into
                oracle/_ref) on the same 2^28 values, on this host, with the
                other configurations' CPU baselines beside it
plus the two passes' practical floors on the same box (copy_ceiling: an
elementwise copy of the same buffers; read_ceiling: one streaming read of the
input), the skewed (Zipf 1.5) variant, the per-kernel breakdown, and the C host
drivers' results: the batched segmented workload (10,000 lognormal segments,
one graph-replayed launch vs one plain launch per segment) and the serving
workload (Poisson arrivals; continuous-batching server vs plain per-request
calls; p50/p99 latency and aggregate Gelem/s).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gelem/s & % HBM peak, 2^28 int32 m=8; batched-segment vs per-launch"
PEAK_HBM_GBS = 8000.0
N = 1 << 28
M = 8
SEED = 0x5EED0000 + 2
BYTES_PER_ELEM = 8  # one 4-byte read + one 4-byte write per element


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# The sources the plain partition kernel is compiled from: the HBM traffic in
# profiles/ was counted for one version of them and of the compiler flags.
KERNEL_SOURCES = ["mdiscriminator_amd/csrc/mdiscr_api.hip",
                  "mdiscriminator_amd/csrc/plain_kernels.hpp",
                  "mdiscriminator_amd/csrc/device_common.hpp",
                  "mdiscriminator_amd/csrc/divmagic.h",
                  "mdiscriminator_amd/csrc/api_internal.hpp"]


def kernel_source_hash():
    import hashlib
    from mdiscriminator_amd import build
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        with open(os.path.join(ROOT, f), "rb") as fh:
            h.update(fh.read())
    h.update(repr(build.HIPCC_FLAGS).encode())
    return h.hexdigest()[:16]


def time_steps(ctx, fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def kernel_times(ctx, fn, reps, names):
    import torch
    ctx.timing(True)
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    ctx.timing(False)
    out = {}
    for nm in names:
        ms, cnt = ctx.timing_read(nm)
        out[nm] = (ms / cnt) if cnt else None
    return out


def cpu_baseline(quick=False):
    """CPU baselines on this host (oracle/cpu_baseline.py, in a child process so
    that OMP_PROC_BIND=close is in place before the OpenMP runtime starts):
    the reference's own SeqMdiscr (compiled from source into oracle/_ref) on
    configs 1 and 2 at full size, the OpenMP restatement of OpenmpMdiscr.h on
    config 2, and the SoA and per-segment restatements on configs 3 and 4."""
    import subprocess
    env = dict(os.environ)
    env["OMP_PROC_BIND"] = "close"
    env.setdefault("OMP_PLACES", "cores")
    cmd = [sys.executable, "-m", "oracle.cpu_baseline"] + (["--quick"] if quick else [])
    try:
        r = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    except subprocess.TimeoutExpired:
        return {"error": "cpu baseline timed out after 900 s"}
    sys.stderr.write(r.stderr)
    if r.returncode != 0:
        return {"error": f"cpu baseline exit {r.returncode}: {r.stderr[-300:]}"}
    res = json.loads(r.stdout.strip().splitlines()[-1])
    head = dict(res["config2_seq"])   # the headline workload's own CPU time
    head["configs"] = {k: v for k, v in res.items() if k.startswith("config")}
    head["host"] = res["host"]
    head["method"] = res["method"]
    return head


def read_ceiling(d_in, stream):
    """The practical ceiling of the count pass: one streaming read of the same
    1 GiB with the count's tile shape and 16-byte non-temporal loads
    (libclkprobe.so), median of 20 event-timed launches on `stream` (a torch
    stream)."""
    import ctypes
    import torch
    probe = ctypes.CDLL(os.path.join(ROOT, "mdiscriminator_amd", "libclkprobe.so"))
    probe.mdiscr_read_probe.restype = ctypes.c_int
    probe.mdiscr_read_probe.argtypes = [ctypes.c_void_p, ctypes.c_ulonglong, ctypes.c_void_p,
                                        ctypes.c_void_p]
    n = d_in.numel()
    sink = torch.zeros(n // 8192, dtype=torch.int32, device=d_in.device)
    times = []
    for i in range(23):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        if probe.mdiscr_read_probe(d_in.data_ptr(), n, sink.data_ptr(), stream.cuda_stream) != 0:
            raise RuntimeError("read probe launch failed")
        e1.record(stream)
        torch.cuda.synchronize()
        if i >= 3:
            times.append(e0.elapsed_time(e1))
    return sorted(times)[len(times) // 2]


def first_steps(ctx, step, stream, steps):
    """The first `steps` steps, each preceded by a shader-clock probe."""
    import ctypes
    import torch
    probe = ctypes.CDLL(os.path.join(ROOT, "mdiscriminator_amd", "libclkprobe.so"))
    khz = ctypes.c_int()
    if probe.mdiscr_wall_clock_khz(ctypes.byref(khz)) != 0 or khz.value <= 0:
        raise RuntimeError("wall clock rate unavailable")
    buf = torch.zeros(2 * steps, dtype=torch.int64, device="cuda:0")
    span = khz.value * 20 // 1000  # 20 us of wall clock
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    for i in range(steps):
        if probe.mdiscr_clk_probe(ctypes.c_void_p(buf.data_ptr()), i, span,
                                  ctypes.c_void_p(stream.cuda_stream)) != 0:
            raise RuntimeError("clock probe launch failed")
        ev[i].record()
        step()
    ev[steps].record()
    torch.cuda.synchronize()
    b = buf.cpu().tolist()
    rows = []
    for i in range(steps):
        ticks, wall = b[2 * i], b[2 * i + 1]
        rows.append({"step": i, "sclk_mhz": round(ticks / (wall / (khz.value * 1e3)) / 1e6, 1),
                     "ms": round(ev[i].elapsed_time(ev[i + 1]), 4)})
        log(f"step {i:3d}: shader clock {rows[-1]['sclk_mhz']:7.1f} MHz, "
            f"step {rows[-1]['ms']:.4f} ms")
    print(json.dumps({"first_steps": rows}), flush=True)


def run_driver(args, timeout, tuned=False):
    """Runs a C host driver of the C ABI and returns its JSON line (or an error).

    tuned: the driver's autotuner writes its choices to a fresh cache file
    (MDISCR_TUNE_CACHE), returned under "tuned" as {bucket: candidate name}."""
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "mdiscriminator_amd", "bin", args[0])
    if not os.path.exists(exe):
        return {"error": f"{args[0]} not built"}
    env = dict(os.environ)
    tdir = tempfile.mkdtemp(prefix="mdiscr_tune_") if tuned else None
    if tuned:
        env["MDISCR_TUNE_CACHE"] = os.path.join(tdir, "tune.txt")
        env.pop("MDISCR_SEG_CONFIG", None)
        env.pop("MDISCR_AUTOTUNE", None)
    try:
        r = subprocess.run([exe] + [str(a) for a in args[1:]], capture_output=True, text=True,
                           timeout=timeout, env=env)
    except subprocess.TimeoutExpired:
        return {"error": f"{args[0]} timed out after {timeout} s"}
    if r.returncode != 0:
        return {"error": f"{args[0]} exit {r.returncode}: {r.stderr[-300:]}"}
    out = json.loads(r.stdout.strip().splitlines()[-1])
    if tuned:
        choices = {}
        path = env["MDISCR_TUNE_CACHE"]
        if os.path.exists(path):
            # lines "<bucket>=<candidate index> <candidate name>"
            with open(path) as fh:
                for ln in fh:
                    bucket, _, rest = ln.strip().rpartition("=")
                    if bucket:
                        choices[bucket] = rest.split(" ", 1)[-1]
            os.remove(path)
        os.rmdir(tdir)
        out["tuned"] = choices
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    # A few steps into this workload the device lowers its shader clock (2.4 to
    # about 1.7 GHz) and raises it again over some 25 steps; the partition
    # kernel runs about 15% slower meanwhile.  The default warm-up is past that.
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-quick", action="store_true",
                    help="CPU baselines on small samples (a smoke run)")
    ap.add_argument("--no-drivers", action="store_true",
                    help="skip the batched-segment and serving drivers")
    ap.add_argument("--serve-requests", type=int, default=50000)
    ap.add_argument("--profile-mode", action="store_true",
                    help="for rocprofv3 counter runs: a few partition steps, then an "
                         "elementwise copy of the same 2^28 uint32 (known 1 GiB read + "
                         "1 GiB write) that calibrates FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--first-steps", type=int, default=0,
                    help="the first N plain steps after the inputs are generated, each "
                         "preceded by a one-lane shader-clock probe (about 20 us): prints "
                         "the clock and the event-timed step per step, then exits (under "
                         "rocprofv3 --kernel-trace, the trace gives each kernel's duration)")
    ap.add_argument("--profile-m", type=int, default=M,
                    help="with --profile-mode: the class count of the plain steps (x & (m-1))")
    ap.add_argument("--profile-records", action="store_true",
                    help="with --profile-mode: also a few SoA partitions of 2^26 "
                         "(key, f32, index) records at m = 16, 32, 64")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (across GPUs "
              f"the partition would shard by an all-to-all of class runs, not measured here)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import mdiscriminator_amd as md
    from mdiscriminator_amd import workloads as W

    torch.cuda.set_device(0)
    ctx = md.Context(0)
    stream = torch.cuda.current_stream()
    log("generating inputs on the GPU")
    d_in = W.uniform_u31_torch(N, SEED, "cuda:0")
    d_zipf = W.zipf8_values_torch(N, SEED, "cuda:0")
    d_out = torch.empty_like(d_in)
    d_cnt = torch.empty(M, dtype=torch.int32, device="cuda:0")
    torch.cuda.synchronize()

    def step():
        ctx.partition_u32(d_in, d_out, md.Mod8, d_counts=d_cnt, stream=stream)

    def step_zipf():
        ctx.partition_u32(d_zipf, d_out, md.Mod8, d_counts=d_cnt, stream=stream)

    if args.first_steps:
        first_steps(ctx, step, stream, args.first_steps)
        ctx.close()
        return

    if args.profile_mode:
        pm = md.Mask(args.profile_m)
        for _ in range(args.steps):
            ctx.partition_u32(d_in, d_out, pm, stream=stream)
        torch.cuda.synchronize()
        if args.profile_records:
            ns = 1 << 26
            cols = [W.uniform_u31_torch(ns, SEED + 1, "cuda:0"),
                    torch.randn(ns, device="cuda:0").view(torch.int32),
                    torch.arange(ns, dtype=torch.int32, device="cuda:0")]
            outs = [torch.empty_like(c) for c in cols]
            for m_ in (16, 32, 64):
                for _ in range(args.steps):
                    ctx.partition_soa(cols, outs, md.Mask(m_), stream=stream)
            torch.cuda.synchronize()
            del cols, outs
            # (uint, uchar, bool) records: keyed on the uint (x & 3), then the
            # select z ? x & 7 : y & 7 (the quad-layout kernel).
            cols = [W.uniform_u31_torch(ns, SEED + 3, "cuda:0"),
                    (W.uniform_u31_torch(ns, SEED + 4, "cuda:0") & 0xFF).to(torch.uint8),
                    (W.uniform_u31_torch(ns, SEED + 5, "cuda:0") & 1).to(torch.uint8)]
            outs = [torch.empty_like(c) for c in cols]
            for _ in range(args.steps):
                ctx.partition_soa(cols, outs, md.Mask(4), 0, stream=stream)
            for _ in range(args.steps):
                ctx.partition_soa_select(cols, outs, md.Mask(8), 2, 0, 1, stream=stream)
            torch.cuda.synchronize()
            del cols, outs
        f_in, f_out = d_in.view(torch.float32), d_out.view(torch.float32)
        for _ in range(args.steps):
            torch.neg(f_in, out=f_out)   # 1 GiB read + 1 GiB write, counter calibration
        torch.cuda.synchronize()
        log("profile mode done")
        ctx.close()
        return

    log(f"timing {args.steps} steps after {args.warmup} warmup")
    dt = time_steps(ctx, step, args.steps, args.warmup)
    value = N * args.steps / dt / 1e9
    ms_step = dt / args.steps * 1e3
    kernels = ["chunk_count", "tile_scan", "partition"]
    kt = kernel_times(ctx, step, max(args.steps, 10), kernels)
    dtz = time_steps(ctx, step_zipf, args.steps, 2)
    ktz = kernel_times(ctx, step_zipf, max(args.steps, 10), kernels)
    part_ms = kt["partition"]
    achieved = N * BYTES_PER_ELEM / (part_ms * 1e-3) / 1e9

    # The practical ceiling for the partition's traffic on this device: an
    # elementwise copy (torch.neg) of the same buffers, 1 GiB read + 1 GiB
    # written, median of 20 event-timed launches.
    f_in, f_out = d_in.view(torch.float32), d_out.view(torch.float32)
    for _ in range(3):
        torch.neg(f_in, out=f_out)
    torch.cuda.synchronize()
    copy_ms = []
    for _ in range(20):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        torch.neg(f_in, out=f_out)
        e1.record()
        torch.cuda.synchronize()
        copy_ms.append(e0.elapsed_time(e1))
    copy_ms = sorted(copy_ms)[len(copy_ms) // 2]
    copy_gbs = N * BYTES_PER_ELEM / (copy_ms * 1e-3) / 1e9
    read_ms = read_ceiling(d_in, stream)
    traffic, traffic_note = None, "no counter profile"
    tpath = os.path.join(ROOT, "profiles", "traffic_partition_2p28_m8.json")
    if os.path.exists(tpath):
        tp = json.load(open(tpath))
        if tp.get("source_hash") == kernel_source_hash():
            traffic = tp.get("bytes_per_launch")
            traffic_note = f"{tp.get('kernel')}, counters of the same kernel sources"
        else:
            traffic_note = ("stale: the partition kernel's sources changed since the counter "
                            f"runs (profile {tp.get('source_hash')}, now {kernel_source_hash()})")

    # The plain path across the class axis of the reference's sweep
    # (cuda_code/run.py:66-82) and a magic-divide Mod<N>: 2^28 uniform keys.
    sweep = {}
    if not args.no_drivers:
        for label, dis in (("m2", md.Mask(2)), ("m4", md.Mask(4)), ("m16", md.Mask(16)),
                           ("m32", md.Mask(32)), ("m64", md.Mask(64)), ("mod7", md.Mod(7)),
                           ("mod33", md.Mod(33))):
            def sweep_step():
                ctx.partition_u32(d_in, d_out, dis, stream=stream)
            dt_m = time_steps(ctx, sweep_step, 10, 2)
            ktm = kernel_times(ctx, sweep_step, 10, kernels)
            ms_m = dt_m / 10 * 1e3
            sweep[label] = {"ms_per_call": round(ms_m, 4), "gelem_s": round(N / ms_m / 1e6, 2),
                            "partition_frac": round(N * BYTES_PER_ELEM / (ktm["partition"] * 1e-3)
                                                    / 1e9 / PEAK_HBM_GBS, 4),
                            "kernels_ms": {k: round(v, 5) for k, v in ktm.items() if v}}

    # The reference's own sweep: 50,000,000 keys, class = x % m (Mod<m>) for
    # m = 2, 4, ..., 16 (cuda_code/run.py:6,66-82 with
    # MdiscrTest_optimized_clean.cu:17-18); uniform keys in [0, 2^31).
    ref_sweep = {}
    if not args.no_drivers:
        n_ref = 50_000_000
        r_in, r_out = d_in[:n_ref], d_out[:n_ref]
        for m_ in range(2, 17, 2):
            dis = md.Mod(m_)

            def ref_step():
                ctx.partition_u32(r_in, r_out, dis, stream=stream)
            dt_r = time_steps(ctx, ref_step, 20, 3)
            ktr = kernel_times(ctx, ref_step, 20, kernels)
            ms_r = dt_r / 20 * 1e3
            ref_sweep[f"mod{m_}"] = {
                "ms_per_call": round(ms_r, 4), "gelem_s": round(n_ref / ms_r / 1e6, 2),
                "partition_frac": round(n_ref * BYTES_PER_ELEM / (ktr["partition"] * 1e-3)
                                        / 1e9 / PEAK_HBM_GBS, 4),
                "kernels_ms": {k: round(v, 5) for k, v in ktr.items() if v}}

    # The CPU configuration on the GPU: 2^20 uniform keys, m = 4 (Mod4), the
    # values SeqMdiscr is timed on in cpu_baseline's config 1, stream-ordered
    # back to back (the call is launch-bound at this size).
    config1 = {}
    if not args.no_drivers:
        n1 = 1 << 20
        x1 = W.uniform_u31_torch(n1, W.SEED_BASE + 1, "cuda:0")
        o1 = torch.empty_like(x1)
        c1 = torch.empty(4, dtype=torch.int32, device="cuda:0")

        def c1_step():
            ctx.partition_u32(x1, o1, md.Mod4, d_counts=c1, stream=stream)
        reps = 500
        dt_c1 = time_steps(ctx, c1_step, reps, 20)
        config1 = {"n": n1, "m": 4, "us_per_call": round(dt_c1 / reps * 1e6, 2),
                   "gelem_s": round(n1 * reps / dt_c1 / 1e9, 2), "calls": reps}
        del x1, o1

    # Calls of at most one tile (one launch each), stream-ordered back to back.
    small = {}
    if not args.no_drivers:
        for n_small in (100, 4096):
            xs = d_in[:n_small]
            os_ = d_out[:n_small]
            cs = torch.empty(8, dtype=torch.int32, device="cuda:0")

            def small_step():
                ctx.partition_u32(xs, os_, md.Mod8, d_counts=cs, stream=stream)
            reps = 2000
            dt_sm = time_steps(ctx, small_step, reps, 50)
            small[f"n{n_small}"] = {"us_per_call": round(dt_sm / reps * 1e6, 2),
                                    "calls": reps}

    # Large-m struct-of-arrays records: 2^26 x (key, float32 payload bits, index).
    soa = {}
    if not args.no_drivers:
        ns = 1 << 26
        k_in = W.uniform_u31_torch(ns, SEED + 1, "cuda:0")
        p_in = torch.randn(ns, generator=torch.Generator(device="cuda:0").manual_seed(SEED),
                           device="cuda:0").view(torch.int32)
        i_in = torch.arange(ns, dtype=torch.int32, device="cuda:0")
        outs = [torch.empty_like(k_in) for _ in range(3)]
        for m_ in (16, 32, 64):
            def soa_step():
                ctx.partition_soa([k_in, p_in, i_in], outs, md.Mask(m_), stream=stream)
            dt_s = time_steps(ctx, soa_step, 10, 2)
            kts = kernel_times(ctx, soa_step, 10, ["soa_count", "soa_scan", "soa_partition"])
            ms_s = dt_s / 10 * 1e3
            soa[f"m{m_}"] = {
                "ms_per_call": round(ms_s, 4), "grec_s": round(ns / ms_s / 1e6, 2),
                "hbm_frac_whole_path": round(ns * 24 / (ms_s * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                "partition_frac": round(ns * 24 / (kts["soa_partition"] * 1e-3) / 1e9
                                        / PEAK_HBM_GBS, 4),
                "kernels_ms": {k: round(v, 5) for k, v in kts.items()}}
        del k_in, p_in, i_in, outs

    # The other element shapes and discriminators, each one call on resident
    # data: parFilter on the 2^28 keys, 2^26 arrays of 4 words (class from word
    # 0), and 2^26 (uint x, uchar y, bool z) records under z ? x % 8 : y % 8.
    general = {}
    if not args.no_drivers:
        def shape_line(fn, n_units, bytes_per_unit, names, part_name):
            dt_g = time_steps(ctx, fn, 10, 2)
            ktg = kernel_times(ctx, fn, 10, names)
            ms_g = dt_g / 10 * 1e3
            part = ktg.get(part_name)
            return {"ms_per_call": round(ms_g, 4), "g_units_s": round(n_units / ms_g / 1e6, 2),
                    "hbm_frac_whole_path": round(n_units * bytes_per_unit / (ms_g * 1e-3) / 1e9
                                                 / PEAK_HBM_GBS, 4),
                    "partition_frac": (round(n_units * bytes_per_unit / (part * 1e-3) / 1e9
                                             / PEAK_HBM_GBS, 4) if part else None),
                    "bytes_per_unit_moved": bytes_per_unit,
                    "kernels_ms": {k: round(v, 5) for k, v in ktg.items() if v}}

        fc = torch.empty(2, dtype=torch.int32, device="cuda:0")
        general["filter_2p28_lt_2p30"] = shape_line(
            lambda: ctx.filter_u32(d_in, d_out, md.Pred.lt(1 << 30), d_counts=fc, stream=stream),
            N, 8, kernels, "partition")
        na, length = 1 << 26, 4
        a_in = W.uniform_u31_torch(na * length, SEED + 2, "cuda:0")
        a_out = torch.empty_like(a_in)
        general["arrays_2p26_len4_m16"] = shape_line(
            lambda: ctx.partition_arrays(a_in, a_out, length, md.Mask(16), stream=stream),
            na, 2 * 4 * length, ["rows_count", "rows_scan", "rows_partition"], "rows_partition")
        del a_in, a_out
        nt = 1 << 26
        tx = W.uniform_u31_torch(nt, SEED + 3, "cuda:0")
        ty = (W.uniform_u31_torch(nt, SEED + 4, "cuda:0") & 0xFF).to(torch.uint8)
        tz = (W.uniform_u31_torch(nt, SEED + 5, "cuda:0") & 1).to(torch.uint8)
        touts = [torch.empty_like(t) for t in (tx, ty, tz)]
        # The reference's own tuple test: (uint, uchar, bool) columns, class
        # from the first (ModTuple4, MdiscrHost_tuples.cu.h:14-29).
        general["ref_tuple_2p26_m4"] = shape_line(
            lambda: ctx.partition_soa([tx, ty, tz], touts, md.Mask(4), 0, stream=stream),
            nt, 2 * 6, ["soa_count", "soa_scan", "soa_partition"], "soa_partition")
        general["select_tuples_2p26_m8"] = shape_line(
            lambda: ctx.partition_soa_select([tx, ty, tz], touts, md.Mask(8), 2, 0, 1,
                                             stream=stream),
            nt, 2 * 6, ["soa_count", "soa_scan", "soa_partition"], "soa_partition")
        # above 16 classes the select runs the ballot-ranked record kernel
        for ms_ in (32, 64):
            general[f"select_tuples_2p26_m{ms_}"] = shape_line(
                lambda ms_=ms_: ctx.partition_soa_select([tx, ty, tz], touts, md.Mask(ms_), 2, 0,
                                                         1, stream=stream),
                nt, 2 * 6, ["soa_count", "soa_scan", "soa_partition"], "soa_partition")
        del tx, ty, tz, touts

    seg = serve = None
    if not args.no_drivers:
        log("batched segmented vs one plain launch per segment (C driver)")
        seg = run_driver(["mdiscr_segbench", 10000, 4096, 1.5, 4, 5], 300, tuned=True)
        log(f"serving {args.serve_requests} requests (C driver)")
        serve = run_driver(["mdiscr_servebench", args.serve_requests, 0.8, 64, 4, 16], 400)

    cpu = None
    if not args.no_cpu:
        log("cpu baselines (configs 1-4) on this host")
        cpu = cpu_baseline(args.cpu_quick)

    whole_gbs = N * BYTES_PER_ELEM / (ms_step * 1e-3) / 1e9
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gelem/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "plain partition, 2^28 uint32 uniform in [0,2^31), m=8, "
                               "class = x & 7 (Mod8)",
                   "n": N, "m": M, "seed": SEED, "parallelism": "1 GPU"},
        "hbm_frac_whole_path": round(whole_gbs / PEAK_HBM_GBS, 4),
        # the same step against the minimum traffic of two passes (the count
        # reads every element before any can be placed): 12 B/elem
        "hbm_frac_whole_path_12B": round(whole_gbs * 1.5 / PEAK_HBM_GBS, 4),
        "roofline": {"bound": "hbm", "kernel": "partition",
                     "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                     "traffic_source": traffic_note,
                     "algorithmic_bytes_per_launch": N * BYTES_PER_ELEM,
                     "avg_launch_ms": round(part_ms, 5)},
        "copy_ceiling": {"kernel": "elementwise copy of the same buffers (torch.neg), "
                                   "1 GiB read + 1 GiB written",
                         "ms": round(copy_ms, 5), "achieved": round(copy_gbs, 1), "unit": "GB/s",
                         "frac_of_peak": round(copy_gbs / PEAK_HBM_GBS, 4),
                         "partition_vs_copy": round(achieved / copy_gbs, 4)},
        # the count pass against a streaming read of the same 1 GiB, and the
        # step against that read plus the copy (the two passes' floor)
        "read_ceiling": {"kernel": "streaming read of the same 1 GiB, 16-byte loads in the "
                                   "count's 8192-word tiles (libclkprobe.so)",
                         "ms": round(read_ms, 5),
                         "achieved": round(N * 4 / (read_ms * 1e-3) / 1e9, 1), "unit": "GB/s",
                         "count_vs_read": (round(read_ms / kt["chunk_count"], 4)
                                           if kt.get("chunk_count") else None),
                         "step_vs_read_plus_copy": round((read_ms + copy_ms) / ms_step, 4)},
        "kernels_ms": {k: round(v, 5) for k, v in kt.items() if v},
        "zipf_variant": {"value": round(N * args.steps / dtz / 1e9, 3), "unit": "Gelem/s",
                         "kernels_ms": {k: round(v, 5) for k, v in ktz.items() if v}},
        "plain_across_m": sweep or None,
        "reference_sweep_5e7": ref_sweep or None,
        "config1_gpu": config1 or None,
        "small_calls": small or None,
        "soa_records_2p26": soa or None,
        "general": general or None,
        "batched_segments": seg,
        "serving": serve,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
