"""Benchmark: KoalaBear coset LDE + Poseidon2 Merkle commit of a 2^20 x 256
trace (log_blowup 1), the single-shard commit of Pico's RISC-V prover
(TwoAdicFriPcs::commit at vm/src/machine/prover.rs:238-244).

One step = one Pcs::commit of a trace already resident in HBM: LDE into a
new 2^21 x 256 buffer, leaf hashing, all compress layers, root copied to the
host, prover data released.  Prints one JSON line (see --help).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "KoalaBear LDE+Merkle commit cells/s at 2^20x256; p50/p99 job latency"

# Poseidon2-W16 KoalaBear work per permutation, textbook Montgomery form:
# 616 modmul + 1416 modadd (8 full rounds x 16 S-boxes x 2 mul; 20 partial
# rounds x (2 + 16); 9 external layers x 72 adds + 128 + 20 x 32) = 5 + 3
# VALU lane-instructions each -> 7328.  Reported beside the count of the
# built kernel (pico_amd/_build/isa_stats.json, loop body per sponge chunk).
TEXTBOOK_OPS_PER_PERM = 616 * 5 + 1416 * 3
# MI355X VALU single-issue ceiling: 256 CU x 4 SIMD x 16 lanes x 2.4 GHz =
# 39.3 T lane-slots/s (the published 157.3 TFLOP/s FP32 vector peak is this
# times 2 for FMA times 2 for packed FP32, which integer code cannot use).
# v_add_u32 / v_sub_u32 / v_mov_b32 / v_and_b32 (and their 16-bit and
# bitwise relatives) issue up to twice per slot and are priced at half a slot:
# the leaf hash's measured duration is below the floor its instruction count
# gives at one slot each, which single issue could not reach.
VALU_PEAK_TSLOTS = 256 * 4 * 16 * 2.4e9 / 1e12


def isa_stats(field="koalabear"):
    path = os.path.join(ROOT, "pico_amd", "_build", "isa_stats.json")
    try:
        with open(path) as f:
            return json.load(f)[field]
    except (OSError, KeyError, ValueError):
        return None


def pmc_traffic(workload, kernel):
    """HBM bytes per dispatch of `kernel` from the committed counter summary
    (FETCH_SIZE and WRITE_SIZE in KiB, each times its copy-kernel factor)."""
    try:
        with open(os.path.join(ROOT, "profiles", "pmc_traffic.json")) as f:
            d = json.load(f)
        k = d["workloads"][workload][kernel]
        return (k["fetch_kib"] * d["fetch_factor"] + k["write_kib"] * d["write_factor"]) * 1024
    except (OSError, KeyError, ValueError, TypeError):
        return None


HBM_PEAK_GBS = 8000.0

LDE_KIND_IDS = {"lde_single": 0, "lde_pass_a": 1, "lde_pass_bc": 2, "lde_pass_d": 3}


def lde_tile_log_e(log_s, kind, variant):
    """log2 of the values per thread of the LDE pass kernel that runs a size
    2^log_s transform: the choice of pico_amd/csrc/ntt.hip (shape_log_e, and
    shape_alt_log_e for tuning variant 1)."""
    coset_loop = kind in (0, 2)
    if variant:
        alt = {7: 3, 8: 5, 9: 4, 10: 5, 11: 6 if coset_loop else 4}.get(log_s, -1)
        if alt >= 0:
            return alt
    if log_s <= 5:
        return log_s
    return {6: 3, 7: 4, 8: 4, 9: 5, 10: 4}.get(log_s, 4 if coset_loop else 6)


def lde_slot_floors(st, field_id, log_h, w, lb):
    """Per LDE pass of an h x w LDE at blowup 2^lb: the issue slots its
    threads execute (static gfx950 instruction counts of the kernel that runs,
    pico_amd/_build/isa_stats.json) and the time they take at the
    single-issue ceiling."""
    import pico_amd
    if not st or "lde_kernels" not in st:
        return {}
    L = log_h
    k1 = 0 if L <= 11 else (L + 1) // 2
    k2 = L - k1
    passes = {"lde_single": L} if L <= 11 else {"lde_pass_a": k1, "lde_pass_bc": k2, "lde_pass_d": k1}
    elems = {"lde_single": 1 << L, "lde_pass_a": 1 << L, "lde_pass_bc": 1 << L, "lde_pass_d": 1 << (L + lb)}
    out = {}
    for fam, log_s in passes.items():
        kind = LDE_KIND_IDS[fam]
        var = pico_amd.tune_get(field_id, kind, log_s, w)
        le = lde_tile_log_e(log_s, kind, var)
        rec = next((r for r in st["lde_kernels"]
                    if r["kind"] == fam and r["log_e"] == le and r["log_p"] == log_s - le), None)
        if rec is None:
            continue
        ncos = (1 << lb) if kind in (0, 2) else 0
        per_thread = rec["slots_outside_loop"] + ncos * rec["slots_per_coset"]
        threads = elems[fam] * w >> le
        out[fam] = {"tile": f"{1 << le}x{1 << (log_s - le)}", "slots_per_element": round(per_thread / (1 << le), 2),
                    "slot_floor_ms": threads * per_thread / (VALU_PEAK_TSLOTS * 1e12) * 1e3}
    return out


def percentile(xs, q):
    s = sorted(xs)
    k = (len(s) - 1) * q / 100.0
    lo, hi = math.floor(k), math.ceil(k)
    return s[lo] + (s[hi] - s[lo]) * (k - lo)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(log_h, w, lb, runs, gpu_root):
    """The whole job on the host: oracle/cpu_baseline.c (the reference's CPU
    commit path restated in Montgomery form, AVX-512 where the host has it,
    else AVX2, OpenMP over every host thread) commits the same seeded
    2^log_h x w trace; one warm-up run, then the median of `runs` timed runs.
    Its root is checked against the GPU's."""
    import numpy as np

    from oracle import oracle

    f = oracle.KOALABEAR
    h = 1 << log_h
    mont = oracle.cpu_to_monty(f, oracle.random_matrix(f, (2 << 32) | 0, h, w))  # input form, untimed
    root, _ = oracle.cpu_pcs_commit_mont(f, mont, lb)  # warm-up
    times, phases = [], []
    for _ in range(runs):
        t0 = time.perf_counter()
        oracle.cpu_pcs_commit_mont(f, mont, lb)
        times.append(time.perf_counter() - t0)
        phases.append(oracle.cpu_last_phase_seconds())
    del mont
    med = float(np.median(times))
    isa = oracle.CPU_ISAS[oracle.cpu_isa()]
    return {
        "value": (h * w) / med,
        "unit": "cells/s",
        "cores": oracle.cpu_num_threads(),
        "kind": "port",
        "sample": (f"full job: Pcs::commit of the same seeded 2^{log_h} x {w} KoalaBear trace at blowup "
                   f"{1 << lb} (C restatement of the reference algorithm in Montgomery form, {isa}, "
                   f"not the reference binary); median of {runs} runs after 1 warm-up"),
        "isa": isa,
        "runs": runs,
        "median_s": med,
        "min_s": min(times),
        "max_s": max(times),
        "lde_s_median": float(np.median([a for a, _ in phases])),
        "merkle_s_median": float(np.median([b for _, b in phases])),
        "nproc": os.cpu_count(),
        "cpu_model": cpu_model(),
        "root_matches_gpu": bool(np.array_equal(oracle.from_monty(f, root), np.asarray(gpu_root))),
    }


def golden_root_check(h, w, lb, root):
    """True/False against the oracle's root of the same seeded job
    (tests/golden/headline_kb_2p20x256.json); None for another shape."""
    try:
        with open(os.path.join(ROOT, "tests", "golden", "headline_kb_2p20x256.json")) as fh:
            g = json.load(fh)
    except (OSError, ValueError):
        return None
    if (g["rows"], g["cols"], g["log_blowup"], g["seed"]) != (h, w, lb, (2 << 32) | 0):
        return None
    return [int(x) for x in root] == g["root"]


def copy_calibration(pico_amd, reps=5):
    """Five 1 GiB device-to-device copies through hipMemcpy (the runtime's
    copy kernel reads 1 GiB and writes 1 GiB per dispatch)."""
    import ctypes
    ctx = pico_amd.Context(0)
    a, b = ctx.matrix(1 << 18, 1024), ctx.matrix(1 << 18, 1024)  # 1 GiB each
    hip = None
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            hip = ctypes.CDLL(name)
            break
        except OSError:
            continue
    if hip is None:
        raise OSError("libamdhip64.so not found")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemcpy.restype = ctypes.c_int
    for _ in range(reps):
        rc = hip.hipMemcpy(b.device_ptr, a.device_ptr, 1 << 30, 3)  # hipMemcpyDeviceToDevice
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed: {rc}")
    pico_amd.device_synchronize()
    a.free()
    b.free()
    ctx.close()
    print(json.dumps({"copy_calibration_bytes_per_dispatch": 1 << 30, "dispatches": reps}), flush=True)


T_START = time.perf_counter()


def progress(msg):
    """A line on stderr per finished section (stdout carries only the JSON)."""
    print(f"[bench {time.perf_counter() - T_START:7.1f} s] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--log-h", type=int, default=20)
    ap.add_argument("--width", type=int, default=256)
    ap.add_argument("--log-blowup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-runs", type=int, default=5, help="timed runs of the CPU baseline (after 1 warm-up)")
    ap.add_argument("--no-workloads", action="store_true",
                    help="skip the shard+FRI, batched-jobs and served-stream sections")
    ap.add_argument("--commit-only", action="store_true",
                    help="skip the end-to-end and Pcs::open legs too (profiled runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only five plain 1 GiB device-to-device copies (the runtime's copy kernel): the "
                         "known-size kernel that scales rocprofv3's FETCH_SIZE / WRITE_SIZE to bytes")
    ap.add_argument("--served-jobs", type=int, default=48)
    ap.add_argument("--rates", type=str, default="5,10,20,50,100,200")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "only one GPU is supported: the commitment path runs on one GPU's "
                          "share (a whole job or matrix); sharding across GPUs is not implemented",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np

    import pico_amd

    if args.copy_calibration:
        copy_calibration(pico_amd)
        return

    # Every C-ABI call returns after its device work has completed, so the
    # timed region is bracketed by completed device work on both sides.  torch
    # is deliberately not imported: its wheel bundles a second HIP runtime,
    # and two runtimes in one process do not share the device.
    sync = pico_amd.device_synchronize

    f = pico_amd.KOALABEAR
    h, w, lb = 1 << args.log_h, args.width, args.log_blowup
    ctx = pico_amd.Context(0)
    trace = ctx.random_matrix(f, (2 << 32) | 0, h, w)  # resident before timing
    pcs = pico_amd.TwoAdicFriPcs(ctx, f, lb)
    dom = pcs.natural_domain_for_degree(h)

    def step():
        root, pd = pcs.commit([(dom, trace)])
        pd.free()
        return root

    for _ in range(args.warmup):
        step()
    lat = []
    sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        root = step()
        lat.append((time.perf_counter() - s0) * 1e3)
    sync()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    cells = h * w
    progress(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step")

    # per-kernel durations: HIP events on the launch stream, separate pass
    ctx.profile(True)
    ctx.profile_reset()
    prof_steps = 2
    for _ in range(prof_steps):
        step()
    ctx.profile(False)
    fams = ["lde_single", "lde_pass_a", "lde_pass_bc", "lde_pass_d", "leaf_hash", "leaf_hash_w8",
            "compress", "compress_inject"]
    kern = {}
    for fam in fams:
        ms, n = ctx.profile_get(fam)
        if n:
            kern[fam] = {"avg_ms": ms / n, "launches_per_step": n / prof_steps,
                         "ms_per_step": ms / prof_steps}
    dominant = max(kern, key=lambda k: kern[k]["ms_per_step"])
    n_rows = h << lb
    perms_leaf = n_rows * ((w + 7) // 8)
    roofline = None
    workload = f"koalabear_pcs_commit_2^{args.log_h}x{w}_blowup{1 << lb}"
    st = isa_stats()
    if dominant == "leaf_hash_w8" and st:
        avg_s = kern[dominant]["avg_ms"] / 1e3
        rate = perms_leaf / avg_s
        slots = st["leaf_hash_w8_slots_per_chunk"]
        achieved = rate * slots / 1e12
        single_issue_s = perms_leaf * st["leaf_hash_w8_valu_per_chunk"] / (VALU_PEAK_TSLOTS * 1e12)
        roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": round(VALU_PEAK_TSLOTS, 3),
                    "unit": "T VALU issue-slots/s", "frac": round(achieved / VALU_PEAK_TSLOTS, 4),
                    "traffic": pmc_traffic(workload, dominant),
                    "kernel": dominant, "avg_ms": round(kern[dominant]["avg_ms"], 4),
                    "perms_per_launch": perms_leaf, "perm_rate_per_s": rate,
                    "slots_per_perm": slots, "valu_instr_per_perm": st["leaf_hash_w8_valu_per_chunk"],
                    "textbook_instr_per_perm": TEXTBOOK_OPS_PER_PERM,
                    "algorithmic_bytes": n_rows * w * 4 + n_rows * 32,
                    "frac_at_one_slot_per_instruction": round(single_issue_s / avg_s, 4)}
        # the same kernel against the HBM roof, to show it is not the bound
        hbm_gbs = roofline["algorithmic_bytes"] / avg_s / 1e9
        roofline["hbm_GB_per_s"] = round(hbm_gbs, 1)
        roofline["hbm_frac"] = round(hbm_gbs / HBM_PEAK_GBS, 4)
    # HBM view of the NTT passes: algorithmic bytes per launch / duration
    by = 4 * w
    pass_bytes = {"lde_pass_a": 2 * h * by, "lde_pass_bc": h * by + n_rows * by,
                  "lde_pass_d": 2 * n_rows * by, "lde_single": h * by + n_rows * by}
    ntt = {}
    floors = lde_slot_floors(st, f, args.log_h, w, lb)
    for fam, b in pass_bytes.items():
        if fam in kern:
            gbs = b / (kern[fam]["avg_ms"] / 1e3) / 1e9
            ntt[fam] = {"bytes": b, "avg_ms": round(kern[fam]["avg_ms"], 4), "GB/s": round(gbs, 1),
                        "frac_hbm": round(gbs / HBM_PEAK_GBS, 4)}
            if fam in floors:
                fl = floors[fam]
                ntt[fam].update({"tile": fl["tile"], "valu_slots_per_element": fl["slots_per_element"],
                                 "valu_slot_floor_ms": round(fl["slot_floor_ms"], 4),
                                 "valu_issue_frac": round(fl["slot_floor_ms"] / kern[fam]["avg_ms"], 4)})
    lde_ms = sum(kern[fam]["avg_ms"] * kern[fam]["launches_per_step"] for fam in pass_bytes if fam in kern)
    if lde_ms:
        # the LDE's algorithmic bytes (read the trace once, write the LDE once)
        # over the time of all its passes
        lde_bytes = h * w * 4 + n_rows * w * 4
        floor = sum(fl["slot_floor_ms"] for fl in floors.values())
        ntt["whole_lde"] = {"algorithmic_bytes": lde_bytes, "ms": round(lde_ms, 4),
                            "GB/s": round(lde_bytes / (lde_ms / 1e3) / 1e9, 1),
                            "frac_hbm": round(lde_bytes / (lde_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                            "valu_slot_floor_ms": round(floor, 4) if floor else None,
                            "valu_issue_frac": round(floor / lde_ms, 4) if floor else None}

    out = {
        "metric": METRIC,
        "value": cells / (ms_per_step / 1e3),
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": workload,
                   "field": "KoalaBear", "rows": h, "cols": w, "log_blowup": lb,
                   "hash": "Poseidon2-W16 sponge(8)/truncated-compress", "lde_rows": n_rows},
        "p50_ms": percentile(lat, 50),
        "p99_ms": percentile(lat, 99),
        "roofline": roofline,
        "roofline_ntt": ntt,
        "kernels": {k: {kk: round(vv, 4) for kk, vv in v.items()} for k, v in kern.items()},
        "lde_tuning": os.environ.get("PICO_AMD_LDE_TUNING", os.path.relpath(pico_amd.TUNING_FILE, ROOT)),
        "root": [int(x) for x in root],
        "root_matches_golden": golden_root_check(h, w, lb, root),
    }
    if not args.commit_only:
        # End to end: the trace starts in (pageable) host memory, is copied to the
        # device and committed.  Reported beside `value`, which is device-resident.
        host_trace = trace.to_host()
        e2e = []
        for _ in range(3):
            sync()
            t0 = time.perf_counter()
            m = ctx.upload(host_trace)
            t1 = time.perf_counter()
            root_e2e, pd = pcs.commit([(dom, m)])
            t2 = time.perf_counter()
            pd.free()
            m.free()
            e2e.append((t1 - t0, t2 - t0))
        h2d_s, tot_s = min(e2e, key=lambda x: x[1])
        out["end_to_end"] = {"ms": tot_s * 1e3, "cells_per_s": cells / tot_s, "h2d_ms": h2d_s * 1e3,
                             "h2d_GB_per_s": host_trace.nbytes / h2d_s / 1e9, "host_memory": "pageable",
                             "root_matches": bool(all(int(a) == int(b) for a, b in zip(root_e2e, root)))}
        # pg_pcs_commit_host straight from the pageable array (staged uploads,
        # column blocks behind the upload)
        e2e = []
        for _ in range(3):
            sync()
            t0 = time.perf_counter()
            root_e2e, pd = pcs.commit_host([(dom, host_trace)], canonical=False)
            e2e.append(time.perf_counter() - t0)
            pd.free()
        tot_s = min(e2e)
        out["end_to_end_overlapped"] = {
            "ms": tot_s * 1e3, "cells_per_s": cells / tot_s, "host_memory": "pageable",
            "root_matches": bool(all(int(a) == int(b) for a, b in zip(root_e2e, root)))}
        # the same with the trace in page-locked memory (pg_host_alloc)
        pinned = ctx.host_matrix(h, w)
        pinned.array[:] = host_trace
        e2e = []
        for _ in range(3):
            sync()
            t0 = time.perf_counter()
            m = ctx.upload(pinned.array)
            t1 = time.perf_counter()
            root_e2e, pd = pcs.commit([(dom, m)])
            t2 = time.perf_counter()
            pd.free()
            m.free()
            e2e.append((t1 - t0, t2 - t0))
        h2d_s, tot_s = min(e2e, key=lambda x: x[1])
        out["end_to_end_pinned"] = {"ms": tot_s * 1e3, "cells_per_s": cells / tot_s, "h2d_ms": h2d_s * 1e3,
                                    "h2d_GB_per_s": host_trace.nbytes / h2d_s / 1e9, "host_memory": "page-locked",
                                    "root_matches": bool(all(int(a) == int(b) for a, b in zip(root_e2e, root)))}
        # pg_pcs_commit_host from page-locked memory: column blocks uploaded,
        # extended and absorbed behind one another
        e2e = []
        for _ in range(3):
            sync()
            t0 = time.perf_counter()
            root_e2e, pd = pcs.commit_host([(dom, pinned.array)], canonical=False)
            e2e.append(time.perf_counter() - t0)
            pd.free()
        tot_s = min(e2e)
        out["end_to_end_pinned_overlapped"] = {
            "ms": tot_s * 1e3, "cells_per_s": cells / tot_s, "host_memory": "page-locked",
            "root_matches": bool(all(int(a) == int(b) for a, b in zip(root_e2e, root)))}
        pinned.free()
        progress("end-to-end legs done")
        del host_trace
        # Pcs::open's device part on the same commitment: values at zeta and
        # g zeta, then the reduced opening over all 2^21 x 256 LDE cells.
        _, pd = pcs.commit([(dom, trace)])
        p = pico_amd.PRIMES[f]
        zeta = np.array([123456789, 987654321, 55555, 77777], np.uint64) % p
        pts = np.stack([zeta, zeta * pico_amd.two_adic_generator(f, args.log_h) % p])
        alpha = np.array([31, 41, 59, 26], np.uint64)
        for rep in range(3):
            ctx.profile(rep == 2)
            ctx.profile_reset()
            t0 = time.perf_counter()
            vals = pcs.open_values(pd, 0, pts)
            t1 = time.perf_counter()
            ro = pcs.reduce_openings([(pd, [pts])], [[vals]], alpha)
            t2 = time.perf_counter()
            for buf in ro.values():
                buf.free()
        ctx.profile(False)
        red_ms = ctx.profile_get("reduce_openings")[0]
        val_ms = ctx.profile_get("open_values")[0]
        red_bytes = n_rows * w * 4 + n_rows * 16
        out["pcs_open"] = {"points": 2, "open_values_ms": (t1 - t0) * 1e3, "open_values_kernel_ms": val_ms,
                           "reduce_openings_ms": (t2 - t1) * 1e3, "reduce_kernel_ms": red_ms,
                           "reduce_bytes": red_bytes, "reduce_GB_per_s": red_bytes / (red_ms / 1e3) / 1e9,
                           "reduce_frac_hbm": red_bytes / (red_ms / 1e3) / 1e9 / HBM_PEAK_GBS}
        # get_evaluations_on_domain over the whole 2^21-row LDE domain (the
        # quotient domain at quotient degree 2): natural-order Montgomery
        # words into a pageable host array, as the ABI hands them to Rust
        from pico_amd._lib import check, lib
        from pico_amd.core import _ptr
        ev = np.zeros((n_rows, w), np.uint32)
        ev_s = []
        for _ in range(3):
            t0 = time.perf_counter()
            check(lib().pg_pcs_evaluations(ctx.handle, pd.handle, 0, n_rows, _ptr(ev)))
            ev_s.append(time.perf_counter() - t0)
        out["evaluations_on_domain"] = {"rows": n_rows, "cols": w, "ms": min(ev_s) * 1e3,
                                        "GB_per_s": ev.nbytes / min(ev_s) / 1e9, "host_memory": "pageable"}
        del ev
        pd.free()
    trace.free()
    if not args.no_workloads:
        from pico_amd import workloads
        workloads.shard_commit_fri(ctx)  # warm tables
        out["shard_commit_fri"] = workloads.shard_commit_fri(ctx)
        progress("shard commit + FRI done")
        out["shard_ingest"] = workloads.shard_ingest(ctx)
        progress("shard ingest done")
        out["chip_tables"] = workloads.chip_tables(ctx)
        progress("chip tables done")
        out["quotient_eval"] = workloads.quotient_eval(ctx)
        out["quotient_eval_lookups"] = workloads.quotient_eval_lookups(ctx)
        progress("quotient evaluation done")
        out["batched_jobs"] = workloads.batched_jobs(ctx)
        progress("batched jobs done")
        jobs = workloads.served_jobs(n=args.served_jobs)
        rates = [float(r) for r in args.rates.split(",")]
        served = {"jobs": [list(j) for j in jobs], "sequential": workloads.served_sequential(ctx, jobs)}
        served["service"] = [workloads.served_stream(ctx, r, jobs) for r in rates]
        progress("served stream (device-resident traces) done")
        # the same jobs arriving in host memory: each job's upload is part of it
        host = workloads.served_host_traces(ctx, jobs)
        served["from_host"] = {"sequential": workloads.served_sequential(ctx, jobs, host_traces=host),
                               "service": [workloads.served_stream(ctx, r, jobs, host_traces=host) for r in rates]}
        del host
        out["served_stream"] = served
        progress("served stream (host traces) done")
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.log_h, w, lb, args.cpu_runs, out["root"])
        progress("CPU baseline done")
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
