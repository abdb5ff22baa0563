"""Summarise rocprofv3 --pmc CSV output into per-kernel HBM traffic.

Inputs are the *_counter_collection.csv files of four runs, FETCH_SIZE and
WRITE_SIZE (KiB per dispatch) for the bench workload and for probes/copy_bw
(1 GiB read + 1 GiB written by one kernel).  The copy kernel's readings give
the factor by which each counter reports true bytes on this device; each
kernel family's mean reading per dispatch times that factor is its traffic.
Writes (merges) the result into the JSON file named by --out, which bench.py
reads for roofline.traffic.

  python3 pico_amd/probes/pmc_summary.py --workload NAME \
      --fetch F.csv --write W.csv --copy-fetch CF.csv --copy-write CW.csv \
      --out profiles/pmc_traffic.json
"""
import argparse
import csv
import json
import os
import re
from collections import defaultdict

KIND_NAMES = {0: "lde_single", 1: "lde_pass_a", 2: "lde_pass_bc", 3: "lde_pass_d"}


def family(kernel):
    m = re.search(r"lde_kernel<.*?, (\d+), (\d+), (\d+)>\(", kernel)
    if m:
        return KIND_NAMES[int(m.group(3))]
    # the 16-lane forms count with their one-thread families, as in bench.py
    lanes = {"compress_lanes_kernel": "compress", "compress_batch_lanes_kernel": "compress_batch",
             "compress_digest_lanes_kernel": "compress_digest", "leaf8_lanes_kernel": "leaf_hash_w8"}
    for name, fam in lanes.items():
        if name + "<" in kernel or name + "(" in kernel:
            return fam
    for name in ("leaf_hash_w8_kernel", "leaf_hash_batch_kernel", "leaf_hash_kernel", "compress_batch_kernel",
                 "compress_digest_kernel", "compress_kernel", "copy_kernel"):
        if name + "<" in kernel or name + "(" in kernel:
            return name[: -len("_kernel")]
    return None


def mean_per_family(path, counter):
    acc = defaultdict(list)
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            fam = family(r["Kernel_Name"])
            if fam:
                acc[fam].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", required=True)
    ap.add_argument("--fetch", required=True)
    ap.add_argument("--write", required=True)
    ap.add_argument("--copy-fetch", required=True)
    ap.add_argument("--copy-write", required=True)
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    gib_kib = float(1 << 20)
    ff = gib_kib / mean_per_family(a.copy_fetch, "FETCH_SIZE")["copy"]
    wf = gib_kib / mean_per_family(a.copy_write, "WRITE_SIZE")["copy"]
    fetch = mean_per_family(a.fetch, "FETCH_SIZE")
    write = mean_per_family(a.write, "WRITE_SIZE")
    d = {}
    if os.path.exists(a.out):
        with open(a.out) as f:
            d = json.load(f)
    d["fetch_factor"] = round(ff, 4)
    d["write_factor"] = round(wf, 4)
    d.setdefault("workloads", {})[a.workload] = {
        k: {"fetch_kib": round(fetch[k], 1), "write_kib": round(write.get(k, 0.0), 1),
            "bytes": round((fetch[k] * ff + write.get(k, 0.0) * wf) * 1024)}
        for k in sorted(fetch)}
    with open(a.out, "w") as f:
        json.dump(d, f, indent=1)
    print(json.dumps(d["workloads"][a.workload], indent=1))


if __name__ == "__main__":
    main()
