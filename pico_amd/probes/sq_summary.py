"""Summarise rocprofv3 --pmc SQ counter runs per kernel family: VALU
instructions per wave and per element, the issue-bound floor those imply at
the VALU ceiling, and the measured time against it.

  python3 pico_amd/probes/sq_summary.py --csv A.csv B.csv --bench BENCH.json \
      --elements lde_pass_a=268435456 ... --command "..." --out profiles/NAME.txt

Each --csv is a *_counter_collection.csv; counters of one family are averaged
over its dispatches.  --bench gives the HIP-event averages (kernels.*.avg_ms)
of a run without the profiler.
"""
import argparse
import csv
import json
import re
from collections import defaultdict

KINDS = {0: "lde_single", 1: "lde_pass_a", 2: "lde_pass_bc", 3: "lde_pass_d"}
VALU_LANE_SLOTS_PER_S = 39.32e12  # 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz


def family(kernel):
    m = re.search(r"lde_kernel<.*?, (\d+), (\d+), (\d+)>\(", kernel)
    if m:
        return KINDS[int(m.group(3))]
    for name in ("leaf_hash_w8", "compress"):
        if name + "_kernel" in kernel:
            return name
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--csv", nargs="+", required=True)
    ap.add_argument("--bench", required=True)
    ap.add_argument("--elements", nargs="*", default=[], help="family=elements per launch")
    ap.add_argument("--command", required=True)
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    acc = defaultdict(lambda: defaultdict(list))
    for path in a.csv:
        for r in csv.DictReader(open(path)):
            fam = family(r["Kernel_Name"])
            if fam:
                acc[fam][r["Counter_Name"]].append(float(r["Counter_Value"]))
    bench = json.loads(open(a.bench).read().strip().splitlines()[-1])
    elems = dict(kv.split("=") for kv in a.elements)
    lines = ["# SQ counters per kernel family (means over the dispatches of one commit)", "#",
             "# commands (from the repository root, on one MI355X):"]
    lines += [f"#   {c}" for c in a.command.split("\n")]
    lines += ["#", "# valu_floor_ms = VALU instructions x 64 lanes / 39.32e12 lane-slots/s (every instruction",
              "# single-issue); issue_frac = floor / measured.  Counters: SQ_WAVES, SQ_INSTS_VALU, SQ_INSTS_VMEM_RD,",
              "# SQ_INSTS_LDS, SQ_LDS_BANK_CONFLICT, SQ_ACTIVE_INST_LDS.", "#"]
    hdr = f"  {'family':<14}{'waves':>9}{'valu/wave':>11}{'vmem/wave':>11}{'lds/wave':>10}{'valu/elem':>11}" \
          f"{'floor_ms':>10}{'meas_ms':>9}{'issue_frac':>12}{'lds_conf':>10}"
    lines.append(hdr)
    for fam in ("lde_pass_a", "lde_pass_bc", "lde_pass_d", "leaf_hash_w8"):
        if fam not in acc:
            continue
        d = {k: sum(v) / len(v) for k, v in acc[fam].items()}
        w = d.get("SQ_WAVES", 0.0)
        valu = d.get("SQ_INSTS_VALU", 0.0)
        floor_ms = valu * 64 / VALU_LANE_SLOTS_PER_S * 1e3
        meas = bench.get("kernels", {}).get(fam, {}).get("avg_ms")
        per_elem = valu * 64 / float(elems[fam]) if fam in elems else float("nan")
        conf = d.get("SQ_LDS_BANK_CONFLICT", 0.0) / max(d.get("SQ_ACTIVE_INST_LDS", 0.0), 1.0)
        lines.append(f"  {fam:<14}{w:>9.0f}{valu / w:>11.0f}{d.get('SQ_INSTS_VMEM_RD', 0) / w:>11.0f}"
                     f"{d.get('SQ_INSTS_LDS', 0) / w:>10.0f}{per_elem:>11.1f}{floor_ms:>10.3f}"
                     f"{(meas or float('nan')):>9.3f}{(floor_ms / meas if meas else float('nan')):>12.3f}{conf:>10.3f}")
    open(a.out, "w").write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
