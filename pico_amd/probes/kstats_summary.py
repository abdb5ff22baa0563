"""Format a rocprofv3 --kernel-trace --stats run of bench.py into the short
text summary kept under profiles/.

  python3 pico_amd/probes/kstats_summary.py --stats KT_kernel_stats.csv \
      --bench BENCH.json --command "<the rocprofv3 command>" --out profiles/NAME.txt

--bench is the JSON line of a bench.py run without the profiler (same box,
same build): its HIP-event per-kernel averages are printed beside the
profiler's, which must agree for the dominant kernel.
"""
import argparse
import csv
import json
import re

FIELDS = {"2130706433u": "KB", "2013265921u": "BB"}


def short(name):
    f = next((v for k, v in FIELDS.items() if k in name), "")
    m = re.search(r"pg::(\w+)<(.*)>\(", name)
    if not m:
        return name.split("(")[0][:60]
    kern, targs = m.group(1), m.group(2)
    rest = re.sub(r"pg::MontyField<[^>]*>", "", targs).strip(", ")
    return f"{kern}<{f}{', ' + rest if rest else ''}>"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--stats", required=True)
    ap.add_argument("--bench", required=True)
    ap.add_argument("--command", required=True)
    ap.add_argument("--title", default="KoalaBear PCS commit of a 2^20 x 256 trace, blowup 2")
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    rows = list(csv.DictReader(open(a.stats)))
    bench = json.loads(open(a.bench).read().strip().splitlines()[-1])
    lines = [f"# rocprofv3 kernel statistics: {a.title}", "#",
             "# command (from the repository root, on one MI355X):"]
    lines += [f"#   {c}" for c in a.command.split("\n")]
    lines += ["#", f"# {'kernel':<58}{'calls':>8}{'avg_ms':>10}{'min_ms':>10}{'max_ms':>10}{'share%':>8}"]
    for r in sorted(rows, key=lambda r: -float(r["Percentage"])):
        lines.append(f"  {short(r['Name']):<58}{int(r['Calls']):>8}{float(r['AverageNs']) / 1e6:>10.4f}"
                     f"{float(r['MinNs']) / 1e6:>10.4f}{float(r['MaxNs']) / 1e6:>10.4f}{float(r['Percentage']):>8.2f}")
    lines += ["#", "# lde_kernel<field, log E, log P, kind>: kind 1 = pass A, 2 = pass BC, 3 = pass D, 0 = single pass.",
              "# bench.py's HIP-event timing in a run without the profiler (same box, same build):"]
    kern = bench.get("kernels", {})
    parts = [f"{k} {v['avg_ms']:.3f} ms" for k, v in kern.items()]
    lines.append("#   " + ", ".join(parts) + f"; {bench['ms_per_step']:.3f} ms per commit")
    open(a.out, "w").write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
