"""Regenerate the shipped LDE tile-shape table on the current GPU:

  python3 pico_amd/tune/make_table.py [--out pico_amd/tune/lde_shapes.txt]

Runs pg_autotune for KoalaBear and BabyBear over heights 2^7 .. 2^22 at one
width per width bucket (8, 32, 128, 256), blowup 2, three timed runs per
shape, and writes the exported table with the decision reports as comments.
The package imports the file at load time (pico_amd.load_tuning)."""
import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import pico_amd  # noqa: E402
from pico_amd import BABYBEAR, KOALABEAR  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "lde_shapes.txt"))
    ap.add_argument("--reps", type=int, default=3)
    a = ap.parse_args()
    pico_amd.tune_import("".join(" ".join(l.split()[:4]) + " 0\n" for l in pico_amd.tune_export().splitlines() if l))
    ctx = pico_amd.Context(0)
    reports = []
    for f, name in ((KOALABEAR, "KoalaBear"), (BABYBEAR, "BabyBear")):
        t0 = time.time()
        rep = pico_amd.autotune(ctx, f, 7, 22, [8, 32, 128, 256], log_blowup=1, reps=a.reps)
        print(f"{name}: {time.time() - t0:.1f} s", flush=True)
        reports.append((name, rep))
    ctx.close()
    lines = ["# LDE tile-shape table (field kind log_s width_bucket variant), written by",
             "# pico_amd/tune/make_table.py on one MI355X: pg_autotune over heights 2^7..2^22,",
             f"# widths 8/32/128/256, blowup 2, {a.reps} timed runs per shape.  Decision reports:"]
    for name, rep in reports:
        lines.append(f"# {name}")
        lines += ["#   " + l for l in rep.splitlines() if l and not l.startswith("#")]
    lines.append(pico_amd.tune_export().rstrip("\n"))
    with open(a.out, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    print("wrote", a.out)


if __name__ == "__main__":
    main()
