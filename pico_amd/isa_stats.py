"""Build-time statistics of this package's own gfx950 kernels.

Compiles merkle.hip to device assembly and counts the VALU instructions in the
loop body of the single-matrix leaf-hash kernel (one sponge chunk = one
Poseidon2 permutation plus its two 16-byte loads), per field, by opcode and by
issue class.  Issue classes, from a per-opcode throughput microbenchmark on MI355X:
  double   v_add_u32, v_sub_u32, v_subrev_u32, v_mov_b32, v_and_b32: up to
           two per lane per clock (measured 64-70 T/s against the 39.3 T/s
           single-issue ceiling), priced at half a slot
  single   every other VALU opcode, one slot (the 32x32-bit multiplies and
           64-bit adds measure at 0.90 of the ceiling, the rest at 0.97)
bench.py turns the slot count into the leaf hash's achieved issue rate.

Usage: python3 isa_stats.py OUT.json   (run from pico_amd/, by the Makefile)
"""
import json
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
KERNELS = {
    "koalabear": "_ZN2pg19leaf_hash_w8_kernelINS_10MontyFieldILj2130706433E",
    "babybear": "_ZN2pg19leaf_hash_w8_kernelINS_10MontyFieldILj2013265921E",
}


def kernel_lines(asm, prefix):
    out, inside = [], False
    for line in asm.splitlines():
        if line.startswith(prefix) and ":" in line.split(";")[0]:
            inside = True
            continue
        if inside:
            out.append(line)
            if "s_endpgm" in line:
                break
    return out


DOUBLE_ISSUE = {"v_add_u32", "v_sub_u32", "v_subrev_u32", "v_mov_b32", "v_and_b32"}


def opcode(line):
    m = re.match(r"^\s+(v_\w+)", line)
    if not m:
        return None
    op = m.group(1)
    for suf in ("_e32", "_e64", "_sdwa", "_dpp"):
        if op.endswith(suf):
            op = op[: -len(suf)]
    return op


def loop_body(lines):
    """Lines of the longest loop: from its header label to its backward branch."""
    labels = {}
    for i, line in enumerate(lines):
        m = re.match(r"^(\.LBB\w+):", line)
        if m:
            labels[m.group(1)] = i
    best = []
    for i, line in enumerate(lines):
        m = re.match(r"^\s+s_cbranch_\w+\s+(\.LBB\w+)", line)
        if m and m.group(1) in labels and labels[m.group(1)] < i:
            body = lines[labels[m.group(1)]:i + 1]
            if len(body) > len(best):
                best = body
    return best


def main(out_path):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    asm_path = os.path.join(HERE, "_build", "merkle_gfx950.s")
    os.makedirs(os.path.dirname(asm_path), exist_ok=True)
    subprocess.run([hipcc, "-O3", "-std=c++17", "--offload-arch=gfx950", "-S", "--cuda-device-only",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                    os.path.join(HERE, "csrc", "merkle.hip"), "-o", asm_path],
                   check=True, stderr=subprocess.DEVNULL)
    asm = open(asm_path).read()
    stats = {}
    for field, prefix in KERNELS.items():
        lines = kernel_lines(asm, prefix)
        hist = {}
        for x in loop_body(lines):
            op = opcode(x)
            if op:
                hist[op] = hist.get(op, 0) + 1
        dbl = sum(n for op, n in hist.items() if op in DOUBLE_ISSUE)
        total = sum(hist.values())
        stats[field] = {
            "leaf_hash_w8_valu_total": sum(1 for x in lines if opcode(x)),
            "leaf_hash_w8_valu_per_chunk": total,
            "leaf_hash_w8_double_issue_per_chunk": dbl,
            "leaf_hash_w8_slots_per_chunk": (total - dbl) + 0.5 * dbl,
            "leaf_hash_w8_opcodes": dict(sorted(hist.items(), key=lambda kv: -kv[1])),
        }
    with open(out_path, "w") as f:
        json.dump(stats, f, indent=1)
    print(json.dumps(stats))


if __name__ == "__main__":
    main(sys.argv[1])
