"""Build-time statistics of this package's own gfx950 kernels.

Compiles merkle.hip to device assembly and counts the VALU instructions in the
loop body of the single-matrix leaf-hash kernel (one sponge chunk = one
Poseidon2 permutation plus its two 16-byte loads), per field.  bench.py uses
the count to express the leaf hash's achieved VALU issue rate.

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


def loop_body_valu(lines):
    """VALU count between the loop header label and its backward branch."""
    labels = {}
    for i, line in enumerate(lines):
        m = re.match(r"^(\.LBB\w+):", line)
        if m:
            labels[m.group(1)] = i
    best = None
    for i, line in enumerate(lines):
        m = re.match(r"^\s+s_cbranch_\w+\s+(\.LBB\w+)", line)
        if m and m.group(1) in labels and labels[m.group(1)] < i:
            body = lines[labels[m.group(1)]:i + 1]
            n = sum(1 for x in body if re.match(r"^\s+v_", x))
            best = max(best or 0, n)
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
        stats[field] = {
            "leaf_hash_w8_valu_total": sum(1 for x in lines if re.match(r"^\s+v_", x)),
            "leaf_hash_w8_valu_per_chunk": loop_body_valu(lines),
        }
    with open(out_path, "w") as f:
        json.dump(stats, f, indent=1)
    print(json.dumps(stats))


if __name__ == "__main__":
    main(sys.argv[1])
