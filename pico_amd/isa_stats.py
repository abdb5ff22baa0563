"""Build-time statistics of this package's own gfx950 kernels.

Compiles merkle.hip to device assembly and counts the VALU instructions one
trip of the chunk loop of the single-matrix leaf-hash kernel executes (one
sponge chunk = one Poseidon2 permutation plus its two 16-byte loads; a loop
nested in it, KoalaBear's partial rounds, counted once per trip), per field,
by opcode and by issue class.  Issue classes, from a per-opcode throughput microbenchmark on MI355X:
  double   v_add_u32, v_sub_u32, v_subrev_u32, v_mov_b32, v_and_b32: up to
           two per lane per clock (measured 64-70 T/s against the 39.3 T/s
           single-issue ceiling), priced at half a slot
  single   every other VALU opcode, one slot (the 32x32-bit multiplies and
           64-bit adds measure at 0.90 of the ceiling, the rest at 0.97)
bench.py turns the slot count into the leaf hash's achieved issue rate.

It also compiles ntt.hip and counts, for every instantiation of the LDE pass
kernel lde_kernel<field, log E, log P, kind>, the VALU instructions and issue
slots one thread executes outside the coset loop and per coset iteration
(passes BC and SINGLE loop over the 2^b cosets; A and D have no such loop).
A thread owns E elements, so slots per element = (outside + cosets x loop)/E;
bench.py prices each LDE pass at that count against the single-issue ceiling.

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


def slots(hist):
    dbl = sum(n for op, n in hist.items() if op in DOUBLE_ISSUE)
    return sum(hist.values()) - 0.5 * dbl


def histogram(lines):
    hist = {}
    for x in lines:
        op = opcode(x)
        if op:
            hist[op] = hist.get(op, 0) + 1
    return hist


def coset_loop(lines):
    """Lines of the loop whose header the compiler marks as a loop of depth 1
    and that holds LDS accesses (the coset loop of passes BC and SINGLE),
    from its first block to its last backward branch (s_branch included)."""
    labels = {}
    for i, line in enumerate(lines):
        m = re.match(r"^(\.LBB\w+):", line)
        if m:
            labels[m.group(1)] = i
    best = []
    for i, line in enumerate(lines):
        m = re.match(r"^\s+s_(?:cbranch_\w+|branch)\s+(\.LBB\w+)", line)
        if m and m.group(1) in labels and labels[m.group(1)] < i:
            body = lines[labels[m.group(1)]:i + 1]
            if any(re.match(r"^\s+ds_", x) for x in body) and len(body) > len(best):
                best = body
    return best


LDE_FIELDS = {"koalabear": "2130706433", "babybear": "2013265921"}
LDE_KINDS = {0: "lde_single", 1: "lde_pass_a", 2: "lde_pass_bc", 3: "lde_pass_d"}


def lde_stats(asm):
    """Per field: one record per lde_kernel<field, log E, log P, kind>."""
    out = {}
    pat = re.compile(r"^(_ZN2pg10lde_kernelINS_10MontyFieldILj(\d+)E\w*?EEELi(\d+)ELi(\d+)ELi(\d)EEE\w*):",
                     re.M)
    for m in pat.finditer(asm):
        name, prime, le, lp, kind = m.group(1), m.group(2), int(m.group(3)), int(m.group(4)), int(m.group(5))
        field = next(f for f, pr in LDE_FIELDS.items() if pr == prime)
        lines = kernel_lines(asm, name)
        total = histogram(lines)
        loop = histogram(coset_loop(lines)) if kind in (0, 2) else {}
        out.setdefault(field, []).append({
            "kind": LDE_KINDS[kind], "log_e": le, "log_p": lp,
            "valu_outside_loop": sum(total.values()) - sum(loop.values()),
            "slots_outside_loop": slots(total) - slots(loop),
            "valu_per_coset": sum(loop.values()), "slots_per_coset": slots(loop)})
    return out


def compile_asm(hipcc, src, asm_path):
    subprocess.run([hipcc, "-O3", "-std=c++17", "--offload-arch=gfx950", "-S", "--cuda-device-only",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                    os.path.join(HERE, "csrc", src), "-o", asm_path],
                   check=True, stderr=subprocess.DEVNULL)
    return open(asm_path).read()


def inner_loop(lines, body):
    """The loop nested in `body` (the compiler's marker "Inner Loop Header:
    Depth=2" to its backward branch), or []."""
    start = next((i for i, l in enumerate(body) if "Loop Header: Depth=2" in l), None)
    if start is None:
        return []
    m = re.match(r"^(\.LBB\w+):", body[start - 1] if not body[start].startswith(".LBB") else body[start])
    label = m.group(1) if m else None
    end = max((i for i, l in enumerate(body) if label and re.match(r"^\s+s_(?:cbranch_\w+|branch)\s+" + re.escape(label) + r"\b", l)),
              default=None)
    return body[start:end + 1] if end is not None else []


def partial_trips():
    """ROUNDS_P / P2_PARTIAL_UNROLL for KoalaBear (20 partial rounds), read
    from poseidon2.h: the trips of the partial-round loop."""
    src = open(os.path.join(HERE, "csrc", "poseidon2.h")).read()
    unroll = int(re.search(r"constexpr int P2_PARTIAL_UNROLL = (\d+);", src).group(1))
    return 20 // unroll


def main(out_path):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    asm = compile_asm(hipcc, "merkle.hip", os.path.join(HERE, "_build", "merkle_gfx950.s"))
    ntt = lde_stats(compile_asm(hipcc, "ntt.hip", os.path.join(HERE, "_build", "ntt_gfx950.s")))
    stats = {}
    for field, prefix in KERNELS.items():
        lines = kernel_lines(asm, prefix)
        body = loop_body(lines)
        hist = histogram(body)
        # a nested loop (KoalaBear's partial rounds) runs `trips` times per chunk
        inner = inner_loop(lines, body)
        trips = partial_trips() if inner and field == "koalabear" else 1
        for op, n in histogram(inner).items():
            hist[op] = hist.get(op, 0) + (trips - 1) * n
        dbl = sum(n for op, n in hist.items() if op in DOUBLE_ISSUE)
        total = sum(hist.values())
        stats[field] = {
            "leaf_hash_w8_valu_total": sum(1 for x in lines if opcode(x)),
            "leaf_hash_w8_valu_per_chunk": total,
            "leaf_hash_w8_double_issue_per_chunk": dbl,
            "leaf_hash_w8_slots_per_chunk": (total - dbl) + 0.5 * dbl,
            "leaf_hash_w8_opcodes": dict(sorted(hist.items(), key=lambda kv: -kv[1])),
            "lde_kernels": ntt.get(field, []),
        }
    with open(out_path, "w") as f:
        json.dump(stats, f, indent=1)
    print(json.dumps({f: {k: v for k, v in d.items() if k != "lde_kernels"} for f, d in stats.items()}))


if __name__ == "__main__":
    main(sys.argv[1])
