// quotient.hip -- the constraint quotient of one chip on its quotient domain
// (what compute_quotient_values, vm/src/machine/utils.rs:201-356, computes on
// the host), evaluated on the device from the committed LDEs.
//
// The committed LDE holds f on GEN <w_N> (N >= |Q|) in bit-reversed row
// order, so its first |Q| rows are f on the quotient coset Q in bit-reversed
// order: row r holds x_i = GEN w_Q^i for i = rev_Q(r), and the next row of x_i
// (f at w_n x_i = x_{i + 2^qdb}) is row rev_Q((i + 2^qdb) mod |Q|).
//
// A workgroup is one wave evaluating QT consecutive rows [r0, r0 + QT), one
// per thread.  Their next rows form one aligned block of QT rows as well: in
// i = rev_Q(r) the thread index occupies the top bits and the block the low
// bits, so adding 2^qdb changes the low bits alike for the whole block and
// carries (uniformly) into the top ones.  Both blocks are contiguous in HBM,
// and the columns a program reads are staged into LDS a window of QW columns
// at a time with coalesced 64-byte row-segment loads (STAGE instructions the
// host compiler places, least recently used window replaced).  Program
// operands then read the staged values, the registers (also in LDS, register
// r of thread t at word r QT + t: a wave's accesses hit 64 distinct banks),
// constants, selectors or the previous result (forwarded in a VGPR).  The
// program is the same for every thread: its instructions are scalar loads
// and the dispatch is a uniform branch.
//
// The asserts fold as acc = acc alpha + c_j (vm/src/machine/folder.rs:
// 222-226), i.e. acc = sum_j c_j alpha^(m-1-j): the host tabulates the powers
// and each assert costs four base-by-extension products.  The result is
// acc / Z_H(x_i), stored at natural index i.
//
// Selectors on the trace domain H = <w_n> (vm/src/compiler/recursion/
// circuit/domain.rs:46-60): Z_H(x) = x^n - 1, is_first_row = Z_H(x) / (x - 1),
// is_last_row = Z_H(x) / (x - w_n^-1), is_transition = x - w_n^-1.  x^n =
// GEN^n w_{2^qdb}^(i mod 2^qdb) takes 2^qdb values, tabulated with their
// inverses on the host.
#include "kernels.h"
#include "field.h"

namespace pg {

using cqptr = const __attribute__((address_space(4))) QInstr*;
using cwptr = const __attribute__((address_space(4))) uint32_t*;

__device__ __forceinline__ uint32_t rev_bits(uint32_t x, uint32_t bits) {
    return bits ? __brev(x) >> (32 - bits) : 0u;
}

template <class F>
__global__ void __launch_bounds__(QT) quotient_kernel(QuotientArgs a) {
    extern __shared__ uint32_t lds[];
    const uint32_t t = threadIdx.x;
    const uint32_t B = a.q < (uint32_t)QT ? a.q : (uint32_t)QT;  // rows of the block (a power of two)
    const uint32_t r0 = blockIdx.x * (uint32_t)QT;
    const bool live = t < B;
    const uint32_t tt = t & (B - 1);  // idle threads (|Q| < QT) shadow a live one
    const uint32_t mask = a.q - 1u;
    const uint32_t i = rev_bits(r0 + tt, a.log_q);
    const uint32_t rn = rev_bits((i + a.next_step) & mask, a.log_q);
    const uint32_t r0n = __builtin_amdgcn_readfirstlane(rn) & ~(B - 1u);
    const uint32_t tn = rn - r0n;  // < B
    const uint32_t zsel = i & a.zh_mask;
    const gcptr zh_tab = (gcptr)a.zh;
    const uint32_t inv_zh = zh_tab[a.zh_mask + 1 + zsel];
    uint32_t first = 0, last = 0, trans = 0;
    if (a.need_selectors) {
        const uint32_t zh = zh_tab[zsel];
        const uint32_t x = fmul<F>(a.shift, ((gcptr)a.tw)[i]);
        first = fmul<F>(zh, finv<F>(fsub<F>(x, F::R1)));
        trans = fsub<F>(x, a.omega_n_inv);
        last = fmul<F>(zh, finv<F>(trans));
    }
    uint32_t* const windows = lds + a.nregs * (uint32_t)QT;
    const cwptr consts = (cwptr)a.consts;
    uint32_t prev = 0;  // the last arithmetic result
    auto fetch = [&](uint32_t code) -> uint32_t {
        const uint32_t kind = code & QO_KIND, idx = code & QO_INDEX;
        if (kind == QO_LDS) return lds[idx + tt];
        if (kind == QO_NEXT) return lds[idx + tn];
        if (kind == QO_CONST) return consts[idx];
        return idx == QO_SEL_PREV ? prev : idx == QO_SEL_FIRST ? first : idx == QO_SEL_LAST ? last : trans;
    };
    Ext4 acc{{0, 0, 0, 0}};
    const cwptr apow = (cwptr)a.apow;
    uint32_t j = 0;  // asserts folded so far
    const cqptr prog = (cqptr)a.prog;
    // the next instruction is fetched while the current one runs (the
    // program carries one padding entry past its end)
    QInstr nxt{prog[0].op, prog[0].dst, prog[0].a, prog[0].b};
    for (uint32_t pc = 0; pc < a.nprog; pc++) {
        const QInstr in_ = nxt;
        nxt = QInstr{prog[pc + 1].op, prog[pc + 1].dst, prog[pc + 1].a, prog[pc + 1].b};
        if (in_.op == QI_STAGE) {
            // window of columns [a, a + QW) of rows [rb, rb + B): thread t
            // loads column a + t % QW of rows t / QW + (QT / QW) k
            const bool pre = in_.b >= 2, next_rows = (in_.b & 1) != 0;
            const gcptr base = (gcptr)(pre ? a.pre : a.main);
            const uint64_t ld = pre ? a.ld_pre : a.ld_main;
            const uint32_t width = pre ? a.w_pre : a.w_main;
            const uint32_t rb = next_rows ? r0n : r0;
            const uint32_t c = t % (uint32_t)QW, u0 = t / (uint32_t)QW;
            const uint32_t col = in_.a + c < width ? in_.a + c : in_.a;  // columns past the width: never read
            uint32_t v[QW];
#pragma unroll
            for (int k = 0; k < QW; k++) {
                const uint32_t u = u0 + (uint32_t)(QT / QW) * k;
                v[k] = base[(uint64_t)(rb + (u < B ? u : 0u)) * ld + col];
            }
            uint32_t* const slot = windows + in_.dst * (uint32_t)QSLOT + c * (uint32_t)QRS;
#pragma unroll
            for (int k = 0; k < QW; k++) slot[u0 + (uint32_t)(QT / QW) * k] = v[k];
            __syncthreads();
            continue;
        }
        const uint32_t x = fetch(in_.a);
        if (in_.op == QI_ASSERT) {
            const cwptr pw = apow + 4u * j++;
#pragma unroll
            for (int e = 0; e < 4; e++) acc.c[e] = fadd<F>(acc.c[e], fmul<F>(x, pw[e]));
            continue;
        }
        uint32_t v;
        if (in_.op == QI_NEG) {
            v = fsub<F>(0u, x);
        } else {
            const uint32_t y = fetch(in_.b);
            v = in_.op == QI_ADD ? fadd<F>(x, y) : in_.op == QI_SUB ? fsub<F>(x, y) : fmul<F>(x, y);
        }
        prev = v;
        if (in_.dst != QO_NO_DST) lds[in_.dst + tt] = v;
    }
    if (live) {
        const Ext4 q = escale<F>(acc, inv_zh);
        reinterpret_cast<GLOBAL u32x4*>((gptr)a.out)[i] = u32x4{q.c[0], q.c[1], q.c[2], q.c[3]};
    }
}

size_t quotient_lds_bytes(uint32_t nregs, uint32_t nslots) {
    const size_t words = (size_t)nregs * QT + (size_t)nslots * QSLOT;
    return (words ? words : 1) * 4;
}

hipError_t launch_quotient(int field, const QuotientArgs& a, hipStream_t stream) {
    if (a.q == 0) return hipSuccess;
    if (a.nregs > QUOTIENT_MAX_REGS || a.nslots > QUOTIENT_SLOTS) return hipErrorInvalidValue;
    const unsigned grid = (unsigned)((a.q + QT - 1) / QT);
    const size_t lds = quotient_lds_bytes(a.nregs, a.nslots);
    if (field == FIELD_KOALABEAR)
        quotient_kernel<KoalaBear><<<grid, QT, lds, stream>>>(a);
    else
        quotient_kernel<BabyBear><<<grid, QT, lds, stream>>>(a);
    return hipGetLastError();
}

}  // namespace pg
