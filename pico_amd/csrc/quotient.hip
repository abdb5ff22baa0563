// quotient.hip -- the constraint quotient of one chip on its quotient domain
// (what compute_quotient_values, vm/src/machine/utils.rs:201-356, computes on
// the host), evaluated on the device from the committed LDEs.
//
// The committed LDE holds f on GEN <w_N> (N >= |Q|) in bit-reversed row
// order, so its first |Q| rows are f on the quotient coset Q in bit-reversed
// order: row r holds x_i = GEN w_Q^i for i = rev_Q(r), and the next row of x_i
// (f at w_n x_i = x_{i + 2^qdb}) is row rev_Q((i + 2^qdb) mod |Q|).
//
// A workgroup is one wave evaluating QB consecutive rows [r0, r0 + QB), QR
// rows per thread (each interpreted instruction does QR rows' work, which
// spreads its scalar dispatch, on the CU's one scalar unit, over them).  The
// next rows form one aligned block of QB rows as well: in i = rev_Q(r) the
// row's offset in the block occupies the top bits and the block the low
// bits, so adding 2^qdb changes the low bits alike for the whole block and
// carries (uniformly) into the top ones.  Both blocks are contiguous in HBM,
// and the columns a program reads are staged into LDS a window of QW columns
// at a time with coalesced loads (STAGE instructions the host compiler
// places, least recently used window replaced).  Program operands then read
// the staged values, the registers (also in LDS, register r of slot row u at
// word r QB + u: a wave's accesses hit 64 distinct banks), constants,
// selectors or the previous result (forwarded in VGPRs).  The program is the
// same for every thread: its instructions are scalar loads and the dispatch
// is a uniform branch.
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
    const uint32_t B = a.q < (uint32_t)QB ? a.q : (uint32_t)QB;  // rows of the block (a power of two)
    const uint32_t r0 = blockIdx.x * (uint32_t)QB;
    const uint32_t mask = a.q - 1u;
    // base of the block of next rows (uniform)
    const uint32_t r0n = rev_bits((rev_bits(r0, a.log_q) + a.next_step) & mask, a.log_q) & ~(B - 1u);
    // Slot row u = t + QT k of the block is row r0 + (u mod B): with |Q| <
    // QB the surplus slots shadow live rows (staged from them, never stored).
    uint32_t nrow[QR], nat[QR], inv_zh[QR], first[QR], last[QR], trans[QR];
    const gcptr zh_tab = (gcptr)a.zh;
#pragma unroll
    for (int k = 0; k < QR; k++) {
        const uint32_t u = (t + (uint32_t)(QT * k)) & (B - 1u);
        const uint32_t i = rev_bits(r0 + u, a.log_q);
        nat[k] = i;
        nrow[k] = rev_bits((i + a.next_step) & mask, a.log_q) - r0n;  // < B
        const uint32_t zsel = i & a.zh_mask;
        inv_zh[k] = zh_tab[a.zh_mask + 1 + zsel];
        first[k] = last[k] = trans[k] = 0;
        if (a.need_selectors) {
            const uint32_t zh = zh_tab[zsel];
            const uint32_t x = fmul<F>(a.shift, ((gcptr)a.tw)[i]);
            first[k] = fmul<F>(zh, finv<F>(fsub<F>(x, F::R1)));
            trans[k] = fsub<F>(x, a.omega_n_inv);
            last[k] = fmul<F>(zh, finv<F>(trans[k]));
        }
    }
    uint32_t* const windows = lds + a.nregs * (uint32_t)QB;
    const cwptr consts = (cwptr)a.consts;
    uint32_t prev[QR] = {};  // the last arithmetic result
    auto fetch = [&](uint32_t code, uint32_t (&v)[QR]) {
        const uint32_t kind = code & QO_KIND, idx = code & QO_INDEX;
        if (kind == QO_LDS) {
#pragma unroll
            for (int k = 0; k < QR; k++) v[k] = lds[idx + t + (uint32_t)(QT * k)];
        } else if (kind == QO_NEXT) {
#pragma unroll
            for (int k = 0; k < QR; k++) v[k] = lds[idx + nrow[k]];
        } else if (kind == QO_CONST) {
            const uint32_t c = consts[idx];
#pragma unroll
            for (int k = 0; k < QR; k++) v[k] = c;
        } else {
#pragma unroll
            for (int k = 0; k < QR; k++)
                v[k] = idx == QO_SEL_PREV ? prev[k] : idx == QO_SEL_FIRST ? first[k] : idx == QO_SEL_LAST ? last[k] : trans[k];
        }
    };
    Ext4 acc[QR];
#pragma unroll
    for (int k = 0; k < QR; k++) acc[k] = Ext4{{0, 0, 0, 0}};
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
            // window of columns [a, a + QW) of the QB slot rows: thread t
            // loads column a + t % QW of slot rows t / QW + (QT / QW) m
            const bool pre = in_.b >= 2, next_rows = (in_.b & 1) != 0;
            const gcptr base = (gcptr)(pre ? a.pre : a.main);
            const uint64_t ld = pre ? a.ld_pre : a.ld_main;
            const uint32_t width = pre ? a.w_pre : a.w_main;
            const uint32_t rb = next_rows ? r0n : r0;
            const uint32_t c = t % (uint32_t)QW, u0 = t / (uint32_t)QW;
            const uint32_t col = in_.a + c < width ? in_.a + c : in_.a;  // columns past the width: never read
            constexpr int NL = QB * QW / QT;  // loads per thread
            uint32_t v[NL];
#pragma unroll
            for (int m = 0; m < NL; m++) {
                const uint32_t u = u0 + (uint32_t)(QT / QW) * m;
                v[m] = base[(uint64_t)(rb + (u & (B - 1u))) * ld + col];
            }
            uint32_t* const slot = windows + in_.dst * (uint32_t)QSLOT + c * (uint32_t)QRS;
#pragma unroll
            for (int m = 0; m < NL; m++) slot[u0 + (uint32_t)(QT / QW) * m] = v[m];
            __syncthreads();
            continue;
        }
        uint32_t x[QR];
        fetch(in_.a, x);
        if (in_.op == QI_ASSERT) {
            const cwptr pw = apow + 4u * j++;
            const uint32_t p0 = pw[0], p1 = pw[1], p2 = pw[2], p3 = pw[3];
#pragma unroll
            for (int k = 0; k < QR; k++) {
                acc[k].c[0] = fadd<F>(acc[k].c[0], fmul<F>(x[k], p0));
                acc[k].c[1] = fadd<F>(acc[k].c[1], fmul<F>(x[k], p1));
                acc[k].c[2] = fadd<F>(acc[k].c[2], fmul<F>(x[k], p2));
                acc[k].c[3] = fadd<F>(acc[k].c[3], fmul<F>(x[k], p3));
            }
            continue;
        }
        if (in_.op == QI_NEG) {
#pragma unroll
            for (int k = 0; k < QR; k++) prev[k] = fsub<F>(0u, x[k]);
        } else {
            uint32_t y[QR];
            fetch(in_.b, y);
            if (in_.op == QI_ADD) {
#pragma unroll
                for (int k = 0; k < QR; k++) prev[k] = fadd<F>(x[k], y[k]);
            } else if (in_.op == QI_SUB) {
#pragma unroll
                for (int k = 0; k < QR; k++) prev[k] = fsub<F>(x[k], y[k]);
            } else {
#pragma unroll
                for (int k = 0; k < QR; k++) prev[k] = fmul<F>(x[k], y[k]);
            }
        }
        if (in_.dst != QO_NO_DST) {
#pragma unroll
            for (int k = 0; k < QR; k++) lds[in_.dst + t + (uint32_t)(QT * k)] = prev[k];
        }
    }
#pragma unroll
    for (int k = 0; k < QR; k++) {
        if (t + (uint32_t)(QT * k) < B) {
            const Ext4 q = escale<F>(acc[k], inv_zh[k]);
            reinterpret_cast<GLOBAL u32x4*>((gptr)a.out)[nat[k]] = u32x4{q.c[0], q.c[1], q.c[2], q.c[3]};
        }
    }
}

size_t quotient_lds_bytes(uint32_t nregs, uint32_t nslots) {
    const size_t words = (size_t)nregs * QB + (size_t)nslots * QSLOT;
    return (words ? words : 1) * 4;
}

hipError_t launch_quotient(int field, const QuotientArgs& a, hipStream_t stream) {
    if (a.q == 0) return hipSuccess;
    if (a.nregs > QUOTIENT_MAX_REGS || a.nslots > QUOTIENT_SLOTS) return hipErrorInvalidValue;
    const unsigned grid = (unsigned)((a.q + QB - 1) / QB);
    const size_t lds = quotient_lds_bytes(a.nregs, a.nslots);
    if (field == FIELD_KOALABEAR)
        quotient_kernel<KoalaBear><<<grid, QT, lds, stream>>>(a);
    else
        quotient_kernel<BabyBear><<<grid, QT, lds, stream>>>(a);
    return hipGetLastError();
}

}  // namespace pg
