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
// rows per thread (each interpreted instruction does QR rows' work).  The
// next rows form one aligned block of QB rows as well: in i = rev_Q(r) the
// row's offset in the block occupies the top bits and the block the low
// bits, so adding 2^qdb changes the low bits alike for the whole block and
// carries (uniformly) into the top ones.  Both blocks are contiguous in HBM,
// and the columns a program reads are staged into LDS a window of QW columns
// at a time with coalesced loads (STAGE instructions the host compiler
// places, least recently used window replaced).  Every operand is then an
// LDS word (a staged column, a register, a selector, a constant) read at the
// row, at its next row's slot or broadcast, or the previous result held in
// VGPRs; the host folds the operand kinds into the opcode, so each
// instruction runs one straight-line handler.  The program is the same for
// every thread: it is read in blocks of QT instructions by vector loads (the
// next block in flight) and each instruction is taken from its lane with
// readlane.
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
#include <type_traits>

#include "kernels.h"
#include "field.h"

namespace pg {

__device__ __forceinline__ uint32_t rev_bits(uint32_t x, uint32_t bits) {
    return bits ? __brev(x) >> (32 - bits) : 0u;
}

template <uint32_t N>
using IC = std::integral_constant<uint32_t, N>;

// Row block of this workgroup.  The next rows of block b are block
// rev(rev(b) + 2^qdb) (in bit-reversed order the block index is the low bits
// of the natural index), thousands of blocks away in launch order.  Blocks
// are therefore numbered by L = rev(b), whose next block is L + 2^qdb, and
// logical blocks are dealt to the 8 XCDs (round-robin by launch index) in
// contiguous ranges: a block's next rows are then staged by a neighbour on
// the same XCD at about the same time, so the second read of each line can
// hit that XCD's L2, and neighbours' output elements share lines.
__device__ __forceinline__ uint32_t quotient_block(uint32_t nblocks, uint32_t log_nblocks) {
    const uint32_t b = blockIdx.x, x = b & 7u, y = b >> 3, q = nblocks >> 3, r = nblocks & 7u;
    const uint32_t logical = x * q + (x < r ? x : r) + y;
    return rev_bits(logical, log_nblocks);
}

template <class F>
__global__ void __launch_bounds__(QT) quotient_kernel(QuotientArgs a) {
    extern __shared__ uint32_t lds[];
    const uint32_t t = threadIdx.x;
    const uint32_t B = a.q < (uint32_t)QB ? a.q : (uint32_t)QB;  // rows of the block (a power of two)
    const uint32_t log_nblocks = a.q > (uint32_t)QB ? (uint32_t)a.log_q - (uint32_t)__builtin_ctz((uint32_t)QB) : 0u;
    const uint32_t r0 = quotient_block(gridDim.x, log_nblocks) * (uint32_t)QB;
    const uint32_t mask = a.q - 1u;
    // base of the block of next rows (uniform)
    const uint32_t r0n = rev_bits((rev_bits(r0, a.log_q) + a.next_step) & mask, a.log_q) & ~(B - 1u);
    const uint32_t wbase = a.nregs * (uint32_t)QB;              // windows
    const uint32_t sbase = wbase + a.nslots * (uint32_t)QSLOT;  // selectors
    const uint32_t cbase = sbase + (a.need_selectors ? 3u * QB : 0u);  // constants
    // Slot row u = t + QT k of the block is row r0 + (u mod B): with |Q| <
    // QB the surplus slots shadow live rows (staged from them, never stored).
    uint32_t nrow[QR], nat[QR], inv_zh[QR];
    const gcptr zh_tab = (gcptr)a.zh;
#pragma unroll
    for (int k = 0; k < QR; k++) {
        const uint32_t u = (t + (uint32_t)(QT * k)) & (B - 1u);
        const uint32_t i = rev_bits(r0 + u, a.log_q);
        nat[k] = i;
        nrow[k] = rev_bits((i + a.next_step) & mask, a.log_q) - r0n;  // < B
        const uint32_t zsel = i & a.zh_mask;
        inv_zh[k] = zh_tab[a.zh_mask + 1 + zsel];
        if (a.need_selectors) {
            const uint32_t zh = zh_tab[zsel];
            const uint32_t x = fmul<F>(a.shift, ((gcptr)a.tw)[i]);
            const uint32_t trans = fsub<F>(x, a.omega_n_inv);
            uint32_t* const sel = lds + sbase + t + (uint32_t)(QT * k);
            sel[0] = fmul<F>(zh, finv<F>(fsub<F>(x, F::R1)));  // is_first_row
            sel[QB] = fmul<F>(zh, finv<F>(trans));               // is_last_row
            sel[2 * QB] = trans;                                  // is_transition
        }
    }
    {  // the constant table
        const gcptr consts = (gcptr)a.consts;
        for (uint32_t w = t; w < a.nconsts; w += QT) lds[cbase + w] = consts[w];
    }
    __syncthreads();
    uint32_t prev[QR] = {};  // the last arithmetic result
    // operand of kind K at LDS word idx
    auto load = [&](auto kc, uint32_t idx, uint32_t (&v)[QR]) {
        constexpr uint32_t K = decltype(kc)::value;
        if constexpr (K == QK_ROW) {
            const uint32_t* const p = lds + idx + t;
#pragma unroll
            for (int k = 0; k < QR; k++) v[k] = p[QT * k];
        } else if constexpr (K == QK_NEXT) {
#pragma unroll
            for (int k = 0; k < QR; k++) v[k] = lds[idx + nrow[k]];
        } else if constexpr (K == QK_BCAST) {
            const uint32_t c = lds[idx];
#pragma unroll
            for (int k = 0; k < QR; k++) v[k] = c;
        } else {
#pragma unroll
            for (int k = 0; k < QR; k++) v[k] = prev[k];
        }
    };
    auto binop = [&](auto oc, auto kac, auto kbc, uint32_t ia, uint32_t ib) {
        constexpr uint32_t OP = decltype(oc)::value;
        uint32_t x[QR], y[QR];
        load(kac, ia, x);
        load(kbc, ib, y);
#pragma unroll
        for (int k = 0; k < QR; k++)
            prev[k] = OP == 0 ? fadd<F>(x[k], y[k]) : OP == 1 ? fsub<F>(x[k], y[k]) : fmul<F>(x[k], y[k]);
    };
    auto negop = [&](auto kac, uint32_t ia) {
        uint32_t x[QR];
        load(kac, ia, x);
#pragma unroll
        for (int k = 0; k < QR; k++) prev[k] = fsub<F>(0u, x[k]);
    };
    Ext4 acc[QR];
#pragma unroll
    for (int k = 0; k < QR; k++) acc[k] = Ext4{{0, 0, 0, 0}};
    // The program and the powers of alpha are read in blocks of QT entries,
    // one per lane, with vector loads (the next block in flight while the
    // current one runs), and each entry is taken from its lane with
    // readlane: scalar loads would share their wait counter with the LDS
    // reads, so every operand read would also wait for them.
    const GLOBAL u32x4* const pg = reinterpret_cast<const GLOBAL u32x4*>((gcptr)a.prog);
    const GLOBAL u32x4* const pw = reinterpret_cast<const GLOBAL u32x4*>((gcptr)a.apow);
    const uint32_t nasserts = a.nasserts;
    auto prog_block = [&](uint32_t blk) {  // entries past the end read the padding entry
        const uint32_t e = blk * QT + t;
        return pg[e < a.nprog ? e : a.nprog];
    };
    auto pow_block = [&](uint32_t blk) {
        const uint32_t e = blk * QT + t;
        return pw[e < nasserts ? e : (nasserts ? nasserts - 1 : 0)];
    };
    u32x4 pcur = prog_block(0), pnext = prog_block(1);
    u32x4 wcur = pow_block(0), wnext = pow_block(1);
    uint32_t j = 0;  // asserts folded so far
    auto assertop = [&](auto kac, uint32_t ia) {
        uint32_t x[QR];
        load(kac, ia, x);
        const uint32_t ju = j % QT;
        if (ju == 0 && j) {
            wcur = wnext;
            wnext = pow_block(j / QT + 1);
        }
        j++;
        const uint32_t p0 = (uint32_t)__builtin_amdgcn_readlane((int)wcur[0], (int)ju),
                       p1 = (uint32_t)__builtin_amdgcn_readlane((int)wcur[1], (int)ju),
                       p2 = (uint32_t)__builtin_amdgcn_readlane((int)wcur[2], (int)ju),
                       p3 = (uint32_t)__builtin_amdgcn_readlane((int)wcur[3], (int)ju);
#pragma unroll
        for (int k = 0; k < QR; k++) {
            acc[k].c[0] = fadd<F>(acc[k].c[0], fmul<F>(x[k], p0));
            acc[k].c[1] = fadd<F>(acc[k].c[1], fmul<F>(x[k], p1));
            acc[k].c[2] = fadd<F>(acc[k].c[2], fmul<F>(x[k], p2));
            acc[k].c[3] = fadd<F>(acc[k].c[3], fmul<F>(x[k], p3));
        }
    };
    uint32_t* const windows = lds + wbase;
    for (uint32_t pc = 0; pc < a.nprog; pc++) {
        const uint32_t u = pc % QT;
        if (u == 0 && pc) {
            pcur = pnext;
            pnext = prog_block(pc / QT + 1);
        }
        const uint32_t op = (uint32_t)__builtin_amdgcn_readlane((int)pcur[0], (int)u);
        const uint32_t dst = (uint32_t)__builtin_amdgcn_readlane((int)pcur[1], (int)u);
        const uint32_t ia = (uint32_t)__builtin_amdgcn_readlane((int)pcur[2], (int)u);
        const uint32_t ib = (uint32_t)__builtin_amdgcn_readlane((int)pcur[3], (int)u);
        if (op < QOP_NEG) {
            switch (op) {
#define QBIN(o) \
    case o: binop(IC<(o) / 16>{}, IC<((o) / 4) % 4>{}, IC<(o) % 4>{}, ia, ib); break;
#define QBIN4(o) QBIN(o) QBIN(o + 1) QBIN(o + 2) QBIN(o + 3)
                QBIN4(0) QBIN4(4) QBIN4(8) QBIN4(12) QBIN4(16) QBIN4(20)
                QBIN4(24) QBIN4(28) QBIN4(32) QBIN4(36) QBIN4(40) QBIN4(44)
#undef QBIN4
#undef QBIN
                default: break;
            }
        } else if (op < QOP_ASSERT) {
            switch (op - QOP_NEG) {
                case QK_ROW: negop(IC<QK_ROW>{}, ia); break;
                case QK_NEXT: negop(IC<QK_NEXT>{}, ia); break;
                case QK_BCAST: negop(IC<QK_BCAST>{}, ia); break;
                default: negop(IC<QK_PREV>{}, ia); break;
            }
        } else if (op < QOP_STAGE) {
            switch (op - QOP_ASSERT) {
                case QK_ROW: assertop(IC<QK_ROW>{}, ia); break;
                case QK_NEXT: assertop(IC<QK_NEXT>{}, ia); break;
                case QK_BCAST: assertop(IC<QK_BCAST>{}, ia); break;
                default: assertop(IC<QK_PREV>{}, ia); break;
            }
            continue;
        } else {
            // window of columns [ia, ia + QW) of the QB slot rows: thread t
            // loads column ia + t % QW of slot rows t / QW + (QT / QW) m
            const uint32_t which = ib >> 1;  // main, preprocessed, permutation
            const bool next_rows = (ib & 1) != 0;
            const gcptr base = (gcptr)(which == 0 ? a.main : which == 1 ? a.pre : a.perm);
            const uint64_t ld = which == 0 ? a.ld_main : which == 1 ? a.ld_pre : a.ld_perm;
            const uint32_t width = which == 0 ? a.w_main : which == 1 ? a.w_pre : a.w_perm;
            const uint32_t rb = next_rows ? r0n : r0;
            const uint32_t c = t % (uint32_t)QW, u0 = t / (uint32_t)QW;
            const uint32_t col = ia + c < width ? ia + c : ia;  // columns past the width: never read
            constexpr int NL = QB * QW / QT;  // loads per thread
            uint32_t v[NL];
#pragma unroll
            for (int m = 0; m < NL; m++) {
                const uint32_t uu = u0 + (uint32_t)(QT / QW) * m;
                v[m] = base[(uint64_t)(rb + (uu & (B - 1u))) * ld + col];
            }
            uint32_t* const slot = windows + dst * (uint32_t)QSLOT + c * (uint32_t)QRS;
#pragma unroll
            for (int m = 0; m < NL; m++) slot[u0 + (uint32_t)(QT / QW) * m] = v[m];
            __syncthreads();
            continue;
        }
        if (dst != QO_NO_DST) {
            uint32_t* const p = lds + dst + t;
#pragma unroll
            for (int k = 0; k < QR; k++) p[QT * k] = prev[k];
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

size_t quotient_lds_bytes(uint32_t nregs, uint32_t nslots, bool selectors, uint32_t nconsts) {
    const size_t words = (size_t)nregs * QB + (size_t)nslots * QSLOT + (selectors ? 3u * QB : 0u) + nconsts;
    return (words ? words : 1) * 4;
}

hipError_t launch_quotient(int field, const QuotientArgs& a, hipStream_t stream) {
    if (a.q == 0) return hipSuccess;
    if (a.nregs > QUOTIENT_MAX_REGS || a.nslots > QUOTIENT_SLOTS) return hipErrorInvalidValue;
    const unsigned grid = (unsigned)((a.q + QB - 1) / QB);
    const size_t lds = quotient_lds_bytes(a.nregs, a.nslots, a.need_selectors != 0, a.nconsts);
    if (lds > 65536) return hipErrorInvalidValue;
    if (field == FIELD_KOALABEAR)
        quotient_kernel<KoalaBear><<<grid, QT, lds, stream>>>(a);
    else
        quotient_kernel<BabyBear><<<grid, QT, lds, stream>>>(a);
    return hipGetLastError();
}

}  // namespace pg
