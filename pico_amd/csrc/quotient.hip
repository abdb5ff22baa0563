// quotient.hip -- the constraint quotient of one chip on its quotient domain
// (what compute_quotient_values, vm/src/machine/utils.rs:201-356, computes on
// the host), evaluated on the device from the committed LDEs.
//
// One thread per point x_i = GEN w_Q^i of the quotient coset Q (natural
// order, |Q| = n 2^qdb).  The committed LDE holds f on GEN <w_N> (N >= |Q|)
// in bit-reversed row order, so its first |Q| rows are f on Q in
// bit-reversed order: the local row of x_i is row rev_Q(i) and the next row
// (f at w_n x_i = x_{i + 2^qdb}) is row rev_Q((i + 2^qdb) mod |Q|).
//
// The constraints arrive as a register program compiled on the host from the
// AIR's symbolic expressions (the program is the same for every thread, so
// instructions are scalar loads and the dispatch is a uniform branch).
// Registers live in LDS, register r of thread t at word r * QT + t, so a
// wave's 64 accesses to one register hit 64 distinct banks.  Each assert
// folds acc = acc * alpha + value (vm/src/machine/folder.rs:222-226); the
// result is acc / Z_H(x_i).
//
// Selectors on the trace domain H = <w_n> (vm/src/compiler/recursion/
// circuit/domain.rs:46-60): Z_H(x) = x^n - 1, is_first_row = Z_H(x) / (x - 1),
// is_last_row = Z_H(x) / (x - w_n^-1), is_transition = x - w_n^-1.  x^n =
// GEN^n w_{2^qdb}^(i mod 2^qdb) takes 2^qdb values, tabulated with their
// inverses on the host.
#include "kernels.h"
#include "field.h"

namespace pg {

constexpr int QT = 64;  // threads per workgroup (one wave)

template <class F>
__global__ void __launch_bounds__(QT) quotient_kernel(QuotientArgs a) {
    extern __shared__ uint32_t regs[];
    const uint32_t tid = threadIdx.x;
    const uint64_t i = (uint64_t)blockIdx.x * QT + tid;
    if (i >= a.q) return;  // no barriers below: each thread touches only its own registers
    const uint32_t ii = (uint32_t)i, mask = a.q - 1u;
    const uint32_t r_loc = a.log_q ? __brev(ii) >> (32 - a.log_q) : 0u;
    const uint32_t inext = (ii + a.next_step) & mask;
    const uint32_t r_nxt = a.log_q ? __brev(inext) >> (32 - a.log_q) : 0u;
    const gcptr main_loc = (gcptr)a.main + (uint64_t)r_loc * a.ld_main;
    const gcptr main_nxt = (gcptr)a.main + (uint64_t)r_nxt * a.ld_main;
    const gcptr pre_loc = (gcptr)a.pre + (uint64_t)r_loc * a.ld_pre;
    const gcptr pre_nxt = (gcptr)a.pre + (uint64_t)r_nxt * a.ld_pre;
    const uint32_t zsel = ii & a.zh_mask;
    const gcptr zh_tab = (gcptr)a.zh;
    const uint32_t zh = zh_tab[zsel], inv_zh = zh_tab[a.zh_mask + 1 + zsel];
    uint32_t first = 0, last = 0, trans = 0;
    if (a.need_selectors) {
        const uint32_t x = fmul<F>(a.shift, ((gcptr)a.tw)[ii]);
        first = fmul<F>(zh, finv<F>(fsub<F>(x, F::R1)));
        trans = fsub<F>(x, a.omega_n_inv);
        last = fmul<F>(zh, finv<F>(trans));
    }
    Ext4 acc{{0, 0, 0, 0}};
    // the program is read-only for the launch and the same for every lane:
    // constant address space, so its loads are scalar
    const __attribute__((address_space(4))) QInstr* prog = (const __attribute__((address_space(4))) QInstr*)a.prog;
    for (uint32_t pc = 0; pc < a.nprog; pc++) {
        const QInstr in_{prog[pc].op, prog[pc].dst, prog[pc].a, prog[pc].b};
        uint32_t v = 0;
        switch (in_.op) {
            case QI_MAIN: v = (in_.a ? main_nxt : main_loc)[in_.b]; break;
            case QI_PRE: v = (in_.a ? pre_nxt : pre_loc)[in_.b]; break;
            case QI_CONST: v = in_.a; break;
            case QI_FIRST: v = first; break;
            case QI_LAST: v = last; break;
            case QI_TRANS: v = trans; break;
            case QI_ADD: v = fadd<F>(regs[in_.a * QT + tid], regs[in_.b * QT + tid]); break;
            case QI_SUB: v = fsub<F>(regs[in_.a * QT + tid], regs[in_.b * QT + tid]); break;
            case QI_NEG: v = fsub<F>(0u, regs[in_.a * QT + tid]); break;
            case QI_MUL: v = fmul<F>(regs[in_.a * QT + tid], regs[in_.b * QT + tid]); break;
            case QI_ASSERT: {
                acc = emul<F>(acc, a.alpha);
                acc.c[0] = fadd<F>(acc.c[0], regs[in_.a * QT + tid]);
                continue;
            }
            default: break;
        }
        regs[in_.dst * QT + tid] = v;
    }
    const Ext4 q = escale<F>(acc, inv_zh);
    reinterpret_cast<GLOBAL u32x4*>((gptr)a.out)[i] = u32x4{q.c[0], q.c[1], q.c[2], q.c[3]};
}

hipError_t launch_quotient(int field, const QuotientArgs& a, hipStream_t stream) {
    if (a.q == 0) return hipSuccess;
    if (a.nregs > QUOTIENT_MAX_REGS) return hipErrorInvalidValue;
    const unsigned grid = (unsigned)((a.q + QT - 1) / QT);
    const size_t lds = (size_t)(a.nregs ? a.nregs : 1) * QT * 4;
    if (field == FIELD_KOALABEAR)
        quotient_kernel<KoalaBear><<<grid, QT, lds, stream>>>(a);
    else
        quotient_kernel<BabyBear><<<grid, QT, lds, stream>>>(a);
    return hipGetLastError();
}

}  // namespace pg
