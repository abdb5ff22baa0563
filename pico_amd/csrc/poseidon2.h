// poseidon2.h -- Poseidon2 width-16 permutation over KoalaBear / BabyBear,
// one permutation state per thread (16 x u32 in VGPRs), Montgomery form.
//
// Structure (vm/src/primitives/mod.rs:305-346 with the p3 Poseidon2 layout;
// restated in-tree at gnark/poseidon2/poseidon2_koalabear.go:71-196):
//   external linear layer; 4 full rounds (RC rows 0..3); ROUNDS_P partial
//   rounds (RC rows 4.., column 0 only); 4 full rounds (the next 4 rows).
//   full round:    s_i += rc_i; s_i = s_i^d; external layer
//   partial round: s_0 += rc;   s_0 = s_0^d; internal layer
//   external layer: M4 on each 4-block, then s_i += sum_{j = i mod 4} s_j
//   internal layer: S = sum s; s_i = s_i * diag_i + S
// The constants are data shared with the reference: the raw round-constant
// table RC_16_30_U32 (mod.rs:37-218, reduced mod p) and the internal
// diagonals (consts.rs:110-154).
#pragma once
#include "field.h"

namespace pg {

constexpr uint32_t RC16_RAW[30][16] = {
#include "rc16.inc"
};

template <class F>
struct P2Const {
    uint32_t ext_rc[8][16];
    uint32_t int_rc[F::ROUNDS_P];
    // pre_rc[r] = M_E^{-1} ext_rc[r]: adding it before an external layer is
    // the same as adding ext_rc[r] after it (M_E is linear and invertible)
    uint32_t pre_rc[8][16];
    uint32_t diag[16];
};

// External-layer matrix: out_{4b+i} = sum_j M4[i][j] (s_{4b+j} + sum_b' s_{4b'+j}),
// i.e. block (b, b') is 2 M4 on the diagonal and M4 elsewhere.
constexpr uint32_t M4_COEF[4][4] = {{2, 3, 1, 1}, {1, 2, 3, 1}, {1, 1, 2, 3}, {3, 1, 1, 2}};

// Inverse of the 16 x 16 external matrix mod p by Gauss-Jordan elimination
// (compile time; the matrix is MDS, so every pivot search succeeds).
constexpr void ext_matrix_inverse(uint32_t p, uint64_t (&inv)[16][16]) {
    uint64_t a[16][16] = {};
    for (int r = 0; r < 16; r++)
        for (int c = 0; c < 16; c++) {
            a[r][c] = M4_COEF[r % 4][c % 4] * ((r / 4 == c / 4) ? 2u : 1u);
            inv[r][c] = r == c;
        }
    for (int col = 0; col < 16; col++) {
        int piv = col;
        while (a[piv][col] % p == 0) piv++;
        for (int c = 0; c < 16; c++) {
            uint64_t t = a[col][c];
            a[col][c] = a[piv][c];
            a[piv][c] = t;
            t = inv[col][c];
            inv[col][c] = inv[piv][c];
            inv[piv][c] = t;
        }
        const uint64_t pinv = cpow_canon(p, (uint32_t)(a[col][col] % p), p - 2);
        for (int c = 0; c < 16; c++) {
            a[col][c] = a[col][c] * pinv % p;
            inv[col][c] = inv[col][c] * pinv % p;
        }
        for (int r = 0; r < 16; r++) {
            if (r == col) continue;
            const uint64_t f = a[r][col] % p;
            for (int c = 0; c < 16; c++) {
                a[r][c] = (a[r][c] + (p - f) * a[col][c]) % p;
                inv[r][c] = (inv[r][c] + (p - f) * inv[col][c]) % p;
            }
        }
    }
}

template <class F>
constexpr P2Const<F> make_p2const() {
    P2Const<F> k{};
    constexpr uint32_t p = F::P;
    for (int r = 0; r < 4; r++)
        for (int i = 0; i < 16; i++) {
            k.ext_rc[r][i] = cmonty<F>(RC16_RAW[r][i] % p);
            k.ext_rc[4 + r][i] = cmonty<F>(RC16_RAW[4 + F::ROUNDS_P + r][i] % p);
        }
    for (int r = 0; r < F::ROUNDS_P; r++) k.int_rc[r] = cmonty<F>(RC16_RAW[4 + r][0] % p);
    uint64_t inv[16][16] = {};
    ext_matrix_inverse(p, inv);
    for (int r = 0; r < 8; r++) {
        const int row = r < 4 ? r : 4 + F::ROUNDS_P + (r - 4);
        for (int i = 0; i < 16; i++) {
            uint64_t acc = 0;
            for (int j = 0; j < 16; j++) acc = (acc + inv[i][j] * (RC16_RAW[row][j] % p)) % p;
            k.pre_rc[r][i] = cmonty<F>((uint32_t)acc);
        }
    }
    // canonical diagonal values; KoalaBear consts.rs:136-154, BabyBear :110-127
    uint32_t d[16] = {};
    if (F::P == KoalaBear::P) {
        const uint32_t v[16] = {p - 2, 1, 2, (p + 1) >> 1, 3, 4, (p - 1) >> 1, p - 3, p - 4,
                                p - ((p - 1) >> 8), p - ((p - 1) >> 3), p - 127, (p - 1) >> 8,
                                (p - 1) >> 3, (p - 1) >> 4, 127};
        for (int i = 0; i < 16; i++) d[i] = v[i];
    } else {
        const uint32_t v[16] = {p - 2, 1, 2, (p + 1) >> 1, 3, 4, (p - 1) >> 1, p - 3, p - 4,
                                p - ((p - 1) >> 8), p - ((p - 1) >> 2), p - ((p - 1) >> 3), p - 15,
                                (p - 1) >> 8, (p - 1) >> 4, 15};
        for (int i = 0; i < 16; i++) d[i] = v[i];
    }
    for (int i = 0; i < 16; i++) k.diag[i] = cmonty<F>(d[i]);
    return k;
}

template <class F>
struct P2 {
    static constexpr P2Const<F> K = make_p2const<F>();
};

// Exclusive upper bound of a Montgomery product (a b + m p) / 2^32 of
// factors below `a` and `b` before its final subtraction (compile time).
constexpr uint64_t redc_bound(uint64_t a, uint64_t b, uint64_t p) {
    return (uint64_t)(((unsigned __int128)a * b + ((unsigned __int128)p << 32) + 0xffffffffu) >> 32);
}

template <class F>
PG_HD uint32_t sbox(uint32_t x) {
    // x < p.  fmul_lazy leaves a product unreduced; fmul's single final
    // subtraction suffices while its product bound stays below 2p.
    constexpr uint64_t p = F::P;
    constexpr uint64_t b2 = redc_bound(p, p, p);  // x^2 unreduced
    static_assert(redc_bound(b2, p, p) <= 2 * p, "x^2 x reduces with one subtraction");
    if constexpr (F::SBOX == 3) {
        return fmul<F>(fmul_lazy<F>(x, x), x);
    } else {
        // x^4 from the unreduced x^2 stays a 32-bit word, and x^4 x^3
        // reduces with one subtraction (BabyBear: x^2 < 1.47p, x^4 < 2.02p)
        constexpr uint64_t b4 = redc_bound(b2, b2, p);
        static_assert(b4 <= (1ull << 32) && redc_bound(b4, p, p) <= 2 * p, "x^7 bounds");
        static_assert((unsigned __int128)b2 * b2 + ((unsigned __int128)p << 32) <= ((unsigned __int128)1 << 64),
                      "x^2 x^2 + m p fits 64 bits");
        const uint32_t x2 = fmul_lazy<F>(x, x);
        const uint32_t x3 = fmul<F>(x2, x);
        return fmul<F>(fmul_lazy<F>(x2, x2), x3);
    }
}

// y = M4 x with M4 = [[2,3,1,1],[1,2,3,1],[1,1,2,3],[3,1,1,2]] in 11 field
// additions: with s = a+b+c+d,
//   y0 = (s + b) + (a + b),  y1 = (s + b) + 2c,
//   y2 = (s + d) + (c + d),  y3 = (s + d) + 2a.
template <class F>
PG_HD void m4(uint32_t& a, uint32_t& b, uint32_t& c, uint32_t& d) {
    const uint32_t ab = fadd<F>(a, b), cd = fadd<F>(c, d), s = fadd<F>(ab, cd);
    const uint32_t sb = fadd<F>(s, b), sd = fadd<F>(s, d);
    const uint32_t y0 = fadd<F>(sb, ab);
    const uint32_t y1 = fadd<F>(sb, fadd<F>(c, c));
    const uint32_t y2 = fadd<F>(sd, cd);
    const uint32_t y3 = fadd<F>(sd, fadd<F>(a, a));
    a = y0;
    b = y1;
    c = y2;
    d = y3;
}

template <class F>
PG_HD void external_layer(uint32_t (&s)[16]) {
#pragma unroll
    for (int b = 0; b < 16; b += 4) m4<F>(s[b], s[b + 1], s[b + 2], s[b + 3]);
    uint32_t t[4];
#pragma unroll
    for (int i = 0; i < 4; i++) t[i] = fadd<F>(fadd<F>(s[i], s[4 + i]), fadd<F>(s[8 + i], s[12 + i]));
#pragma unroll
    for (int i = 0; i < 16; i++) s[i] = fadd<F>(s[i], t[i & 3]);
}

// x mod p for x < 70 p (KoalaBear): q = x >> 31 <= 69 and x - q p lies in
// [0, 2^31 + 69 (2^24 - 1)), below 2p, so one conditional subtraction ends it.
template <class F>
PG_HD uint32_t reduce_wide(uint64_t x) {
    static_assert(F::P == KoalaBear::P, "bound derived for KoalaBear");
    const uint32_t q = (uint32_t)(x >> 31);
    return freduce<F>((uint32_t)x - q * F::P);
}

// KoalaBear external layer with unreduced 64-bit sums (one instruction each on
// gfx950).  Inputs may be below 2p (an element plus a pre_rc constant); every
// output combines the 16 inputs with coefficients summing to 35, so it stays
// below 70p and is reduced once.
template <class F>
PG_HD void external_layer_wide(uint32_t (&s)[16]) {
    uint64_t y[16];
#pragma unroll
    for (int b = 0; b < 16; b += 4) {
        const uint64_t a = s[b], bb = s[b + 1], c = s[b + 2], d = s[b + 3];
        const uint64_t ab = a + bb, cd = c + d, t = ab + cd, sb = t + bb, sd = t + d;
        y[b] = sb + ab;
        y[b + 1] = sb + (c + c);
        y[b + 2] = sd + cd;
        y[b + 3] = sd + (a + a);
    }
    uint64_t col[4];
#pragma unroll
    for (int i = 0; i < 4; i++) col[i] = (y[i] + y[4 + i]) + (y[8 + i] + y[12 + i]);
#pragma unroll
    for (int i = 0; i < 16; i++) s[i] = reduce_wide<F>(y[i] + col[i & 3]);
}

// Sum of 16 reduced elements.  Pairs are added in 32 bits (< 2p), the rest in
// 64 bits (< 16p < 2^35); then q = T >> 31 <= 15 and T - q p lies in
// [0, 2^31 + 15 (2^31 - p)), below 2p for KoalaBear (one conditional
// subtraction) and below 3p for BabyBear (two).
template <class F>
PG_HD uint32_t sum16(const uint32_t (&s)[16]) {
    uint32_t l1[8];
#pragma unroll
    for (int i = 0; i < 8; i++) l1[i] = s[2 * i] + s[2 * i + 1];
    uint64_t l2[4];
#pragma unroll
    for (int i = 0; i < 4; i++) l2[i] = (uint64_t)l1[2 * i] + l1[2 * i + 1];
    const uint64_t t = (l2[0] + l2[1]) + (l2[2] + l2[3]);
    const uint32_t q = (uint32_t)(t >> 31);
    uint32_t r = (uint32_t)t - q * F::P;
    r = freduce<F>(r);
    if constexpr ((1ull << 31) + 15ull * ((1ull << 31) - F::P) >= 2ull * F::P) r = freduce<F>(r);
    return r;
}

// S + c x mod p for a small constant c, x and S below p.  v = c x + S < (c+1) p
// is one 64-bit multiply-add; with q = v >> 31 <= c, v - q p lies in
// [0, 2^31 + c (2^31 - p)), which fits 32 bits for c < 2^31 / (2^31 - p)
// (128 for KoalaBear, 16 for BabyBear) and needs one conditional subtraction
// when it is below 2p, two otherwise.  The low word of v + q (2^32 - p) is
// v - q p.
template <class F, uint32_t C>
PG_HD uint32_t fmadd_small(uint32_t x, uint32_t S) {
    static_assert((uint64_t)C * ((1ull << 31) - F::P) < (1ull << 31), "bound needs c < 2^31 / (2^31 - p)");
    const uint64_t v = (uint64_t)x * C + S;
    const uint32_t q = (uint32_t)(v >> 31);
    uint32_t r = (uint32_t)v + q * (0u - F::P);
    r = freduce<F>(r);
    if constexpr ((1ull << 31) + (uint64_t)C * ((1ull << 31) - F::P) > 2ull * F::P) r = freduce<F>(r);
    return r;
}

// S - c x = S + c (p - x)
template <class F, uint32_t C>
PG_HD uint32_t fmsub_small(uint32_t x, uint32_t S) {
    return fmadd_small<F, C>(F::P - x, S);
}

// x 2^-K mod p for x < p, where p = 1 (mod 2^K) (KoalaBear: K <= 24,
// BabyBear: K <= 27).  With m = -x mod 2^K, x + m p = x + m = 0 (mod 2^K),
// and x + m p < 2^K p, so (x + m p) / 2^K is the product, already below p:
// a masked negation, one 64-bit multiply-add and a funnel shift.  The map is
// linear, so it applies to Montgomery representatives as they are.
template <class F, int K>
PG_HD uint32_t fdiv_pow2(uint32_t x) {
    static_assert(K >= 1 && K < 32 && ((F::P - 1u) & ((1u << K) - 1u)) == 0, "needs p = 1 mod 2^K");
    const uint32_t m = (0u - x) & ((1u << K) - 1u);
    const uint64_t v = (uint64_t)m * F::P + x;
    return (uint32_t)(v >> K);
}

// s_i <- s_i * diag_i + S.  Both fields' diagonals (consts.rs:110-154) are
//   -2, 1, 2, 1/2, 3, 4, -1/2, -3, -4, then +-2^-k entries and +-c
// with c = 127 (KoalaBear: entries 11 = -127, 15 = 127) or c = 15 (BabyBear:
// 12 = -15, 15 = 15).  The entries -2 .. 1/2 and -1/2 use additions and a
// halving, the small integers a 64-bit multiply-add (fmadd_small), the
// powers of two fdiv_pow2.  As integers the table holds p - (p-1)/2^k
// (= +2^-k) and (p-1)/2^k (= -2^-k):
//   KoalaBear  9: +2^-8  10: +2^-3  12: -2^-8  13: -2^-3  14: -2^-4
//   BabyBear   9: +2^-8  10: +2^-2  11: +2^-3  13: -2^-8  14: -2^-4
template <class F>
PG_HD void internal_layer(uint32_t (&s)[16]) {
    constexpr bool KB = F::P == KoalaBear::P;
    constexpr uint32_t CBIG = KB ? 127u : 15u;
    const uint32_t sum = sum16<F>(s);
    s[0] = fsub<F>(sum, fadd<F>(s[0], s[0]));
    s[1] = fadd<F>(sum, s[1]);
    s[2] = fadd<F>(sum, fadd<F>(s[2], s[2]));
    s[3] = fadd<F>(sum, fhalve<F>(s[3]));
    s[4] = fmadd_small<F, 3>(s[4], sum);
    s[5] = fmadd_small<F, 4>(s[5], sum);
    s[6] = fsub<F>(sum, fhalve<F>(s[6]));
    s[7] = fmsub_small<F, 3>(s[7], sum);
    s[8] = fmsub_small<F, 4>(s[8], sum);
    if constexpr (KB) {
        s[11] = fmsub_small<F, CBIG>(s[11], sum);
        s[9] = fadd<F>(fdiv_pow2<F, 8>(s[9]), sum);
        s[10] = fadd<F>(fdiv_pow2<F, 3>(s[10]), sum);
        s[12] = fsub<F>(sum, fdiv_pow2<F, 8>(s[12]));
        s[13] = fsub<F>(sum, fdiv_pow2<F, 3>(s[13]));
        s[14] = fsub<F>(sum, fdiv_pow2<F, 4>(s[14]));
    } else {
        s[12] = fmsub_small<F, CBIG>(s[12], sum);
        s[9] = fadd<F>(fdiv_pow2<F, 8>(s[9]), sum);
        s[10] = fadd<F>(fdiv_pow2<F, 2>(s[10]), sum);
        s[11] = fadd<F>(fdiv_pow2<F, 3>(s[11]), sum);
        s[13] = fsub<F>(sum, fdiv_pow2<F, 8>(s[13]));
        s[14] = fsub<F>(sum, fdiv_pow2<F, 4>(s[14]));
    }
    s[15] = fmadd_small<F, CBIG>(s[15], sum);
}

// Internal layer with one Montgomery reduction per element, on lazy inputs
// (s_0 < p, the others < 2p).  With S = sum s_j and S_m = S R (mod p) below
// 2p:
//   s_i <- REDC(Dm_i s_i + S_m),   Dm_i = D_i R (the Montgomery diagonal),
// REDC(v) = v 2^-32 mod p.  v <= (p - 1)(2p - 1) + 2p - 1 < p 2^32, so the
// result lies in [0, 2p) and is D_i s_i + S (mod p) for every entry alike:
// three issue slots per element (a 64-bit multiply-add, a low multiply and
// a 64-bit multiply-add).  The outputs stay lazy (< 2p).
template <class F>
PG_HD void internal_layer_redc(uint32_t (&s)[16]) {
    uint64_t t = s[0];
#pragma unroll
    for (int i = 1; i < 16; i++) t = mad_wide<1>(s[i], t);  // < 31 p < 2^36
    uint32_t sm;
    if constexpr (F::P == KoalaBear::P) {
        // q = t >> 31 <= 30 and t - q p = (t - q 2^31) + q (2^31 - p) < 2^31 + 30 (2^24 - 1) < 2p;
        // then S R = REDC(S_r R^2), S_r R^2 < 2p p < p 2^32
        const uint32_t q = (uint32_t)(t >> 31);
        sm = fmul_lazy<F>((uint32_t)t - q * F::P, F::R2);
    } else {
        // BabyBear's 2^31 - p is too large for that bound: u = REDC(t) =
        // S R^-1 (t < p 2^32, u < 2p), then S R = REDC(u R^3)
        constexpr uint32_t R3 = (uint32_t)((uint64_t)F::R2 * F::R1 % F::P);
        const uint32_t m = (uint32_t)t * F::MU;
        const uint32_t u = (uint32_t)((t + (uint64_t)m * F::P) >> 32);
        sm = fmul_lazy<F>(u, R3);
    }
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint64_t v = mad_u32(s[i], P2<F>::K.diag[i], sm);
        const uint32_t m = (uint32_t)v * F::MU;
        s[i] = (uint32_t)((v + (uint64_t)m * F::P) >> 32);
    }
}

// KoalaBear partial rounds per trip of their loop (the leaf-hash statistics
// count the loop body this many times fewer than the rounds)
constexpr int P2_PARTIAL_UNROLL = 10;

template <class F>
PG_HD void permute(uint32_t (&s)[16]) {
    if constexpr (F::P == KoalaBear::P) {
        // Same rounds.  The constants of a full round are added (as pre_rc,
        // unreduced: the result stays below 2p) before the external layer
        // that precedes the round, so the layer's single reduction covers them.
#pragma unroll
        for (int i = 0; i < 16; i++) s[i] += P2<F>::K.pre_rc[0][i];
        external_layer_wide<F>(s);
#pragma unroll
        for (int r = 0; r < 4; r++) {
#pragma unroll
            for (int i = 0; i < 16; i++) {
                s[i] = sbox<F>(s[i]);
                if (r < 3) s[i] += P2<F>::K.pre_rc[r + 1][i];
            }
            external_layer_wide<F>(s);
        }
        // partial rounds: the internal layer leaves every element below 2p;
        // s_0 is reduced before its round constant and S-box.  They run as
        // a loop of ROUNDS_P / P2_PARTIAL_UNROLL trips of that many unrolled
        // rounds (all 20 unrolled measured 2% slower: the longer code).
        static_assert(F::ROUNDS_P % P2_PARTIAL_UNROLL == 0, "whole trips");
#pragma clang loop unroll(disable)
        for (int r0 = 0; r0 < F::ROUNDS_P; r0 += P2_PARTIAL_UNROLL) {
#pragma clang loop unroll(full)
            for (int r = 0; r < P2_PARTIAL_UNROLL; r++) {
                s[0] = sbox<F>(fadd<F>(freduce<F>(s[0]), P2<F>::K.int_rc[r0 + r]));
                internal_layer_redc<F>(s);
            }
        }
#pragma unroll
        for (int i = 0; i < 16; i++) s[i] = fadd<F>(freduce<F>(s[i]), P2<F>::K.ext_rc[4][i]);
#pragma unroll
        for (int r = 4; r < 8; r++) {
#pragma unroll
            for (int i = 0; i < 16; i++) {
                s[i] = sbox<F>(s[i]);
                if (r < 7) s[i] += P2<F>::K.pre_rc[r + 1][i];
            }
            external_layer_wide<F>(s);
        }
        return;
    }
    external_layer<F>(s);
#pragma unroll
    for (int r = 0; r < 4; r++) {
#pragma unroll
        for (int i = 0; i < 16; i++) s[i] = sbox<F>(fadd<F>(s[i], P2<F>::K.ext_rc[r][i]));
        external_layer<F>(s);
    }
#pragma unroll
    for (int r = 0; r < F::ROUNDS_P; r++) {
        s[0] = sbox<F>(fadd<F>(freduce<F>(s[0]), P2<F>::K.int_rc[r]));
        internal_layer_redc<F>(s);
    }
#pragma unroll
    for (int i = 0; i < 16; i++) s[i] = freduce<F>(s[i]);
#pragma unroll
    for (int r = 4; r < 8; r++) {
#pragma unroll
        for (int i = 0; i < 16; i++) s[i] = sbox<F>(fadd<F>(s[i], P2<F>::K.ext_rc[r][i]));
        external_layer<F>(s);
    }
}

}  // namespace pg
