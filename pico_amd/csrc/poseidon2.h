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
    uint32_t diag[16];
};

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

template <class F>
PG_HD uint32_t sbox(uint32_t x) {
    uint32_t x2 = fmul<F>(x, x);
    if constexpr (F::SBOX == 3) {
        return fmul<F>(x2, x);
    } else {
        uint32_t x3 = fmul<F>(x2, x);
        uint32_t x4 = fmul<F>(x2, x2);
        return fmul<F>(x4, x3);
    }
}

// y = M4 x with M4 = [[2,3,1,1],[1,2,3,1],[1,1,2,3],[3,1,1,2]]
template <class F>
PG_HD void m4(uint32_t& a, uint32_t& b, uint32_t& c, uint32_t& d) {
    uint32_t ab = fadd<F>(a, b), cd = fadd<F>(c, d), s = fadd<F>(ab, cd);
    uint32_t y0 = fadd<F>(fadd<F>(s, ab), b);           // 2a+3b+ c+ d
    uint32_t y1 = fadd<F>(fadd<F>(s, b), fadd<F>(c, c)); //  a+2b+3c+ d
    uint32_t y2 = fadd<F>(fadd<F>(s, cd), d);           //  a+ b+2c+3d
    uint32_t y3 = fadd<F>(fadd<F>(s, d), fadd<F>(a, a)); // 3a+ b+ c+2d
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

template <class F>
PG_HD void internal_layer(uint32_t (&s)[16]) {
    uint32_t sum = s[0];
#pragma unroll
    for (int i = 1; i < 16; i++) sum = fadd<F>(sum, s[i]);
#pragma unroll
    for (int i = 0; i < 16; i++) s[i] = fadd<F>(fmul<F>(s[i], P2<F>::K.diag[i]), sum);
}

template <class F>
PG_HD void permute(uint32_t (&s)[16]) {
    external_layer<F>(s);
#pragma unroll
    for (int r = 0; r < 4; r++) {
#pragma unroll
        for (int i = 0; i < 16; i++) s[i] = sbox<F>(fadd<F>(s[i], P2<F>::K.ext_rc[r][i]));
        external_layer<F>(s);
    }
#pragma unroll
    for (int r = 0; r < F::ROUNDS_P; r++) {
        s[0] = sbox<F>(fadd<F>(s[0], P2<F>::K.int_rc[r]));
        internal_layer<F>(s);
    }
#pragma unroll
    for (int r = 4; r < 8; r++) {
#pragma unroll
        for (int i = 0; i < 16; i++) s[i] = sbox<F>(fadd<F>(s[i], P2<F>::K.ext_rc[r][i]));
        external_layer<F>(s);
    }
}

}  // namespace pg
