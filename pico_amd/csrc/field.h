// field.h -- KoalaBear / BabyBear arithmetic in Montgomery form (R = 2^32)
// for gfx950 kernels and for the host-side challenger.
//
// In-memory representation follows p3 MontyField31 (x stored as x*2^32 mod p),
// so device buffers can alias the reference's RowMajorMatrix<KoalaBear> data
// without conversion.  Primes: KoalaBear p = 2^31 - 2^24 + 1
// (gnark/koalabear/koalabear.go:20), BabyBear p = 2^31 - 2^27 + 1.
// Extension: F[x]/(x^4 - W), W = 3 (KoalaBear, vm/src/primitives/consts.rs:134),
// W = 11 (BabyBear, consts.rs:107).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PG_HD __host__ __device__ __forceinline__

namespace pg {

constexpr uint32_t inv_mod_2_32(uint32_t p) {
    uint32_t x = p;  // Newton iteration: x <- x (2 - p x), doubling correct bits
    for (int i = 0; i < 5; i++) x *= 2u - p * x;
    return x;
}

template <uint32_t P_, int TWO_ADICITY_, uint32_t GEN_, uint32_t W_, int RP_, int SBOX_>
struct MontyField {
    static constexpr uint32_t P = P_;
    static constexpr uint32_t MU = 0u - inv_mod_2_32(P_);  // -p^-1 mod 2^32
    static constexpr uint32_t R1 = (uint32_t)((1ull << 32) % P_);  // Montgomery one
    static constexpr uint32_t R2 = (uint32_t)(((uint64_t)R1 * R1) % P_);
    static constexpr int TWO_ADICITY = TWO_ADICITY_;
    static constexpr uint32_t GEN = GEN_;       // canonical multiplicative generator
    static constexpr uint32_t W = W_;           // extension non-residue (canonical)
    static constexpr int ROUNDS_P = RP_;        // Poseidon2 partial rounds
    static constexpr int SBOX = SBOX_;          // Poseidon2 S-box degree
};

// GEN = 3 / 31: the recursion builder's generator()
// (vm/src/compiler/recursion/ir/utils.rs:13-19).
// KoalaBear: 8 full + 20 partial rounds, x^3 (consts.rs:191-203).
using KoalaBear = MontyField<0x7f000001u, 24, 3u, 3u, 20, 3>;
// BabyBear: 8 full + 13 partial rounds, x^7 (consts.rs:176-186).
using BabyBear = MontyField<0x78000001u, 27, 31u, 11u, 13, 7>;

enum FieldId : int { FIELD_KOALABEAR = 0, FIELD_BABYBEAR = 1 };

// ---- base field --------------------------------------------------------
// Inputs in [0, p).  p < 2^31 so a + b never overflows 32 bits; the unsigned
// min picks the reduced value (r - p wraps above r when r < p).
template <class F>
PG_HD uint32_t fadd(uint32_t a, uint32_t b) {
    uint32_t r = a + b, t = r - F::P;
    return t < r ? t : r;
}
template <class F>
PG_HD uint32_t fsub(uint32_t a, uint32_t b) {
    uint32_t r = a - b, t = r + F::P;
    return t < r ? t : r;
}
template <class F>
PG_HD uint32_t fneg(uint32_t a) {
    return a == 0 ? 0 : F::P - a;
}
// Montgomery product a*b*2^-32 mod p.  t = a*b < p^2; m = t*(-p^-1) mod 2^32;
// (t + m p) / 2^32 < 2p fits 32 bits after one conditional subtraction.
template <class F>
PG_HD uint32_t fmul(uint32_t a, uint32_t b) {
    uint64_t t = (uint64_t)a * b;
    uint32_t m = (uint32_t)t * F::MU;
    uint32_t u = (uint32_t)((t + (uint64_t)m * F::P) >> 32);
    uint32_t s = u - F::P;
    return s < u ? s : u;
}
// Montgomery product without the final subtraction.  For a < 2p, b < p the
// result is below (2p^2 + 2^32 p) / 2^32 < 2p; for a, b < p below 1.5p.
template <class F>
PG_HD uint32_t fmul_lazy(uint32_t a, uint32_t b) {
    uint64_t t = (uint64_t)a * b;
    uint32_t m = (uint32_t)t * F::MU;
    return (uint32_t)((t + (uint64_t)m * F::P) >> 32);
}
// t + k x for 32-bit x and a small constant k.  On the device this is one
// v_mad_u64_u32: written as a 64-bit add of the zero-extended x, the compiler
// spends a move clearing the high word of a register pair per addend.
template <uint32_t K>
PG_HD uint64_t mad_wide(uint32_t x, uint64_t t) {
#ifdef __HIP_DEVICE_COMPILE__
    uint64_t r, carry;  // the carry-out (an SGPR pair) is zero here: t + k x < 2^64
    asm("v_mad_u64_u32 %0, %1, %2, %3, %4" : "=v"(r), "=s"(carry) : "v"(x), "i"(K), "v"(t));
    (void)carry;
    return r;
#else
    return t + (uint64_t)K * x;
#endif
}
// a b + t for 32-bit a, b: one v_mad_u64_u32 on the device, where the
// compiler turns a product by a constant power of two into a 64-bit shift
// and a 64-bit add (v_lshl_add_u64 shifts by at most 7).  b must be the same
// in every lane of a wave (a constant): it is passed in a scalar register.
PG_HD uint64_t mad_u32(uint32_t a, uint32_t b, uint64_t t) {
#ifdef __HIP_DEVICE_COMPILE__
    uint64_t r, carry;  // the carry-out (an SGPR pair) is zero here: a b + t < 2^64 for t < 2^33
    asm("v_mad_u64_u32 %0, %1, %2, %3, %4" : "=v"(r), "=s"(carry) : "v"(a), "s"(b), "v"(t));
    (void)carry;
    return r;
#else
    return (uint64_t)a * b + t;
#endif
}
// [0, 2p) -> [0, p)
template <class F>
PG_HD uint32_t freduce(uint32_t u) {
    uint32_t s = u - F::P;
    return s < u ? s : u;
}
// Lazy sum of products of reduced values (a, b < p < 2^31, so a b < 2^62):
// each product is one 64-bit multiply-add into t.  fold() rewrites
// t = hi 2^32 + lo as hi R1 + lo (R1 = 2^32 mod p), below 2^32 (R1 + 1),
// which leaves room for three more products.  value() folds once more and
// applies Montgomery reduction (valid for t < p 2^32): for Montgomery inputs
// it is the Montgomery form of sum a b, the word a chain of fmul and fadd
// gives.  Per product about 1.5 issue slots instead of ~7.
template <class F>
struct LazyDot {
    static_assert((unsigned __int128)(((uint64_t)F::R1 + 1) << 32) + 3 * (unsigned __int128)(F::P - 1) * (F::P - 1) <
                      ((unsigned __int128)1 << 64),
                  "three products after a fold must fit 64 bits");
    static_assert((((uint64_t)F::R1 + 1) << 32) <= (uint64_t)F::P << 32, "a folded sum must be below p 2^32");
    uint64_t t = 0;
    PG_HD void add(uint32_t a, uint32_t b) { t += (uint64_t)a * b; }
    PG_HD void fold() { t = (uint64_t)(uint32_t)(t >> 32) * F::R1 + (uint32_t)t; }
    PG_HD uint32_t value() const {
        const uint64_t u = (uint64_t)(uint32_t)(t >> 32) * F::R1 + (uint32_t)t;
        const uint32_t m = (uint32_t)u * F::MU;
        return freduce<F>((uint32_t)((u + (uint64_t)m * F::P) >> 32));
    }
};
// a / 2 for a < p (linear, so any representation): (a >> 1) + (a odd) (p+1)/2
template <class F>
PG_HD uint32_t fhalve(uint32_t a) {
    const uint32_t odd = 0u - (a & 1u);
    return (a >> 1) + (odd & ((F::P + 1) >> 1));
}
template <class F>
PG_HD uint32_t to_monty(uint32_t a) {
    return fmul<F>(a, F::R2);
}
template <class F>
PG_HD uint32_t from_monty(uint32_t a) {
    return fmul<F>(a, 1u);
}
template <class F>
PG_HD uint32_t fpow(uint32_t a, uint64_t e) {  // Montgomery in, Montgomery out
    uint32_t r = F::R1;
    while (e) {
        if (e & 1) r = fmul<F>(r, a);
        a = fmul<F>(a, a);
        e >>= 1;
    }
    return r;
}
template <class F>
PG_HD uint32_t finv(uint32_t a) {
    return fpow<F>(a, F::P - 2);
}
// canonical-valued constants at compile time
constexpr uint32_t cpow_canon(uint32_t p, uint32_t a, uint64_t e) {
    uint64_t r = 1, x = a % p;
    while (e) {
        if (e & 1) r = (r * x) % p;
        x = (x * x) % p;
        e >>= 1;
    }
    return (uint32_t)r;
}
template <class F>
constexpr uint32_t cmonty(uint32_t canon) {
    return (uint32_t)(((uint64_t)(canon % F::P) << 32) % F::P);
}
// two_adic_generator(k) = GEN^((p-1) >> k), canonical
template <class F>
constexpr uint32_t two_adic_gen_canon(int log_n) {
    return cpow_canon(F::P, F::GEN, (uint64_t)(F::P - 1) >> log_n);
}

// ---- degree-4 extension (Montgomery coefficients c0..c3) -----------------
struct Ext4 {
    uint32_t c[4];
};
template <class F>
PG_HD Ext4 eadd(const Ext4& a, const Ext4& b) {
    Ext4 r;
    for (int i = 0; i < 4; i++) r.c[i] = fadd<F>(a.c[i], b.c[i]);
    return r;
}
template <class F>
PG_HD Ext4 esub(const Ext4& a, const Ext4& b) {
    Ext4 r;
    for (int i = 0; i < 4; i++) r.c[i] = fsub<F>(a.c[i], b.c[i]);
    return r;
}
template <class F>
PG_HD Ext4 escale(const Ext4& a, uint32_t s) {
    Ext4 r;
    for (int i = 0; i < 4; i++) r.c[i] = fmul<F>(a.c[i], s);
    return r;
}
// Schoolbook product with x^4 = W.  W*y for W = 3 or 11 is done by additions
// (linear, so it is representation independent).
template <class F>
PG_HD uint32_t mul_by_w(uint32_t y) {
    uint32_t y2 = fadd<F>(y, y);
    if (F::W == 3) return fadd<F>(y2, y);
    uint32_t y4 = fadd<F>(y2, y2), y8 = fadd<F>(y4, y4);
    return fadd<F>(fadd<F>(y8, y2), y);  // 11 = 8 + 2 + 1
}
template <class F>
PG_HD Ext4 emul(const Ext4& a, const Ext4& b) {
    const uint32_t* x = a.c;
    const uint32_t* y = b.c;
    uint32_t c0 = fmul<F>(x[0], y[0]);
    uint32_t c1 = fadd<F>(fmul<F>(x[0], y[1]), fmul<F>(x[1], y[0]));
    uint32_t c2 = fadd<F>(fadd<F>(fmul<F>(x[0], y[2]), fmul<F>(x[1], y[1])), fmul<F>(x[2], y[0]));
    uint32_t c3 = fadd<F>(fadd<F>(fmul<F>(x[0], y[3]), fmul<F>(x[1], y[2])),
                          fadd<F>(fmul<F>(x[2], y[1]), fmul<F>(x[3], y[0])));
    uint32_t c4 = fadd<F>(fadd<F>(fmul<F>(x[1], y[3]), fmul<F>(x[2], y[2])), fmul<F>(x[3], y[1]));
    uint32_t c5 = fadd<F>(fmul<F>(x[2], y[3]), fmul<F>(x[3], y[2]));
    uint32_t c6 = fmul<F>(x[3], y[3]);
    Ext4 r;
    r.c[0] = fadd<F>(c0, mul_by_w<F>(c4));
    r.c[1] = fadd<F>(c1, mul_by_w<F>(c5));
    r.c[2] = fadd<F>(c2, mul_by_w<F>(c6));
    r.c[3] = c3;
    return r;
}

// Inverse in F[x]/(x^4 - W).  With y = x^2, a = E(y) + x O(y) and
// a(-x) = E - x O, the product a a(-x) = E^2 - y O^2 = B0 + B1 y lies in
// F[y]/(y^2 - W), where (B0 + B1 y)^-1 = (B0 - B1 y) / (B0^2 - W B1^2).  So
// a^-1 = a(-x) (B0 - B1 y) / N: one base-field inversion.  a = 0 gives 0.
template <class F>
PG_HD Ext4 einv(const Ext4& a) {
    const uint32_t a0 = a.c[0], a1 = a.c[1], a2 = a.c[2], a3 = a.c[3];
    // B0 = a0^2 + W a2^2 - 2 W a1 a3,  B1 = 2 a0 a2 - a1^2 - W a3^2
    const uint32_t a1a3 = fmul<F>(a1, a3);
    const uint32_t b0 = fsub<F>(fadd<F>(fmul<F>(a0, a0), mul_by_w<F>(fmul<F>(a2, a2))),
                                mul_by_w<F>(fadd<F>(a1a3, a1a3)));
    const uint32_t a0a2 = fmul<F>(a0, a2);
    const uint32_t b1 = fsub<F>(fsub<F>(fadd<F>(a0a2, a0a2), fmul<F>(a1, a1)), mul_by_w<F>(fmul<F>(a3, a3)));
    const uint32_t n = fsub<F>(fmul<F>(b0, b0), mul_by_w<F>(fmul<F>(b1, b1)));
    const uint32_t ninv = finv<F>(n);
    const uint32_t c0 = fmul<F>(b0, ninv), c1 = fmul<F>(fsub<F>(0u, b1), ninv);  // (B0 - B1 y) / N
    // a(-x) = a0 - a1 x + a2 x^2 - a3 x^3 times c0 + c1 x^2
    const uint32_t m1 = fsub<F>(0u, a1), m3 = fsub<F>(0u, a3);
    Ext4 r;
    r.c[0] = fadd<F>(fmul<F>(a0, c0), mul_by_w<F>(fmul<F>(a2, c1)));
    r.c[1] = fadd<F>(fmul<F>(m1, c0), mul_by_w<F>(fmul<F>(m3, c1)));
    r.c[2] = fadd<F>(fmul<F>(a2, c0), fmul<F>(a0, c1));
    r.c[3] = fadd<F>(fmul<F>(m3, c0), fmul<F>(m1, c1));
    return r;
}
template <class F>
PG_HD Ext4 epow(Ext4 a, uint64_t e) {
    Ext4 r{{F::R1, 0u, 0u, 0u}};
    while (e) {
        if (e & 1) r = emul<F>(r, a);
        a = emul<F>(a, a);
        e >>= 1;
    }
    return r;
}

// reverse the low `bits` bits of x (host and device; kernels use __brev)
PG_HD uint32_t bit_rev(uint32_t x, int bits) {
    x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
    x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
    x = ((x >> 4) & 0x0f0f0f0fu) | ((x & 0x0f0f0f0fu) << 4);
    x = ((x >> 8) & 0x00ff00ffu) | ((x & 0x00ff00ffu) << 8);
    x = (x >> 16) | (x << 16);
    return bits == 0 ? 0u : x >> (32 - bits);
}

}  // namespace pg
