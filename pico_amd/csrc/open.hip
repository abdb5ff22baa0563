// open.hip -- the device part of TwoAdicFriPcs::open: values of committed
// matrices at out-of-domain points, and the reduced-opening vectors that FRI
// folds.
//
// Opened values.  A matrix committed at blowup 2^b holds f on GEN * <w_N> in
// bit-reversed row order; its first h = N / 2^b rows are f on the coset
// s <w_h> (s = GEN) in bit-reversed order.  For z off the coset
//   f(z) = (z^h - s^h) / (h s^(h-1)) * sum_j f(s w^j) w^j / (z - s w^j),
// a weighted column sum over those h rows (the barycentric form; f(z) is
// unique, so any correct evaluation is bit-exact).
//
// Reduced openings (circuit/fri.rs:118-241 is the verifier side).  For each
// LDE log-height, every row i with x_i = GEN * w_N^{rev(i)} accumulates, over
// the matrices of that height in round / matrix / point order,
//   alpha^off * (sum_k alpha^k p_k(z) - sum_k alpha^k M[i][k]) / (z - x_i),
// where off counts the columns already reduced at that height (one counter
// per height, advanced by the width per point).  The host folds
// alpha^off * sum_k alpha^k p_k(z) into one constant per (matrix, point);
// the kernel forms r_i = sum_k alpha^k M[i][k] once per matrix and row.
#include "field.h"
#include "kernels.h"

namespace pg {

// Ext4 values in global memory are stored as native 4-vectors (copyable
// across address spaces); these convert at the load / store.
__device__ __forceinline__ Ext4 ext_of(u32x4 v) { return Ext4{{v.x, v.y, v.z, v.w}}; }
__device__ __forceinline__ u32x4 v4_of(const Ext4& e) { return u32x4{e.c[0], e.c[1], e.c[2], e.c[3]}; }
typedef GLOBAL u32x4* gext;
typedef const GLOBAL u32x4* gcext;

// c_r = w^j / (z - s w^j), j = rev_h(r): the weight of committed row r
template <class F>
__global__ void bary_weights_kernel(Ext4 z, uint32_t shift, int log_h, gcptr __restrict__ tw,
                                    gext __restrict__ out) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= (1u << log_h)) return;
    const uint32_t j = log_h ? __brev(r) >> (32 - log_h) : 0u;
    const uint32_t wj = tw[j];
    Ext4 d = z;
    d.c[0] = fsub<F>(d.c[0], fmul<F>(shift, wj));
    out[r] = v4_of(escale<F>(einv<F>(d), wj));
}

// partial[chunk][col] = sum over rows r of the chunk of M[r][col] * c_r
template <class F>
__global__ void __launch_bounds__(256) column_dot_kernel(gcptr __restrict__ m, uint64_t ld, uint32_t w,
                                                         gcext __restrict__ c, uint32_t rows,
                                                         uint32_t chunk_rows, uint64_t nthreads,
                                                         gext __restrict__ partial) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nthreads) return;
    const uint32_t col = (uint32_t)(t % w), chunk = (uint32_t)(t / w);
    const uint32_t r0 = chunk * chunk_rows, r1 = r0 + chunk_rows < rows ? r0 + chunk_rows : rows;
    uint32_t acc[4] = {0, 0, 0, 0};
    for (uint32_t r = r0; r < r1; r++) {
        const uint32_t v = m[(uint64_t)r * ld + col];
        const Ext4 cr = ext_of(c[r]);
#pragma unroll
        for (int q = 0; q < 4; q++) acc[q] = fadd<F>(acc[q], fmul<F>(v, cr.c[q]));
    }
    partial[t] = u32x4{acc[0], acc[1], acc[2], acc[3]};
}

// out[col] = factor * sum over chunks of partial[chunk][col]
template <class F>
__global__ void column_sum_kernel(gcext __restrict__ partial, uint32_t nchunks, uint32_t w,
                                  Ext4 factor, gext __restrict__ out) {
    const uint32_t col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= w) return;
    Ext4 s{{0, 0, 0, 0}};
    for (uint32_t k = 0; k < nchunks; k++) s = eadd<F>(s, ext_of(partial[(uint64_t)k * w + col]));
    out[col] = v4_of(emul<F>(s, factor));
}

// Reduced opening of one log-height: thread i owns LDE row i.
template <class F>
__global__ void __launch_bounds__(256) reduce_rows_kernel(const RoMat* __restrict__ mats, int nmats,
                                                          const Ext4* __restrict__ pts, int npts,
                                                          gcext __restrict__ alpha_pows,
                                                          uint32_t shift, int log_n, gcptr __restrict__ tw,
                                                          gext __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (1u << log_n)) return;
    const uint32_t x = fmul<F>(shift, tw[log_n ? __brev(i) >> (32 - log_n) : 0u]);
    Ext4 inv[RO_MAX_POINTS];
#pragma unroll
    for (int q = 0; q < RO_MAX_POINTS; q++) {
        if (q < npts) {
            Ext4 d = pts[q];
            d.c[0] = fsub<F>(d.c[0], x);
            inv[q] = einv<F>(d);
        } else {
            inv[q] = Ext4{{0, 0, 0, 0}};
        }
    }
    Ext4 ro{{0, 0, 0, 0}};
    for (int k = 0; k < nmats; k++) {
        const RoMat* mk = mats + k;  // read in place: uniform, scalar loads
        const uint32_t width = mk->width;
        const gcptr row = (gcptr)mk->data + (uint64_t)i * mk->ld;
        uint32_t acc[4] = {0, 0, 0, 0};
        uint32_t col = 0;
        if (mk->vec4) {  // 16-byte aligned rows: four columns per load
            for (; col + 4 <= width; col += 4) {
                const u32x4 v = *reinterpret_cast<const GLOBAL u32x4*>(row + col);
                const uint32_t vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const Ext4 ap = ext_of(alpha_pows[col + u]);
#pragma unroll
                    for (int q = 0; q < 4; q++) acc[q] = fadd<F>(acc[q], fmul<F>(vv[u], ap.c[q]));
                }
            }
        }
        for (; col < width; col++) {
            const uint32_t v = row[col];
            const Ext4 ap = ext_of(alpha_pows[col]);
#pragma unroll
            for (int q = 0; q < 4; q++) acc[q] = fadd<F>(acc[q], fmul<F>(v, ap.c[q]));
        }
        const Ext4 r = Ext4{{acc[0], acc[1], acc[2], acc[3]}};
        const int np = mk->npts;
#pragma unroll
        for (int q = 0; q < 2; q++) {
            if (q >= np) break;
            // (A - alpha^off r) / (z - x); the inverse is picked by comparing
            // against each fixed slot, which keeps inv[] in registers
            const int pt = mk->pt[q];
            Ext4 iv = inv[0];
#pragma unroll
            for (int u = 1; u < RO_MAX_POINTS; u++)
                if (pt == u) iv = inv[u];
            const Ext4 term = esub<F>(mk->a[q], emul<F>(mk->aoff[q], r));
            ro = eadd<F>(ro, emul<F>(term, iv));
        }
    }
    out[i] = v4_of(ro);
}

static unsigned blocks_for(uint64_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

template <class F>
static hipError_t open_values_t(const uint32_t* m, uint64_t ld, uint32_t w, int log_h, const uint32_t* tw,
                                const Ext4& z, const Ext4& factor, Ext4* weights, Ext4* partial,
                                uint32_t nchunks, uint32_t chunk_rows, Ext4* out, hipStream_t s) {
    const uint32_t h = 1u << log_h;
    bary_weights_kernel<F><<<blocks_for(h, 256), 256, 0, s>>>(z, to_monty<F>(F::GEN), log_h, (gcptr)tw,
                                                                (gext)weights);
    const uint64_t nt = (uint64_t)nchunks * w;
    column_dot_kernel<F><<<blocks_for(nt, 256), 256, 0, s>>>((gcptr)m, ld, w, (gcext)weights, h, chunk_rows, nt,
                                                              (gext)partial);
    column_sum_kernel<F><<<blocks_for(w, 256), 256, 0, s>>>((gcext)partial, nchunks, w, factor, (gext)out);
    return hipGetLastError();
}

hipError_t launch_open_values(int field, const uint32_t* m, uint64_t ld, uint32_t w, int log_h,
                              const uint32_t* tw, const Ext4& z, const Ext4& factor, Ext4* weights,
                              Ext4* partial, uint32_t nchunks, uint32_t chunk_rows, Ext4* out,
                              hipStream_t stream) {
    if (w == 0 || nchunks == 0) return hipSuccess;
    return field == FIELD_KOALABEAR
               ? open_values_t<KoalaBear>(m, ld, w, log_h, tw, z, factor, weights, partial, nchunks, chunk_rows,
                                          out, stream)
               : open_values_t<BabyBear>(m, ld, w, log_h, tw, z, factor, weights, partial, nchunks, chunk_rows,
                                         out, stream);
}

hipError_t launch_reduce_rows(int field, const RoMat* mats, int nmats, const Ext4* pts, int npts,
                              const Ext4* alpha_pows, int log_n, const uint32_t* tw, Ext4* out,
                              hipStream_t stream) {
    const uint64_t n = 1ull << log_n;
    if (field == FIELD_KOALABEAR)
        reduce_rows_kernel<KoalaBear><<<blocks_for(n, 256), 256, 0, stream>>>(
            mats, nmats, pts, npts, (gcext)alpha_pows, to_monty<KoalaBear>(KoalaBear::GEN), log_n,
            (gcptr)tw, (gext)out);
    else
        reduce_rows_kernel<BabyBear><<<blocks_for(n, 256), 256, 0, stream>>>(
            mats, nmats, pts, npts, (gcext)alpha_pows, to_monty<BabyBear>(BabyBear::GEN), log_n,
            (gcptr)tw, (gext)out);
    return hipGetLastError();
}

}  // namespace pg
