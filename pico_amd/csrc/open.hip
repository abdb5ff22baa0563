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

// partial[q][chunk][col] = sum over rows r of the chunk of M[r][col] c_q[r]
// for every point q at once (the matrix is read once for all points)
constexpr int OV_MAX_POINTS = 4;
template <class F, int NP>
__global__ void __launch_bounds__(256) column_dot_kernel(gcptr __restrict__ m, uint64_t ld, uint32_t w,
                                                         gcext __restrict__ c, uint32_t rows,
                                                         uint32_t chunk_rows, uint64_t nthreads,
                                                         gext __restrict__ partial) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nthreads) return;
    const uint32_t col = (uint32_t)(t % w), chunk = (uint32_t)(t / w);
    const uint32_t r0 = chunk * chunk_rows, r1 = r0 + chunk_rows < rows ? r0 + chunk_rows : rows;
    uint32_t acc[NP][4] = {};
    for (uint32_t r = r0; r < r1; r++) {
        const uint32_t v = m[(uint64_t)r * ld + col];
#pragma unroll
        for (int q = 0; q < NP; q++) {
            const Ext4 cr = ext_of(c[(uint64_t)q * rows + r]);
#pragma unroll
            for (int e = 0; e < 4; e++) acc[q][e] = fadd<F>(acc[q][e], fmul<F>(v, cr.c[e]));
        }
    }
#pragma unroll
    for (int q = 0; q < NP; q++) partial[(uint64_t)q * nthreads + t] = u32x4{acc[q][0], acc[q][1], acc[q][2], acc[q][3]};
}

// out[col] = factor * sum over chunks of partial[chunk][col]: one 256-thread
// block per column, a strided pass then a tree in LDS
template <class F>
__global__ void __launch_bounds__(256) column_sum_kernel(gcext __restrict__ partial, uint32_t nchunks, uint32_t w,
                                                         Ext4 factor, gext __restrict__ out) {
    __shared__ uint32_t red[4][256];
    const uint32_t col = blockIdx.x, t = threadIdx.x;
    Ext4 s{{0, 0, 0, 0}};
    for (uint32_t k = t; k < nchunks; k += 256) s = eadd<F>(s, ext_of(partial[(uint64_t)k * w + col]));
#pragma unroll
    for (int q = 0; q < 4; q++) red[q][t] = s.c[q];
    __syncthreads();
    for (uint32_t half = 128; half > 0; half >>= 1) {
        if (t < half)
#pragma unroll
            for (int q = 0; q < 4; q++) red[q][t] = fadd<F>(red[q][t], red[q][t + half]);
        __syncthreads();
    }
    if (t == 0) out[col] = v4_of(emul<F>(Ext4{{red[0][0], red[1][0], red[2][0], red[3][0]}}, factor));
}

// Reduced opening of one log-height: a 256-thread block owns 256 rows,
// thread t row base + t.  The rows' columns pass through LDS in tiles of
// 256 rows x 16 columns, loaded so that each wave reads 4 rows x 64
// contiguous bytes (coalesced) and read back one row per thread (a row
// stride of 17 words keeps the banks distinct).
constexpr int RO_ROWS = 256, RO_COLS = 16;
template <class F>
__global__ void __launch_bounds__(RO_ROWS) reduce_rows_kernel(const RoMat* __restrict__ mats, int nmats,
                                                              const Ext4* __restrict__ pts, int npts,
                                                              gcext __restrict__ alpha_pows, uint32_t shift,
                                                              int log_n, gcptr __restrict__ tw,
                                                              gext __restrict__ out) {
    __shared__ uint32_t tile[RO_ROWS * (RO_COLS + 1)];
    const uint32_t n = 1u << log_n, t = threadIdx.x;
    const uint32_t base = blockIdx.x * RO_ROWS, i = base + t;
    const bool live = i < n;
    const uint32_t ir = live ? i : n - 1;
    const uint32_t x = fmul<F>(shift, tw[log_n ? __brev(ir) >> (32 - log_n) : 0u]);
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
    // loader coordinates: thread t loads column t % 16 of rows t / 16 + 16 k
    const uint32_t lc = t % RO_COLS, lr = t / RO_COLS;
    Ext4 ro{{0, 0, 0, 0}};
    for (int k = 0; k < nmats; k++) {
        const RoMat* mk = mats + k;  // read in place: uniform, scalar loads
        const uint32_t width = mk->width;
        const uint64_t ld = mk->ld;
        const gcptr m = (gcptr)mk->data;
        uint32_t acc[4] = {0, 0, 0, 0};
        for (uint32_t c0 = 0; c0 < width; c0 += RO_COLS) {
            const uint32_t col = c0 + lc;
#pragma unroll
            for (int kk = 0; kk < RO_ROWS / (RO_ROWS / RO_COLS); kk++) {
                const uint32_t r = lr + (RO_ROWS / RO_COLS) * kk;
                const uint32_t row = base + r < n ? base + r : n - 1;
                tile[r * (RO_COLS + 1) + lc] = col < width ? m[(uint64_t)row * ld + col] : 0u;
            }
            __syncthreads();
            const uint32_t cn = width - c0 < (uint32_t)RO_COLS ? width - c0 : (uint32_t)RO_COLS;
            for (uint32_t c = 0; c < cn; c++) {
                const uint32_t v = tile[t * (RO_COLS + 1) + c];
                const Ext4 ap = ext_of(alpha_pows[c0 + c]);
#pragma unroll
                for (int q = 0; q < 4; q++) acc[q] = fadd<F>(acc[q], fmul<F>(v, ap.c[q]));
            }
            __syncthreads();
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
    if (live) out[i] = v4_of(ro);
}

static unsigned blocks_for(uint64_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

template <class F>
static hipError_t open_values_t(const uint32_t* m, uint64_t ld, uint32_t w, int log_h, const uint32_t* tw,
                                const Ext4* z, const Ext4* factor, int npts, Ext4* weights, Ext4* partial,
                                uint32_t nchunks, uint32_t chunk_rows, Ext4* out, hipStream_t s) {
    const uint32_t h = 1u << log_h;
    for (int q = 0; q < npts; q++)
        bary_weights_kernel<F><<<blocks_for(h, 256), 256, 0, s>>>(z[q], to_monty<F>(F::GEN), log_h, (gcptr)tw,
                                                                    (gext)(weights + (uint64_t)q * h));
    const uint64_t nt = (uint64_t)nchunks * w;
    switch (npts) {
        case 1: column_dot_kernel<F, 1><<<blocks_for(nt, 256), 256, 0, s>>>((gcptr)m, ld, w, (gcext)weights, h,
                                                                          chunk_rows, nt, (gext)partial); break;
        case 2: column_dot_kernel<F, 2><<<blocks_for(nt, 256), 256, 0, s>>>((gcptr)m, ld, w, (gcext)weights, h,
                                                                          chunk_rows, nt, (gext)partial); break;
        case 3: column_dot_kernel<F, 3><<<blocks_for(nt, 256), 256, 0, s>>>((gcptr)m, ld, w, (gcext)weights, h,
                                                                          chunk_rows, nt, (gext)partial); break;
        default: column_dot_kernel<F, 4><<<blocks_for(nt, 256), 256, 0, s>>>((gcptr)m, ld, w, (gcext)weights, h,
                                                                           chunk_rows, nt, (gext)partial); break;
    }
    for (int q = 0; q < npts; q++)
        column_sum_kernel<F><<<w, 256, 0, s>>>((gcext)(partial + (uint64_t)q * nt), nchunks, w, factor[q],
                                               (gext)(out + (uint64_t)q * w));
    return hipGetLastError();
}

hipError_t launch_open_values(int field, const uint32_t* m, uint64_t ld, uint32_t w, int log_h,
                              const uint32_t* tw, const Ext4* z, const Ext4* factor, int npts, Ext4* weights,
                              Ext4* partial, uint32_t nchunks, uint32_t chunk_rows, Ext4* out,
                              hipStream_t stream) {
    if (w == 0 || nchunks == 0 || npts <= 0) return hipSuccess;
    if (npts > OV_MAX_POINTS) return hipErrorInvalidValue;
    return field == FIELD_KOALABEAR
               ? open_values_t<KoalaBear>(m, ld, w, log_h, tw, z, factor, npts, weights, partial, nchunks,
                                          chunk_rows, out, stream)
               : open_values_t<BabyBear>(m, ld, w, log_h, tw, z, factor, npts, weights, partial, nchunks,
                                         chunk_rows, out, stream);
}

hipError_t launch_reduce_rows(int field, const RoMat* mats, int nmats, const Ext4* pts, int npts,
                              const Ext4* alpha_pows, int log_n, const uint32_t* tw, Ext4* out,
                              hipStream_t stream) {
    const uint64_t n = 1ull << log_n;
    if (field == FIELD_KOALABEAR)
        reduce_rows_kernel<KoalaBear><<<blocks_for(n, RO_ROWS), RO_ROWS, 0, stream>>>(
            mats, nmats, pts, npts, (gcext)alpha_pows, to_monty<KoalaBear>(KoalaBear::GEN), log_n,
            (gcptr)tw, (gext)out);
    else
        reduce_rows_kernel<BabyBear><<<blocks_for(n, RO_ROWS), RO_ROWS, 0, stream>>>(
            mats, nmats, pts, npts, (gcext)alpha_pows, to_monty<BabyBear>(BabyBear::GEN), log_n,
            (gcptr)tw, (gext)out);
    return hipGetLastError();
}

}  // namespace pg
