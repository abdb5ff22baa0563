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

bool open_values_vec4(const uint32_t* m, uint64_t ld, uint32_t w, int npts);

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
    LazyDot<F> acc[NP][4];
    auto fold_all = [&]() {
#pragma unroll
        for (int q = 0; q < NP; q++)
#pragma unroll
            for (int e = 0; e < 4; e++) acc[q][e].fold();
    };
    // Rows go in groups of UR whose loads (values and weights) are all
    // issued before the products that use them, so each thread keeps UR
    // memory round trips in flight instead of one.  A fold follows every
    // third product of an accumulator.
    constexpr int UR = 4;
    uint32_t r = r0;
    for (; r + UR <= r1; r += UR) {
        uint32_t v[UR];
        Ext4 cr[UR][NP];
#pragma unroll
        for (int u = 0; u < UR; u++) {
            v[u] = m[(uint64_t)(r + u) * ld + col];
#pragma unroll
            for (int q = 0; q < NP; q++) cr[u][q] = ext_of(c[(uint64_t)q * rows + r + u]);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < UR; u++) {
#pragma unroll
            for (int q = 0; q < NP; q++)
#pragma unroll
                for (int e = 0; e < 4; e++) acc[q][e].add(v[u], cr[u][q].c[e]);
            if (u % 3 == 2 || u == UR - 1) fold_all();
        }
    }
    for (uint32_t k = 0; r < r1; r++, k++) {  // fewer than UR rows left (short matrices)
        const uint32_t v = m[(uint64_t)r * ld + col];
#pragma unroll
        for (int q = 0; q < NP; q++) {
            const Ext4 cr = ext_of(c[(uint64_t)q * rows + r]);
#pragma unroll
            for (int e = 0; e < 4; e++) acc[q][e].add(v, cr.c[e]);
        }
        if (k % 3 == 2) fold_all();
    }
#pragma unroll
    for (int q = 0; q < NP; q++)
        partial[(uint64_t)q * nthreads + t] =
            u32x4{acc[q][0].value(), acc[q][1].value(), acc[q][2].value(), acc[q][3].value()};
}

// The same for four adjacent columns per thread (16-byte loads; each weight
// load serves four columns), for matrices whose rows are 16-byte aligned
// (w % 4 == 0, aligned base and pitch) and at most two points.  Three rows
// are loaded per step, then every accumulator is folded (LazyDot allows
// three products between folds); measured over a 2^20 x 256 trace with two
// points: 0.339 ms against 0.343 with two rows and 0.368 with one.
constexpr int DOT4_UR = 3;
static_assert(DOT4_UR >= 1 && DOT4_UR <= 3, "LazyDot holds at most three products between folds");
template <class F, int NP>
__global__ void __launch_bounds__(256) column_dot4_kernel(gcptr __restrict__ m, uint64_t ld, uint32_t w,
                                                          gcext __restrict__ c, uint32_t rows,
                                                          uint32_t chunk_rows, uint64_t nthreads4,
                                                          uint64_t nthreads, gext __restrict__ partial) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nthreads4) return;
    const uint32_t w4 = w / 4;
    const uint32_t col = (uint32_t)(t % w4) * 4, chunk = (uint32_t)(t / w4);
    const uint32_t r0 = chunk * chunk_rows, r1 = r0 + chunk_rows < rows ? r0 + chunk_rows : rows;
    LazyDot<F> acc[NP][4][4];  // point, column, coefficient
    auto fold_all = [&]() {
#pragma unroll
        for (int q = 0; q < NP; q++)
#pragma unroll
            for (int j = 0; j < 4; j++)
#pragma unroll
                for (int e = 0; e < 4; e++) acc[q][j][e].fold();
    };
    constexpr int UR = DOT4_UR;
    uint32_t r = r0;
    for (; r + UR <= r1; r += UR) {
        u32x4 v[UR];
        Ext4 cr[UR][NP];
#pragma unroll
        for (int u = 0; u < UR; u++) {
            v[u] = *reinterpret_cast<const GLOBAL u32x4*>(m + (uint64_t)(r + u) * ld + col);
#pragma unroll
            for (int q = 0; q < NP; q++) cr[u][q] = ext_of(c[(uint64_t)q * rows + r + u]);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < UR; u++) {
            const uint32_t vv[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
#pragma unroll
            for (int q = 0; q < NP; q++)
#pragma unroll
                for (int j = 0; j < 4; j++)
#pragma unroll
                    for (int e = 0; e < 4; e++) acc[q][j][e].add(vv[j], cr[u][q].c[e]);
        }
        fold_all();  // UR products since the last fold
    }
    for (; r < r1; r++) {  // fewer than UR rows left
        const u32x4 v = *reinterpret_cast<const GLOBAL u32x4*>(m + (uint64_t)r * ld + col);
        const uint32_t vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int q = 0; q < NP; q++) {
            const Ext4 cr = ext_of(c[(uint64_t)q * rows + r]);
#pragma unroll
            for (int j = 0; j < 4; j++)
#pragma unroll
                for (int e = 0; e < 4; e++) acc[q][j][e].add(vv[j], cr.c[e]);
        }
    }
    const uint64_t out_t = (uint64_t)chunk * w + col;  // the scalar kernel's partial index
#pragma unroll
    for (int q = 0; q < NP; q++)
#pragma unroll
        for (int j = 0; j < 4; j++)
            partial[(uint64_t)q * nthreads + out_t + j] = u32x4{acc[q][j][0].value(), acc[q][j][1].value(),
                                                                acc[q][j][2].value(), acc[q][j][3].value()};
}

// Column sums of the partials in two stages.  Stage 1: block (x, y) adds,
// for the 64 columns x*64 .. x*64+63, the chunks of slice y; its 256 threads
// are 64 columns by 4 interleaved chunk lanes, so each wave reads 64
// adjacent columns' partials (1 KB contiguous) per step; the 4 lanes meet in
// LDS.  Stage 2: per column, the slices' sums times the barycentric factor.
constexpr int SUM_UNROLL = 8;  // partial sums loaded per round trip

template <class F>
__global__ void __launch_bounds__(256) column_slice_sum_kernel(gcext __restrict__ partial, uint32_t nchunks,
                                                               uint32_t w, uint32_t nslices,
                                                               gext __restrict__ sums) {
    __shared__ uint32_t red[4][4][64];
    const uint32_t cl = threadIdx.x % 64, kl = threadIdx.x / 64;
    const uint32_t col = blockIdx.x * 64 + cl, y = blockIdx.y;
    const uint32_t per = (nchunks + nslices - 1) / nslices;
    const uint32_t k0 = y * per, k1 = k0 + per < nchunks ? k0 + per : nchunks;
    Ext4 s{{0, 0, 0, 0}};
    if (col < w) {
        // SUM_UNROLL loads in flight per round trip: one at a time, the ~64
        // sequential loads of a slice made the kernel latency-bound (36 us)
        for (uint32_t k = k0 + kl; k < k1; k += 4 * SUM_UNROLL) {
            Ext4 v[SUM_UNROLL];
#pragma unroll
            for (int u = 0; u < SUM_UNROLL; u++) {
                const uint32_t kk = k + 4 * u;
                v[u] = kk < k1 ? ext_of(partial[(uint64_t)kk * w + col]) : Ext4{{0, 0, 0, 0}};
            }
#pragma unroll
            for (int u = 0; u < SUM_UNROLL; u++) s = eadd<F>(s, v[u]);
        }
    }
#pragma unroll
    for (int e = 0; e < 4; e++) red[kl][e][cl] = s.c[e];
    __syncthreads();
    if (kl == 0 && col < w) {
#pragma unroll
        for (int u = 1; u < 4; u++)
#pragma unroll
            for (int e = 0; e < 4; e++) s.c[e] = fadd<F>(s.c[e], red[u][e][cl]);
        sums[(uint64_t)y * w + col] = v4_of(s);
    }
}

template <class F>
__global__ void __launch_bounds__(256) column_finish_kernel(gcext __restrict__ sums, uint32_t nslices, uint32_t w,
                                                            Ext4 factor, gext __restrict__ out) {
    const uint32_t col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= w) return;
    Ext4 s{{0, 0, 0, 0}};
    for (uint32_t y = 0; y < nslices; y += SUM_UNROLL) {
        Ext4 v[SUM_UNROLL];
#pragma unroll
        for (int u = 0; u < SUM_UNROLL; u++)
            v[u] = y + u < nslices ? ext_of(sums[(uint64_t)(y + u) * w + col]) : Ext4{{0, 0, 0, 0}};
#pragma unroll
        for (int u = 0; u < SUM_UNROLL; u++) s = eadd<F>(s, v[u]);
    }
    out[col] = v4_of(emul<F>(s, factor));
}

// Reduced opening of one log-height: a 64-thread block owns 64 rows, thread
// t row base + t.  The rows' columns pass through LDS in tiles of 64 rows x
// 32 columns, loaded so that each wave reads 2 rows x 128 contiguous bytes
// (whole cache lines) and read back one row per thread (a row stride of 33
// words keeps the banks distinct).  Measured on MI355X over a 2^21 x 256
// LDE: 0.448 ms, against 0.452 (128 x 16), 0.455 (128 x 32), 0.474 (256 x
// 16), 0.615 (128 x 8) and 1.0 ms and more with 64-column tiles.
constexpr int RO_ROWS = 64, RO_COLS = 32;
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
        LazyDot<F> acc[4];
        uint32_t since_fold = 0;
        // the next tile's values are loaded into registers while the current
        // one is reduced, so a tile's HBM round trip overlaps the arithmetic
        constexpr int LPT = RO_ROWS / (RO_ROWS / RO_COLS);  // loads per thread per tile
        uint32_t nxt[LPT];
        auto load_tile = [&](uint32_t c0) {
            const uint32_t col = c0 + lc < width ? c0 + lc : width - 1;  // clamped; masked at the store
#pragma unroll
            for (int kk = 0; kk < LPT; kk++) {
                const uint32_t r = lr + (RO_ROWS / RO_COLS) * kk;
                const uint32_t row = base + r < n ? base + r : n - 1;
                nxt[kk] = m[(uint64_t)row * ld + col];
            }
        };
        if (width) load_tile(0);
        for (uint32_t c0 = 0; c0 < width; c0 += RO_COLS) {
            const bool col_live = c0 + lc < width;
#pragma unroll
            for (int kk = 0; kk < LPT; kk++)
                tile[(lr + (RO_ROWS / RO_COLS) * kk) * (RO_COLS + 1) + lc] = col_live ? nxt[kk] : 0u;
            __syncthreads();
            if (c0 + RO_COLS < width) load_tile(c0 + RO_COLS);
            const uint32_t cn = width - c0 < (uint32_t)RO_COLS ? width - c0 : (uint32_t)RO_COLS;
            for (uint32_t c = 0; c < cn; c++) {
                const uint32_t v = tile[t * (RO_COLS + 1) + c];
                const Ext4 ap = ext_of(alpha_pows[c0 + c]);
#pragma unroll
                for (int q = 0; q < 4; q++) acc[q].add(v, ap.c[q]);
                if (++since_fold == 3) {  // the same in every lane
                    since_fold = 0;
#pragma unroll
                    for (int q = 0; q < 4; q++) acc[q].fold();
                }
            }
            __syncthreads();
        }
        const Ext4 r = Ext4{{acc[0].value(), acc[1].value(), acc[2].value(), acc[3].value()}};
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
                                const Ext4* z, const Ext4* factor, int npts, Ext4* weights, int weights_ready,
                                Ext4* partial, uint32_t nchunks, uint32_t chunk_rows, Ext4* sums, uint32_t nslices,
                                Ext4* out, hipStream_t s) {
    const uint32_t h = 1u << log_h;
    if (!weights_ready)  // else computed for these points and this height by an earlier call
        for (int q = 0; q < npts; q++)
            bary_weights_kernel<F><<<blocks_for(h, 256), 256, 0, s>>>(z[q], to_monty<F>(F::GEN), log_h, (gcptr)tw,
                                                                        (gext)(weights + (uint64_t)q * h));
    const uint64_t nt = (uint64_t)nchunks * w;
    if (open_values_vec4(m, ld, w, npts)) {
        const uint64_t nt4 = nt / 4;
        if (npts == 1)
            column_dot4_kernel<F, 1><<<blocks_for(nt4, 256), 256, 0, s>>>((gcptr)m, ld, w, (gcext)weights, h,
                                                                         chunk_rows, nt4, nt, (gext)partial);
        else
            column_dot4_kernel<F, 2><<<blocks_for(nt4, 256), 256, 0, s>>>((gcptr)m, ld, w, (gcext)weights, h,
                                                                         chunk_rows, nt4, nt, (gext)partial);
    } else switch (npts) {
        case 1: column_dot_kernel<F, 1><<<blocks_for(nt, 256), 256, 0, s>>>((gcptr)m, ld, w, (gcext)weights, h,
                                                                          chunk_rows, nt, (gext)partial); break;
        case 2: column_dot_kernel<F, 2><<<blocks_for(nt, 256), 256, 0, s>>>((gcptr)m, ld, w, (gcext)weights, h,
                                                                          chunk_rows, nt, (gext)partial); break;
        case 3: column_dot_kernel<F, 3><<<blocks_for(nt, 256), 256, 0, s>>>((gcptr)m, ld, w, (gcext)weights, h,
                                                                          chunk_rows, nt, (gext)partial); break;
        default: column_dot_kernel<F, 4><<<blocks_for(nt, 256), 256, 0, s>>>((gcptr)m, ld, w, (gcext)weights, h,
                                                                           chunk_rows, nt, (gext)partial); break;
    }
    for (int q = 0; q < npts; q++) {
        Ext4* sq = sums + (uint64_t)q * nslices * w;
        column_slice_sum_kernel<F><<<dim3(blocks_for(w, 64), nslices), 256, 0, s>>>(
            (gcext)(partial + (uint64_t)q * nt), nchunks, w, nslices, (gext)sq);
        column_finish_kernel<F><<<blocks_for(w, 256), 256, 0, s>>>((gcext)sq, nslices, w, factor[q],
                                                                    (gext)(out + (uint64_t)q * w));
    }
    return hipGetLastError();
}

bool open_values_vec4(const uint32_t* m, uint64_t ld, uint32_t w, int npts) {
    return npts <= 2 && w % 4 == 0 && ld % 4 == 0 && (uintptr_t)m % 16 == 0;
}

hipError_t launch_open_values(int field, const uint32_t* m, uint64_t ld, uint32_t w, int log_h,
                              const uint32_t* tw, const Ext4* z, const Ext4* factor, int npts, Ext4* weights,
                              int weights_ready, Ext4* partial, uint32_t nchunks, uint32_t chunk_rows, Ext4* sums,
                              uint32_t nslices, Ext4* out, hipStream_t stream) {
    if (w == 0 || nchunks == 0 || npts <= 0) return hipSuccess;
    if (nslices == 0 || nslices > nchunks || nslices > 65535) return hipErrorInvalidValue;
    if (npts > OV_MAX_POINTS) return hipErrorInvalidValue;
    return field == FIELD_KOALABEAR
               ? open_values_t<KoalaBear>(m, ld, w, log_h, tw, z, factor, npts, weights, weights_ready, partial,
                                          nchunks, chunk_rows, sums, nslices, out, stream)
               : open_values_t<BabyBear>(m, ld, w, log_h, tw, z, factor, npts, weights, weights_ready, partial,
                                         nchunks, chunk_rows, sums, nslices, out, stream);
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
