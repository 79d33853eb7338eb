// Grouped block-contraction kernel for gfx950 (MI355X), FP64.
//
// One launch processes every (block, output tile, K chunk) task of a plan:
//   gather A(k,m), B(k,n) tiles through signed index tables into LDS
//   (replaces contract/collect_block_impr.f:2-415),
//   FP64 MFMA 16x16x4 accumulation (replaces dgemm('t','n') at
//   contract/contr_blk1blk2_blocked_mm.f:465,523),
//   signed scatter-accumulate of the tile into the result list
//   (replaces contract/scatter_block.f:2-174).
// Workgroup: 256 threads = 4 wave64, tile 64x64, K step 16; each wave owns a
// 32x32 sub-tile = 2x2 MFMA tiles.
#include <hip/hip_runtime.h>
#include "contract_types.h"

namespace gecco {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 16, NT = 256;

__device__ __forceinline__ double gather(const double *__restrict__ base,
                                         const int32_t *first, const int32_t *second,
                                         int64_t ld, int table, int k, int x) {
    int32_t v;
    uint32_t neg;
    int64_t addr;
    if (table) {
        v = first[(int64_t)k + ld * (int64_t)x];
        if (v == 0) return 0.0;
        neg = (uint32_t)v >> 31;
        addr = (int64_t)(v < 0 ? -v : v) - 1;
    } else {
        int32_t r = first[k], c = second[x];
        if (r == 0 || c == 0) return 0.0;
        neg = ((uint32_t)r ^ (uint32_t)c) >> 31;
        addr = (int64_t)(r < 0 ? -r : r) + (int64_t)(c < 0 ? -c : c) - 2;
    }
    double val = base[addr];
    return neg ? -val : val;
}

__global__ __launch_bounds__(NT) void grouped_contract_f64(
    const DevBlock *__restrict__ blocks, const DevTask *__restrict__ tasks, int ntasks,
    double *const *__restrict__ bufs) {
    __shared__ double As[BK][BM + 1];
    __shared__ double Bs[BK][BN + 1];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int lrow = lane >> 4, lcol = lane & 15;

    for (int t = blockIdx.x; t < ntasks; t += gridDim.x) {
        const DevTask task = tasks[t];
        const DevBlock &b = blocks[task.blk];
        const double *__restrict__ pa = bufs[b.buf_a] + b.base_a;
        const double *__restrict__ pb = bufs[b.buf_b] + b.base_b;
        double *pc = bufs[b.buf_c] + b.base_c;
        const int M = b.M, N = b.N;
        const int m0 = task.m0, n0 = task.n0, k1 = task.k1;
        const DevOperand oa = b.A, ob = b.B;

        v4d acc[2][2];
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

        for (int k0 = task.k0; k0 < k1; k0 += BK) {
            double ra[4], rb[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int e = tid + NT * r;
                int kk, mm;
                if (oa.fast_first) { kk = e & (BK - 1); mm = e >> 4; }
                else { mm = e & (BM - 1); kk = e >> 6; }
                const int k = k0 + kk, m = m0 + mm;
                ra[r] = (k < k1 && m < M)
                            ? gather(pa, oa.first, oa.second, oa.ld, oa.table, k, m) : 0.0;
                int kb, nn;
                if (ob.fast_first) { kb = e & (BK - 1); nn = e >> 4; }
                else { nn = e & (BN - 1); kb = e >> 6; }
                const int kq = k0 + kb, n = n0 + nn;
                rb[r] = (kq < k1 && n < N)
                            ? gather(pb, ob.first, ob.second, ob.ld, ob.table, kq, n) : 0.0;
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int e = tid + NT * r;
                if (oa.fast_first) As[e & (BK - 1)][e >> 4] = ra[r];
                else As[e >> 6][e & (BM - 1)] = ra[r];
                if (ob.fast_first) Bs[e & (BK - 1)][e >> 4] = rb[r];
                else Bs[e >> 6][e & (BN - 1)] = rb[r];
            }
            __syncthreads();
#pragma unroll
            for (int kq = 0; kq < BK / 4; ++kq) {
                const int kr = kq * 4 + lrow;
                double a0 = As[kr][wm * 32 + lcol];
                double a1 = As[kr][wm * 32 + 16 + lcol];
                double b0 = Bs[kr][wn * 32 + lcol];
                double b1 = Bs[kr][wn * 32 + 16 + lcol];
                acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
                acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
                acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
                acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
            }
        }

        // signed scatter-accumulate; output element (row, col) of a 16x16
        // f64 MFMA tile sits in lane col + 16*(row%4), register row/4
        const DevOperand oc = b.C;
        const double xfac = b.xfac;
        const int atomic = b.atomic;
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
#pragma unroll
            for (int ni = 0; ni < 2; ++ni)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int m = m0 + wm * 32 + mi * 16 + r * 4 + lrow;
                    const int n = n0 + wn * 32 + ni * 16 + lcol;
                    if (m >= M || n >= N) continue;
                    int32_t v;
                    int64_t addr;
                    bool neg;
                    if (oc.table) {
                        v = oc.first[(int64_t)m + oc.ld * (int64_t)n];
                        if (v == 0) continue;
                        neg = v < 0;
                        addr = (int64_t)(v < 0 ? -v : v) - 1;
                    } else {
                        int32_t rr = oc.first[m], cc = oc.second[n];
                        if (rr == 0 || cc == 0) continue;
                        neg = (rr < 0) != (cc < 0);
                        addr = (int64_t)(rr < 0 ? -rr : rr) + (int64_t)(cc < 0 ? -cc : cc) - 2;
                    }
                    double val = xfac * acc[mi][ni][r];
                    if (neg) val = -val;
                    if (atomic) unsafeAtomicAdd(pc + addr, val);
                    else pc[addr] += val;
                }
        __syncthreads();
    }
}

// Spin-flip symmetrisation of a result list (contract/sym_ab_blk.f:354-377):
// for each precomputed pair (i1 <= i2, factor f = +/-relfac2):
//   v = prefac*(x[i1] + f*x[i2]);  x[i1] = v;  x[i2] = f*v
__global__ void sym_ab_pairs_f64(double *__restrict__ x, const int64_t *__restrict__ i1,
                                 const int64_t *__restrict__ i2,
                                 const int8_t *__restrict__ fsign, int64_t npair,
                                 double prefac) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < npair;
         p += (int64_t)gridDim.x * blockDim.x) {
        const int64_t a = i1[p], c = i2[p];
        const double f = (double)fsign[p];
        const double v = prefac * (x[a] + f * x[c]);
        x[a] = v;
        if (c != a) x[c] = f * v;
    }
}

}  // namespace gecco

extern "C" int gecco_launch_grouped(const gecco::DevBlock *blocks, const gecco::DevTask *tasks,
                                    int ntasks, double *const *bufs, int grid,
                                    hipStream_t stream) {
    if (ntasks <= 0) return 0;
    hipLaunchKernelGGL(gecco::grouped_contract_f64, dim3(grid), dim3(gecco::NT), 0, stream,
                       blocks, tasks, ntasks, bufs);
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

extern "C" int gecco_launch_sym_ab(double *x, const int64_t *i1, const int64_t *i2,
                                   const int8_t *fsign, int64_t npair, double prefac,
                                   hipStream_t stream) {
    if (npair <= 0) return 0;
    int64_t g = (npair + 255) / 256;
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(gecco::sym_ab_pairs_f64, dim3((unsigned)g), dim3(256), 0, stream, x, i1,
                       i2, fsign, npair, prefac);
    return hipGetLastError() == hipSuccess ? 0 : 3;
}
