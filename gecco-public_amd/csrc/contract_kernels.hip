// Grouped block-contraction kernel for gfx950 (MI355X), FP64 or FP32 lists.
//
// One launch processes every (block, output tile, K chunk) task of a plan:
//   gather A(k,m), B(k,n) tiles through signed index tables into LDS
//   (replaces contract/collect_block_impr.f:2-415),
//   MFMA 16x16x4 accumulation in the list precision (replaces dgemm('t','n') at
//   contract/contr_blk1blk2_blocked_mm.f:465,523),
//   signed scatter-accumulate of the tile into the result list
//   (replaces contract/scatter_block.f:2-174).
// Workgroup: 256 threads = 4 wave64, tile 64x64, K step 16; each wave owns a
// 32x32 sub-tile = 2x2 MFMA tiles.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include "contract_types.h"

namespace gecco {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// Element types: lists in float64 or float32 (a plan's precision).  MFMA
// 16x16x4 in the element type; accumulators of 4 elements per lane.
template <typename T> struct Acc;
template <> struct Acc<double> { typedef v4d type; };
template <> struct Acc<float> { typedef v4f type; };
__device__ __forceinline__ v4d mfma16(double a, double b, v4d c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4f mfma16(float a, float b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
// Row of accumulator register r in lane group g = lane / 16 of a 16 x 16
// MFMA tile (the column is lane % 16): f64 16x16x4 puts row 4 r + g in
// register r, f32 16x16x4 row 4 g + r.
typedef short v4s __attribute__((ext_vector_type(4)));
// bfloat16 of a float (nearest, ties to even) and back
__device__ __forceinline__ unsigned short bf16_of(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_val(unsigned short h) { return __uint_as_float((uint32_t)h << 16); }
template <typename T> __device__ __forceinline__ int mrow(int r, int g);
template <> __device__ __forceinline__ int mrow<double>(int r, int g) { return 4 * r + g; }
template <> __device__ __forceinline__ int mrow<float>(int r, int g) { return 4 * g + r; }

// Global-memory loads, stores and atomics through global-address-space
// pointers.  Operand, result and index pointers are read from device tables,
// so the compiler would otherwise emit flat instructions, which also count
// against the LDS wait counter: every wait for an LDS read would then wait
// for all outstanding prefetch loads and atomics as well.
typedef const double __attribute__((address_space(1))) *gcdptr;
typedef double __attribute__((address_space(1))) *gdptr;
typedef const float __attribute__((address_space(1))) *gcfptr;
typedef float __attribute__((address_space(1))) *gfptr;
typedef const int32_t __attribute__((address_space(1))) *gciptr;
__device__ __forceinline__ double ldg(const double *p) { return *(gcdptr)p; }
__device__ __forceinline__ float ldg(const float *p) { return *(gcfptr)p; }
__device__ __forceinline__ int32_t ldg(const int32_t *p) { return *(gciptr)p; }
__device__ __forceinline__ void stg(double *p, double v) { *(gdptr)p = v; }
__device__ __forceinline__ void stg(float *p, float v) { *(gfptr)p = v; }
__device__ __forceinline__ void atomic_add_g(double *p, double v) {
    (void)__builtin_amdgcn_global_atomic_fadd_f64((gdptr)p, v);
}
__device__ __forceinline__ void atomic_add_g(float *p, float v) {
    (void)__builtin_amdgcn_global_atomic_fadd_f32((gfptr)p, v);
}

constexpr int BM = 64, BN = 64, BK = 16, NT = 256;
constexpr int LPT = BK * BM / NT;   // elements of each operand tile per thread

// Per-thread view of one gathered operand tile (BK x 64).  Thread tid loads
// LPT elements; with consecutive-k memory order (fast_first) they share one k
// and have LPT different m/n, otherwise they share one m/n and have LPT k.  The
// per-m/n half of a separable address is fixed for a task and computed once;
// affine halves need no index load at all.
template <typename T>
struct OpView {
    const T *base;
    const int32_t *first;
    int64_t ld, f0, f1;
    int table, faff;
    int kk[LPT];
    int xi[LPT];
    int64_t cadd[LPT];   // separable: offset of the m/n half; table: unused
    uint32_t cneg[LPT];  // sign bit of the m/n half
    int cval[LPT];       // element column valid and not forbidden
};

template <typename T>
__device__ __forceinline__ void op_init(OpView<T> &v, const DevOperand &o, const T *base, int x0,
                                        int X, int tid) {
    v.base = base;
    v.first = o.first;
    v.ld = o.ld;
    v.table = o.table;
    v.faff = o.first_affine;
    v.f0 = o.f0;
    v.f1 = o.f1;
#pragma unroll
    for (int r = 0; r < LPT; ++r) {
        // fast_first: BK consecutive threads walk k; else 64 threads walk m/n
        if (o.fast_first) { v.kk[r] = tid % BK; v.xi[r] = tid / BK + (NT / BK) * r; }
        else { v.kk[r] = tid / BM + (NT / BM) * r; v.xi[r] = tid % BM; }
        const int x = x0 + v.xi[r];
        v.xi[r] = x;
        v.cadd[r] = 0;
        v.cneg[r] = 0;
        v.cval[r] = x < X;
        if (x < X && !o.table) {
            if (o.second_affine) {
                v.cadd[r] = o.s0 + o.s1 * (int64_t)x;
            } else {
                const int32_t c = ldg(o.second + x);
                v.cval[r] = c != 0;
                v.cneg[r] = (uint32_t)c >> 31;
                v.cadd[r] = (int64_t)(c < 0 ? -c : c) - 1;
            }
        }
    }
}

// Issue the loads of one tile; the sign and the zero of forbidden or
// out-of-range elements are applied at the LDS store (op_fix), so the loads
// stay in flight while the MFMAs of the previous tile run.
template <typename T>
__device__ __forceinline__ void op_load(const OpView<T> &v, int k0, int k1, T raw[LPT],
                                        uint32_t &flags) {
    flags = 0;
#pragma unroll
    for (int r = 0; r < LPT; ++r) {
        const int k = k0 + v.kk[r];
        bool ok = (k < k1) && v.cval[r];
        int64_t addr = 0;
        uint32_t neg = 0;
        if (v.table) {
            const int32_t rv = ok ? ldg(v.first + ((int64_t)k + v.ld * (int64_t)v.xi[r])) : 0;
            ok = ok && rv != 0;
            neg = (uint32_t)rv >> 31;
            addr = (int64_t)(rv < 0 ? -rv : rv) - 1;
        } else if (v.faff) {
            addr = v.f0 + v.f1 * (int64_t)k + v.cadd[r];
            neg = v.cneg[r];
        } else {
            const int32_t rv = ok ? ldg(v.first + k) : 0;
            ok = ok && rv != 0;
            neg = ((uint32_t)rv >> 31) ^ v.cneg[r];
            addr = (int64_t)(rv < 0 ? -rv : rv) - 1 + v.cadd[r];
        }
        raw[r] = ldg(v.base + (ok ? addr : 0));
        flags |= ((uint32_t)ok << r) | (neg << (r + LPT));
    }
}

template <typename T>
__device__ __forceinline__ T op_fix(T raw, uint32_t flags, int r) {
    const T v = ((flags >> (r + LPT)) & 1u) ? -raw : raw;
    return ((flags >> r) & 1u) ? v : (T)0;
}

// Instance groups: a task of group g >= 0 does nothing while skip[g] != 0.
__device__ __forceinline__ bool skipped(const int32_t *__restrict__ skip, int grp) {
    return skip != nullptr && grp >= 0 && ldg(skip + grp) != 0;
}

// Norm screening: the kernel call of block b is skipped when the squared
// norm of its A or B operand block fell below 1e-28 (zero: per operand-block
// flags of the plan's norm launch; nullptr when the plan does not screen).
__device__ __forceinline__ bool screened(const int32_t *__restrict__ zero, const DevBlock &b) {
    return zero != nullptr && ((b.zero_a >= 0 && ldg(zero + b.zero_a) != 0) ||
                               (b.zero_b >= 0 && ldg(zero + b.zero_b) != 0));
}

template <typename T>
__global__ __launch_bounds__(NT) void grouped_contract(
    const DevBlock *__restrict__ blocks, const DevTask *__restrict__ tasks, int ntasks,
    T *const *__restrict__ bufs, const int32_t *__restrict__ skip, const int32_t *__restrict__ zero) {
    typedef typename Acc<T>::type AccT;
    __shared__ T As[BK][BM + 1];
    __shared__ T Bs[BK][BN + 1];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int lrow = lane >> 4, lcol = lane & 15;

    for (int t = blockIdx.x; t < ntasks; t += gridDim.x) {
        const DevTask task = tasks[t];
        if (skipped(skip, task.grp)) continue;
        const DevBlock &b = blocks[task.blk];
        if (screened(zero, b)) continue;
        const int M = b.M, N = b.N;
        const int m0 = task.m0, n0 = task.n0, k1 = task.k1;
        OpView<T> va, vb;
        op_init(va, b.A, (const T *)(bufs[b.buf_a] + b.base_a), m0, M, tid);
        op_init(vb, b.B, (const T *)(bufs[b.buf_b] + b.base_b), n0, N, tid);

        AccT acc[2][2];
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) acc[i][j] = AccT{0, 0, 0, 0};
        // bit 0: second 16-row half of the wave tile overlaps the block,
        // bit 1: second 16-column half does (the first halves always do or
        // the whole wave tile is padding, whose results are discarded)
        const uint32_t live = (uint32_t)(m0 + wm * 32 + 16 < M) | ((uint32_t)(n0 + wn * 32 + 16 < N) << 1);

        T ra[LPT], rb[LPT];
        uint32_t fa, fb;
        op_load(va, task.k0, k1, ra, fa);
        op_load(vb, task.k0, k1, rb, fb);
        for (int k0 = task.k0; k0 < k1; k0 += BK) {
            __syncthreads();
#pragma unroll
            for (int r = 0; r < LPT; ++r) {
                As[va.kk[r]][va.xi[r] - m0] = op_fix(ra[r], fa, r);
                Bs[vb.kk[r]][vb.xi[r] - n0] = op_fix(rb[r], fb, r);
            }
            __syncthreads();
            if (k0 + BK < k1) {   // next tile in flight during the MFMAs
                op_load(va, k0 + BK, k1, ra, fa);
                op_load(vb, k0 + BK, k1, rb, fb);
            }
#pragma unroll
            for (int kq = 0; kq < BK / 4; ++kq) {
                const int kr = kq * 4 + lrow;
                T a0 = As[kr][wm * 32 + lcol];
                T a1 = As[kr][wm * 32 + 16 + lcol];
                T b0 = Bs[kr][wn * 32 + lcol];
                T b1 = Bs[kr][wn * 32 + 16 + lcol];
                acc[0][0] = mfma16(a0, b0, acc[0][0]);
                if (live & 2u) acc[0][1] = mfma16(a0, b1, acc[0][1]);
                if (live & 1u) acc[1][0] = mfma16(a1, b0, acc[1][0]);
                if (live == 3u) acc[1][1] = mfma16(a1, b1, acc[1][1]);
            }
        }

        // signed scatter-accumulate; register r of lane l holds row
        // mrow<T>(r, l / 16), column l % 16 of its 16x16 MFMA tile
        const DevOperand oc = b.C;
        T *pc = bufs[b.buf_c] + b.base_c;
        const T xfac = (T)b.xfac;
        int64_t addr[2][2][4];   // -1: padding or forbidden
        T val[2][2][4];
#pragma unroll
        for (int ni = 0; ni < 2; ++ni) {
            const int n = n0 + wn * 32 + ni * 16 + lcol;
            const int32_t cc = (n < N && !oc.table)
                                   ? (oc.second_affine ? (int32_t)(1 + oc.s0 + oc.s1 * (int64_t)n)
                                                       : ldg(oc.second + n))
                                   : 1;
#pragma unroll
            for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int m = m0 + wm * 32 + mi * 16 + mrow<T>(r, lrow);
                    int64_t ad = -1;
                    bool neg = false;
                    if (m < M && n < N) {
                        if (oc.table) {
                            const int32_t v = ldg(oc.first + ((int64_t)m + oc.ld * (int64_t)n));
                            neg = v < 0;
                            ad = (int64_t)(v < 0 ? -v : v) - 1;
                        } else {
                            const int32_t rr = oc.first_affine
                                                   ? (int32_t)(1 + oc.f0 + oc.f1 * (int64_t)m)
                                                   : ldg(oc.first + m);
                            neg = (rr < 0) != (cc < 0);
                            ad = (rr == 0 || cc == 0)
                                     ? -1
                                     : (int64_t)(rr < 0 ? -rr : rr) + (int64_t)(cc < 0 ? -cc : cc) - 2;
                        }
                    }
                    const T v = xfac * acc[mi][ni][r];
                    addr[mi][ni][r] = ad;
                    val[mi][ni][r] = neg ? -v : v;
                }
        }
        if (b.atomic) {
#pragma unroll
            for (int ni = 0; ni < 2; ++ni)
#pragma unroll
                for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        if (addr[mi][ni][r] >= 0) atomic_add_g(pc + addr[mi][ni][r], val[mi][ni][r]);
        } else {
            // the task owns its result elements: loads first, then stores
            T old[2][2][4];
#pragma unroll
            for (int ni = 0; ni < 2; ++ni)
#pragma unroll
                for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        old[mi][ni][r] = addr[mi][ni][r] >= 0 ? ldg(pc + addr[mi][ni][r]) : (T)0;
#pragma unroll
            for (int ni = 0; ni < 2; ++ni)
#pragma unroll
                for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        if (addr[mi][ni][r] >= 0) stg(pc + addr[mi][ni][r], old[mi][ni][r] + val[mi][ni][r]);
        }
    }
}

// Dense blocks whose three index maps are all affine (no forbidden
// elements, all signs +, constant strides -- e.g. the particle-particle
// ladder) are strided GEMMs: addresses are base + stride_k*k + stride_x*x.
// Tile TBM x TBN, K step TBK, 256 threads = 2x2 waves, each wave owns
// (TBM/2) x (TBN/2) = TM x TN MFMA 16x16 tiles.  LDS is double-buffered: one
// barrier per K step; the global loads of step s+1 are in flight while the
// MFMAs of step s run.
//
// Operand loads are raw buffer loads: a descriptor (scalar registers) spans
// exactly the operand block, each element has a 32-bit byte offset that
// advances by a constant per K step, and the hardware returns 0 for any
// offset at or beyond the block's extent.  An element outside the block's
// rows/columns starts at offset 0x80000000; the planner keeps operand
// extents at or below 2^30 bytes, so with all its K steps added the offset
// stays in [2^31, 2^32) and out of range.  Elements beyond the block's K (the
// last step of a block) use offset 0x80000000 directly.  Masking thus costs
// no branch, and a full step one add per element.
constexpr uint32_t BUF_OOB = 0x80000000u;
constexpr int BUF_FLAGS = 0x00020000;   // raw buffer, gfx9-family descriptor word 3

__device__ __forceinline__ void buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off, double &v) {
    v = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ void buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off, float &v) {
    v = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}

template <int X, int TBK, typename T>
struct AffView {
    static constexpr int ES = (int)sizeof(T);
    static constexpr int L = TBK * X / NT;  // elements per thread
    __amdgpu_buffer_rsrc_t rsrc;
    uint32_t off[L];  // byte offset of element r at the current K step
    uint32_t step;    // bytes per K step
    int kk0, kstep;   // k of element r relative to the step start: kk0 + kstep*r
    int xi0, xstep;   // m/n (relative to the tile) of element r: xi0 + xstep*r
    int so, ss;       // LDS index of element r in a tile with rows of X + 1: so + ss*r
    bool ff;          // the operand's fast_first: which of the two layouts above

    // operand block at blk (element (k, x) at o.f0 + o.s0 + o.f1*k + o.s1*x,
    // k < K, x < Xn); the first step starts at k = kstart
    __device__ __forceinline__ void init(const T *blk, const DevOperand &o, int K, int x0,
                                         int Xn, int kstart, int tid) {
        ff = o.fast_first != 0;
        if (ff) {
            kk0 = tid % TBK; kstep = 0;
            xi0 = tid / TBK; xstep = NT / TBK;
        } else {
            kk0 = tid / X; kstep = NT / X;
            xi0 = tid % X; xstep = 0;
        }
        so = kk0 * (X + 1) + xi0;
        ss = kstep * (X + 1) + xstep;
        const int64_t extent = (o.f1 * (int64_t)(K - 1) + o.s1 * (int64_t)(Xn - 1) + 1) * ES;
        rsrc = __builtin_amdgcn_make_buffer_rsrc((void *)(blk + o.f0 + o.s0), (short)0, (int)extent,
                                                 BUF_FLAGS);
        step = (uint32_t)(o.f1 * TBK * ES);
#pragma unroll
        for (int r = 0; r < L; ++r) {
            const int x = x0 + xi0 + xstep * r, k = kstart + kk0 + kstep * r;
            off[r] = x < Xn ? (uint32_t)((o.f1 * (int64_t)k + o.s1 * (int64_t)x) * ES) : BUF_OOB;
        }
    }
    // loads the step starting at block-local k0 (the range ends at k1) and
    // advances to the next step
    __device__ __forceinline__ void load(int k0, int k1, T out[L]) {
        if (k0 + TBK <= k1) {
#pragma unroll
            for (int r = 0; r < L; ++r) {
                buf_load(rsrc, off[r], out[r]);
                off[r] += step;
            }
        } else {
#pragma unroll
            for (int r = 0; r < L; ++r) {
                buf_load(rsrc, k0 + kk0 + kstep * r < k1 ? off[r] : BUF_OOB, out[r]);
                off[r] += step;
            }
        }
    }
    // store of data loaded under layout fl (a step loaded before the view
    // moved on to the next block of a chain, whose layout may differ)
    template <int LD>
    __device__ __forceinline__ static void store_as(T *lds, const T in[L], bool fl, int tid) {
        const int k0 = fl ? tid % TBK : tid / X, ks = fl ? 0 : NT / X;
        const int x0 = fl ? tid / TBK : tid % X, xs = fl ? NT / TBK : 0;
#pragma unroll
        for (int r = 0; r < L; ++r) lds[(k0 + ks * r) * LD + x0 + xs * r] = in[r];
    }
    template <int LD>
    __device__ __forceinline__ void store(T *lds, const T in[L]) const {
        if constexpr (LD == X + 1) {   // the index precomputed at init
#pragma unroll
            for (int r = 0; r < L; ++r) lds[so + ss * r] = in[r];
        } else {
#pragma unroll
            for (int r = 0; r < L; ++r) lds[(kk0 + kstep * r) * LD + xi0 + xstep * r] = in[r];
        }
    }
};

// Element offset (-1: forbidden) and sign of result element (m, n) through
// the result map oc (caff: both halves affine).
__device__ __forceinline__ int64_t result_offset(const DevOperand &oc, bool caff, int m, int n,
                                                 bool &neg) {
    neg = false;
    if (caff) return oc.f0 + oc.f1 * (int64_t)m + oc.s0 + oc.s1 * (int64_t)n;
    if (oc.table) {
        const int32_t t = ldg(oc.first + ((int64_t)m + oc.ld * (int64_t)n));
        neg = t < 0;
        return (int64_t)(t < 0 ? -t : t) - 1;
    }
    int64_t ar, ac;
    bool nr = false, nc = false, ok = true;
    if (oc.first_affine) {
        ar = oc.f0 + oc.f1 * (int64_t)m;
    } else {
        const int32_t t = ldg(oc.first + m);
        ok = t != 0; nr = t < 0; ar = (int64_t)(t < 0 ? -t : t) - 1;
    }
    if (oc.second_affine) {
        ac = oc.s0 + oc.s1 * (int64_t)n;
    } else {
        const int32_t t = ldg(oc.second + n);
        ok = ok && t != 0; nc = t < 0; ac = (int64_t)(t < 0 ? -t : t) - 1;
    }
    neg = nr != nc;
    return ok ? ar + ac : -1;
}

// Tasks of the strided-GEMM launch whose tile is at most 16 x 16 (results
// with a short side of at most 16 and a long contraction: energies, the
// one-index intermediates and singles terms of coupled-cluster equations).
// Each wave runs MFMA 16x16x4 over its own part of the contraction range,
// with the operands loaded from memory straight into the MFMA lane layout
// (lane: k = lane / 16, m or n = lane % 16) and the next step's loads in
// flight during the current step's MFMAs; no LDS traffic and no barrier per
// step.  The four partial tiles are summed in LDS (1024 elements) at the end.
// A step covers 16 * SKINNY_U consecutive k: 4 * SKINNY_U per wave.
constexpr int SKINNY_U = 8;

// A 1 x 1 result (a full contraction, e.g. the energy): a dot product along
// the task's range, every thread on its own k (DOT_U loads in flight per
// operand), then summed over the workgroup.  The MFMA layout would keep 60
// of every 64 lanes idle here.
constexpr int DOT_U = 8;
static_assert(NT == 256, "dot_task sums four waves");

template <typename T>
__device__ __forceinline__ void dot_task(const DevTask &task, const DevBlock *__restrict__ blocks,
                                         T *const *__restrict__ bufs, T *lds, int tid, T *raw) {
    const DevBlock &b = blocks[task.blk];
    T s = (T)0;
    int blk = task.blk, kb = 0, kl = task.k0;
    for (;;) {
        const DevBlock &cb = blocks[blk];
        const int kh = min(cb.K, task.k1 - kb);
        const DevOperand &oa = cb.A, &ob = cb.B;
        const T *__restrict__ qa = bufs[cb.buf_a] + cb.base_a + oa.f0 + oa.s0;
        const T *__restrict__ qb = bufs[cb.buf_b] + cb.base_b + ob.f0 + ob.s0;
        const int64_t fa = oa.f1, fb = ob.f1;
        for (int k0 = kl + tid; k0 < kh; k0 += NT * DOT_U) {
            T a[DOT_U], bv[DOT_U];
#pragma unroll
            for (int u = 0; u < DOT_U; ++u) {
                const int k = k0 + u * NT;
                a[u] = k < kh ? ldg(qa + fa * (int64_t)k) : (T)0;
                bv[u] = k < kh ? ldg(qb + fb * (int64_t)k) : (T)0;
            }
#pragma unroll
            for (int u = 0; u < DOT_U; ++u) s = fma(a[u], bv[u], s);
        }
        kb += cb.K;
        blk = cb.next;
        if (blk < 0 || kb >= task.k1) break;
        kl = 0;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    __syncthreads();   // LDS of the previous task is no longer read
    if ((tid & 63) == 0) lds[tid >> 6] = s;
    __syncthreads();
    if (tid == 0 && raw) {
        stg(raw + task.lo, (T)b.xfac * ((lds[0] + lds[1]) + (lds[2] + lds[3])));
    } else if (tid == 0) {
        const T sum = (lds[0] + lds[1]) + (lds[2] + lds[3]);
        const DevOperand &oc = b.C;
        const bool caff = !oc.table && oc.first_affine && oc.second_affine;
        bool neg;
        const int64_t a = result_offset(oc, caff, task.m0, task.n0, neg);
        if (a >= 0) {
            T *pc = bufs[b.buf_c] + b.base_c;
            const T x = (T)b.xfac * sum;
            const T v = neg ? -x : x;
            if (b.atomic) atomic_add_g(pc + a, v);
            else stg(pc + a, ldg(pc + a) + v);
        }
    }
}

template <typename T>
__device__ __forceinline__ void skinny_task(const DevTask &task, const DevBlock *__restrict__ blocks,
                                            T *const *__restrict__ bufs, T *lds, int tid, T *raw) {
    typedef typename Acc<T>::type AccT;
    if (blocks[task.blk].M == 1 && blocks[task.blk].N == 1) {
        dot_task(task, blocks, bufs, lds, tid, raw);
        return;
    }
    const int lane = tid & 63, wave = tid >> 6;
    const int km = lane >> 4, xl = lane & 15;
    const DevBlock &b = blocks[task.blk];
    const int M = b.M, N = b.N;
    const int m = task.m0 + xl, n = task.n0 + xl;
    const bool mok = m < M, nok = n < N;
    const int koff = 4 * SKINNY_U * wave + km;   // k of MFMA u: step start + koff + 4u
    AccT acc = AccT{0, 0, 0, 0};
    // range [task.k0, task.k1) from the start of block task.blk, along its chain
    int blk = task.blk, kb = 0, kl = task.k0;
    for (;;) {
        const DevBlock &cb = blocks[blk];
        const int kh = min(cb.K, task.k1 - kb);
        const DevOperand &oa = cb.A, &ob = cb.B;
        const T *__restrict__ qa =
            bufs[cb.buf_a] + cb.base_a + oa.f0 + oa.s0 + (mok ? oa.s1 * (int64_t)m : 0);
        const T *__restrict__ qb =
            bufs[cb.buf_b] + cb.base_b + ob.f0 + ob.s0 + (nok ? ob.s1 * (int64_t)n : 0);
        const int64_t fa = oa.f1, fb = ob.f1;
        T a[SKINNY_U], bv[SKINNY_U];
        auto load = [&](int k0, T *ra, T *rb) {
#pragma unroll
            for (int u = 0; u < SKINNY_U; ++u) {
                const int k = k0 + koff + 4 * u;
                const bool ok = k < kh;
                ra[u] = (ok && mok) ? ldg(qa + fa * (int64_t)k) : (T)0;
                rb[u] = (ok && nok) ? ldg(qb + fb * (int64_t)k) : (T)0;
            }
        };
        load(kl, a, bv);
        for (int k0 = kl; k0 < kh; k0 += 16 * SKINNY_U) {
            T an[SKINNY_U], bn[SKINNY_U];
            const bool more = k0 + 16 * SKINNY_U < kh;
            if (more) load(k0 + 16 * SKINNY_U, an, bn);
#pragma unroll
            for (int u = 0; u < SKINNY_U; ++u)
                acc = mfma16(a[u], bv[u], acc);
            if (more) {
#pragma unroll
                for (int u = 0; u < SKINNY_U; ++u) { a[u] = an[u]; bv[u] = bn[u]; }
            }
        }
        kb += cb.K;
        blk = cb.next;
        if (blk < 0 || kb >= task.k1) break;
        kl = 0;
    }
    __syncthreads();   // LDS of the previous task is no longer read
#pragma unroll
    for (int r = 0; r < 4; ++r) lds[wave * 256 + r * 64 + lane] = acc[r];
    __syncthreads();
    // thread tid owns tile element (r, lane) = (tid / 64, tid % 64):
    // m = m0 + mrow(r, lane / 16), n = n0 + lane % 16
    const T sum = lds[tid] + lds[256 + tid] + lds[512 + tid] + lds[768 + tid];
    const int mm = task.m0 + mrow<T>(tid >> 6, lane >> 4), nn = task.n0 + (lane & 15);
    if (raw) {        // fixed-order mode: the unsigned tile into the task's range
        if (mm < M && nn < N) stg(raw + task.lo + (mm - task.m0) * 16 + (nn - task.n0), (T)b.xfac * sum);
    } else if (mm < M && nn < N) {
        const DevOperand &oc = b.C;
        const bool caff = !oc.table && oc.first_affine && oc.second_affine;
        bool neg;
        const int64_t a = result_offset(oc, caff, mm, nn, neg);
        if (a >= 0) {
            T *pc = bufs[b.buf_c] + b.base_c;
            const T x = (T)b.xfac * sum;
            const T v = neg ? -x : x;
            if (b.atomic) atomic_add_g(pc + a, v);
            else stg(pc + a, ldg(pc + a) + v);
        }
    }
}

// Tasks of the strided-GEMM launch with a short contraction (K <= 32) and
// strided A and B, e.g. the rank-1 to rank-25 updates T1 x integral of
// CCSD: MFMA tiles would idle on padding and on per-task latency, so each
// thread computes whole dot products on the vector ALU.  They share the
// launch with the GEMM tiles, whose tail they fill.  A task covers dm rows x dn columns, task.pad = dn | ng << 9 |
// rpt << 18: the first ng * dn <= NT threads form ng row groups of dn
// columns (the others only help to load), dm = rpt * ng with
// K * dm <= DIRECT_A.  Thread
// (nl, ml) owns column n0 + nl and rows m0 + ml + ng * i, i < rpt.  The A
// tile (K x dm) is staged in LDS, the thread's B column in registers.  The
// result map is read n-fastest (table = 2 stores the table transposed).
constexpr int DIRECT_A = 32 * 32;        // LDS elements for the A tile
constexpr int DIRECT_STAGE = 4096;       // LDS elements for a staged result range

template <typename T>
__device__ __forceinline__ void direct_task(const DevTask &task, const DevBlock *__restrict__ blocks,
                                            T *const *__restrict__ bufs, T *lds, int tid, T *raw) {
    T *const As = lds;
    T *const Cs = lds + DIRECT_A;
    const DevBlock &b = blocks[task.blk];
    const int M = b.M, N = b.N, K = b.K;
    const int dn = task.pad & 511, ng = (task.pad >> 9) & 511, rpt = task.pad >> 18;
    const int dm = rpt * ng;
    const int nl = tid % dn, ml = tid / dn;
    const bool active = ml < ng;
    const int m0 = task.m0, n = task.n0 + nl;
    const DevOperand &oa = b.A, &ob = b.B, &oc = b.C;
    const T *__restrict__ pa = bufs[b.buf_a] + b.base_a + oa.f0 + oa.s0;
    const T *__restrict__ pb = bufs[b.buf_b] + b.base_b + ob.f0 + ob.s0;
    T *pc = bufs[b.buf_c] + b.base_c;
    // The A tile (K * dm <= DIRECT_A = 4 NT elements) and the first DB
    // values of the thread's B column: every load in flight at once, issued
    // before the barrier that frees the LDS of the previous task (these tasks
    // are memory-bound; a loop of load-then-store had kept one load in flight
    // per thread and left them waiting on latency).
    constexpr int DB = 16;
    const bool colok = active && n < N;
    const T *__restrict__ pbn = pb + ob.s1 * (int64_t)n;
    T bv[DB];
#pragma unroll
    for (int u = 0; u < DB; ++u) bv[u] = colok && u < K ? ldg(pbn + ob.f1 * (int64_t)u) : (T)0;
    constexpr int AU = DIRECT_A / NT;
    const int KD = K * dm;
    T av[AU];
#pragma unroll
    for (int u = 0; u < AU; ++u) {
        const int e = tid + u * NT;
        av[u] = (T)0;
        if (e < KD) {
            const int k = e / dm, m = m0 + (e - k * dm);
            if (m < M) av[u] = ldg(pa + (oa.f1 * (int64_t)k + oa.s1 * (int64_t)m));
        }
    }
    __syncthreads();   // LDS of the previous task is no longer read
#pragma unroll
    for (int u = 0; u < AU; ++u) {
        const int e = tid + u * NT;
        if (e < KD) As[e] = av[u];
    }
    const int span = task.span;
    for (int i = tid; i < span; i += NT) Cs[i] = (T)0;
    __syncthreads();
    // result column data (separable / affine maps)
    int64_t cadd = 0;
    bool cneg = false, cok = active && n < N;
    if (cok && oc.table == 0) {
        if (oc.second_affine) {
            cadd = oc.s0 + oc.s1 * (int64_t)n;
        } else {
            const int32_t v = ldg(oc.second + n);
            cok = v != 0;
            cneg = v < 0;
            cadd = (int64_t)(v < 0 ? -v : v) - 1;
        }
    }
    // K outer, rpt rows inner: one B value live at a time
    T acc[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) acc[i] = (T)0;
    if (cok) {
        // the B column in rounds of DB loads in flight (the first round is
        // in bv already); per round the rows one after another (a scheduling
        // barrier per row keeps the compiler from hoisting every row's LDS
        // reads of the round at once, which took the kernel past its
        // register budget)
        for (int k0 = 0; k0 < K; k0 += DB) {
            if (k0 > 0) {
#pragma unroll
                for (int u = 0; u < DB; ++u) bv[u] = k0 + u < K ? ldg(pbn + ob.f1 * (int64_t)(k0 + u)) : (T)0;
            }
            const int ku = min(DB, K - k0);
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                if (i < rpt) {
                    const T *ar = As + k0 * dm + ml + ng * i;
#pragma unroll
                    for (int u = 0; u < DB; ++u)
                        if (u < ku) acc[i] = fma(ar[u * dm], bv[u], acc[i]);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    }
    if (raw) {        // fixed-order mode: rows of dn columns in the task's range
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int m = m0 + ml + ng * i;
            if (i < rpt && active && n < N && m < M)
                stg(raw + task.lo + (int64_t)(m - m0) * dn + nl, (T)b.xfac * acc[i]);
        }
        return;
    }
    // results: into the LDS stage (span > 0), else non-returning atomic adds
    // (with a single writer the L2 performs the same IEEE addition as a
    // read-modify-write).  The result-map entries of all the thread's rows
    // are loaded before the first add, so their latencies overlap.
    int64_t adr[8];
    bool ng8[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int m = m0 + ml + ng * i;
        adr[i] = -1;
        ng8[i] = cneg;
        if (!(i < rpt && cok && m < M)) continue;
        if (oc.table) {
            const int32_t v = ldg(oc.first + ((int64_t)n + oc.ld * (int64_t)m));
            ng8[i] = v < 0;
            adr[i] = (int64_t)(v < 0 ? -v : v) - 1;
        } else if (oc.first_affine) {
            adr[i] = oc.f0 + oc.f1 * (int64_t)m + cadd;
        } else {
            const int32_t v = ldg(oc.first + m);
            ng8[i] = cneg != (v < 0);
            adr[i] = v == 0 ? -1 : (int64_t)(v < 0 ? -v : v) - 1 + cadd;
        }
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        if (adr[i] < 0) continue;
        const T x = (T)b.xfac * acc[i];
        const T v = ng8[i] ? -x : x;
        if (span > 0) atomicAdd(Cs + (adr[i] - task.lo), v);
        else atomic_add_g(pc + adr[i], v);
    }
    if (span > 0) {
        __syncthreads();
        T *pr = pc + task.lo;
        if (task.owned) {
            for (int i0 = 0; i0 < span; i0 += 4 * NT) {
                T o[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int i = i0 + u * NT + tid;
                    o[u] = i < span ? ldg(pr + i) : (T)0;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int i = i0 + u * NT + tid;
                    if (i < span) stg(pr + i, o[u] + Cs[i]);
                }
            }
        } else {
            for (int i = tid; i < span; i += NT) {
                const T v = Cs[i];
                if (v != (T)0) atomic_add_g(pr + i, v);
            }
        }
    }
}

// 128 x 128 tiles hold 128 accumulator registers per lane: the launch bound
// asks for 2 waves per SIMD (at most 256 registers per lane in total).
// PF: K steps whose operand loads are in flight while a step's MFMAs run
// (1: the next step; 2: the next two, in two register sets).  PF = 3
// (float lists, TBK = 16): the next step, and the GEMM tiles run as split-
// bfloat16 products -- each operand element x = hi + lo with hi = bf16(x),
// lo = bf16(x - hi) stored k-contiguous in LDS, and the tile product as
// lo.hi + hi.lo + hi.hi on v_mfma_f32_16x16x16_bf16 (float32 accumulation);
// the short and thin tasks of the launch stay float32.
// WM: waves along M (2: each wave a (TBM/2) x (TBN/2) quarter; 4: each wave
// TBM/4 rows across the whole tile width, so that a wave whose rows all lie
// beyond the block -- M = 100 in 64-row tiles: rows 112..127 -- skips its
// MFMAs).
// Waves per SIMD the register allocation aims at: 64 x 64 x 16 tiles with
// one K step in flight fit 6 in float32 (80 registers) and 4 in float64 (104
// registers; without the hint the direct tasks' batched loads would take the
// float64 kernel to 133 registers and 3 waves).
template <int TBM, int TBN, int TBK, typename T, int PF, int WM>
constexpr int AFFINE_WAVES = TBM * TBN >= 128 * 128 ? 2
                           : !(TBM == 64 && TBN == 64 && TBK == 16 && PF == 1) ? 1
                           : sizeof(T) == 4 ? 6 : 4;

// RAW: fixed-order mode (tiles into raw instead of atomic adds), its own
// instantiation so that the other kernels keep their registers.
template <int TBM, int TBN, int TBK, typename T, int PF = 1, int WM = 2, bool RAW = false>
__global__ __launch_bounds__(NT, (AFFINE_WAVES<TBM, TBN, TBK, T, PF, WM>)) void grouped_contract_affine(
    const DevBlock *__restrict__ blocks, const DevTask *__restrict__ tasks, int ntasks,
    T *const *__restrict__ bufs, const int32_t *__restrict__ skip, const int32_t *__restrict__ zero,
    T *__restrict__ raw_arg) {
    T *const raw = RAW ? raw_arg : nullptr;
    typedef typename Acc<T>::type AccT;
    constexpr int LDA = TBM + 1, LDB = TBN + 1;
    constexpr int WN = 4 / WM, RM = TBM / WM, RN = TBN / WN;   // rows and columns per wave
    constexpr int TM = RM / 16, TN = RN / 16;
    static_assert(WM * WN == 4 && TM >= 1 && TN >= 1, "wave layout");
    constexpr int SA = TBK * LDA, SB = TBK * LDB;
    // double-buffered A and B tiles; after the K loop the array holds half
    // of the result tile (epilogue), or the partial tiles of a skinny task;
    // a direct task uses DIRECT_A + DIRECT_STAGE elements
    constexpr int LDS_T = 2 * SA + 2 * SB > (TBM / 2) * (TBN + 1) ? 2 * SA + 2 * SB : (TBM / 2) * (TBN + 1);
    constexpr int LDS_N = LDS_T > DIRECT_A + DIRECT_STAGE ? LDS_T : DIRECT_A + DIRECT_STAGE;
    __shared__ T lds[LDS_N];
    static_assert(LDS_N >= 1024, "LDS too small for the skinny-task reduction");
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave % WM, wn = wave / WM;
    const int lrow = lane >> 4, lcol = lane & 15;

    for (int t = blockIdx.x; t < ntasks; t += gridDim.x) {
        const DevTask task = tasks[t];
        if (skipped(skip, task.grp) || screened(zero, blocks[task.blk])) continue;
        if (task.pad == 1) {   // tile of at most 16 x 16
            skinny_task(task, blocks, bufs, lds, tid, raw);
            continue;
        }
        if (task.pad > 1) {    // short contraction on the vector ALU
            direct_task(task, blocks, bufs, lds, tid, raw);
            continue;
        }
        const DevBlock &b = blocks[task.blk];
        const int M = b.M, N = b.N;
        const int m0 = task.m0, n0 = task.n0;
        // The K steps walk the task's range [task.k0, task.k1), counted from
        // the start of block task.blk and continued along its chain: blocks
        // of one chain share M, N and the result map, and differ in their
        // operands.  (blk, kb): current block and the chain position of its
        // start; [kl, kh): block-local range of the current step's block.
        int blk = task.blk, kb = 0, kcur = b.K;
        int kl = task.k0, kh = min(kcur, task.k1);
        AffView<TBM, TBK, T> va;
        AffView<TBN, TBK, T> vb;
        va.init(bufs[b.buf_a] + b.base_a, b.A, kcur, m0, M, kl, tid);
        vb.init(bufs[b.buf_b] + b.base_b, b.B, kcur, n0, N, kl, tid);

        AccT acc[TM][TN];
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) acc[i][j] = AccT{0, 0, 0, 0};
        // a wave whose 16-row or 16-column MFMA sub-tiles all lie outside
        // the block skips its MFMAs (scalar branch); otherwise every
        // sub-tile is computed, padding rows and columns from the zeros that
        // the masked loads put into LDS (computing only the sub-tiles that
        // reach into the block, behind a scalar branch per MFMA, made the
        // kernel take 163 instead of 126 registers, 3 waves/SIMD instead of
        // 4, and the launch 19% slower)
        const bool wave_live = __builtin_amdgcn_readfirstlane(
            (m0 + wm * RM < M && n0 + wn * RN < N) ? 1 : 0);
        if constexpr (PF == 3) {
            static_assert(sizeof(T) == 4 && TBK == 16, "split-bfloat16 tiles: float lists, TBK 16");
            constexpr int LK = TBK + 4;                       // padded k row (bf16)
            constexpr int HALF = 2 * (TBM + TBN) * LK;        // one buffer: A hi, A lo, B hi, B lo
            static_assert(2 * HALF * 2 <= LDS_N * (int)sizeof(T), "LDS too small for split tiles");
            unsigned short *const L16 = reinterpret_cast<unsigned short *>(lds);
            auto split_store = [&](int c, const T *xa, const T *xb) {
                unsigned short *ah = L16 + c * HALF, *al = ah + TBM * LK;
                unsigned short *bh = al + TBM * LK, *bl = bh + TBN * LK;
#pragma unroll
                for (int r = 0; r < AffView<TBM, TBK, T>::L; ++r) {
                    const int k = va.kk0 + va.kstep * r, x = va.xi0 + va.xstep * r;
                    const unsigned short h = bf16_of(xa[r]);
                    ah[x * LK + k] = h;
                    al[x * LK + k] = bf16_of(xa[r] - bf16_val(h));
                }
#pragma unroll
                for (int r = 0; r < AffView<TBN, TBK, T>::L; ++r) {
                    const int k = vb.kk0 + vb.kstep * r, x = vb.xi0 + vb.xstep * r;
                    const unsigned short h = bf16_of(xb[r]);
                    bh[x * LK + k] = h;
                    bl[x * LK + k] = bf16_of(xb[r] - bf16_val(h));
                }
            };
            T ra[AffView<TBM, TBK, T>::L], rb[AffView<TBN, TBK, T>::L];
            __syncthreads();   // LDS of the previous task is no longer read
            va.load(kl, kh, ra);
            vb.load(kl, kh, rb);
            split_store(0, ra, rb);
            __syncthreads();
            int cur = 0;
            const int kg = 4 * (lane >> 4);                  // this lane's 4 k of the 16
            for (;;) {
                bool more = true;
                kl += TBK;
                if (kl >= kh) {
                    kb += kcur;
                    blk = blocks[blk].next;
                    more = blk >= 0 && kb < task.k1;
                    if (more) {
                        const DevBlock &nb = blocks[blk];
                        kcur = nb.K;
                        kl = 0;
                        kh = min(kcur, task.k1 - kb);
                        va.init(bufs[nb.buf_a] + nb.base_a, nb.A, kcur, m0, M, 0, tid);
                        vb.init(bufs[nb.buf_b] + nb.base_b, nb.B, kcur, n0, N, 0, tid);
                    }
                }
                if (more) {
                    va.load(kl, kh, ra);
                    vb.load(kl, kh, rb);
                }
                if (wave_live) {
                    const unsigned short *ah = L16 + cur * HALF, *al = ah + TBM * LK;
                    const unsigned short *bh = al + TBM * LK, *bl = bh + TBN * LK;
                    v4s a_h[TM], a_l[TM], b_h[TN], b_l[TN];
#pragma unroll
                    for (int i = 0; i < TM; ++i) {
                        const int row = wm * RM + i * 16 + lcol;
                        a_h[i] = *reinterpret_cast<const v4s *>(ah + row * LK + kg);
                        a_l[i] = *reinterpret_cast<const v4s *>(al + row * LK + kg);
                    }
#pragma unroll
                    for (int j = 0; j < TN; ++j) {
                        const int row = wn * RN + j * 16 + lcol;
                        b_h[j] = *reinterpret_cast<const v4s *>(bh + row * LK + kg);
                        b_l[j] = *reinterpret_cast<const v4s *>(bl + row * LK + kg);
                    }
#pragma unroll
                    for (int i = 0; i < TM; ++i)
#pragma unroll
                        for (int j = 0; j < TN; ++j) {
                            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a_l[i], b_h[j], acc[i][j], 0, 0, 0);
                            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a_h[i], b_l[j], acc[i][j], 0, 0, 0);
                            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a_h[i], b_h[j], acc[i][j], 0, 0, 0);
                        }
                }
                if (more) split_store(cur ^ 1, ra, rb);
                __syncthreads();
                cur ^= 1;
                if (!more) break;
            }
        } else if constexpr (PF == 2) {
            // fetch(): loads the step at the current position and advances
            // the position (along the chain); false once the range is done
            bool have = true;
            // fa / fb: the layouts the step was loaded under
            auto fetch = [&](T *xa, T *xb, bool &fa, bool &fb) -> bool {
                if (!have) return false;
                fa = va.ff;
                fb = vb.ff;
                va.load(kl, kh, xa);
                vb.load(kl, kh, xb);
                kl += TBK;
                if (kl >= kh) {
                    kb += kcur;
                    blk = blocks[blk].next;
                    have = blk >= 0 && kb < task.k1;
                    if (have) {
                        const DevBlock &nb = blocks[blk];
                        kcur = nb.K;
                        kl = 0;
                        kh = min(kcur, task.k1 - kb);
                        va.init(bufs[nb.buf_a] + nb.base_a, nb.A, kcur, m0, M, 0, tid);
                        vb.init(bufs[nb.buf_b] + nb.base_b, nb.B, kcur, n0, N, 0, tid);
                    }
                }
                return true;
            };
            auto mma = [&](int c) {
                const T *A = lds + c * SA;
                const T *B = lds + 2 * SA + c * SB;
                if (!wave_live) return;
#pragma unroll
                for (int kq = 0; kq < TBK / 4; ++kq) {
                    const int kr = kq * 4 + lrow;
                    T a[TM], bb[TN];
#pragma unroll
                    for (int i = 0; i < TM; ++i) a[i] = A[kr * LDA + wm * RM + i * 16 + lcol];
#pragma unroll
                    for (int j = 0; j < TN; ++j) bb[j] = B[kr * LDB + wn * RN + j * 16 + lcol];
#pragma unroll
                    for (int i = 0; i < TM; ++i)
#pragma unroll
                        for (int j = 0; j < TN; ++j) acc[i][j] = mfma16(a[i], bb[j], acc[i][j]);
                }
            };
            typedef AffView<TBM, TBK, T> VA;
            typedef AffView<TBN, TBK, T> VB;
            T xa[VA::L], xb[VB::L], ya[VA::L], yb[VB::L];
            bool fxa, fxb, fya, fyb;
            __syncthreads();   // LDS of the previous task is no longer read
            fetch(xa, xb, fxa, fxb);     // the task's range is never empty
            VA::template store_as<LDA>(lds, xa, fxa, tid);
            VB::template store_as<LDB>(lds + 2 * SA, xb, fxb, tid);
            __syncthreads();
            bool okx = fetch(xa, xb, fxa, fxb);
            bool oky = fetch(ya, yb, fya, fyb);
            int cur = 0;
            for (;;) {
                mma(cur);
                if (okx) {
                    VA::template store_as<LDA>(lds + (cur ^ 1) * SA, xa, fxa, tid);
                    VB::template store_as<LDB>(lds + 2 * SA + (cur ^ 1) * SB, xb, fxb, tid);
                }
                __syncthreads();
                cur ^= 1;
                if (!okx) break;
                okx = fetch(xa, xb, fxa, fxb);
                mma(cur);
                if (oky) {
                    VA::template store_as<LDA>(lds + (cur ^ 1) * SA, ya, fya, tid);
                    VB::template store_as<LDB>(lds + 2 * SA + (cur ^ 1) * SB, yb, fyb, tid);
                }
                __syncthreads();
                cur ^= 1;
                if (!oky) break;
                oky = fetch(ya, yb, fya, fyb);
            }
        } else {
        T ra[AffView<TBM, TBK, T>::L], rb[AffView<TBN, TBK, T>::L];
        __syncthreads();   // LDS of the previous task is no longer read
        va.load(kl, kh, ra);
        vb.load(kl, kh, rb);
        va.template store<LDA>(lds, ra);
        vb.template store<LDB>(lds + 2 * SA, rb);
        __syncthreads();
        int cur = 0;
        for (;;) {
            // next step: the rest of this block, else the next chain block
            bool more = true;
            kl += TBK;
            if (kl >= kh) {
                kb += kcur;
                blk = blocks[blk].next;
                more = blk >= 0 && kb < task.k1;
                if (more) {
                    const DevBlock &nb = blocks[blk];
                    kcur = nb.K;
                    kl = 0;
                    kh = min(kcur, task.k1 - kb);
                    va.init(bufs[nb.buf_a] + nb.base_a, nb.A, kcur, m0, M, 0, tid);
                    vb.init(bufs[nb.buf_b] + nb.base_b, nb.B, kcur, n0, N, 0, tid);
                }
            }
            if (more) {
                va.load(kl, kh, ra);
                vb.load(kl, kh, rb);
            }
            const T *A = lds + cur * SA;
            const T *B = lds + 2 * SA + cur * SB;
            if (wave_live) {
                // the LDS reads of k quad kq + 1 are issued before the MFMAs
                // of quad kq, so they return while those MFMAs run
                T a[2][TM], bb[2][TN];
#pragma unroll
                for (int i = 0; i < TM; ++i) a[0][i] = A[lrow * LDA + wm * RM + i * 16 + lcol];
#pragma unroll
                for (int j = 0; j < TN; ++j) bb[0][j] = B[lrow * LDB + wn * RN + j * 16 + lcol];
#pragma unroll
                for (int kq = 0; kq < TBK / 4; ++kq) {
                    const int c = kq & 1;
                    if (kq + 1 < TBK / 4) {
                        const int kr = (kq + 1) * 4 + lrow;
#pragma unroll
                        for (int i = 0; i < TM; ++i) a[c ^ 1][i] = A[kr * LDA + wm * RM + i * 16 + lcol];
#pragma unroll
                        for (int j = 0; j < TN; ++j) bb[c ^ 1][j] = B[kr * LDB + wn * RN + j * 16 + lcol];
                    }
                    // keep the compiler from sinking those reads below the
                    // MFMAs (it would reuse the operand registers)
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int i = 0; i < TM; ++i)
#pragma unroll
                        for (int j = 0; j < TN; ++j)
                            acc[i][j] = mfma16(a[c][i], bb[c][j], acc[i][j]);
                }
            }
            if (more) {
                va.template store<LDA>(lds + (cur ^ 1) * SA, ra);
                vb.template store<LDB>(lds + 2 * SA + (cur ^ 1) * SB, rb);
            }
            __syncthreads();
            cur ^= 1;
            if (!more) break;
        }
        }

        const DevOperand &oc = b.C;
        T *pc = bufs[b.buf_c] + b.base_c;
        const T xfac = (T)b.xfac;
        // The tile goes through LDS as a row-major [row][col] array (row
        // stride TBN + 1 against bank conflicts), one half of the rows --
        // those of the waves whose rows lie in half h -- at a time; then each thread adds
        // elements n-fastest, so the result-map reads are coalesced.  Every
        // element is a non-returning atomic add, also when the task owns its
        // result elements (injective map, one K chunk, no overlapping
        // block): with a single writer the L2 performs the same IEEE
        // addition as a read-modify-write, and the wave does not wait for
        // old values.  Colliding elements of a non-injective map (the two
        // orderings of an antisymmetrised string pair) add atomically.
        constexpr int HR = TBM / 2, LDE = TBN + 1;
        // Thread tid adds column col = tid % TBN of rows r0 + RSTEP*k of each
        // half.  Its column's half of a separable result map is read once;
        // the row halves (or table entries) of all its RPT elements are
        // loaded before the first of its atomic adds, so their latencies
        // overlap instead of adding up element by element.
        constexpr int RSTEP = NT / TBN, RPT = HR / RSTEP;
        static_assert(NT % TBN == 0 && HR % RSTEP == 0, "epilogue thread layout");
        const int col = tid % TBN, r0 = tid / TBN;
        const int n = n0 + col;
        if (raw) {
            // fixed-order mode: the tile goes unsigned into the task's own
            // range of raw (row-major, TBN columns); a later pass sums the
            // tiles of every result element in a fixed order
            for (int h = 0; h < 2; ++h) {
                if (h) __syncthreads();
                if ((wm * RM) / HR == h) {
#pragma unroll
                    for (int i = 0; i < TM; ++i)
#pragma unroll
                        for (int j = 0; j < TN; ++j)
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                                lds[(wm * RM - h * HR + i * 16 + mrow<T>(r, lrow)) * LDE + wn * RN + j * 16 + lcol] =
                                    acc[i][j][r];
                }
                __syncthreads();
#pragma unroll
                for (int k = 0; k < RPT; ++k) {
                    const int row = r0 + RSTEP * k, m = m0 + h * HR + row;
                    if (m < M && n < N)
                        stg(raw + task.lo + (int64_t)(h * HR + row) * TBN + col, xfac * lds[row * LDE + col]);
                }
            }
            continue;
        }
        bool cok = n < N, cneg = false;
        int64_t cn = 0;
        if (cok && !oc.table) {
            if (oc.second_affine) {
                cn = oc.s0 + oc.s1 * (int64_t)n;
            } else {
                const int32_t t = ldg(oc.second + n);
                cok = t != 0; cneg = t < 0; cn = (int64_t)(t < 0 ? -t : t) - 1;
            }
        }
        for (int h = 0; h < 2; ++h) {
            if (h) __syncthreads();   // the previous half has been read
            if ((wm * RM) / HR == h) {
#pragma unroll
                for (int i = 0; i < TM; ++i)
#pragma unroll
                    for (int j = 0; j < TN; ++j)
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            lds[(wm * RM - h * HR + i * 16 + mrow<T>(r, lrow)) * LDE + wn * RN + j * 16 + lcol] =
                                acc[i][j][r];
            }
            const int mh = m0 + h * HR;
            int64_t adr[RPT];
            bool ng[RPT];
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                const int m = mh + r0 + RSTEP * k;
                adr[k] = -1;
                ng[k] = cneg;
                if (!cok || m >= M) continue;
                if (oc.table) {
                    const int32_t t = ldg(oc.first + ((int64_t)m + oc.ld * (int64_t)n));
                    ng[k] = t < 0;
                    adr[k] = t != 0 ? (int64_t)(t < 0 ? -t : t) - 1 : -1;
                } else if (oc.first_affine) {
                    adr[k] = oc.f0 + oc.f1 * (int64_t)m + cn;
                } else {
                    const int32_t t = ldg(oc.first + m);
                    ng[k] = cneg != (t < 0);
                    adr[k] = t != 0 ? (int64_t)(t < 0 ? -t : t) - 1 + cn : -1;
                }
            }
            __syncthreads();   // (h = 0: the K loop ended with a barrier)
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                if (adr[k] < 0) continue;
                const T x = xfac * lds[(r0 + RSTEP * k) * LDE + col];
                atomic_add_g(pc + adr[k], ng[k] ? -x : x);
            }
        }
    }
}

// Streaming launches (gathers, list initialisation, projection) with many
// tasks: workgroup w of G takes tasks w, w + G, w + 2G, ... (per of them;
// strided, so that a contiguous run of live tasks -- one instance's lists
// -- spreads over all workgroups).  Its threads read the task headers
// together and keep the live ones (group not flagged in skip) in LDS, so
// tasks of converged instances cost one parallel read instead of a
// workgroup each.  per = 1 (at most 8192 tasks): one task per workgroup.
// The task bodies hold no barriers and write disjoint elements, so the
// order of the live tasks does not matter.
template <typename TaskT, typename F>
__device__ __forceinline__ void for_live_tasks(const TaskT *__restrict__ tasks, int ntasks, int per,
                                               const int32_t *__restrict__ skip, F &&f) {
    if (per == 1) {
        const TaskT t = tasks[blockIdx.x];
        if (!skipped(skip, t.grp)) f(t);
        return;
    }
    __shared__ int live[256];
    __shared__ int nlive;
    if (threadIdx.x == 0) nlive = 0;
    __syncthreads();
    const int mine = (int)blockIdx.x + (int)threadIdx.x * (int)gridDim.x;
    if ((int)threadIdx.x < per && mine < ntasks && !skipped(skip, tasks[mine].grp))
        live[atomicAdd(&nlive, 1)] = mine;
    __syncthreads();
    const int nl = nlive;
    for (int i = 0; i < nl; ++i) {
        const TaskT t = tasks[live[i]];
        f(t);
    }
}

static inline int tasks_per_group(int ntasks) {
    return ntasks <= 8192 ? 1 : min(256, (ntasks + 8191) / 8192);
}

// Signed gather of operands into scratch (one launch for all jobs of a
// plan); HBM-bound: one read of each allowed source element, one write.
template <typename T>
__global__ __launch_bounds__(256) void gather_operands(
    const DevGather *__restrict__ jobs, const DevGatherTask *__restrict__ tasks, int ntasks,
    T *const *__restrict__ bufs, T *__restrict__ scratch,
    const int32_t *__restrict__ skip, int per) {
  for_live_tasks(tasks, ntasks, per, skip, [&](const DevGatherTask &t) {
    const DevGather &g = jobs[t.job];
    const T *__restrict__ src = bufs[g.src_buf] + g.src_base;
    T *__restrict__ dst = scratch + g.dst_base;
    const DevOperand &o = g.src;
    const uint32_t inner = g.kfast ? (uint32_t)g.K : (uint32_t)g.X;
    // signed source offset of element e (-1: forbidden); U elements per
    // thread and round, so their index and data loads are in flight together
    auto source = [&](uint32_t e, bool &neg) -> int64_t {
        const uint32_t q = e / inner, rmd = e - q * inner;
        const uint32_t k = g.kfast ? rmd : q, x = g.kfast ? q : rmd;
        if (o.table) {
            const int32_t v = ldg(o.first + ((int64_t)k + o.ld * (int64_t)x));
            neg = v < 0;
            return v == 0 ? -1 : (int64_t)(v < 0 ? -v : v) - 1;
        }
        int64_t a1, a2;
        bool n1 = false, n2 = false, ok = true;
        if (o.first_affine) {
            a1 = o.f0 + o.f1 * (int64_t)k;
        } else {
            const int32_t v = ldg(o.first + k);
            ok = v != 0; n1 = v < 0; a1 = (int64_t)(v < 0 ? -v : v) - 1;
        }
        if (o.second_affine) {
            a2 = o.s0 + o.s1 * (int64_t)x;
        } else {
            const int32_t v = ldg(o.second + x);
            ok = ok && v != 0; n2 = v < 0; a2 = (int64_t)(v < 0 ? -v : v) - 1;
        }
        neg = n1 != n2;
        return ok ? a1 + a2 : -1;
    };
    constexpr int U = 4;
    for (uint32_t e0 = (uint32_t)t.e0 + threadIdx.x; e0 < (uint32_t)t.e1; e0 += U * blockDim.x) {
        int64_t addr[U];
        bool neg[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t e = e0 + u * blockDim.x;
            addr[u] = e < (uint32_t)t.e1 ? source(e, neg[u]) : -1;
        }
        T val[U];
#pragma unroll
        for (int u = 0; u < U; ++u) val[u] = addr[u] >= 0 ? ldg(src + addr[u]) : (T)0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t e = e0 + u * blockDim.x;
            if (e < (uint32_t)t.e1) stg(dst + e, neg[u] ? -val[u] : val[u]);
        }
    }
  });
}

// List initialisation (copy with factor, or zero) ahead of the contractions.
template <typename T>
__global__ __launch_bounds__(256) void init_lists(const DevInit *__restrict__ jobs,
                                                      const DevInitTask *__restrict__ tasks,
                                                      T *const *__restrict__ bufs,
                                                      const int32_t *__restrict__ skip, int ntasks,
                                                      int per) {
  for_live_tasks(tasks, ntasks, per, skip, [&](const DevInitTask &t) {
    const DevInit &j = jobs[t.job];
    T *__restrict__ dst = bufs[j.dst] + j.off_dst;
    // U elements per thread and round, their loads in flight together
    constexpr int U = 4;
    if (j.src >= 0) {
        const T *__restrict__ src = bufs[j.src] + j.off_src;
        const T f = (T)j.factor;
        for (int64_t e0 = t.e0 + threadIdx.x; e0 < t.e1; e0 += U * blockDim.x) {
            T x[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int64_t e = e0 + u * blockDim.x;
                x[u] = e < t.e1 ? ldg(src + e) : (T)0;
            }
            if (j.accumulate) {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int64_t e = e0 + u * blockDim.x;
                    if (e < t.e1 && x[u] != (T)0) atomic_add_g(dst + e, f * x[u]);
                }
            } else {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int64_t e = e0 + u * blockDim.x;
                    if (e < t.e1) stg(dst + e, f * x[u]);
                }
            }
        }
    } else {
        for (int64_t e = t.e0 + threadIdx.x; e < t.e1; e += blockDim.x) stg(dst + e, (T)0);
    }
  });
}

// Partial trace rows: dst[row] += fact * sum_k sign_k * src[col_k] over the
// row's entries (CSR, one thread per row, fixed order: no atomics).
template <typename T>
__global__ __launch_bounds__(256) void trace_rows(const T *__restrict__ src,
                                                      T *__restrict__ dst,
                                                      const int64_t *__restrict__ rowptr,
                                                      const int64_t *__restrict__ rowdst,
                                                      const int64_t *__restrict__ col,
                                                      const int8_t *__restrict__ sgn, int64_t nrows,
                                                      double fact) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows;
         r += (int64_t)gridDim.x * blockDim.x) {
        T acc = (T)0;
        for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
            const T v = src[col[k]];
            acc += sgn[k] < 0 ? -v : v;
        }
        dst[rowdst[r]] += (T)fact * acc;
    }
}

// All spin-flip symmetrisation steps of a plan in one launch (lists of one
// plan are distinct buffers or disjoint ranges; pairs of one list are
// disjoint because the flip is an involution).
template <typename T>
__global__ __launch_bounds__(256) void sym_ab_batched(const DevSymStep *__restrict__ steps,
                                                          const DevSymTask *__restrict__ tasks,
                                                          T *const *__restrict__ bufs,
                                                          const int32_t *__restrict__ i1,
                                                          const int32_t *__restrict__ i2,
                                                          const int8_t *__restrict__ fsign,
                                                          const int32_t *__restrict__ skip, int ntasks,
                                                          int per) {
  for_live_tasks(tasks, ntasks, per, skip, [&](const DevSymTask &t) {
    const DevSymStep &st = steps[t.step];
    T *__restrict__ x = bufs[st.buf] + st.base;
    const T prefac = (T)st.prefac;
    for (int64_t p = t.p0 + threadIdx.x; p < t.p1; p += blockDim.x) {
        const int64_t a = ldg(i1 + p), c = ldg(i2 + p);
        const T f = (T)fsign[p];
        const T v = prefac * (ldg(x + a) + f * ldg(x + c));
        stg(x + a, v);
        if (c != a) stg(x + c, f * v);
    }
  });
}

}  // namespace gecco

// Launchers: elem = 8 (float64 lists) or 4 (float32); bufs holds the
// plan's buffer table of that element type.
template <typename T> static T *const *tbl(void *const *b) { return reinterpret_cast<T *const *>(b); }

extern "C" int gecco_launch_grouped(const gecco::DevBlock *blocks, const gecco::DevTask *tasks,
                                    int ntasks, void *const *bufs, int grid,
                                    const int32_t *skip, const int32_t *zero, int elem,
                                    hipStream_t stream) {
    if (ntasks <= 0) return 0;
    if (elem == 4)
        hipLaunchKernelGGL(gecco::grouped_contract<float>, dim3(grid), dim3(gecco::NT), 0, stream,
                           blocks, tasks, ntasks, tbl<float>(bufs), skip, zero);
    else
        hipLaunchKernelGGL(gecco::grouped_contract<double>, dim3(grid), dim3(gecco::NT), 0, stream,
                           blocks, tasks, ntasks, tbl<double>(bufs), skip, zero);
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

// Tile variants of the affine kernel: {TBM, TBN, TBK}.
extern "C" int gecco_affine_variant_shape(int v, int *bm, int *bn, int *bk) {
    static const int shapes[][3] = {{64, 64, 32}, {64, 64, 16}, {64, 128, 16},
                                    {128, 64, 16}, {128, 128, 16}, {128, 128, 8},
                                    {64, 64, 16}, {64, 64, 32}, {64, 64, 16}, {64, 64, 16}};
    if (v < 0 || v >= (int)(sizeof shapes / sizeof shapes[0])) return 1;
    *bm = shapes[v][0];
    *bn = shapes[v][1];
    *bk = shapes[v][2];
    return 0;
}

template <typename T>
static int launch_affine(const gecco::DevBlock *blocks, const gecco::DevTask *tasks, int ntasks,
                         T *const *bufs, int grid, int variant, const int32_t *skip,
                         const int32_t *zero, T *raw, hipStream_t stream) {
    using namespace gecco;
    if (raw && variant != 1 && variant != 9) return 1;   // fixed-order mode: tile variants 1 and 9
    switch (variant) {
    case 0: hipLaunchKernelGGL((grouped_contract_affine<64, 64, 32, T>), dim3(grid), dim3(NT), 0, stream, blocks, tasks, ntasks, bufs, skip, zero, raw); break;
    case 1:
        if (raw)
            hipLaunchKernelGGL((grouped_contract_affine<64, 64, 16, T, 1, 2, true>), dim3(grid), dim3(NT), 0, stream, blocks, tasks, ntasks, bufs, skip, zero, raw);
        else
            hipLaunchKernelGGL((grouped_contract_affine<64, 64, 16, T>), dim3(grid), dim3(NT), 0, stream, blocks, tasks, ntasks, bufs, skip, zero, raw);
        break;
    case 2: hipLaunchKernelGGL((grouped_contract_affine<64, 128, 16, T>), dim3(grid), dim3(NT), 0, stream, blocks, tasks, ntasks, bufs, skip, zero, raw); break;
    case 3: hipLaunchKernelGGL((grouped_contract_affine<128, 64, 16, T>), dim3(grid), dim3(NT), 0, stream, blocks, tasks, ntasks, bufs, skip, zero, raw); break;
    case 4: hipLaunchKernelGGL((grouped_contract_affine<128, 128, 16, T>), dim3(grid), dim3(NT), 0, stream, blocks, tasks, ntasks, bufs, skip, zero, raw); break;
    case 5: hipLaunchKernelGGL((grouped_contract_affine<128, 128, 8, T>), dim3(grid), dim3(NT), 0, stream, blocks, tasks, ntasks, bufs, skip, zero, raw); break;
    // two K steps of loads in flight
    case 6: hipLaunchKernelGGL((grouped_contract_affine<64, 64, 16, T, 2>), dim3(grid), dim3(NT), 0, stream, blocks, tasks, ntasks, bufs, skip, zero, raw); break;
    case 7: hipLaunchKernelGGL((grouped_contract_affine<64, 64, 32, T, 2>), dim3(grid), dim3(NT), 0, stream, blocks, tasks, ntasks, bufs, skip, zero, raw); break;
    // 64 x 64 x 16 with the waves stacked along M
    case 9:
        if (raw)
            hipLaunchKernelGGL((grouped_contract_affine<64, 64, 16, T, 1, 4, true>), dim3(grid), dim3(NT), 0, stream, blocks, tasks, ntasks, bufs, skip, zero, raw);
        else
            hipLaunchKernelGGL((grouped_contract_affine<64, 64, 16, T, 1, 4>), dim3(grid), dim3(NT), 0, stream, blocks, tasks, ntasks, bufs, skip, zero, raw);
        break;
    // split-bfloat16 GEMM tiles (float lists only)
    case 8:
        if constexpr (sizeof(T) == 4) {
            hipLaunchKernelGGL((grouped_contract_affine<64, 64, 16, T, 3>), dim3(grid), dim3(NT), 0, stream, blocks, tasks, ntasks, bufs, skip, zero, raw);
            break;
        } else {
            return 1;
        }
    default: return 1;
    }
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

extern "C" int gecco_launch_grouped_affine(const gecco::DevBlock *blocks,
                                           const gecco::DevTask *tasks, int ntasks,
                                           void *const *bufs, int grid, int variant,
                                           const int32_t *skip, const int32_t *zero, void *raw, int elem,
                                           hipStream_t stream) {
    if (ntasks <= 0) return 0;
    if (elem == 4)
        return launch_affine(blocks, tasks, ntasks, tbl<float>(bufs), grid, variant, skip, zero, (float *)raw, stream);
    return launch_affine(blocks, tasks, ntasks, tbl<double>(bufs), grid, variant, skip, zero, (double *)raw, stream);
}


extern "C" int gecco_direct_stage_cap(void) { return gecco::DIRECT_STAGE; }

namespace gecco {
// Norm screening, first pass: partial[t] = sum of x^2 over chunk t (one
// workgroup per chunk, fixed-order block reduction).
template <typename T>
__global__ __launch_bounds__(256) void norm_chunks(const DevNormTask *__restrict__ tasks,
                                                   T *const *__restrict__ bufs, double *__restrict__ partial) {
    const DevNormTask t = tasks[blockIdx.x];
    const T *__restrict__ x = bufs[t.buf] + t.off;
    double acc = 0.0;
    for (int64_t e = t.e0 + threadIdx.x; e < t.e1; e += blockDim.x) {
        const double v = (double)ldg(x + e);
        acc += v * v;
    }
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Second pass: zero[j] = 1 when the chunks [rows[j], rows[j+1]) of operand
// block j sum to less than 1e-28 (contract/contr_op1op2_wmaps_c.f:1652-1655).
__global__ __launch_bounds__(256) void norm_flags(const double *__restrict__ partial,
                                                  const int64_t *__restrict__ rows, int64_t njobs,
                                                  int32_t *__restrict__ zero) {
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= njobs) return;
    double s = 0.0;
    for (int64_t t = rows[j]; t < rows[j + 1]; ++t) s += partial[t];
    zero[j] = s < 1e-28 ? 1 : 0;
}
}  // namespace gecco

namespace gecco {
// Singlet spin projection: thread g replaces the elements of group g by
// P_g times them (P_g row-major n x n in mats at gmat[g]; n <= 32).
template <typename T>
__global__ __launch_bounds__(256) void spin_project(T *__restrict__ x, const int64_t *__restrict__ gptr,
                                                    const int64_t *__restrict__ elem,
                                                    const int64_t *__restrict__ gmat,
                                                    const double *__restrict__ mats, int64_t ngroups) {
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (g >= ngroups) return;
    const int64_t e0 = gptr[g];
    const int n = (int)(gptr[g + 1] - e0);
    const double *__restrict__ P = mats + gmat[g];
    double v[32];
    for (int i = 0; i < n && i < 32; ++i) v[i] = (double)x[elem[e0 + i]];
    for (int i = 0; i < n && i < 32; ++i) {
        double s = 0.0;
        for (int j = 0; j < n && j < 32; ++j) s += P[i * n + j] * v[j];
        x[elem[e0 + i]] = (T)s;
    }
}
}  // namespace gecco

extern "C" int gecco_launch_spin_project(void *x, const int64_t *gptr, const int64_t *elem, const int64_t *gmat,
                                         const double *mats, int64_t ngroups, int elem_size, hipStream_t stream) {
    if (ngroups <= 0) return 0;
    const unsigned grid = (unsigned)((ngroups + 255) / 256);
    if (elem_size == 4)
        hipLaunchKernelGGL(gecco::spin_project<float>, dim3(grid), dim3(256), 0, stream, (float *)x, gptr, elem,
                           gmat, mats, ngroups);
    else
        hipLaunchKernelGGL(gecco::spin_project<double>, dim3(grid), dim3(256), 0, stream, (double *)x, gptr, elem,
                           gmat, mats, ngroups);
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

extern "C" int gecco_launch_norms(const gecco::DevNormTask *tasks, int64_t ntasks, const int64_t *rows,
                                  int64_t njobs, void *const *bufs, double *partial, int32_t *zero, int elem,
                                  hipStream_t stream) {
    if (njobs <= 0) return 0;
    if (ntasks > 0) {
        if (elem == 4)
            hipLaunchKernelGGL(gecco::norm_chunks<float>, dim3((unsigned)ntasks), dim3(256), 0, stream, tasks,
                               tbl<float>(bufs), partial);
        else
            hipLaunchKernelGGL(gecco::norm_chunks<double>, dim3((unsigned)ntasks), dim3(256), 0, stream, tasks,
                               tbl<double>(bufs), partial);
    }
    hipLaunchKernelGGL(gecco::norm_flags, dim3((unsigned)((njobs + 255) / 256)), dim3(256), 0, stream, partial,
                       rows, njobs, zero);
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

extern "C" int gecco_launch_gather(const gecco::DevGather *jobs, const gecco::DevGatherTask *tasks,
                                   int ntasks, void *const *bufs, void *scratch,
                                   const int32_t *skip, int elem, hipStream_t stream) {
    if (ntasks <= 0) return 0;
    const int per = gecco::tasks_per_group(ntasks), grid = (ntasks + per - 1) / per;
    if (elem == 4)
        hipLaunchKernelGGL(gecco::gather_operands<float>, dim3(grid), dim3(256), 0, stream, jobs, tasks,
                           ntasks, tbl<float>(bufs), (float *)scratch, skip, per);
    else
        hipLaunchKernelGGL(gecco::gather_operands<double>, dim3(grid), dim3(256), 0, stream, jobs, tasks,
                           ntasks, tbl<double>(bufs), (double *)scratch, skip, per);
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

extern "C" int gecco_launch_init(const gecco::DevInit *jobs, const gecco::DevInitTask *tasks,
                                 int ntasks, void *const *bufs, const int32_t *skip, int elem,
                                 hipStream_t stream) {
    if (ntasks <= 0) return 0;
    const int per = gecco::tasks_per_group(ntasks), grid = (ntasks + per - 1) / per;
    if (elem == 4)
        hipLaunchKernelGGL(gecco::init_lists<float>, dim3(grid), dim3(256), 0, stream, jobs, tasks,
                           tbl<float>(bufs), skip, ntasks, per);
    else
        hipLaunchKernelGGL(gecco::init_lists<double>, dim3(grid), dim3(256), 0, stream, jobs, tasks,
                           tbl<double>(bufs), skip, ntasks, per);
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

extern "C" int gecco_launch_trace(const void *src, void *dst, const int64_t *rowptr,
                                  const int64_t *rowdst, const int64_t *col, const int8_t *sgn,
                                  int64_t nrows, double fact, int elem, hipStream_t stream) {
    if (nrows <= 0) return 0;
    const int64_t grid = std::min<int64_t>((nrows + 255) / 256, 65536);
    if (elem == 4)
        hipLaunchKernelGGL(gecco::trace_rows<float>, dim3((unsigned)grid), dim3(256), 0, stream,
                           (const float *)src, (float *)dst, rowptr, rowdst, col, sgn, nrows, fact);
    else
        hipLaunchKernelGGL(gecco::trace_rows<double>, dim3((unsigned)grid), dim3(256), 0, stream,
                           (const double *)src, (double *)dst, rowptr, rowdst, col, sgn, nrows, fact);
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

extern "C" int gecco_launch_sym_batched(const gecco::DevSymStep *steps,
                                        const gecco::DevSymTask *tasks, int ntasks,
                                        void *const *bufs, const int32_t *i1, const int32_t *i2,
                                        const int8_t *fsign, const int32_t *skip, int elem,
                                        hipStream_t stream) {
    if (ntasks <= 0) return 0;
    const int per = gecco::tasks_per_group(ntasks), grid = (ntasks + per - 1) / per;
    if (elem == 4)
        hipLaunchKernelGGL(gecco::sym_ab_batched<float>, dim3(grid), dim3(256), 0, stream, steps, tasks,
                           tbl<float>(bufs), i1, i2, fsign, skip, ntasks, per);
    else
        hipLaunchKernelGGL(gecco::sym_ab_batched<double>, dim3(grid), dim3(256), 0, stream, steps, tasks,
                           tbl<double>(bufs), i1, i2, fsign, skip, ntasks, per);
    return hipGetLastError() == hipSuccess ? 0 : 3;
}

// ---------------------------------------------------------------- live tasks
// Live-task lists of a plan with instance groups: the records (tasks of one
// launch class, `words` 32-bit words each, group id at word gw) whose
// group's skip flag is 0 -- or whose group is negative -- copied in their
// order.  Three launches: live records per chunk of CCHUNK, one
// workgroup's exclusive scan of the chunk counts (and the total), the
// stable scatter (wave ballots give each live record its rank).
namespace gecco {

constexpr int CCHUNK = 1024;   // records per workgroup, 4 per thread

__device__ __forceinline__ bool record_live(const uint32_t *__restrict__ src, int64_t i, int words,
                                            int gw, const int32_t *__restrict__ skip) {
    const int32_t g = (int32_t)src[i * words + gw];
    return g < 0 || skip[g] == 0;
}

__global__ __launch_bounds__(256) void compact_count(const uint32_t *__restrict__ src, int64_t n,
                                                     int words, int gw, const int32_t *__restrict__ skip,
                                                     int32_t *__restrict__ chunk) {
    int c = 0;
    for (int r = 0; r < CCHUNK / 256; ++r) {
        const int64_t i = (int64_t)blockIdx.x * CCHUNK + r * 256 + threadIdx.x;
        c += __syncthreads_count(i < n && record_live(src, i, words, gw, skip));
    }
    if (threadIdx.x == 0) chunk[blockIdx.x] = c;
}

__global__ __launch_bounds__(256) void compact_scan(int32_t *__restrict__ chunk, int nchunk,
                                                    int64_t *__restrict__ count) {
    __shared__ int s[256];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int b0 = 0; b0 < nchunk; b0 += 256) {
        const int b = b0 + (int)threadIdx.x;
        const int v = b < nchunk ? chunk[b] : 0;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1) {
            const int x = (int)threadIdx.x >= o ? s[threadIdx.x - o] : 0;
            __syncthreads();
            s[threadIdx.x] += x;
            __syncthreads();
        }
        const int incl = s[threadIdx.x];
        if (b < nchunk) chunk[b] = carry + incl - v;   // exclusive offset of chunk b
        __syncthreads();
        if (threadIdx.x == 255) carry += incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) *count = carry;
}

__global__ __launch_bounds__(256) void compact_scatter(const uint32_t *__restrict__ src,
                                                       uint32_t *__restrict__ dst, int64_t n, int words,
                                                       int gw, const int32_t *__restrict__ skip,
                                                       const int32_t *__restrict__ chunk) {
    __shared__ int wsum[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int base = chunk[blockIdx.x];
    for (int r = 0; r < CCHUNK / 256; ++r) {
        const int64_t i = (int64_t)blockIdx.x * CCHUNK + r * 256 + threadIdx.x;
        const bool live = i < n && record_live(src, i, words, gw, skip);
        const uint64_t m = __ballot64(live);
        if (lane == 0) wsum[wave] = __popcll(m);
        __syncthreads();
        int off = base;
        for (int w = 0; w < wave; ++w) off += wsum[w];
        if (live) {
            const int64_t j = off + __popcll(m & ((1ull << lane) - 1));
            for (int q = 0; q < words; ++q) dst[j * words + q] = src[i * words + q];
        }
        const int tot = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
        base += tot;
    }
}

}  // namespace gecco

// chunk: ceil(n / 1024) int32 of device scratch; *count (device) receives
// the number of live records.  n fits int32 ranks (the plans' task counts do).
extern "C" int gecco_launch_compact(const void *src, void *dst, int64_t n, int rec_bytes, int grp_byte,
                                    const int32_t *skip, int32_t *chunk, int64_t *count,
                                    hipStream_t stream) {
    if (n <= 0 || n > INT32_MAX || rec_bytes % 4 || grp_byte % 4 || grp_byte >= rec_bytes) return 1;
    const int64_t nchunk = (n + gecco::CCHUNK - 1) / gecco::CCHUNK;
    const int words = rec_bytes / 4, gw = grp_byte / 4;
    hipLaunchKernelGGL(gecco::compact_count, dim3((unsigned)nchunk), dim3(256), 0, stream,
                       (const uint32_t *)src, n, words, gw, skip, chunk);
    hipLaunchKernelGGL(gecco::compact_scan, dim3(1), dim3(256), 0, stream, chunk, (int)nchunk, count);
    hipLaunchKernelGGL(gecco::compact_scatter, dim3((unsigned)nchunk), dim3(256), 0, stream,
                       (const uint32_t *)src, (uint32_t *)dst, n, words, gw, skip, chunk);
    return hipGetLastError() == hipSuccess ? 0 : 3;
}
