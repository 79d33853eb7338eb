// Amplitude solver of the MI355X engine: diagonal preconditioning and Pulay
// DIIS on device-resident amplitude vectors, and the C ABI declared in
// include/gecco_hip_solver.h.
//
// One step restates, for a vector function g(t) = diag*t + rest:
//   r = g / diag, |diag| <= 1e-10 -> 1e-10      (baserout/diavc.f:31-35)
//   subspace += (r, v = t - r)                  (solve/optc_diis_sbsp_add.f:150-205)
//   B(i, j) = <r_i|r_j>, oldest dropped when full (solve/optc_update_redsp1.f)
//   [B -1; -1 0] (w, l) = (0, -1)               (solve/optc_diis_extr.f:66-105)
//   t = sum_i w_i v_i                           (solve/optc_expand_vec.f)
// Device work per step: one fused pass (g, r, v, ||g||^2 and the new row of
// B), a one-workgroup reduction in fixed order, one expansion pass.  The
// synchronous step solves the (m+1) x (m+1) system on the host after a
// single synchronisation; the asynchronous step keeps the subspace
// bookkeeping and the solve on the device (no synchronisation, the same
// launches with the same arguments every iteration, so an iteration can be
// captured and replayed as a graph) and reports norms through a ring of
// records in pinned host memory.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gecco_hip_solver.h"

namespace gecco {
// error state shared with the contraction engine (plan.cpp)
int set_error(int code, const char *msg);
void clear_error();
}  // namespace gecco

namespace {

constexpr int MAXSUB = 16;       // largest DIIS subspace
constexpr int NTH = 256;
constexpr int NGROUP = 1024;     // workgroups of the fused pass
static_assert((MAXSUB - 1) * (MAXSUB - 1) <= NTH, "one thread per element of the shifted B");
constexpr double DIA_THRESH = 1e-10;

struct Slots {
    int m;                        // vectors in the subspace after the step
    int s[MAXSUB];                // slot of the logical vector j (j = m-1 is the new one)
};

struct Weights {
    int m;
    int s[MAXSUB];
    double w[MAXSUB];
};

// g = diag*t + rest; r = g/diag; v = t - r into slot s[m-1]; per-workgroup
// partial sums of <r_j|r> (j < m) and <g|g>, in part[blockIdx.x*(MAXSUB+1)+.]
// M = sl.m vectors after the step (compile-time: the per-element loop over
// the old vectors has no branches)
template <int M>
__device__ __forceinline__ void prepare_m(const double *__restrict__ t,
                                          const double *__restrict__ rest,
                                          const double *__restrict__ diag,
                                          const double *__restrict__ scale,
                                          double *__restrict__ rsub, double *__restrict__ vsub,
                                          int64_t n, const Slots &sl, double acc[MAXSUB + 1]) {
    double *__restrict__ rn = rsub + (int64_t)sl.s[M - 1] * n;
    double *__restrict__ vn = vsub + (int64_t)sl.s[M - 1] * n;
    const double *__restrict__ old[M > 1 ? M - 1 : 1];
#pragma unroll
    for (int j = 0; j < M - 1; ++j) old[j] = rsub + (int64_t)sl.s[j] * n;
    for (int64_t i = (int64_t)blockIdx.x * NTH + threadIdx.x; i < n; i += (int64_t)gridDim.x * NTH) {
        const double ti = t[i], d = diag[i];
        const double g = d * ti + rest[i];
        const double dv = fabs(d) <= DIA_THRESH ? DIA_THRESH : d;
        double r = g / dv;
        if (scale) r *= scale[i];     // learned per-amplitude factor
        rn[i] = r;
        vn[i] = ti - r;
        acc[MAXSUB] += g * g;
#pragma unroll
        for (int j = 0; j < M - 1; ++j) acc[j] += r * old[j][i];
        acc[M - 1] += r * r;
    }
}

__device__ __forceinline__ void prepare_body(const double *__restrict__ t,
                                             const double *__restrict__ rest,
                                             const double *__restrict__ diag,
                                             const double *__restrict__ scale,
                                             double *__restrict__ rsub, double *__restrict__ vsub,
                                             int64_t n, const Slots &sl, double *__restrict__ part) {
    double acc[MAXSUB + 1];
#pragma unroll
    for (int j = 0; j <= MAXSUB; ++j) acc[j] = 0.0;
    switch (sl.m) {
#define GECCO_PREP(M) case M: prepare_m<M>(t, rest, diag, scale, rsub, vsub, n, sl, acc); break;
        GECCO_PREP(1) GECCO_PREP(2) GECCO_PREP(3) GECCO_PREP(4) GECCO_PREP(5) GECCO_PREP(6)
        GECCO_PREP(7) GECCO_PREP(8) GECCO_PREP(9) GECCO_PREP(10) GECCO_PREP(11) GECCO_PREP(12)
        GECCO_PREP(13) GECCO_PREP(14) GECCO_PREP(15) GECCO_PREP(16)
#undef GECCO_PREP
        default: break;
    }
    __shared__ double red[NTH / 64][MAXSUB + 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j <= MAXSUB; ++j) {
        double v = acc[j];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
        if (lane == 0) red[wave][j] = v;
    }
    __syncthreads();
    if (threadIdx.x <= MAXSUB) {
        double v = 0.0;
        for (int w = 0; w < NTH / 64; ++w) v += red[w][threadIdx.x];
        part[(int64_t)blockIdx.x * (MAXSUB + 1) + threadIdx.x] = v;
    }
}

__global__ __launch_bounds__(NTH) void diis_prepare_f64(
    const double *__restrict__ t, const double *__restrict__ rest,
    const double *__restrict__ diag, const double *__restrict__ scale,
    double *__restrict__ rsub, double *__restrict__ vsub,
    int64_t n, Slots sl, double *__restrict__ part) {
    prepare_body(t, rest, diag, scale, rsub, vsub, n, sl, part);
}

// out[j] = sum over workgroups of part[g][j] in workgroup order; out[MAXSUB+1]
// = energy scalar (or 0)
__global__ __launch_bounds__(NTH) void diis_reduce_f64(const double *__restrict__ part, int ngroup,
                                                       const double *__restrict__ energy,
                                                       double *__restrict__ out) {
    __shared__ double red[NTH / 64][MAXSUB + 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int j = 0; j <= MAXSUB; ++j) {
        double v = 0.0;
        for (int g = threadIdx.x; g < ngroup; g += NTH) v += part[(int64_t)g * (MAXSUB + 1) + j];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
        if (lane == 0) red[wave][j] = v;
    }
    __syncthreads();
    if (threadIdx.x <= MAXSUB) {
        double v = 0.0;
        for (int w = 0; w < NTH / 64; ++w) v += red[w][threadIdx.x];
        out[threadIdx.x] = v;
    }
    if (threadIdx.x == 0) out[MAXSUB + 1] = energy ? energy[0] : 0.0;
}

// t = sum_j w_j v_j (M = wt.m compile-time)
template <int M>
__device__ __forceinline__ void expand_m(double *__restrict__ t, const double *__restrict__ vsub,
                                         int64_t n, const Weights &wt) {
    const double *__restrict__ v[M];
#pragma unroll
    for (int j = 0; j < M; ++j) v[j] = vsub + (int64_t)wt.s[j] * n;
    for (int64_t i = (int64_t)blockIdx.x * NTH + threadIdx.x; i < n; i += (int64_t)gridDim.x * NTH) {
        double x = 0.0;
#pragma unroll
        for (int j = 0; j < M; ++j) x += wt.w[j] * v[j][i];
        t[i] = x;
    }
}

__device__ __forceinline__ void expand_body(double *__restrict__ t, const double *__restrict__ vsub,
                                            int64_t n, const Weights &wt) {
    switch (wt.m) {
#define GECCO_EXP(M) case M: expand_m<M>(t, vsub, n, wt); break;
        GECCO_EXP(1) GECCO_EXP(2) GECCO_EXP(3) GECCO_EXP(4) GECCO_EXP(5) GECCO_EXP(6)
        GECCO_EXP(7) GECCO_EXP(8) GECCO_EXP(9) GECCO_EXP(10) GECCO_EXP(11) GECCO_EXP(12)
        GECCO_EXP(13) GECCO_EXP(14) GECCO_EXP(15) GECCO_EXP(16)
#undef GECCO_EXP
        default: break;
    }
}

__global__ __launch_bounds__(NTH) void diis_expand_f64(double *__restrict__ t,
                                                       const double *__restrict__ vsub, int64_t n,
                                                       Weights wt) {
    expand_body(t, vsub, n, wt);
}

// ---------------------------------------------------------------- async
// Device-resident bookkeeping of the asynchronous step.  The state always
// describes the coming step: m counts the vectors including the one the
// coming step adds, whose slot is order[m-1] (reset: m = 1, order[0] = 0).
constexpr int RING = 64;   // status records kept in pinned host memory

struct DevState {
    int m;                        // vectors after the coming step
    int order[MAXSUB];            // slot of logical vector j, oldest first
    int skip;                     // 1: the last step leaves t unchanged
    int mw;                       // vectors of the last step's extrapolation
    int wslot[MAXSUB];            // their slots
    int64_t steps;                // asynchronous steps taken
    double gn_last;               // ||g|| of the last step that evaluated it
    double b[MAXSUB * MAXSUB];    // <r_i|r_j>, logical order, stride MAXSUB
    double bw[MAXSUB * MAXSUB];   // B of the last step's extrapolation
    double w[MAXSUB];             // weights of the last step
};

struct StatusRec {                // one per step, written by the device
    double gnorm, energy;
    int64_t step;                 // index of the step (-1: none yet)
    int64_t converged;            // 1 converged, 0 extrapolated, -1 singular system
};

__global__ __launch_bounds__(NTH) void diis_prepare_async_f64(
    const double *__restrict__ t, const double *__restrict__ rest,
    const double *__restrict__ diag, const double *__restrict__ scale,
    double *__restrict__ rsub, double *__restrict__ vsub,
    int64_t n, const DevState *__restrict__ st, double *__restrict__ part,
    const int32_t *__restrict__ flag) {
    if (flag && *flag) return;    // converged before this step: nothing to add
    Slots sl;
    sl.m = st->m;
    for (int j = 0; j < MAXSUB; ++j) sl.s[j] = j < sl.m ? st->order[j] : 0;
    prepare_body(t, rest, diag, scale, rsub, vsub, n, sl, part);
}

// Reduction in workgroup order (as diis_reduce_f64), the new row of B, the
// augmented system [B -1; -1 0] (w, l) = (0, -1) by Gaussian elimination
// with partial pivoting (the host step's algorithm; each pivot column's
// search, row swap, factors and row updates in parallel over the
// workgroup), the status record, and the slot of the next step's vector:
// a free one, or the oldest one's when full, whose row and column of B are
// dropped (optc_update_redsp1 'shift').
__device__ __forceinline__ void finish_body(const double *__restrict__ part, int ngroup,
                                            const double *__restrict__ energy,
                                            DevState *__restrict__ st, int maxsub, double tol,
                                            StatusRec *ring, int32_t *__restrict__ flag) {
    // a flag set before this step (the instance converged and its residual
    // is no longer re-evaluated): repeat the converged record, leave the
    // subspace and t as they are
    if (flag && *flag) {
        if (threadIdx.x == 0) {
            StatusRec &rec = ring[st->steps % RING];
            rec.gnorm = st->gn_last;
            rec.energy = energy ? energy[0] : 0.0;
            rec.converged = 1;
            __threadfence_system();
            rec.step = st->steps;
            st->steps += 1;
            st->skip = 1;
        }
        return;
    }
    constexpr int KM = MAXSUB + 1;
    __shared__ double red[NTH / 64][MAXSUB + 1];
    __shared__ double tot[MAXSUB + 1];
    __shared__ double a[KM * KM], rhs[KM], fac[KM];
    __shared__ int piv, code_s;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // partial sums of this thread, all loads in flight together
    constexpr int PER = (NGROUP + NTH - 1) / NTH;
    double v[MAXSUB + 1];
#pragma unroll
    for (int j = 0; j <= MAXSUB; ++j) v[j] = 0.0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int g = tid + q * NTH;
        if (g < ngroup) {
#pragma unroll
            for (int j = 0; j <= MAXSUB; ++j) v[j] += part[(int64_t)g * (MAXSUB + 1) + j];
        }
    }
#pragma unroll
    for (int j = 0; j <= MAXSUB; ++j) {
        double x = v[j];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
        if (lane == 0) red[wave][j] = x;
    }
    __syncthreads();
    const int m = st->m;
    const int k = m + 1;
    if (tid <= MAXSUB) {
        double x = 0.0;
        for (int w = 0; w < NTH / 64; ++w) x += red[w][tid];
        tot[tid] = x;
    }
    __syncthreads();
    if (tid < m) {
        st->b[(m - 1) * MAXSUB + tid] = tot[tid];
        st->b[tid * MAXSUB + m - 1] = tot[tid];
    }
    const double gn = sqrt(tot[MAXSUB]);
    const bool conv = gn < tol;
    if (tid == 0) code_s = conv ? 1 : 0;
    // augmented matrix (row-major, k x k) and right-hand side
    for (int e = tid; e < k * k; e += NTH) {
        const int i = e / k, j = e - i * k;
        a[e] = (i < m && j < m) ? (i == m - 1 ? tot[j] : j == m - 1 ? tot[i] : st->b[i * MAXSUB + j])
                                : (i == j ? 0.0 : -1.0);
        if (i < m && j < m) st->bw[i * MAXSUB + j] = a[e];
    }
    if (tid < k) rhs[tid] = tid == m ? -1.0 : 0.0;
    __syncthreads();
    if (!conv) {
        for (int c = 0; c < k; ++c) {
            if (tid == 0) {
                int p = c;
                for (int r = c + 1; r < k; ++r)
                    if (fabs(a[r * k + c]) > fabs(a[p * k + c])) p = r;
                piv = p;
                if (a[p * k + c] == 0.0) code_s = -1;
            }
            __syncthreads();
            if (code_s != 0) break;
            const int p = piv;
            if (p != c && tid <= k) {
                if (tid < k) {
                    const double x = a[p * k + tid];
                    a[p * k + tid] = a[c * k + tid];
                    a[c * k + tid] = x;
                } else {
                    const double x = rhs[p];
                    rhs[p] = rhs[c];
                    rhs[c] = x;
                }
            }
            __syncthreads();
            if (tid > c && tid < k) fac[tid] = a[tid * k + c] / a[c * k + c];
            __syncthreads();
            // rows r > c, columns j >= c, and the right-hand side (j = k)
            const int nc = k - c + 1;
            for (int e = tid; e < (k - c - 1) * nc; e += NTH) {
                const int r = c + 1 + e / nc, j = c + e % nc;
                const double f = fac[r];
                if (f == 0.0) continue;
                if (j < k) a[r * k + j] -= f * a[c * k + j];
                else rhs[r] -= f * rhs[c];
            }
            __syncthreads();
        }
    }
    const int64_t code = code_s;
    // the next step's vector: a free slot, or the oldest one's when the
    // subspace is full, whose row and column of B are dropped -- read here
    // by the whole workgroup, written after the barrier below (a serial
    // copy of B by one thread waited on each load in turn)
    const bool full = m == maxsub;
    const int mm = full ? m - 1 : m;
    const int slot0 = st->order[0];
    int o_next = 0, bi = 0, bj = 0;
    double b_next = 0.0;
    if (full) {
        if (tid + 1 < m) o_next = st->order[tid + 1];
        if (tid < mm * mm) {      // mm <= MAXSUB - 1, so mm * mm < NTH
            bi = tid / mm;
            bj = tid - bi * mm;
            b_next = st->b[(bi + 1) * MAXSUB + bj + 1];
        }
    }
    if (tid == 0) {
        if (code == 0) {
            for (int r = k - 1; r >= 0; --r) {
                double x = rhs[r];
                for (int j = r + 1; j < k; ++j) x -= a[r * k + j] * rhs[j];
                rhs[r] = x / a[r * k + r];
            }
        }
        for (int j = 0; j < m; ++j) {
            st->w[j] = code == 0 ? rhs[j] : 0.0;
            st->wslot[j] = st->order[j];
        }
        st->mw = m;
        st->skip = code != 0 ? 1 : 0;
        StatusRec &rec = ring[st->steps % RING];
        rec.gnorm = gn;
        st->gn_last = gn;
        rec.energy = energy ? energy[0] : 0.0;
        rec.converged = code;
        if (flag) *flag = code == 1 ? 1 : 0;
        __threadfence_system();
        rec.step = st->steps;
        st->steps += 1;
    }
    __syncthreads();   // every read of order and B above is done
    if (full) {
        if (tid + 1 < m) st->order[tid] = o_next;
        if (tid < mm * mm) st->b[bi * MAXSUB + bj] = b_next;
    }
    if (tid == 0) {
        st->order[mm] = full ? slot0 : mm;
        st->m = mm + 1;
    }
}

__global__ __launch_bounds__(NTH) void diis_finish_f64(const double *__restrict__ part, int ngroup,
                                                       const double *__restrict__ energy,
                                                       DevState *__restrict__ st, int maxsub,
                                                       double tol, StatusRec *ring,
                                                       int32_t *__restrict__ flag) {
    finish_body(part, ngroup, energy, st, maxsub, tol, ring, flag);
}

// Grouped asynchronous steps of many subspaces (independent problems):
// blockIdx.y (prepare, expand) or blockIdx.x (finish) selects the member.
struct GroupMember {
    double *t;
    const double *rest, *diag, *scale, *energy;
    double *rsub, *vsub, *part;
    DevState *st;
    StatusRec *ring;
    int32_t *flag;
    int64_t n;
    int maxsub, pad;
};

__global__ __launch_bounds__(NTH) void diis_prepare_group_f64(const GroupMember *__restrict__ mem) {
    const GroupMember &g = mem[blockIdx.y];
    if (g.flag && *g.flag) return;
    Slots sl;
    sl.m = g.st->m;
    for (int j = 0; j < MAXSUB; ++j) sl.s[j] = j < sl.m ? g.st->order[j] : 0;
    prepare_body(g.t, g.rest, g.diag, g.scale, g.rsub, g.vsub, g.n, sl, g.part);
}

__global__ __launch_bounds__(NTH) void diis_finish_group_f64(const GroupMember *__restrict__ mem,
                                                             int ngroup, double tol) {
    const GroupMember &g = mem[blockIdx.x];
    finish_body(g.part, ngroup, g.energy, g.st, g.maxsub, tol, g.ring, g.flag);
}

__global__ __launch_bounds__(NTH) void diis_expand_group_f64(const GroupMember *__restrict__ mem) {
    const GroupMember &g = mem[blockIdx.y];
    if (g.st->skip) return;
    Weights wt;
    wt.m = g.st->mw;
    for (int j = 0; j < MAXSUB; ++j) {
        wt.s[j] = j < wt.m ? g.st->wslot[j] : 0;
        wt.w[j] = j < wt.m ? g.st->w[j] : 0.0;
    }
    expand_body(g.t, g.vsub, g.n, wt);
}

__global__ __launch_bounds__(NTH) void diis_expand_async_f64(double *__restrict__ t,
                                                             const double *__restrict__ vsub,
                                                             int64_t n,
                                                             const DevState *__restrict__ st) {
    if (st->skip) return;
    Weights wt;
    wt.m = st->mw;
    for (int j = 0; j < MAXSUB; ++j) {
        wt.s[j] = j < wt.m ? st->wslot[j] : 0;
        wt.w[j] = j < wt.m ? st->w[j] : 0.0;
    }
    expand_body(t, vsub, n, wt);
}

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    return gecco::set_error(code, buf);
}

int check(hipError_t e, const char *what) {
    if (e != hipSuccess) return fail(GECCO_HIP_EDEVICE, "%s: %s", what, hipGetErrorString(e));
    return 0;
}

// Solve a x = b (dense, row-major, size k) by Gaussian elimination with
// partial pivoting; returns false if a pivot vanishes.  The reference
// factorises the packed symmetric indefinite matrix with LINPACK dspco/dspsl
// (math/linpack.f); both give the unique solution of the augmented system.
bool solve_dense(std::vector<double> a, std::vector<double> &b, int k) {
    for (int c = 0; c < k; ++c) {
        int p = c;
        for (int r = c + 1; r < k; ++r)
            if (std::fabs(a[r * k + c]) > std::fabs(a[p * k + c])) p = r;
        if (a[p * k + c] == 0.0) return false;
        if (p != c) {
            for (int j = 0; j < k; ++j) std::swap(a[p * k + j], a[c * k + j]);
            std::swap(b[p], b[c]);
        }
        for (int r = c + 1; r < k; ++r) {
            const double f = a[r * k + c] / a[c * k + c];
            if (f == 0.0) continue;
            for (int j = c; j < k; ++j) a[r * k + j] -= f * a[c * k + j];
            b[r] -= f * b[c];
        }
    }
    for (int r = k - 1; r >= 0; --r) {
        double s = b[r];
        for (int j = r + 1; j < k; ++j) s -= a[r * k + j] * b[j];
        b[r] = s / a[r * k + r];
    }
    return true;
}

}  // namespace

struct gecco_hip_diis {
    int64_t n = 0;
    int maxsub = 0;
    double *rsub = nullptr, *vsub = nullptr;   // maxsub slots of n doubles each
    double *part = nullptr;                    // NGROUP x (MAXSUB+1) partial sums
    double *red = nullptr;                     // MAXSUB+2 reduced values (device)
    double *host = nullptr;                    // pinned copy of red
    std::vector<int> order;                    // slot of logical vector j, oldest first
    std::vector<double> bmat;                  // <r_i|r_j>, logical order, maxsub x maxsub
    std::vector<double> w;                     // weights of the last step
    DevState *dstate = nullptr;                // bookkeeping of the asynchronous step
    StatusRec *ring = nullptr;                 // pinned status records (device-written)
    int32_t *flag = nullptr;                   // optional convergence flag (device)
    const double *scale = nullptr;             // optional per-amplitude step factor (device)
    bool async = false;                        // steps so far were asynchronous
};

extern "C" {

int gecco_hip_diis_create(int64_t n, int64_t maxsub, gecco_hip_diis **out) {
    gecco::clear_error();
    if (!out || n <= 0 || maxsub < 1 || maxsub > MAXSUB)
        return fail(GECCO_HIP_EINVAL, "diis_create: need n > 0 and 1 <= maxsub <= %d", MAXSUB);
    auto *d = new gecco_hip_diis();
    d->n = n;
    d->maxsub = (int)maxsub;
    d->bmat.assign((size_t)maxsub * maxsub, 0.0);
    if (check(hipMalloc(&d->rsub, (size_t)maxsub * n * sizeof(double)), "hipMalloc(diis r)") ||
        check(hipMalloc(&d->vsub, (size_t)maxsub * n * sizeof(double)), "hipMalloc(diis v)") ||
        check(hipMalloc(&d->part, (size_t)NGROUP * (MAXSUB + 1) * sizeof(double)), "hipMalloc(diis partials)") ||
        check(hipMalloc(&d->red, (MAXSUB + 2) * sizeof(double)), "hipMalloc(diis reduction)") ||
        check(hipHostMalloc(&d->host, (MAXSUB + 2) * sizeof(double)), "hipHostMalloc(diis)") ||
        check(hipMalloc(&d->dstate, sizeof(DevState)), "hipMalloc(diis state)") ||
        check(hipHostMalloc(&d->ring, RING * sizeof(StatusRec), hipHostMallocMapped | hipHostMallocCoherent),
              "hipHostMalloc(diis status)")) {
        gecco_hip_diis_destroy(d);
        return GECCO_HIP_ENOMEM;
    }
    if (gecco_hip_diis_reset(d)) {
        gecco_hip_diis_destroy(d);
        return GECCO_HIP_EDEVICE;
    }
    *out = d;
    return 0;
}

int gecco_hip_diis_destroy(gecco_hip_diis *d) {
    if (!d) return 0;
    if (d->rsub) (void)hipFree(d->rsub);
    if (d->vsub) (void)hipFree(d->vsub);
    if (d->part) (void)hipFree(d->part);
    if (d->red) (void)hipFree(d->red);
    if (d->host) (void)hipHostFree(d->host);
    if (d->dstate) (void)hipFree(d->dstate);
    if (d->ring) (void)hipHostFree(d->ring);
    delete d;
    return 0;
}

int gecco_hip_diis_reset(gecco_hip_diis *d) {
    gecco::clear_error();
    if (!d) return fail(GECCO_HIP_EINVAL, "null diis");
    d->order.clear();
    std::fill(d->bmat.begin(), d->bmat.end(), 0.0);
    d->w.clear();
    d->async = false;
    // device state from a zeroed host copy (synchronous, no memset)
    DevState zero;
    std::memset(&zero, 0, sizeof zero);
    zero.m = 1;              // the first step adds the vector in slot 0
    if (check(hipMemcpy(d->dstate, &zero, sizeof zero, hipMemcpyHostToDevice), "hipMemcpy(diis state)"))
        return GECCO_HIP_EDEVICE;
    for (int r = 0; r < RING; ++r) {
        d->ring[r].step = -1;
        d->ring[r].gnorm = d->ring[r].energy = 0.0;
        d->ring[r].converged = 0;
    }
    return 0;
}

int gecco_hip_diis_step(gecco_hip_diis *d, double *t, const double *rest, const double *diag,
                        const double *energy_dev, double tol, double *gnorm, double *energy,
                        int64_t *converged, void *stream) {
    gecco::clear_error();
    if (!d || !t || !rest || !diag) return fail(GECCO_HIP_EINVAL, "diis_step: null argument");
    if (d->async) return fail(GECCO_HIP_ESTATE, "diis_step: subspace holds asynchronous steps (reset first)");
    hipStream_t s = (hipStream_t)stream;
    const int ms = d->maxsub;
    // slot of the new vector: a free one, or the oldest one's when full
    const bool full = (int)d->order.size() == ms;
    int slot;
    if (full) {
        slot = d->order.front();
        d->order.erase(d->order.begin());
        // drop the oldest row/column of B (optc_update_redsp1 'shift')
        for (int i = 0; i < ms - 1; ++i)
            for (int j = 0; j < ms - 1; ++j) d->bmat[i * ms + j] = d->bmat[(i + 1) * ms + j + 1];
    } else {
        slot = (int)d->order.size();
    }
    d->order.push_back(slot);
    const int m = (int)d->order.size();
    Slots sl;
    std::memset(&sl, 0, sizeof sl);
    sl.m = m;
    for (int j = 0; j < m; ++j) sl.s[j] = d->order[j];
    const int ng = (int)std::min<int64_t>(NGROUP, std::max<int64_t>(1, (d->n + NTH - 1) / NTH));
    hipLaunchKernelGGL(diis_prepare_f64, dim3(ng), dim3(NTH), 0, s, t, rest, diag, d->scale, d->rsub,
                       d->vsub, d->n, sl, d->part);
    hipLaunchKernelGGL(diis_reduce_f64, dim3(1), dim3(NTH), 0, s, d->part, ng, energy_dev, d->red);
    if (check(hipGetLastError(), "diis kernels") ||
        check(hipMemcpyAsync(d->host, d->red, (MAXSUB + 2) * sizeof(double), hipMemcpyDeviceToHost, s),
              "hipMemcpyAsync(diis)") ||
        check(hipStreamSynchronize(s), "hipStreamSynchronize(diis)"))
        return GECCO_HIP_EDEVICE;
    const double gn = std::sqrt(d->host[MAXSUB]);
    if (gnorm) *gnorm = gn;
    if (energy) *energy = d->host[MAXSUB + 1];
    for (int j = 0; j < m; ++j) {
        d->bmat[(m - 1) * ms + j] = d->host[j];
        d->bmat[j * ms + m - 1] = d->host[j];
    }
    if (converged) *converged = gn < tol ? 1 : 0;
    if (gn < tol) {
        d->w.assign(m, 0.0);
        return 0;
    }
    // augmented system [B -1; -1 0] (w, l) = (0, -1)
    const int k = m + 1;
    std::vector<double> a((size_t)k * k, 0.0), b(k, 0.0);
    for (int i = 0; i < m; ++i) {
        for (int j = 0; j < m; ++j) a[i * k + j] = d->bmat[i * ms + j];
        a[i * k + m] = -1.0;
        a[m * k + i] = -1.0;
    }
    b[m] = -1.0;
    if (!solve_dense(a, b, k))
        return fail(GECCO_HIP_EINVAL, "DIIS matrix is singular (subspace of %d vectors)", m);
    Weights wt;
    std::memset(&wt, 0, sizeof wt);
    wt.m = m;
    d->w.assign(b.begin(), b.begin() + m);
    for (int j = 0; j < m; ++j) {
        wt.s[j] = d->order[j];
        wt.w[j] = b[j];
    }
    hipLaunchKernelGGL(diis_expand_f64, dim3(ng), dim3(NTH), 0, s, t, d->vsub, d->n, wt);
    return check(hipGetLastError(), "diis expand");
}

int gecco_hip_diis_step_async(gecco_hip_diis *d, double *t, const double *rest, const double *diag,
                              const double *energy_dev, double tol, void *stream) {
    gecco::clear_error();
    if (!d || !t || !rest || !diag) return fail(GECCO_HIP_EINVAL, "diis_step_async: null argument");
    if (!d->order.empty()) return fail(GECCO_HIP_ESTATE, "diis_step_async: subspace holds synchronous steps (reset first)");
    hipStream_t s = (hipStream_t)stream;
    d->async = true;
    const int ng = (int)std::min<int64_t>(NGROUP, std::max<int64_t>(1, (d->n + NTH - 1) / NTH));
    hipLaunchKernelGGL(diis_prepare_async_f64, dim3(ng), dim3(NTH), 0, s, t, rest, diag, d->scale,
                       d->rsub, d->vsub, d->n, d->dstate, d->part, d->flag);
    hipLaunchKernelGGL(diis_finish_f64, dim3(1), dim3(NTH), 0, s, d->part, ng, energy_dev, d->dstate,
                       d->maxsub, tol, d->ring, d->flag);
    hipLaunchKernelGGL(diis_expand_async_f64, dim3(ng), dim3(NTH), 0, s, t, d->vsub, d->n, d->dstate);
    return check(hipGetLastError(), "diis async kernels");
}

struct gecco_hip_diis_group {
    std::vector<gecco_hip_diis *> members;
    GroupMember *d_mem = nullptr;
    int ng = 1;
};

int gecco_hip_diis_group_create(gecco_hip_diis *const *d, int64_t count, double *const *t,
                                const double *const *rest, const double *const *diag,
                                const double *const *energy, gecco_hip_diis_group **out) {
    gecco::clear_error();
    if (!d || !t || !rest || !diag || !out || count <= 0 || count > 65535)
        return fail(GECCO_HIP_EINVAL, "diis_group_create: bad arguments");
    std::vector<GroupMember> mem((size_t)count);
    int64_t nmax = 1;
    for (int64_t k = 0; k < count; ++k) {
        gecco_hip_diis *x = d[k];
        if (!x || !t[k] || !rest[k] || !diag[k])
            return fail(GECCO_HIP_EINVAL, "diis_group_create: member %lld incomplete", (long long)k);
        if (!x->order.empty())
            return fail(GECCO_HIP_ESTATE, "diis_group_create: member %lld holds synchronous steps", (long long)k);
        GroupMember &g = mem[(size_t)k];
        std::memset(&g, 0, sizeof g);
        g.t = t[k];
        g.rest = rest[k];
        g.diag = diag[k];
        g.scale = x->scale;
        g.energy = energy ? energy[k] : nullptr;
        g.rsub = x->rsub;
        g.vsub = x->vsub;
        g.part = x->part;
        g.st = x->dstate;
        g.ring = x->ring;
        g.flag = x->flag;
        g.n = x->n;
        g.maxsub = x->maxsub;
        nmax = std::max<int64_t>(nmax, x->n);
    }
    auto *grp = new gecco_hip_diis_group();
    grp->members.assign(d, d + count);
    grp->ng = (int)std::min<int64_t>(NGROUP, std::max<int64_t>(1, (nmax + NTH - 1) / NTH));
    if (check(hipMalloc(&grp->d_mem, mem.size() * sizeof(GroupMember)), "hipMalloc(diis group)") ||
        check(hipMemcpy(grp->d_mem, mem.data(), mem.size() * sizeof(GroupMember), hipMemcpyHostToDevice),
              "hipMemcpy(diis group)")) {
        if (grp->d_mem) (void)hipFree(grp->d_mem);
        delete grp;
        return GECCO_HIP_EDEVICE;
    }
    *out = grp;
    return 0;
}

int gecco_hip_diis_group_destroy(gecco_hip_diis_group *g) {
    if (!g) return 0;
    if (g->d_mem) (void)hipFree(g->d_mem);
    delete g;
    return 0;
}

int gecco_hip_diis_group_step_async(gecco_hip_diis_group *g, double tol, void *stream) {
    gecco::clear_error();
    if (!g) return fail(GECCO_HIP_EINVAL, "null diis group");
    for (auto *x : g->members) {
        if (!x->order.empty()) return fail(GECCO_HIP_ESTATE, "diis group member holds synchronous steps");
        x->async = true;
    }
    hipStream_t s = (hipStream_t)stream;
    const unsigned cnt = (unsigned)g->members.size();
    hipLaunchKernelGGL(diis_prepare_group_f64, dim3(g->ng, cnt), dim3(NTH), 0, s, g->d_mem);
    hipLaunchKernelGGL(diis_finish_group_f64, dim3(cnt), dim3(NTH), 0, s, g->d_mem, g->ng, tol);
    hipLaunchKernelGGL(diis_expand_group_f64, dim3(g->ng, cnt), dim3(NTH), 0, s, g->d_mem);
    return check(hipGetLastError(), "diis group kernels");
}

int gecco_hip_diis_set_scale(gecco_hip_diis *d, const double *scale) {
    gecco::clear_error();
    if (!d) return fail(GECCO_HIP_EINVAL, "null diis");
    if (scale) {
        hipPointerAttribute_t at;
        if (hipPointerGetAttributes(&at, scale) != hipSuccess || at.type != hipMemoryTypeDevice) {
            (void)hipGetLastError();
            return fail(GECCO_HIP_EINVAL, "diis_set_scale: scale must be device memory");
        }
    }
    d->scale = scale;
    return 0;
}

int gecco_hip_diis_set_flag(gecco_hip_diis *d, int32_t *flag) {
    gecco::clear_error();
    if (!d) return fail(GECCO_HIP_EINVAL, "null diis");
    if (flag) {
        hipPointerAttribute_t at;
        if (hipPointerGetAttributes(&at, flag) != hipSuccess || at.type != hipMemoryTypeDevice) {
            (void)hipGetLastError();
            return fail(GECCO_HIP_EINVAL, "diis_set_flag: flag must be device memory");
        }
    }
    d->flag = flag;
    return 0;
}

int gecco_hip_diis_status(const gecco_hip_diis *d, int64_t step, double *gnorm, double *energy,
                          int64_t *converged) {
    gecco::clear_error();
    if (!d || step < 0) return fail(GECCO_HIP_EINVAL, "diis_status: bad arguments");
    const volatile StatusRec &rec = d->ring[step % RING];
    if (rec.step != step)
        return fail(GECCO_HIP_ESTATE, "diis_status: step %lld not completed or no longer kept",
                    (long long)step);
    if (gnorm) *gnorm = rec.gnorm;
    if (energy) *energy = rec.energy;
    if (converged) *converged = rec.converged;
    return 0;
}

int gecco_hip_diis_weights(const gecco_hip_diis *d, double *w, int64_t nw, int64_t *m) {
    if (!d) return fail(GECCO_HIP_EINVAL, "null diis");
    if (d->async) {
        DevState st;
        if (check(hipMemcpy(&st, d->dstate, sizeof st, hipMemcpyDeviceToHost), "hipMemcpy(diis state)"))
            return GECCO_HIP_EDEVICE;
        if (m) *m = st.mw;
        if (w)
            for (int64_t j = 0; j < std::min<int64_t>(nw, st.mw); ++j) w[j] = st.w[j];
        return 0;
    }
    if (m) *m = (int64_t)d->w.size();
    if (w)
        for (int64_t j = 0; j < std::min<int64_t>(nw, (int64_t)d->w.size()); ++j) w[j] = d->w[j];
    return 0;
}

int gecco_hip_diis_matrix(const gecco_hip_diis *d, double *b, int64_t nb, int64_t *m) {
    if (!d) return fail(GECCO_HIP_EINVAL, "null diis");
    if (d->async) {
        DevState st;
        if (check(hipMemcpy(&st, d->dstate, sizeof st, hipMemcpyDeviceToHost), "hipMemcpy(diis state)"))
            return GECCO_HIP_EDEVICE;
        // the matrix of the last step's extrapolation (the state already
        // describes the next step)
        const int held = st.mw;
        if (m) *m = held;
        if (b && nb >= (int64_t)held * held)
            for (int i = 0; i < held; ++i)
                for (int j = 0; j < held; ++j) b[i * held + j] = st.bw[i * MAXSUB + j];
        return 0;
    }
    const int64_t k = (int64_t)d->order.size();
    if (m) *m = k;
    if (b && nb >= k * k)
        for (int64_t i = 0; i < k; ++i)
            for (int64_t j = 0; j < k; ++j) b[i * k + j] = d->bmat[i * d->maxsub + j];
    return 0;
}

}  // extern "C"
