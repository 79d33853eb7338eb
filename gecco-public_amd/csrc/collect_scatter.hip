// The route-3 collect and scatter steps as separate C-ABI calls with the
// reference's Fortran argument lists (contract/collect_block_impr.f:2-17,
// contract/scatter_block.f:2-20): the host works out, from the batch bounds,
// string lengths and string maps, which list element (with which sign) each
// element of the batch block comes from or goes to; the GPU then moves the
// data (a signed gather into the block, or a signed atomic scatter-add into
// the list).  Host arrays are staged through device memory; device pointers
// are used in place.  The strided-GEMM plans never call these: they resolve
// the same maps once per plan and gather or scatter inside their kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <vector>

#include "../../include/gecco_hip.h"

namespace gecco {
int set_error(int code, const char *msg);
void clear_error();
}

namespace {

typedef int64_t I;

// Component indices (1-based) of compound string index i (0-based) whose
// components, fastest first, have lengths lstr[0..n) and sit at the
// positions ireo[0..n) (1-based) of istr; positions not named keep 1.
void decompose(I i, const I *lstr, const I *ireo, I n, I *istr, I nall) {
    for (I c = 0; c < nall; ++c) istr[c] = 1;
    for (I c = 0; c < n; ++c) {
        istr[ireo[c] - 1] = i % lstr[c] + 1;
        i /= lstr[c];
    }
}

inline I sgn(I x) { return x < 0 ? -1 : 1; }
inline I iabs(I x) { return x < 0 ? -x : x; }

struct Shape {
    I kc_st0, kc_nd0, kc_len, ka_st0, ka_nd0, ka_len, k_len;
    I lc_st0, lc_nd0, lc_len, la_st0, la_nd0, la_len, l_len;
    const I *map_kl_c, *map_kl_a;
    I ncblk_op, nablk_op;
    const I *ldim_opc, *ldim_opa;
    I ncblk_k, nablk_k, ncblk_l, nablk_l;
    const I *lstr_kc, *lstr_ka, *lstr_lc, *lstr_la;
    const I *nstr_kc, *nstr_ka, *nstr_lc, *nstr_la;
    const I *ireo_kc, *ireo_ka, *ireo_lc, *ireo_la;
};

constexpr I MAXC = 32;   // components per side (GeCCo: a few)

// collect_block_impr: src[kldx] = +-(1 + list offset) of the element that
// block element kldx (0-based, k fastest) takes, 0 where it stays zero.
// The loops follow contract/collect_block_impr.f:98-409: KA outer (with
// kca_trp all of 1..ka_len, the K batch then running over KA within each
// KC), LA, LC, KC innermost; the KA x LA map gives the A part of the
// address and sign, the LC x KC map the C part; a zero map entry leaves the
// element (or the whole LA string) at zero.
int collect_index(const Shape &s, bool trp, std::vector<I> &src) {
    if (s.ncblk_op > MAXC || s.nablk_op > MAXC) return 1;
    src.assign((size_t)std::max<I>(0, s.k_len * s.l_len), 0);
    I ka_c[MAXC], la_c[MAXC], lc_c[MAXC], kc_c[MAXC];
    const I ka_st = trp ? 1 : s.ka_st0, ka_nd = trp ? s.ka_len : s.ka_nd0;
    I kdx0 = trp ? (s.ka_len > 0 ? (s.ka_len - s.ka_st0 + 1) % s.ka_len : 0) : 0;
    for (I ka = ka_st; ka <= ka_nd; ++ka) {
        I kc_st, kc_nd;
        if (!trp) {
            kc_st = ka == s.ka_st0 ? s.kc_st0 : 1;
            kc_nd = ka == s.ka_nd0 ? s.kc_nd0 : s.kc_len;
        } else {
            kc_st = s.kc_st0 + (ka < s.ka_st0 ? 1 : 0);
            kc_nd = s.kc_nd0 - (ka > s.ka_nd0 ? 1 : 0);
        }
        if (trp && ka == s.ka_st0) kdx0 = 0;
        const I kdx_st = kdx0;
        kdx0 += trp ? 1 : kc_nd - kc_st + 1;
        decompose(ka - 1, s.lstr_ka, s.ireo_ka, s.nablk_k, ka_c, s.nablk_op);
        I ldx0 = 0;
        for (I la = s.la_st0; la <= s.la_nd0; ++la) {
            const I lc_st = la == s.la_st0 ? s.lc_st0 : 1;
            const I lc_nd = la == s.la_nd0 ? s.lc_nd0 : s.lc_len;
            I ldx = ldx0;
            ldx0 += lc_nd - lc_st + 1;
            decompose(la - 1, s.lstr_la, s.ireo_la, s.nablk_l, la_c, s.nablk_op);
            I ioff = 0, sa = 1, a0 = 1;
            bool ok = true;
            for (I c = 0; c < s.nablk_op && ok; ++c) {
                const I e = s.map_kl_a[ioff + (la_c[c] - 1) * s.nstr_ka[c] + ka_c[c] - 1];
                if (e == 0) ok = false;
                sa *= sgn(e);
                a0 += (iabs(e) - 1) * s.ldim_opa[c];
                ioff += s.nstr_ka[c] * s.nstr_la[c];
            }
            if (!ok) continue;
            for (I lc = lc_st; lc <= lc_nd; ++lc) {
                ++ldx;
                decompose(lc - 1, s.lstr_lc, s.ireo_lc, s.ncblk_l, lc_c, s.ncblk_op);
                I kldx = (ldx - 1) * s.k_len + kdx_st;
                I inc = 1;
                if (trp) {
                    kldx = (ldx - 1) * s.k_len + kdx_st + 1 - s.ka_len;
                    inc = s.ka_len;
                }
                for (I kc = kc_st; kc <= kc_nd; ++kc) {
                    kldx += inc;
                    I sg = sa, at = a0;
                    if (s.ncblk_op == 1) {
                        // the reference's one-component loop indexes the map by kc itself
                        const I e = s.map_kl_c[(lc_c[0] - 1) * s.nstr_kc[0] + kc - 1];
                        if (e == 0) continue;
                        sg *= sgn(e);
                        at += (iabs(e) - 1) * s.ldim_opc[0];
                    } else if (s.ncblk_op > 1) {
                        decompose(kc - 1, s.lstr_kc, s.ireo_kc, s.ncblk_k, kc_c, s.ncblk_op);
                        I io = 0;
                        bool okc = true;
                        for (I c = 0; c < s.ncblk_op && okc; ++c) {
                            const I e = s.map_kl_c[io + (lc_c[c] - 1) * s.nstr_kc[c] + kc_c[c] - 1];
                            if (e == 0) okc = false;
                            sg *= sgn(e);
                            at += (iabs(e) - 1) * s.ldim_opc[c];
                            io += s.nstr_kc[c] * s.nstr_lc[c];
                        }
                        if (!okc) continue;
                    }
                    if (kldx < 1 || kldx > (I)src.size()) return 2;
                    src[(size_t)(kldx - 1)] = sg * at;
                }
            }
        }
    }
    return 0;
}

// scatter_block: pairs (kldx, +-(1 + list offset)): list element += sign *
// block element kldx (contract/scatter_block.f:56-170; here the KA x LA map
// is indexed KA-slow, and K runs over the batch in order).
int scatter_index(const Shape &s, std::vector<I> &blk, std::vector<I> &dst) {
    if (s.ncblk_op > MAXC || s.nablk_op > MAXC) return 1;
    blk.clear();
    dst.clear();
    I ka_c[MAXC], la_c[MAXC], lc_c[MAXC], kc_c[MAXC];
    I kdx0 = 0;
    for (I ka = s.ka_st0; ka <= s.ka_nd0; ++ka) {
        const I kc_st = ka == s.ka_st0 ? s.kc_st0 : 1;
        const I kc_nd = ka == s.ka_nd0 ? s.kc_nd0 : s.kc_len;
        const I kdx_st = kdx0;
        kdx0 += kc_nd - kc_st + 1;
        decompose(ka - 1, s.lstr_ka, s.ireo_ka, s.nablk_k, ka_c, s.nablk_op);
        I ldx0 = 0;
        for (I la = s.la_st0; la <= s.la_nd0; ++la) {
            const I lc_st = la == s.la_st0 ? s.lc_st0 : 1;
            const I lc_nd = la == s.la_nd0 ? s.lc_nd0 : s.lc_len;
            I ldx = ldx0;
            ldx0 += lc_nd - lc_st + 1;
            decompose(la - 1, s.lstr_la, s.ireo_la, s.nablk_l, la_c, s.nablk_op);
            I ioff = 0, sa = 1, a0 = 1;
            bool ok = true;
            for (I c = 0; c < s.nablk_op && ok; ++c) {
                const I e = s.map_kl_a[ioff + (ka_c[c] - 1) * s.nstr_la[c] + la_c[c] - 1];
                if (e == 0) ok = false;
                sa *= sgn(e);
                a0 += (iabs(e) - 1) * s.ldim_opa[c];
                ioff += s.nstr_ka[c] * s.nstr_la[c];
            }
            if (!ok) continue;
            for (I lc = lc_st; lc <= lc_nd; ++lc) {
                ++ldx;
                decompose(lc - 1, s.lstr_lc, s.ireo_lc, s.ncblk_l, lc_c, s.ncblk_op);
                I kldx = (ldx - 1) * s.k_len + kdx_st;
                for (I kc = kc_st; kc <= kc_nd; ++kc) {
                    ++kldx;
                    decompose(kc - 1, s.lstr_kc, s.ireo_kc, s.ncblk_k, kc_c, s.ncblk_op);
                    I io = 0, sg = sa, at = a0;
                    bool okc = true;
                    for (I c = 0; c < s.ncblk_op && okc; ++c) {
                        const I e = s.map_kl_c[io + (lc_c[c] - 1) * s.nstr_kc[c] + kc_c[c] - 1];
                        if (e == 0) okc = false;
                        sg *= sgn(e);
                        at += (iabs(e) - 1) * s.ldim_opc[c];
                        io += s.nstr_kc[c] * s.nstr_lc[c];
                    }
                    if (!okc) continue;
                    blk.push_back(kldx);
                    dst.push_back(sg * at);
                }
            }
        }
    }
    return 0;
}

__global__ __launch_bounds__(256) void gather_signed(double *__restrict__ blk, const double *__restrict__ op,
                                                     const int64_t *__restrict__ src, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = src[i];
        blk[i] = s == 0 ? 0.0 : (s < 0 ? -op[-s - 1] : op[s - 1]);
    }
}

__global__ __launch_bounds__(256) void scatter_signed(double *__restrict__ op, const double *__restrict__ blk,
                                                      const int64_t *__restrict__ from,
                                                      const int64_t *__restrict__ to, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t d = to[i];
        const double v = blk[from[i] - 1];
        atomicAdd(op + ((d < 0 ? -d : d) - 1), d < 0 ? -v : v);
    }
}

bool on_device(const void *p) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice;
}

// growing device buffers for staged host arrays and index lists
std::mutex g_mu;
void *g_buf[4] = {nullptr, nullptr, nullptr, nullptr};
size_t g_cap[4] = {0, 0, 0, 0};

bool stage(int i, size_t bytes, void **out) {
    bytes = std::max<size_t>(bytes, 8);
    if (g_cap[i] < bytes) {
        if (g_buf[i]) (void)hipFree(g_buf[i]);
        g_buf[i] = nullptr;
        g_cap[i] = 0;
        if (hipMalloc(&g_buf[i], bytes) != hipSuccess) return false;
        g_cap[i] = bytes;
    }
    *out = g_buf[i];
    return true;
}

int grid_of(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)); }

Shape shape_of(const I *kc_st0, const I *kc_nd0, const I *kc_len, const I *ka_st0, const I *ka_nd0,
               const I *ka_len, const I *k_len, const I *lc_st0, const I *lc_nd0, const I *lc_len,
               const I *la_st0, const I *la_nd0, const I *la_len, const I *l_len, const I *map_kl_c,
               const I *map_kl_a, const I *ncblk_op, const I *nablk_op, const I *ldim_opc,
               const I *ldim_opa, const I *ncblk_k, const I *nablk_k, const I *ncblk_l,
               const I *nablk_l, const I *lstr_kc, const I *lstr_ka, const I *lstr_lc,
               const I *lstr_la, const I *nstr_kc, const I *nstr_ka, const I *nstr_lc,
               const I *nstr_la, const I *ireo_kc, const I *ireo_ka, const I *ireo_lc,
               const I *ireo_la) {
    return Shape{*kc_st0, *kc_nd0, *kc_len, *ka_st0, *ka_nd0, *ka_len, *k_len,
                 *lc_st0, *lc_nd0, *lc_len, *la_st0, *la_nd0, *la_len, *l_len,
                 map_kl_c, map_kl_a, *ncblk_op, *nablk_op, ldim_opc, ldim_opa,
                 *ncblk_k, *nablk_k, *ncblk_l, *nablk_l, lstr_kc, lstr_ka, lstr_lc, lstr_la,
                 nstr_kc, nstr_ka, nstr_lc, nstr_la, ireo_kc, ireo_ka, ireo_lc, ireo_la};
}

}  // namespace

extern "C" {

void collect_block_impr_(double *xblk, const double *xop, const I *kc_st0, const I *kc_nd0,
                         const I *kc_len, const I *ka_st0, const I *ka_nd0, const I *ka_len,
                         const I *k_len, const I *kca_trp, const I *lc_st0, const I *lc_nd0,
                         const I *lc_len, const I *la_st0, const I *la_nd0, const I *la_len,
                         const I *l_len, const I *map_kl_c, const I *map_kl_a, const I *ncblk_op,
                         const I *nablk_op, const I *ldim_opc, const I *ldim_opa, const I *ncblk_k,
                         const I *nablk_k, const I *ncblk_l, const I *nablk_l, const I *lstr_kc,
                         const I *lstr_ka, const I *lstr_lc, const I *lstr_la, const I *nstr_kc,
                         const I *nstr_ka, const I *nstr_lc, const I *nstr_la, const I *ireo_kc,
                         const I *ireo_ka, const I *ireo_lc, const I *ireo_la) {
    gecco::clear_error();
    const Shape s = shape_of(kc_st0, kc_nd0, kc_len, ka_st0, ka_nd0, ka_len, k_len, lc_st0, lc_nd0, lc_len,
                             la_st0, la_nd0, la_len, l_len, map_kl_c, map_kl_a, ncblk_op, nablk_op,
                             ldim_opc, ldim_opa, ncblk_k, nablk_k, ncblk_l, nablk_l, lstr_kc, lstr_ka,
                             lstr_lc, lstr_la, nstr_kc, nstr_ka, nstr_lc, nstr_la, ireo_kc, ireo_ka,
                             ireo_lc, ireo_la);
    std::vector<I> src;
    if (collect_index(s, *kca_trp != 0, src)) {
        gecco::set_error(GECCO_HIP_EINVAL, "collect_block_impr: batch or components out of range");
        return;
    }
    const int64_t n = (int64_t)src.size();
    if (n == 0) return;
    I ext = 0;
    for (I v : src) ext = std::max(ext, iabs(v));
    std::lock_guard<std::mutex> lock(g_mu);
    double *d_blk = xblk, *d_op = const_cast<double *>(xop);
    int64_t *d_src = nullptr;
    const bool hb = !on_device(xblk), ho = !on_device(xop);
    if (!stage(0, (size_t)n * 8, (void **)&d_src) || (hb && !stage(1, (size_t)n * 8, (void **)&d_blk)) ||
        (ho && !stage(2, (size_t)ext * 8, (void **)&d_op))) {
        gecco::set_error(GECCO_HIP_EDEVICE, "collect_block_impr: device memory");
        return;
    }
    bool ok = hipMemcpy(d_src, src.data(), (size_t)n * 8, hipMemcpyHostToDevice) == hipSuccess;
    if (ok && ho && ext > 0) ok = hipMemcpy(d_op, xop, (size_t)ext * 8, hipMemcpyHostToDevice) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(gather_signed, dim3(grid_of(n)), dim3(256), 0, 0, d_blk, d_op, d_src, n);
        ok = hipGetLastError() == hipSuccess;
    }
    if (ok && hb) ok = hipMemcpy(xblk, d_blk, (size_t)n * 8, hipMemcpyDeviceToHost) == hipSuccess;
    if (ok && !hb) ok = hipStreamSynchronize(0) == hipSuccess;
    if (!ok) gecco::set_error(GECCO_HIP_EDEVICE, "collect_block_impr: device copy or launch failed");
}

void scatter_block_(double *xop, const double *xblk, const I *kc_st0, const I *kc_nd0, const I *kc_len,
                    const I *ka_st0, const I *ka_nd0, const I *ka_len, const I *k_len, const I *lc_st0,
                    const I *lc_nd0, const I *lc_len, const I *la_st0, const I *la_nd0, const I *la_len,
                    const I *l_len, const I *map_kl_c, const I *map_kl_a, const I *ncblk_op,
                    const I *nablk_op, const I *ldim_opc, const I *ldim_opa, const I *ncblk_k,
                    const I *nablk_k, const I *ncblk_l, const I *nablk_l, const I *lstr_kc,
                    const I *lstr_ka, const I *lstr_lc, const I *lstr_la, const I *nstr_kc,
                    const I *nstr_ka, const I *nstr_lc, const I *nstr_la, const I *ireo_kc,
                    const I *ireo_ka, const I *ireo_lc, const I *ireo_la, I *istr_kc, I *istr_ka,
                    I *istr_lc, I *istr_la) {
    (void)istr_kc;
    (void)istr_ka;
    (void)istr_lc;
    (void)istr_la;
    gecco::clear_error();
    const Shape s = shape_of(kc_st0, kc_nd0, kc_len, ka_st0, ka_nd0, ka_len, k_len, lc_st0, lc_nd0, lc_len,
                             la_st0, la_nd0, la_len, l_len, map_kl_c, map_kl_a, ncblk_op, nablk_op,
                             ldim_opc, ldim_opa, ncblk_k, nablk_k, ncblk_l, nablk_l, lstr_kc, lstr_ka,
                             lstr_lc, lstr_la, nstr_kc, nstr_ka, nstr_lc, nstr_la, ireo_kc, ireo_ka,
                             ireo_lc, ireo_la);
    std::vector<I> from, to;
    if (scatter_index(s, from, to)) {
        gecco::set_error(GECCO_HIP_EINVAL, "scatter_block: components out of range");
        return;
    }
    const int64_t n = (int64_t)from.size();
    if (n == 0) return;
    I ext = 0, nb = 0;
    for (size_t i = 0; i < from.size(); ++i) {
        ext = std::max(ext, iabs(to[i]));
        nb = std::max(nb, from[i]);
    }
    std::lock_guard<std::mutex> lock(g_mu);
    double *d_op = xop, *d_blk = const_cast<double *>(xblk);
    int64_t *d_idx = nullptr;
    const bool ho = !on_device(xop), hb = !on_device(xblk);
    if (!stage(0, (size_t)n * 16, (void **)&d_idx) || (ho && !stage(2, (size_t)ext * 8, (void **)&d_op)) ||
        (hb && !stage(1, (size_t)nb * 8, (void **)&d_blk))) {
        gecco::set_error(GECCO_HIP_EDEVICE, "scatter_block: device memory");
        return;
    }
    bool ok = hipMemcpy(d_idx, from.data(), (size_t)n * 8, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(d_idx + n, to.data(), (size_t)n * 8, hipMemcpyHostToDevice) == hipSuccess;
    if (ok && ho) ok = hipMemcpy(d_op, xop, (size_t)ext * 8, hipMemcpyHostToDevice) == hipSuccess;
    if (ok && hb) ok = hipMemcpy(d_blk, xblk, (size_t)nb * 8, hipMemcpyHostToDevice) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(scatter_signed, dim3(grid_of(n)), dim3(256), 0, 0, d_op, d_blk, d_idx, d_idx + n, n);
        ok = hipGetLastError() == hipSuccess;
    }
    if (ok && ho) ok = hipMemcpy(xop, d_op, (size_t)ext * 8, hipMemcpyDeviceToHost) == hipSuccess;
    if (ok && !ho) ok = hipStreamSynchronize(0) == hipSuccess;
    if (!ok) gecco::set_error(GECCO_HIP_EDEVICE, "scatter_block: device copy or launch failed");
}

}  // extern "C"
