// Host side of the MI355X block-contraction engine: resolution of GeCCo's
// string maps into signed index tables, task tiling of the grouped launch,
// and the C ABI declared in include/gecco_hip.h.
//
// The index arithmetic restates what the reference does element by element
// inside its kernels:
//   * compound string index -> component strings, C before A, first
//     component fastest (contract/collect_block_impr.f:120-290,
//     contract/scatter_block.f:67-141),
//   * component map lookup  map(str1 + n1*str2), entry sign*(1+index), 0 =
//     forbidden (contract/collect_block_impr.f:255-264,381-390),
//   * in-block address 1 + sum idx*ldim with ldim from
//     contract/set_op_ldim_c.f:237-308,
//   * map placement per target component from map_info
//     (contract/set_strmapdim_c2.f:176-235),
// but does it once per plan instead of once per call, so the kernel reads a
// ready-made signed offset per row/column (separable case) or per element.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/gecco_hip.h"
#include "contract_types.h"

// launchers (contract_kernels.hip); elem: 8 = float64 lists, 4 = float32
extern "C" int gecco_launch_grouped(const gecco::DevBlock *blocks, const gecco::DevTask *tasks,
                                    int ntasks, void *const *bufs, int grid,
                                    const int32_t *skip, const int32_t *zero, int elem,
                                    hipStream_t stream);
extern "C" int gecco_launch_grouped_affine(const gecco::DevBlock *blocks,
                                           const gecco::DevTask *tasks, int ntasks,
                                           void *const *bufs, int grid, int variant,
                                           const int32_t *skip, const int32_t *zero, void *raw, int elem,
                                           hipStream_t stream);
extern "C" int gecco_affine_variant_shape(int v, int *bm, int *bn, int *bk);
extern "C" int gecco_launch_spin_project(void *x, const int64_t *gptr, const int64_t *elem, const int64_t *gmat,
                                         const double *mats, int64_t ngroups, int elem_size, hipStream_t stream);
extern "C" int gecco_launch_norms(const gecco::DevNormTask *tasks, int64_t ntasks, const int64_t *rows,
                                  int64_t njobs, void *const *bufs, double *partial, int32_t *zero, int elem,
                                  hipStream_t stream);
extern "C" int gecco_direct_stage_cap(void);
extern "C" int gecco_launch_sym_batched(const gecco::DevSymStep *steps,
                                        const gecco::DevSymTask *tasks, int ntasks,
                                        void *const *bufs, const int32_t *i1, const int32_t *i2,
                                        const int8_t *fsign, const int32_t *skip, int elem,
                                        hipStream_t stream);
extern "C" int gecco_launch_trace(const void *src, void *dst, const int64_t *rowptr,
                                  const int64_t *rowdst, const int64_t *col, const int8_t *sgn,
                                  int64_t nrows, double fact, int elem, hipStream_t stream);
extern "C" int gecco_launch_init(const gecco::DevInit *jobs, const gecco::DevInitTask *tasks,
                                 int ntasks, void *const *bufs, const int32_t *skip, int elem,
                                 hipStream_t stream);
extern "C" int gecco_launch_compact(const void *src, void *dst, int64_t n, int rec_bytes, int grp_byte,
                                    const int32_t *skip, int32_t *chunk, int64_t *count,
                                    hipStream_t stream);
extern "C" int gecco_launch_gather(const gecco::DevGather *jobs, const gecco::DevGatherTask *tasks,
                                   int ntasks, void *const *bufs, void *scratch,
                                   const int32_t *skip, int elem, hipStream_t stream);

namespace gecco {

static thread_local int g_err = 0;
static thread_local std::string g_msg;

static int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = code;
    g_msg = buf;
    return code;
}

static int ok() { return 0; }

static const int64_t HPVXSEQ[4] = {4, 2, 3, 1};  // include/hpvxseq.h:2-4

static void set_op_ldim_c(std::vector<int64_t> &ldc, std::vector<int64_t> &lda,
                          const int64_t *hc, const int64_t *ha, const int64_t *nstr,
                          int64_t nc, int64_t na, bool transpose) {
    ldc.assign(nc, 0);
    lda.assign(na, 0);
    int64_t ld = 1;
    for (int ih = 0; ih < 4; ++ih) {
        const int64_t h = HPVXSEQ[ih];
        if (!transpose) {
            for (int64_t i = 0; i < nc; ++i)
                if (hc[i] == h) { ldc[i] = ld; ld *= nstr[i]; }
            for (int64_t i = 0; i < na; ++i)
                if (ha[i] == h) { lda[i] = ld; ld *= nstr[nc + i]; }
        } else {
            for (int64_t i = na - 1; i >= 0; --i)
                if (ha[i] == h) { lda[i] = ld; ld *= nstr[nc + i]; }
            for (int64_t i = nc - 1; i >= 0; --i)
                if (hc[i] == h) { ldc[i] = ld; ld *= nstr[i]; }
        }
    }
}

static int64_t prod(const int64_t *a, int64_t n) {
    int64_t p = 1;
    for (int64_t i = 0; i < n; ++i) p *= a[i];
    return p;
}

// A target component of the operand being addressed: its string comes from
// (at most) one component of each index space.
struct Target {
    int side1 = -1, comp1 = -1;   // source 1 (fastest in the map)
    int side2 = -1, comp2 = -1;   // source 2
    int64_t n1 = 1, n2 = 1;
    const int64_t *map = nullptr;
    int64_t ldim = 0;
};

struct Space {
    std::vector<int64_t> len;  // component lengths, first fastest
    int64_t total() const {
        int64_t p = 1;
        for (auto l : len) p *= l;
        return p;
    }
};

// Parse one map_info array (contract/set_strmapdim_c2.f:203-232) for ntgt
// target components.  src1/src2 give the index space side and the component
// offset of the source lists inside that space.
static int add_targets(std::vector<Target> &out, const int64_t *mi, int64_t ntgt,
                       const int64_t *map, const std::vector<int64_t> &ldim, int side1,
                       int64_t off1, int64_t nsrc1, int side2, int64_t off2, int64_t nsrc2,
                       const Space *spaces) {
    int64_t k = 0, mapoff = 0;
    for (int64_t t = 0; t < ntgt; ++t) {
        Target tg;
        int64_t n = mi[k++];
        if (n > 1) return fail(GECCO_HIP_EINVAL, "map_info: %lld sources merged into one component (one-vertex operators only)", (long long)n);
        if (n == 1) {
            int64_t c = mi[k++];
            if (c < 1 || c > nsrc1) return fail(GECCO_HIP_EINVAL, "map_info: bad source index");
            tg.side1 = side1;
            tg.comp1 = (int)(off1 + c - 1);
            tg.n1 = spaces[side1].len[tg.comp1];
        }
        n = mi[k++];
        if (n > 1) return fail(GECCO_HIP_EINVAL, "map_info: %lld sources merged into one component (one-vertex operators only)", (long long)n);
        if (n == 1) {
            int64_t c = mi[k++];
            if (c < 1 || c > nsrc2) return fail(GECCO_HIP_EINVAL, "map_info: bad source index");
            tg.side2 = side2;
            tg.comp2 = (int)(off2 + c - 1);
            tg.n2 = spaces[side2].len[tg.comp2];
        }
        if (tg.side1 < 0 && tg.side2 < 0) return fail(GECCO_HIP_EINVAL, "map_info: component without source");
        tg.map = map + mapoff;
        tg.ldim = ldim[t];
        mapoff += tg.n1 * tg.n2;
        out.push_back(tg);
    }
    return ok();
}

struct Resolved {
    bool table = false;
    std::vector<int32_t> first, second;  // separable vectors, or table in `first`
    int64_t F = 0, S = 0;
    int64_t maxaddr = -1;                // largest element offset touched
};

// Evaluate the signed offset of (f, s) for a subset of targets.
// Returns false if the element is forbidden.
static inline bool eval(const std::vector<Target> &tg, const int64_t *ci[2], int64_t &addr,
                        int &sgn, int onlyside) {
    addr = 0;
    sgn = 1;
    for (const auto &t : tg) {
        if (onlyside >= 0) {
            int side = t.side1 >= 0 ? t.side1 : t.side2;
            if (side != onlyside) continue;
        }
        int64_t a = t.side1 >= 0 ? ci[t.side1][t.comp1] : 0;
        int64_t b = t.side2 >= 0 ? ci[t.side2][t.comp2] : 0;
        int64_t m = t.map[a + t.n1 * b];
        if (m == 0) return false;
        if (m < 0) { sgn = -sgn; m = -m; }
        addr += (m - 1) * t.ldim;
    }
    return true;
}

static void decompose(int64_t idx, const Space &sp, int64_t *out) {
    for (size_t c = 0; c < sp.len.size(); ++c) {
        out[c] = idx % sp.len[c];
        idx /= sp.len[c];
    }
}

static int resolve(const Space &fs, const Space &ss, const std::vector<Target> &tg,
                   int64_t oplen, Resolved &r) {
    // targets refer to sides 0 (first space) and 1 (second space)
    bool sep = true;
    for (const auto &t : tg)
        if (t.side1 >= 0 && t.side2 >= 0 && t.side1 != t.side2) sep = false;
    r.F = fs.total();
    r.S = ss.total();
    r.table = !sep;
    int64_t cf[16] = {0}, cs[16] = {0};
    const int64_t *ci[2] = {cf, cs};
    if (fs.len.size() > 16 || ss.len.size() > 16) return fail(GECCO_HIP_EINVAL, "too many components");
    if (sep) {
        r.first.resize(r.F);
        r.second.resize(r.S);
        int64_t mf = 0, ms = 0;
        for (int64_t f = 0; f < r.F; ++f) {
            decompose(f, fs, cf);
            int64_t a;
            int s;
            if (!eval(tg, ci, a, s, 0)) { r.first[f] = 0; continue; }
            if (a + 1 > INT32_MAX) return fail(GECCO_HIP_EINVAL, "block too large for 32-bit offsets");
            r.first[f] = (int32_t)(s * (a + 1));
            mf = std::max(mf, a);
        }
        for (int64_t s2 = 0; s2 < r.S; ++s2) {
            decompose(s2, ss, cs);
            int64_t a;
            int s;
            if (!eval(tg, ci, a, s, 1)) { r.second[s2] = 0; continue; }
            if (a + 1 > INT32_MAX) return fail(GECCO_HIP_EINVAL, "block too large for 32-bit offsets");
            r.second[s2] = (int32_t)(s * (a + 1));
            ms = std::max(ms, a);
        }
        r.maxaddr = mf + ms;
    } else {
        if (r.F * r.S > (int64_t)1 << 31) return fail(GECCO_HIP_EINVAL, "index table too large");
        r.first.resize(r.F * r.S);
        int64_t mx = -1;
        for (int64_t s2 = 0; s2 < r.S; ++s2) {
            decompose(s2, ss, cs);
            for (int64_t f = 0; f < r.F; ++f) {
                decompose(f, fs, cf);
                int64_t a;
                int s;
                if (!eval(tg, ci, a, s, -1)) { r.first[f + r.F * s2] = 0; continue; }
                if (a + 1 > INT32_MAX) return fail(GECCO_HIP_EINVAL, "block too large for 32-bit offsets");
                r.first[f + r.F * s2] = (int32_t)(s * (a + 1));
                mx = std::max(mx, a);
            }
        }
        r.maxaddr = mx;
    }
    if (r.maxaddr >= oplen)
        return fail(GECCO_HIP_EINVAL, "string map addresses element %lld of a block of length %lld",
                    (long long)r.maxaddr, (long long)oplen);
    return ok();
}

static inline int64_t absv(int32_t v) { return v < 0 ? -(int64_t)v : (int64_t)v; }

// A table whose entries factor as entry(f, s) = first[f] (+) second[s] --
// zero pattern an outer product of a row and a column pattern, sign a
// product of row and column signs, offset a sum of row and column offsets --
// is separable even when its targets mix both index spaces (e.g. the string
// pair of an alpha and a beta electron: no Pauli exclusion, offset
// idx(a) + n_a * idx(b)).  Converts such a table in place.
static void try_separate(Resolved &r) {
    if (!r.table || r.F == 0 || r.S == 0) return;
    const int64_t F = r.F, S = r.S;
    const std::vector<int32_t> &t = r.first;
    int64_t f0 = -1, s0 = -1;
    for (int64_t s = 0; s < S && f0 < 0; ++s)
        for (int64_t f = 0; f < F; ++f)
            if (t[f + F * s]) { f0 = f; s0 = s; break; }
    std::vector<int32_t> fv(F, 0), sv(S, 0);
    if (f0 < 0) {   // nothing allowed
        r.first.swap(fv);
        r.second.swap(sv);
        r.table = false;
        return;
    }
    const int32_t ref = t[f0 + F * s0];
    const int64_t aref = absv(ref) - 1;
    // column offsets relative to the smallest one in row f0
    int64_t cmin = INT64_MAX;
    for (int64_t s = 0; s < S; ++s)
        if (t[f0 + F * s]) cmin = std::min(cmin, absv(t[f0 + F * s]) - 1);
    for (int64_t s = 0; s < S; ++s) {
        const int32_t v = t[f0 + F * s];
        if (!v) continue;
        const int64_t a = absv(v) - 1 - cmin;
        if (a + 1 > INT32_MAX) return;
        sv[s] = (int32_t)((v < 0) == (ref < 0) ? a + 1 : -(a + 1));
    }
    for (int64_t f = 0; f < F; ++f) {
        const int32_t v = t[f + F * s0];
        if (!v) continue;
        const int64_t a = absv(v) - 1 - aref + cmin;
        if (a < 0 || a + 1 > INT32_MAX) return;
        fv[f] = (int32_t)(v < 0 ? -(a + 1) : a + 1);
    }
    // row signs were taken relative to column s0 whose own sign is
    // sign(sv[s0]) * sign(ref) = +; verify every entry
    for (int64_t s = 0; s < S; ++s)
        for (int64_t f = 0; f < F; ++f) {
            const int32_t v = t[f + F * s];
            const int32_t a = fv[f], b = sv[s];
            if (!a || !b) {
                if (v) return;
                continue;
            }
            if (!v) return;
            const bool neg = (a < 0) != (b < 0);
            if ((v < 0) != neg || absv(v) != absv(a) + absv(b) - 1) return;
        }
    r.first.swap(fv);
    r.second.swap(sv);
    r.table = false;
}

// address of first-index step f=1 relative to f=0 (separable), or of the
// table entry (1,0) relative to (0,0); used to pick coalescing directions
static bool first_is_contiguous(const Resolved &r) {
    if (r.F < 2) return false;
    if (r.table) {
        for (int64_t s = 0; s < std::min<int64_t>(r.S, 4); ++s) {
            int32_t a = r.first[0 + r.F * s], b = r.first[1 + r.F * s];
            if (a != 0 && b != 0) return absv(b) - absv(a) == 1;
        }
        return false;
    }
    return r.first[0] != 0 && r.first[1] != 0 && absv(r.first[1]) - absv(r.first[0]) == 1;
}

// entry i == 1 + a0 + a1*i for all i, and every entry positive: no forbidden
// element (0) and no sign flip (a decreasing run such as [1, 0] or [2, 1, 0,
// -1] fits the formula but holds forbidden or negative entries)
static bool affine_of(const std::vector<int32_t> &v, int64_t &a0, int64_t &a1) {
    if (v.empty() || v[0] <= 0) return false;
    a0 = (int64_t)v[0] - 1;
    a1 = v.size() > 1 ? (int64_t)v[1] - (int64_t)v[0] : 0;
    for (size_t i = 0; i < v.size(); ++i)
        if (v[i] <= 0 || (int64_t)v[i] != 1 + a0 + a1 * (int64_t)i) return false;
    return true;
}

// Range [amin, amax] of the element offsets a tile (rows [m0, m1), columns
// [n0, n1) in kernel orientation) writes, and how many elements it writes.
static int64_t tile_range(const Resolved &c, int64_t m0, int64_t m1, int64_t n0, int64_t n1,
                          int64_t &amin, int64_t &amax) {
    amin = INT64_MAX;
    amax = -1;
    int64_t count = 0;
    if (c.table) {
        for (int64_t n = n0; n < n1; ++n)
            for (int64_t m = m0; m < m1; ++m) {
                const int32_t v = c.first[m + c.F * n];
                if (!v) continue;
                const int64_t a = absv(v) - 1;
                amin = std::min(amin, a);
                amax = std::max(amax, a);
                ++count;
            }
        return count;
    }
    int64_t fmin = INT64_MAX, fmax = -1, fc = 0, smin = INT64_MAX, smax = -1, sc = 0;
    for (int64_t m = m0; m < m1; ++m)
        if (c.first[m]) {
            const int64_t a = absv(c.first[m]) - 1;
            fmin = std::min(fmin, a); fmax = std::max(fmax, a); ++fc;
        }
    for (int64_t n = n0; n < n1; ++n)
        if (c.second[n]) {
            const int64_t a = absv(c.second[n]) - 1;
            smin = std::min(smin, a); smax = std::max(smax, a); ++sc;
        }
    if (!fc || !sc) return 0;
    amin = fmin + smin;
    amax = fmax + smax;
    return fc * sc;
}

static bool result_injective(const Resolved &r, int64_t len) {
    std::vector<uint8_t> seen(len, 0);
    if (r.table) {
        for (auto v : r.first) {
            if (!v) continue;
            int64_t a = absv(v) - 1;
            if (seen[a]) return false;
            seen[a] = 1;
        }
        return true;
    }
    for (int64_t s = 0; s < r.S; ++s) {
        if (!r.second[s]) continue;
        for (int64_t f = 0; f < r.F; ++f) {
            if (!r.first[f]) continue;
            int64_t a = absv(r.first[f]) + absv(r.second[s]) - 2;
            if (seen[a]) return false;
            seen[a] = 1;
        }
    }
    return true;
}

struct HostOperand {
    Resolved r;
    int fast_first = 0;
    int64_t off_first = -1, off_second = -1;  // element offsets in the index pool
    int64_t off_first_t = -1;                  // transposed table (result of direct blocks)
};

struct HostBlock {
    double xfac;
    int32_t buf1, buf2, buf12;
    int64_t off1, off2, off12, len12;
    // lengths of the operand blocks at off1 / off2 (the whole Ms/IRREP
    // distribution block of each operand: the range the norm screening
    // looks at)
    int64_t len1 = 0, len2 = 0;
    int64_t M, N, K;    // kernel orientation
    bool swap;          // A=Op2, B=Op1 when the result is contiguous along EX1
    bool injective;
    bool strided = false;             // A and B strided (natively or materialised)
    bool direct = false;              // runs on the short-K vector-ALU kernel
    bool skinny = false;              // strided-GEMM launch, 16 x 16 wave-split-K tiles
    bool mat_a = false, mat_b = false;  // A / B materialised into scratch
    int32_t next = -1;                  // next block of its chain (finalize)
    int32_t group = 0;                  // instance group (skip flags)
    HostOperand A, B, C;
};

}  // namespace gecco

struct SymStep {
    int32_t buf = 0;
    int32_t group = 0;      // instance group (skip flags)
    double prefac = 0.5;
    std::vector<int64_t> i1, i2;
    std::vector<int8_t> f;
    int64_t npair = 0;
};

// Partial trace of a list into another (CSR rows of signed sources).
struct TraceStep {
    int32_t buf_src = 0, buf_dst = 0;
    int64_t off_src = 0, off_dst = 0;
    double fact = 1.0;
    std::vector<int64_t> rowptr, rowdst, col;
    std::vector<int8_t> sgn;
    int64_t nrows = 0, nnz = 0;
    int64_t *d_rowptr = nullptr, *d_rowdst = nullptr, *d_col = nullptr;
    int8_t *d_sgn = nullptr;
};

// Singlet spin projection of a list (contract/spin_prj_blk.f): the list's
// elements fall into groups with the same spatial orbitals and different
// spins; every group is replaced by P_g x_g, P_g the projector onto the
// singlet operators of the group (one of a table of distinct projectors).
struct SpinStep {
    int32_t buf = 0, group = 0;
    int64_t off = 0;
    std::vector<int64_t> gptr, elem, gmat;   // CSR groups; offset of P_g in mats
    std::vector<double> mats;                // n x n row-major projectors
    int64_t ngroups = 0;
    int64_t *d_gptr = nullptr, *d_elem = nullptr, *d_gmat = nullptr;
    double *d_mats = nullptr;
};

struct gecco_hip_plan {
    std::vector<gecco::HostBlock> blocks;
    std::vector<TraceStep> traces;                // after the contractions
    // fixed-order mode (gecco_hip_plan_set_deterministic): the strided-GEMM
    // launch writes every task's unsigned tile into d_raw, and the CSR sums
    // of dtraces (one per result slot, contributions in task order) add them
    // into the result lists
    bool deterministic = false;
    void *d_raw = nullptr;
    int64_t raw_elems = 0;
    std::vector<TraceStep> dtraces;
    std::vector<SpinStep> spins;                  // after the spin-flip projections
    std::vector<gecco::DevInit> inits;            // list initialisation, first launch
    std::vector<int32_t> init_group;              // instance group of each init
    gecco::DevInit *d_inits = nullptr;
    gecco::DevInitTask *d_itasks = nullptr;
    int64_t nitasks = 0;
    std::vector<SymStep> syms;
    // all symmetrisation pairs of the plan, one launch (int32 offsets
    // relative to each step's base element)
    gecco::DevSymStep *d_symsteps = nullptr;
    gecco::DevSymTask *d_symtasks = nullptr;
    int32_t *d_sym_i1 = nullptr, *d_sym_i2 = nullptr;
    int8_t *d_sym_f = nullptr;
    int64_t nsymtasks = 0;
    // buffer slots whose contents stay fixed between executes (integrals):
    // their gathers run on the first execute and after a slot changes
    std::vector<uint8_t> constant_slot;
    int64_t ngtasks_static = 0;
    bool static_done = false;
    bool finalized = false;
    gecco::DevBlock *d_blocks = nullptr;
    gecco::DevTask *d_tasks = nullptr;
    int32_t *d_index = nullptr;
    void **d_bufs = nullptr;
    // task classes, in this order in d_tasks: strided-GEMM launch (GEMM
    // tiles, 16 x 16 wave-split-K tiles and direct short-contraction tasks),
    // gather kernel
    int64_t ntasks = 0, ntasks_affine = 0;
    // operand materialisation (signed gather into plan-owned scratch)
    gecco::DevGather *d_gjobs = nullptr;
    gecco::DevGatherTask *d_gtasks = nullptr;
    void *d_scratch = nullptr;
    int64_t ngjobs = 0, ngtasks = 0, scratch_elems = 0;
    int64_t materialize_min = 1;    // min(M, N) from which gathers are materialised
    bool chain = true;              // link strided-GEMM blocks with a common result tile
    // norm screening (contract/contr_op1op2_wmaps_c.f:1652-1655,1767-1770):
    // a kernel call whose operand block has a squared norm below 1e-28 is
    // skipped; the norms of the distinct operand blocks (jobs) are summed in
    // chunks (partial) and compared (zero) before the contraction launches
    bool screen = false;
    gecco::DevNormTask *d_ntasks = nullptr;
    double *d_npartial = nullptr;
    int64_t *d_nrow = nullptr;              // chunk range of each job (CSR)
    int32_t *d_nzero = nullptr;             // 1: the job's block is screened out
    int64_t nnorm_tasks = 0, nnorm_jobs = 0;
    int64_t skinny_max = 16;        // min(M, N) up to which 16 x 16 wave-split-K tiles are used
    int64_t skinny_kchunk = 2048;   // contraction length of one such task
    int64_t nchained = 0;           // blocks linked behind a chain head
    int64_t nstaged = 0;            // tasks with an LDS-staged result range
    double gather_bytes = 0;
    int affine_variant = -1;        // tile variant; -1: 1 for float64 plans, 9 for float32
    int64_t target_tasks = 2048;
    int64_t min_kchunk = 256;
    int64_t max_kchunk = 1024;      // longest contraction range of one strided-GEMM task
    // order of the strided-GEMM launch's tasks: 0 longest first; 1 patches
    // of tiles that share operand panels kept on one XCD (workgroup w runs
    // on XCD w mod 8), patches spread over the XCDs by cost
    int xcd_order = 1;
    int64_t nbufs_max = 0;
    int grid = 0;
    double flops = 0, payload = 0, index_bytes = 0;
    std::vector<void *> last_bufs;
    // optional HIP-event timing of the launches of execute (per launch class:
    // 0 strided-GEMM kernel, 1 gather kernel, 2 spin-flip symmetrisation,
    // 3 operand materialisation, 4 unused: the direct short-contraction tasks
    // run in the strided-GEMM launch)
    bool timing = false;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
    double ms[5] = {0, 0, 0, 0, 0};
    int64_t launches[5] = {0, 0, 0, 0, 0};
    double flops_affine = 0, flops_general = 0, flops_direct = 0;
    int64_t direct_kmax = 32;   // K bound of the direct kernel (0 disables it)
    // instance groups: items recorded after set_group(g) belong to group g;
    // with skip flags set, every launch skips the tasks of group g while
    // skip[g] != 0 (device memory, read when each task starts)
    int32_t cur_group = 0, max_group = 0;
    // keep the host index data after finalize (gecco_hip_plan_retune)
    bool keep_host = false;
    // element size of the lists (8: float64, 4: float32)
    int elem = 8;
    const int32_t *d_skip = nullptr;
    int64_t nskip = 0;
    // live-task lists (gecco_hip_plan_compact): copies of the task lists
    // with only the tasks of groups whose skip flag was 0 when they were
    // made, in their order, so that the launches need no workgroups for
    // converged instances; the launches use them while `compacted`
    bool compacted = false;
    gecco::DevTask *d_ltasks = nullptr;          // strided-GEMM at [0, lt[0]), gather kernel at [na, na + lt[1])
    gecco::DevInitTask *d_litasks = nullptr;     // [0, lt[2])
    gecco::DevSymTask *d_lsymtasks = nullptr;    // [0, lt[3])
    gecco::DevGatherTask *d_lgtasks = nullptr;   // constant-slot gathers at [0, lt[4]), others at [ngtasks_static, + lt[5])
    int32_t *d_lchunk = nullptr;
    int64_t *d_lcount = nullptr;
    int64_t lt[6] = {0, 0, 0, 0, 0, 0};
    // device memory from a caller-given region instead of hipMalloc (the
    // allocation-free entry point): base, size, bytes handed out
    char *arena = nullptr;
    size_t arena_bytes = 0, arena_used = 0;
};

namespace gecco {

static int build_block(const gecco_hip_block_args *a, HostBlock &hb) {
    const int64_t ncc = a->ncblk_cnt, nca = a->nablk_cnt;
    const int64_t n1c = a->ncblk_ex1, n1a = a->nablk_ex1;
    const int64_t n2c = a->ncblk_ex2, n2a = a->nablk_ex2;
    for (int64_t v : {ncc, nca, n1c, n1a, n2c, n2a, a->ncblk_op1, a->nablk_op1, a->ncblk_op2,
                      a->nablk_op2, a->ncblk_op1op2, a->nablk_op1op2})
        if (v < 0 || v > 8) return fail(GECCO_HIP_EINVAL, "component count out of range");
    Space cnt, ex1, ex2;
    cnt.len.assign(a->lstr_cnt, a->lstr_cnt + ncc + nca);
    ex1.len.assign(a->lstr_ex1, a->lstr_ex1 + n1c + n1a);
    ex2.len.assign(a->lstr_ex2, a->lstr_ex2 + n2c + n2a);
    for (auto *sp : {&cnt, &ex1, &ex2})
        for (auto l : sp->len)
            if (l <= 0) return fail(GECCO_HIP_EINVAL, "zero or negative string length");
    const int64_t len1 = prod(a->lstrop1, a->ncblk_op1 + a->nablk_op1);
    const int64_t len2 = prod(a->lstrop2, a->ncblk_op2 + a->nablk_op2);
    const int64_t len12 = prod(a->lstrop1op2, a->ncblk_op1op2 + a->nablk_op1op2);

    std::vector<int64_t> l1c, l1a, l2c, l2a, l12c, l12a;
    set_op_ldim_c(l1c, l1a, a->hpvxop1c, a->hpvxop1a, a->lstrop1, a->ncblk_op1, a->nablk_op1,
                  a->tra_op1 != 0);
    set_op_ldim_c(l2c, l2a, a->hpvxop2c, a->hpvxop2a, a->lstrop2, a->ncblk_op2, a->nablk_op2,
                  a->tra_op2 != 0);
    set_op_ldim_c(l12c, l12a, a->hpvxop1op2c, a->hpvxop1op2a, a->lstrop1op2, a->ncblk_op1op2,
                  a->nablk_op1op2, a->tra_op1op2 != 0);

    // Op1(CNT, EX1): C <- (CNT C, EX1 C), A <- (CNT A, EX1 A)
    std::vector<Target> t1, t2, t12;
    {
        const Space sp[2] = {cnt, ex1};
        if (add_targets(t1, a->map_info_1c, a->ncblk_op1, a->map_ex1cntc, l1c, 0, 0, ncc, 1, 0, n1c, sp)) return g_err;
        if (add_targets(t1, a->map_info_1a, a->nablk_op1, a->map_ex1cnta, l1a, 0, ncc, nca, 1, n1c, n1a, sp)) return g_err;
    }
    // Op2(CNT, EX2): C <- (CNT A, EX2 C), A <- (CNT C, EX2 A)   ("CNT^+")
    {
        const Space sp[2] = {cnt, ex2};
        if (add_targets(t2, a->map_info_2c, a->ncblk_op2, a->map_ex2cntc, l2c, 0, ncc, nca, 1, 0, n2c, sp)) return g_err;
        if (add_targets(t2, a->map_info_2a, a->nablk_op2, a->map_ex2cnta, l2a, 0, 0, ncc, 1, n2c, n2a, sp)) return g_err;
    }
    // Res(EX1, EX2): C <- (EX1 C, EX2 C), A <- (EX2 A, EX1 A)
    {
        const Space sp[2] = {ex1, ex2};
        if (add_targets(t12, a->map_info_12c, a->ncblk_op1op2, a->map_ex1ex2c, l12c, 0, 0, n1c, 1, 0, n2c, sp)) return g_err;
        if (add_targets(t12, a->map_info_12a, a->nablk_op1op2, a->map_ex1ex2a, l12a, 1, n2c, n2a, 0, n1c, n1a, sp)) return g_err;
    }
    Resolved r1, r2, r12;
    if (resolve(cnt, ex1, t1, len1, r1)) return g_err;
    if (resolve(cnt, ex2, t2, len2, r2)) return g_err;
    if (resolve(ex1, ex2, t12, len12, r12)) return g_err;
    try_separate(r1);
    try_separate(r2);
    try_separate(r12);

    hb.xfac = a->xfac;
    hb.len12 = len12;
    hb.injective = result_injective(r12, len12);
    hb.swap = first_is_contiguous(r12);
    hb.K = cnt.total();
    int64_t lop1 = 1, lop2 = 1;
    for (int64_t c = 0; c < a->ncblk_op1 + a->nablk_op1; ++c) lop1 *= a->lstrop1[c];
    for (int64_t c = 0; c < a->ncblk_op2 + a->nablk_op2; ++c) lop2 *= a->lstrop2[c];
    hb.len1 = hb.swap ? lop2 : lop1;
    hb.len2 = hb.swap ? lop1 : lop2;
    if (!hb.swap) {
        hb.M = ex1.total(); hb.N = ex2.total();
        hb.buf1 = a->buf_op1; hb.off1 = a->off_op1;
        hb.buf2 = a->buf_op2; hb.off2 = a->off_op2;
        hb.A.r = std::move(r1); hb.B.r = std::move(r2); hb.C.r = std::move(r12);
    } else {
        hb.M = ex2.total(); hb.N = ex1.total();
        hb.buf1 = a->buf_op2; hb.off1 = a->off_op2;
        hb.buf2 = a->buf_op1; hb.off2 = a->off_op1;
        hb.A.r = std::move(r2); hb.B.r = std::move(r1);
        // result in kernel orientation (first = EX2, second = EX1)
        Resolved &c = hb.C.r;
        c = std::move(r12);
        if (!c.table) {
            std::swap(c.first, c.second);
        } else {
            std::vector<int32_t> tt(c.F * c.S);
            for (int64_t s = 0; s < c.S; ++s)
                for (int64_t f = 0; f < c.F; ++f) tt[s + c.S * f] = c.first[f + c.F * s];
            c.first.swap(tt);
        }
        std::swap(c.F, c.S);
    }
    hb.buf12 = a->buf_op1op2;
    hb.off12 = a->off_op1op2;
    hb.A.fast_first = hb.A.r.table ? 1 : (first_is_contiguous(hb.A.r) ? 1 : 0);
    hb.B.fast_first = hb.B.r.table ? 1 : (first_is_contiguous(hb.B.r) ? 1 : 0);
    if (hb.M > INT32_MAX || hb.N > INT32_MAX || hb.K > INT32_MAX)
        return fail(GECCO_HIP_EINVAL, "block dimension exceeds 32 bits");
    return ok();
}

static int check(hipError_t e, const char *what) {
    if (e != hipSuccess) return fail(GECCO_HIP_EDEVICE, "%s: %s", what, hipGetErrorString(e));
    return ok();
}

// Device memory of a plan: hipMalloc, or, for a plan that owns none, a
// bump allocation inside a caller-given region (256-byte aligned pieces).
static int dmalloc(gecco_hip_plan *p, void **ptr, size_t bytes, const char *what) {
    if (p->arena) {
        const size_t off = (p->arena_used + 255) & ~(size_t)255;
        if (off + bytes > p->arena_bytes)
            return fail(GECCO_HIP_EINVAL, "%s: %zu bytes needed beyond %zu in use exceed the %zu-byte "
                        "scratch of the allocation-free entry point", what, bytes, off, p->arena_bytes);
        *ptr = p->arena + off;
        p->arena_used = off + bytes;
        return ok();
    }
    return check(hipMalloc(ptr, std::max<size_t>(bytes, 1)), what);
}

static void dfree(gecco_hip_plan *p, void *ptr) {
    if (!p->arena && ptr) (void)hipFree(ptr);
}

static void plan_free_device(gecco_hip_plan *p) {
    if (p->d_blocks) dfree(p, p->d_blocks);
    if (p->d_tasks) dfree(p, p->d_tasks);
    if (p->d_ntasks) dfree(p, p->d_ntasks);
    if (p->d_npartial) dfree(p, p->d_npartial);
    if (p->d_nrow) dfree(p, p->d_nrow);
    if (p->d_nzero) dfree(p, p->d_nzero);
    p->d_ntasks = nullptr;
    p->d_npartial = nullptr;
    p->d_nrow = nullptr;
    p->d_nzero = nullptr;
    p->nnorm_tasks = p->nnorm_jobs = 0;
    if (p->d_index) dfree(p, p->d_index);
    if (p->d_bufs) dfree(p, p->d_bufs);
    for (auto &sp : p->spins) {
        if (sp.d_gptr) dfree(p, sp.d_gptr);
        if (sp.d_elem) dfree(p, sp.d_elem);
        if (sp.d_gmat) dfree(p, sp.d_gmat);
        if (sp.d_mats) dfree(p, sp.d_mats);
        sp.d_gptr = sp.d_elem = sp.d_gmat = nullptr;
        sp.d_mats = nullptr;
    }
    for (auto &tr : p->traces) {
        if (tr.d_rowptr) dfree(p, tr.d_rowptr);
        if (tr.d_rowdst) dfree(p, tr.d_rowdst);
        if (tr.d_col) dfree(p, tr.d_col);
        if (tr.d_sgn) dfree(p, tr.d_sgn);
        tr.d_rowptr = tr.d_rowdst = tr.d_col = nullptr;
        tr.d_sgn = nullptr;
    }
    for (auto &tr : p->dtraces) {
        if (tr.d_rowptr) dfree(p, tr.d_rowptr);
        if (tr.d_rowdst) dfree(p, tr.d_rowdst);
        if (tr.d_col) dfree(p, tr.d_col);
        if (tr.d_sgn) dfree(p, tr.d_sgn);
    }
    p->dtraces.clear();
    if (p->d_raw) dfree(p, p->d_raw);
    p->d_raw = nullptr;
    p->raw_elems = 0;
    if (p->d_inits) dfree(p, p->d_inits);
    if (p->d_itasks) dfree(p, p->d_itasks);
    p->d_inits = nullptr;
    p->d_itasks = nullptr;
    if (p->d_gjobs) dfree(p, p->d_gjobs);
    if (p->d_gtasks) dfree(p, p->d_gtasks);
    if (p->d_scratch) dfree(p, p->d_scratch);
    p->d_gjobs = nullptr;
    p->d_gtasks = nullptr;
    p->d_scratch = nullptr;
    if (p->d_symsteps) dfree(p, p->d_symsteps);
    if (p->d_symtasks) dfree(p, p->d_symtasks);
    if (p->d_sym_i1) dfree(p, p->d_sym_i1);
    if (p->d_sym_i2) dfree(p, p->d_sym_i2);
    if (p->d_sym_f) dfree(p, p->d_sym_f);
    p->d_symsteps = nullptr;
    p->d_symtasks = nullptr;
    p->d_sym_i1 = p->d_sym_i2 = nullptr;
    p->d_sym_f = nullptr;
    if (p->d_ltasks) dfree(p, p->d_ltasks);
    if (p->d_litasks) dfree(p, p->d_litasks);
    if (p->d_lsymtasks) dfree(p, p->d_lsymtasks);
    if (p->d_lgtasks) dfree(p, p->d_lgtasks);
    if (p->d_lchunk) dfree(p, p->d_lchunk);
    if (p->d_lcount) dfree(p, p->d_lcount);
    p->d_ltasks = nullptr;
    p->d_litasks = nullptr;
    p->d_lsymtasks = nullptr;
    p->d_lgtasks = nullptr;
    p->d_lchunk = nullptr;
    p->d_lcount = nullptr;
    p->compacted = false;
    p->d_blocks = nullptr;
    p->d_tasks = nullptr;
    p->d_index = nullptr;
    p->d_bufs = nullptr;
}

static void plan_release_host(gecco_hip_plan *p) {
    for (auto &b : p->blocks)
        for (HostOperand *o : {&b.A, &b.B, &b.C}) {
            std::vector<int32_t>().swap(o->r.first);
            std::vector<int32_t>().swap(o->r.second);
        }
    for (auto &y : p->syms) {
        std::vector<int64_t>().swap(y.i1);
        std::vector<int64_t>().swap(y.i2);
        std::vector<int8_t>().swap(y.f);
    }
    for (auto &tr : p->traces) {
        std::vector<int64_t>().swap(tr.rowptr);
        std::vector<int64_t>().swap(tr.rowdst);
        std::vector<int64_t>().swap(tr.col);
        std::vector<int8_t>().swap(tr.sgn);
    }
    p->keep_host = false;
}

// XCD-grouped launch order of n tasks given in cost order (longest first):
// patch[t] = patch of task t (tiles that share operand panels), cost[t] its
// length.  A workgroup w of a launch with one workgroup per task runs on XCD
// w mod 8, each XCD with its own L2.  The patches go, longest task first, to
// the XCD with the least work so far (cost + 1 per task); each XCD keeps its
// patches in that order, a patch's tasks together and in their given order;
// the launch takes the XCDs' lists in turn (position w: the next task of XCD
// w mod 8, or of the list with most tasks left once that one is used up).
// perm[w] = the task at position w.
static void xcd_task_order(int64_t n, const int32_t *patch, const int64_t *cost, int64_t *perm) {
    constexpr int NX = 8;
    const size_t n0 = (size_t)n;
    std::vector<std::pair<int32_t, size_t>> byp(n0);   // (patch, task) with tasks in cost order
    for (size_t t = 0; t < n0; ++t) byp[t] = {patch[t], t};
    std::stable_sort(byp.begin(), byp.end(), [](const std::pair<int32_t, size_t> &x,
                                                 const std::pair<int32_t, size_t> &y) {
        return x.first < y.first;
    });
    struct Patch { size_t first, last; int64_t cost, top; };
    std::vector<Patch> pts;
    for (size_t a = 0; a < n0;) {
        size_t e = a;
        Patch q{a, a, 0, 0};
        while (e < n0 && byp[e].first == byp[a].first) {
            const int64_t c = cost[byp[e].second] + 1;
            q.cost += c;
            q.top = std::max(q.top, c);
            ++e;
        }
        q.last = e;
        pts.push_back(q);
        a = e;
    }
    std::stable_sort(pts.begin(), pts.end(), [](const Patch &x, const Patch &y) {
        return x.top != y.top ? x.top > y.top : x.cost > y.cost;
    });
    std::vector<std::vector<size_t>> lists(NX);
    int64_t load[NX] = {0};
    for (auto &q : pts) {
        int x = 0;
        for (int y = 1; y < NX; ++y)
            if (load[y] < load[x]) x = y;
        load[x] += q.cost;
        for (size_t e = q.first; e < q.last; ++e) lists[(size_t)x].push_back(byp[e].second);
    }
    size_t at[NX] = {0};
    for (size_t w = 0; w < n0; ++w) {
        int x = (int)(w % NX);
        if (at[x] == lists[(size_t)x].size()) {
            for (int y = 0; y < NX; ++y)
                if (lists[(size_t)y].size() - at[y] > lists[(size_t)x].size() - at[x]) x = y;
        }
        perm[w] = (int64_t)lists[(size_t)x][at[x]++];
    }
}

// Kernel orientation turned around: A and B (and M and N) exchanged, the
// result map transposed -- the same product C(m, n) = sum_k A(k, m) B(k, n)
// read as C^T(n, m) = sum_k B(k, n) A(k, m).
static void flip_orientation(HostBlock &b) {
    std::swap(b.M, b.N);
    std::swap(b.buf1, b.buf2);
    std::swap(b.off1, b.off2);
    std::swap(b.len1, b.len2);
    std::swap(b.A, b.B);
    Resolved &c = b.C.r;
    if (!c.table) {
        std::swap(c.first, c.second);
    } else {
        std::vector<int32_t> tt(c.F * c.S);
        for (int64_t s = 0; s < c.S; ++s)
            for (int64_t f = 0; f < c.F; ++f) tt[s + c.S * f] = c.first[f + c.F * s];
        c.first.swap(tt);
    }
    std::swap(c.F, c.S);
    b.swap = !b.swap;
}

// Whether a block's operands reach the strided-GEMM launch: each is a
// strided view of its list (affine index maps; *sa / *sb), or is gathered
// into dense scratch first (min(M, N) >= materialize_min).  The kernels
// address an operand block with unsigned 32-bit byte offsets: non-negative
// strides, extents up to 2^30 bytes.
static bool strided_operands(const gecco_hip_plan *p, const HostBlock &b, bool *sa_out = nullptr,
                             bool *sb_out = nullptr) {
    const int64_t max_extent = (int64_t)1 << 30;
    auto strided = [&](const HostOperand &o, int64_t X) {
        int64_t f0, f1, s0, s1;
        if (o.r.table || !affine_of(o.r.first, f0, f1) || !affine_of(o.r.second, s0, s1)) return false;
        return f1 >= 0 && s1 >= 0 && (f1 * (b.K - 1) + s1 * (X - 1) + 1) * p->elem <= max_extent;
    };
    const bool sa = strided(b.A, b.M), sb = strided(b.B, b.N);
    const bool big = p->materialize_min > 0 && std::min(b.M, b.N) >= p->materialize_min;
    const bool fa = b.K * b.M * p->elem <= max_extent, fb = b.K * b.N * p->elem <= max_extent;
    if (sa_out) *sa_out = sa;
    if (sb_out) *sb_out = sb;
    return (sa || (big && fa)) && (sb || (big && fb));
}

static int plan_finalize(gecco_hip_plan *p) {
    if (p->finalized) return fail(GECCO_HIP_ESTATE, "plan already finalized");
    // Short contractions (the direct tasks below) with at most 8 result
    // columns but more rows: turned so that the long side runs along the
    // lanes.  A direct task's thread owns one column, loads that column of
    // B with all its K values in flight and computes the task's few rows
    // from the A tile in LDS; with the long side along the rows, each task
    // would stage only DIRECT_A / K rows of the long operand and spend its
    // time on latency (e.g. T1 x <oo||vv>: 2500 x 5 results, K = 25).
    for (auto &b : p->blocks)
        if (b.N < b.M && b.N <= 8 && b.K <= p->direct_kmax && strided_operands(p, b)) flip_orientation(b);
    // index pool
    std::vector<int32_t> pool;
    auto put = [&](const std::vector<int32_t> &v) {
        int64_t off = (int64_t)pool.size();
        pool.insert(pool.end(), v.begin(), v.end());
        while (pool.size() % 64) pool.push_back(0);
        return off;
    };
    int64_t nb_max = 0;
    double flops = 0, payload = 0;
    // range overlap detection among result blocks of the plan
    std::vector<std::pair<std::pair<int64_t, int64_t>, size_t>> ranges;
    for (size_t i = 0; i < p->blocks.size(); ++i) {
        auto &b = p->blocks[i];
        for (HostOperand *o : {&b.A, &b.B, &b.C}) {
            o->off_first = put(o->r.first);
            o->off_second = o->r.table ? -1 : put(o->r.second);
        }
        nb_max = std::max<int64_t>(nb_max, std::max(b.buf1, std::max(b.buf2, b.buf12)) + 1);
        flops += 2.0 * (double)b.M * (double)b.N * (double)b.K;
        payload += (double)p->elem * ((double)b.A.r.maxaddr + 1 + (double)b.B.r.maxaddr + 1 + 2.0 * (double)b.len12);
        ranges.push_back({{(int64_t)b.buf12 * ((int64_t)1 << 40) + b.off12,
                           (int64_t)b.buf12 * ((int64_t)1 << 40) + b.off12 + b.len12}, i});
    }
    for (auto &y : p->syms) nb_max = std::max<int64_t>(nb_max, y.buf + 1);
    for (auto &j : p->inits) nb_max = std::max<int64_t>(nb_max, std::max(j.dst, j.src) + 1);
    for (auto &tr : p->traces) nb_max = std::max<int64_t>(nb_max, std::max(tr.buf_src, tr.buf_dst) + 1);
    for (auto &sp : p->spins) nb_max = std::max<int64_t>(nb_max, sp.buf + 1);
    // Blocks whose operands A and B are strided views (affine index maps)
    // run on the strided-GEMM kernel with the tile variant p->affine_variant;
    // its epilogue scatters through any result map.  A block with gathered
    // operands joins them when min(M, N) >= p->materialize_min: each such
    // operand is first written densely into scratch by the gather launch
    // (one read and one write of the operand, instead of one gathered read
    // per result tile).  The remaining blocks use the 64x64 gather kernel.
    std::vector<uint8_t> is_aff(p->blocks.size(), 0);
    for (size_t i = 0; i < p->blocks.size(); ++i) {
        auto &b = p->blocks[i];
        bool sa, sb;
        b.strided = strided_operands(p, b, &sa, &sb);
        b.mat_a = b.strided && !sa;
        b.mat_b = b.strided && !sb;
        // short contractions with a thin result side (or a very short
        // contraction) run as direct tasks of the strided-GEMM launch
        // (direct_task in contract_kernels.hip)
        b.direct = b.strided && b.K <= p->direct_kmax &&
                   (std::min(b.M, b.N) <= 32 || b.K <= 8);
        is_aff[i] = b.strided && !b.direct ? 1 : 0;
        // results with a short side of at most skinny_max: 16 x 16 tiles
        // whose contraction range is split between the waves (skinny_task)
        b.skinny = is_aff[i] && std::min(b.M, b.N) <= p->skinny_max;
    }
    // Chains: strided-GEMM blocks that add into the same result block
    // through the same result map -- e.g. the IRREP/Ms distributions of the
    // contraction indices of a ring term, each a kernel call of the
    // reference with its own operand blocks -- are linked (DevBlock::next).
    // One task runs the GEMM over their concatenated contraction ranges and
    // writes its result tile once, instead of one read-modify-write (or
    // atomic) pass per block.
    std::vector<int32_t> chain_next(p->blocks.size(), -1);
    std::vector<size_t> chain_head(p->blocks.size());
    {
        std::unordered_map<uint64_t, std::vector<size_t>> heads;
        std::vector<size_t> tail(p->blocks.size());
        auto same_result = [&](const HostBlock &x, const HostBlock &y) {
            return x.buf12 == y.buf12 && x.off12 == y.off12 && x.M == y.M && x.N == y.N &&
                   x.group == y.group &&
                   x.xfac == y.xfac && x.injective == y.injective && x.C.r.table == y.C.r.table &&
                   x.C.r.first == y.C.r.first && x.C.r.second == y.C.r.second;
        };
        p->nchained = 0;
        for (size_t i = 0; i < p->blocks.size(); ++i) {
            chain_head[i] = i;
            tail[i] = i;
            if (!is_aff[i] || !p->chain || p->screen) continue;   // screening skips single calls
            const HostBlock &b = p->blocks[i];
            uint64_t h = 1469598103934665603ull;
            auto mix = [&](uint64_t v) { h = (h ^ v) * 1099511628211ull; };
            mix((uint64_t)b.buf12); mix((uint64_t)b.off12); mix((uint64_t)b.M); mix((uint64_t)b.N);
            uint64_t xb;
            std::memcpy(&xb, &b.xfac, 8);
            mix(xb);
            mix((uint64_t)b.C.r.table);
            for (auto v : b.C.r.first) mix((uint64_t)(uint32_t)v);
            for (auto v : b.C.r.second) mix((uint64_t)(uint32_t)v);
            auto &cand = heads[h];
            bool linked = false;
            for (size_t c : cand)
                if (same_result(p->blocks[c], b)) {
                    chain_head[i] = c;
                    chain_next[tail[c]] = (int32_t)i;
                    tail[c] = i;
                    linked = true;
                    ++p->nchained;
                    break;
                }
            if (!linked) cand.push_back(i);
        }
    }
    // range overlap among result blocks: a chain counts once
    std::vector<uint8_t> overlap(p->blocks.size(), 0);
    {
        std::vector<std::pair<std::pair<int64_t, int64_t>, size_t>> hr;
        for (auto &r : ranges)
            if (chain_head[r.second] == r.second) hr.push_back(r);
        std::sort(hr.begin(), hr.end());
        int64_t maxend = -1;
        size_t maxend_i = 0;
        for (size_t r = 0; r < hr.size(); ++r) {
            if (r > 0 && hr[r].first.first < maxend) {
                overlap[hr[r].second] = 1;
                overlap[maxend_i] = 1;
            }
            if (hr[r].first.second > maxend) {
                maxend = hr[r].first.second;
                maxend_i = hr[r].second;
            }
        }
        for (size_t i = 0; i < p->blocks.size(); ++i) overlap[i] = overlap[chain_head[i]];
    }
    int abm = 64, abn = 64, abk = 32;
    // float32 64 x 64 tiles with the waves stacked along M were as fast or
    // faster on every CCSD level (the 100-row ladder blocks: 3.5%)
    if (p->affine_variant < 0) p->affine_variant = p->elem == 4 ? 9 : 1;
    if (gecco_affine_variant_shape(p->affine_variant, &abm, &abn, &abk))
        return fail(GECCO_HIP_EINVAL, "unknown tile variant %d", p->affine_variant);
    if (p->affine_variant == 8 && p->elem != 4)
        return fail(GECCO_HIP_EINVAL, "split-bfloat16 tiles (variant 8) need a float32 plan");
    // class of a task's block: 0 strided-GEMM launch, 3 gather kernel; tile
    // shape per block
    std::vector<int> cls(p->blocks.size());
    std::vector<int64_t> tm(p->blocks.size()), tn(p->blocks.size());
    // direct tasks: columns dn | row groups ng << 9 | rows per thread << 18
    std::vector<int32_t> dpack(p->blocks.size(), 0);
    for (size_t i = 0; i < p->blocks.size(); ++i) {
        auto &b = p->blocks[i];
        if (b.direct) {
            cls[i] = 0;
            // columns per task: all of N up to 256 threads, at most 64 when
            // the result map is not injective (colliding elements then tend
            // to share a task whose result range fits the LDS stage); as
            // many row groups as fit the 256 threads and the K x rows A tile
            // of 1024 doubles, up to 8 rows per thread
            const int64_t dn = std::min<int64_t>(b.N, b.injective ? 256 : 64);
            const int64_t ng = std::min<int64_t>(256 / dn, std::max<int64_t>(1, 1024 / b.K));
            const int64_t rpt = std::min<int64_t>(8, 1024 / (b.K * ng));
            tn[i] = dn;
            tm[i] = rpt * ng;
            dpack[i] = (int32_t)(dn | (ng << 9) | (rpt << 18));
        } else if (b.skinny) {
            cls[i] = 0; tm[i] = 16; tn[i] = 16;
        } else if (is_aff[i]) {
            cls[i] = 0; tm[i] = abm; tn[i] = abn;
        } else {
            cls[i] = 3; tm[i] = 64; tn[i] = 64;
        }
    }
    auto tile_m = [&](size_t i) { return tm[i]; };
    auto tile_n = [&](size_t i) { return tn[i]; };
    // split K so that the launch has about p->target_tasks tasks
    int64_t ntiles = 0;
    for (size_t i = 0; i < p->blocks.size(); ++i) {
        auto &b = p->blocks[i];
        if (chain_head[i] != i) continue;
        ntiles += ((b.M + tile_m(i) - 1) / tile_m(i)) * ((b.N + tile_n(i) - 1) / tile_n(i));
    }
    const int64_t want = p->target_tasks;
    int64_t split = ntiles > 0 ? (want + ntiles - 1) / ntiles : 1;
    std::vector<DevTask> tasks;
    std::vector<DevBlock> dblocks(p->blocks.size());
    std::vector<int64_t> nsplit(p->blocks.size(), 1);
    std::vector<int64_t> patch;     // patch id of every task
    int64_t npatch = 0;
    for (size_t i = 0; i < p->blocks.size(); ++i) {
        auto &b = p->blocks[i];
        const bool head = chain_head[i] == i;
        // contraction length of the chain
        int64_t Kc = 0;
        for (int64_t j = (int64_t)i; head && j >= 0; j = chain_next[(size_t)j]) Kc += p->blocks[(size_t)j].K;
        // parallel split for the whole plan, and a bound on the serial K
        // extent of one task (rank-reducing blocks, e.g. full contractions);
        // direct blocks are never split
        int64_t ns = 1, chunk = Kc;
        if (!head) {
            ns = nsplit[chain_head[i]];
        } else if (Kc == 0) {
            ns = 0;                       // empty contraction range: no tasks
        } else if (b.skinny) {
            ns = (Kc + p->skinny_kchunk - 1) / p->skinny_kchunk;
            chunk = (Kc + ns - 1) / ns;
        } else if (!b.direct) {
            ns = std::min<int64_t>(split, std::max<int64_t>(1, Kc / p->min_kchunk));
            ns = std::max<int64_t>(ns, (Kc + p->max_kchunk - 1) / p->max_kchunk);
            chunk = (Kc + ns - 1) / ns;
            chunk = ((chunk + 31) / 32) * 32;
            ns = (Kc + chunk - 1) / chunk;
        }
        nsplit[i] = ns;
        DevBlock &d = dblocks[i];
        d.xfac = b.xfac;
        d.base_a = b.off1;
        d.base_b = b.off2;
        d.base_c = b.off12;
        d.buf_a = b.buf1;
        d.buf_b = b.buf2;
        d.buf_c = b.buf12;
        d.M = (int32_t)b.M;
        d.N = (int32_t)b.N;
        d.K = (int32_t)b.K;
        d.atomic = (ns > 1 || !b.injective || overlap[i]) ? 1 : 0;
        d.next = b.next = chain_next[i];
        d.zero_a = d.zero_b = -1;
        if (!head) continue;
        // patches of PM x PN tiles with one contraction range share their
        // A panels (PN times) and B panels (PM times); about 16 tiles each
        const int64_t mt = (b.M + tile_m(i) - 1) / tile_m(i), nt = (b.N + tile_n(i) - 1) / tile_n(i);
        const int64_t PN = std::min<int64_t>(nt, 4), PM = std::min<int64_t>(mt, std::max<int64_t>(1, 16 / PN));
        const int64_t cm = (mt + PM - 1) / PM, cn = (nt + PN - 1) / PN;
        // a task's range [k0, k1) counts from the start of the chain block
        // it begins in
        for (int64_t m0 = 0; m0 < b.M; m0 += tile_m(i))
            for (int64_t n0 = 0; n0 < b.N; n0 += tile_n(i))
                for (int64_t s = 0; s < ns; ++s) {
                    const int64_t c0 = s * chunk, c1 = std::min(Kc, c0 + chunk);
                    int64_t j = (int64_t)i, start = 0;
                    while (start + p->blocks[(size_t)j].K <= c0) {
                        start += p->blocks[(size_t)j].K;
                        j = chain_next[(size_t)j];
                    }
                    tasks.push_back({(int32_t)j, (int32_t)m0, (int32_t)n0, (int32_t)(c0 - start),
                                     (int32_t)(c1 - start)});
                    patch.push_back(npatch + (s * cm + m0 / tile_m(i) / PM) * cn + n0 / tile_n(i) / PN);
                }
        npatch += ns * cm * cn;
    }
    // Result range of every task.  Tasks of the direct kernel whose result
    // map is a table or not injective stage their tile in LDS when the range
    // fits (lds_cap doubles) and is at least 1/4 dense; a task owns its range
    // when no other task's range intersects it.
    {
        const int64_t lds_cap_direct = gecco_direct_stage_cap();
        struct Iv { int64_t lo, hi; size_t t; };
        std::vector<Iv> ivs;
        ivs.reserve(tasks.size());
        for (size_t t = 0; t < tasks.size(); ++t) {
            DevTask &task = tasks[t];
            const size_t i = (size_t)task.blk;
            const HostBlock &b = p->blocks[i];
            task.span = 0;
            task.owned = 0;
            // direct tasks: columns per task; strided-GEMM launch: 1 for a
            // 16 x 16 wave-split-K task
            task.pad = b.direct ? dpack[i] : b.skinny ? 1 : 0;
            task.grp = b.group;
            task.pad2 = (int32_t)patch[t];   // until the launch order is set
            task.lo = 0;
            const int64_t m1 = std::min<int64_t>(b.M, task.m0 + tile_m(i));
            const int64_t n1 = std::min<int64_t>(b.N, task.n0 + tile_n(i));
            int64_t amin, amax;
            const int64_t cnt = tile_range(b.C.r, task.m0, m1, task.n0, n1, amin, amax);
            if (!cnt) continue;
            const int64_t key = (int64_t)b.buf12 << 44;
            ivs.push_back({key + b.off12 + amin, key + b.off12 + amax, t});
            const int64_t span = amax - amin + 1;
            if (b.direct && (b.C.r.table || !b.injective) && span <= lds_cap_direct &&
                4 * cnt >= span) {
                task.span = (int32_t)span;
                task.lo = amin;
            }
        }
        std::sort(ivs.begin(), ivs.end(), [](const Iv &x, const Iv &y) { return x.lo < y.lo; });
        for (size_t c0 = 0; c0 < ivs.size();) {
            size_t c1 = c0 + 1;
            int64_t hi = ivs[c0].hi;
            while (c1 < ivs.size() && ivs[c1].lo <= hi) hi = std::max(hi, ivs[c1++].hi);
            if (c1 == c0 + 1) tasks[ivs[c0].t].owned = 1;
            c0 = c1;
        }
        p->nstaged = 0;
        for (auto &t : tasks) p->nstaged += t.span > 0;
    }
    // by class, then longest first (a wave-split-K task covers 8 times the
    // contraction length of a tile task in about the same time)
    auto cost = [&](const DevTask &x) {
        return (int64_t)(x.k1 - x.k0) / (p->blocks[(size_t)x.blk].skinny ? 8 : 1);
    };
    std::stable_sort(tasks.begin(), tasks.end(), [&](const DevTask &x, const DevTask &y) {
        if (cls[x.blk] != cls[y.blk]) return cls[x.blk] < cls[y.blk];
        return cost(x) > cost(y);
    });
    int64_t ncls[4] = {0, 0, 0, 0};
    for (auto &t : tasks) ncls[cls[t.blk]]++;
    // XCD-grouped order of the strided-GEMM launch (one workgroup per task,
    // dispatched round-robin over the 8 XCDs, each with its own L2): the
    // patches go, longest task first, to the XCD with the least work so far
    // (cost = contraction length), each XCD keeps its patches in that order
    // with a patch's tasks together, and the launch takes the XCDs' lists in
    // turn so that workgroup w gets the next task of XCD w mod 8.
    if (p->xcd_order && ncls[0] > 0) {
        const size_t n0 = (size_t)ncls[0];
        std::vector<int32_t> pid(n0);
        std::vector<int64_t> tc(n0);
        for (size_t t = 0; t < n0; ++t) { pid[t] = tasks[t].pad2; tc[t] = cost(tasks[t]); }
        std::vector<int64_t> perm(n0);
        xcd_task_order((int64_t)n0, pid.data(), tc.data(), perm.data());
        std::vector<DevTask> ord(n0);
        for (size_t w = 0; w < n0; ++w) ord[w] = tasks[(size_t)perm[w]];
        std::copy(ord.begin(), ord.end(), tasks.begin());
    }
    for (auto &t : tasks) t.pad2 = 0;
    if (p->deterministic) {
        // every task of the strided-GEMM launch gets its own range of d_raw
        // (tile rows x tile columns, row-major); each result element's
        // contributions are summed in task order by one CSR row
        if (ncls[3] > 0)
            return fail(GECCO_HIP_EINVAL, "fixed-order plans need every kernel call on the strided-GEMM launch "
                                          "(materialised operands)");
        struct Contrib { int64_t dst, src; int8_t sgn; };
        std::map<int32_t, std::vector<Contrib>> per_slot;
        int64_t off = 0;
        for (size_t t = 0; t < (size_t)ncls[0]; ++t) {
            DevTask &task = tasks[t];
            const size_t i = (size_t)task.blk;
            const HostBlock &b = p->blocks[i];
            const int64_t tw = tile_n(i), th = tile_m(i);
            task.span = 0;
            task.owned = 0;
            task.lo = off;
            const Resolved &c = b.C.r;
            std::vector<Contrib> &v = per_slot[b.buf12];
            const int64_t m1 = std::min<int64_t>(b.M, task.m0 + th), n1 = std::min<int64_t>(b.N, task.n0 + tw);
            for (int64_t m = task.m0; m < m1; ++m)
                for (int64_t n = task.n0; n < n1; ++n) {
                    int64_t a;
                    int sg = 1;
                    if (c.table) {
                        const int32_t e = c.first[m + c.F * n];
                        if (!e) continue;
                        sg = e < 0 ? -1 : 1;
                        a = absv(e) - 1;
                    } else {
                        const int32_t e1 = c.first[m], e2 = c.second[n];
                        if (!e1 || !e2) continue;
                        sg = (e1 < 0) != (e2 < 0) ? -1 : 1;
                        a = absv(e1) + absv(e2) - 2;
                    }
                    v.push_back({b.off12 + a, off + (m - task.m0) * tw + (n - task.n0), (int8_t)sg});
                }
            off += th * tw;
        }
        p->raw_elems = off;
        for (auto &kv : per_slot) {
            std::vector<Contrib> &v = kv.second;
            std::stable_sort(v.begin(), v.end(), [](const Contrib &x, const Contrib &y) { return x.dst < y.dst; });
            TraceStep tr;
            tr.buf_src = -1;
            tr.buf_dst = kv.first;
            tr.rowptr.push_back(0);
            for (size_t k = 0; k < v.size(); ++k) {
                if (k == 0 || v[k].dst != v[k - 1].dst) {
                    if (k) tr.rowptr.push_back((int64_t)k);
                    tr.rowdst.push_back(v[k].dst);
                }
                tr.col.push_back(v[k].src);
                tr.sgn.push_back(v[k].sgn);
            }
            tr.rowptr.push_back((int64_t)v.size());
            tr.nrows = (int64_t)tr.rowdst.size();
            tr.nnz = (int64_t)v.size();
            if (!tr.nrows) continue;
            std::vector<Contrib>().swap(v);
            p->dtraces.push_back(std::move(tr));
        }
        if (dmalloc(p, &p->d_raw, (size_t)std::max<int64_t>(1, off) * p->elem, "hipMalloc(raw)")) return g_err;
        for (auto &tr : p->dtraces)
            if (dmalloc(p, (void **)&tr.d_rowptr, (tr.nrows + 1) * 8, "hipMalloc(fixed-order sums)") ||
                dmalloc(p, (void **)&tr.d_rowdst, tr.nrows * 8, "hipMalloc(fixed-order sums)") ||
                dmalloc(p, (void **)&tr.d_col, tr.nnz * 8, "hipMalloc(fixed-order sums)") ||
                dmalloc(p, (void **)&tr.d_sgn, tr.nnz, "hipMalloc(fixed-order sums)") ||
                check(hipMemcpy(tr.d_rowptr, tr.rowptr.data(), (tr.nrows + 1) * 8, hipMemcpyHostToDevice), "hipMemcpy(fixed-order sums)") ||
                check(hipMemcpy(tr.d_rowdst, tr.rowdst.data(), tr.nrows * 8, hipMemcpyHostToDevice), "hipMemcpy(fixed-order sums)") ||
                check(hipMemcpy(tr.d_col, tr.col.data(), tr.nnz * 8, hipMemcpyHostToDevice), "hipMemcpy(fixed-order sums)") ||
                check(hipMemcpy(tr.d_sgn, tr.sgn.data(), tr.nnz, hipMemcpyHostToDevice), "hipMemcpy(fixed-order sums)"))
                return g_err;
        for (auto &tr : p->dtraces) {
            std::vector<int64_t>().swap(tr.rowptr);
            std::vector<int64_t>().swap(tr.rowdst);
            std::vector<int64_t>().swap(tr.col);
            std::vector<int8_t>().swap(tr.sgn);
        }
    }
    p->flops_affine = p->flops_general = p->flops_direct = 0;
    for (size_t i = 0; i < p->blocks.size(); ++i) {
        const double f = 2.0 * (double)p->blocks[i].M * (double)p->blocks[i].N * (double)p->blocks[i].K;
        (p->blocks[i].direct ? p->flops_direct : cls[i] == 3 ? p->flops_general : p->flops_affine) += f;
    }
    // direct blocks read a tabled result map n-fastest
    for (size_t i = 0; i < p->blocks.size(); ++i) {
        auto &b = p->blocks[i];
        if (!b.direct || !b.C.r.table) continue;
        const Resolved &c = b.C.r;
        std::vector<int32_t> tt((size_t)(c.F * c.S));
        for (int64_t n = 0; n < c.S; ++n)
            for (int64_t m = 0; m < c.F; ++m) tt[n + c.S * m] = c.first[m + c.F * n];
        b.C.off_first_t = put(tt);
    }
    if (pool.empty()) pool.push_back(0);
    int32_t *d_index = nullptr;
    if (dmalloc(p, (void **)&d_index, pool.size() * sizeof(int32_t), "hipMalloc(index)")) return g_err;
    if (check(hipMemcpy(d_index, pool.data(), pool.size() * sizeof(int32_t), hipMemcpyHostToDevice), "hipMemcpy(index)")) { dfree(p, d_index); return g_err; }
    for (size_t i = 0; i < p->blocks.size(); ++i) {
        auto &b = p->blocks[i];
        DevBlock &d = dblocks[i];
        HostOperand *hs[3] = {&b.A, &b.B, &b.C};
        DevOperand *ds[3] = {&d.A, &d.B, &d.C};
        for (int q = 0; q < 3; ++q) {
            ds[q]->first = d_index + hs[q]->off_first;
            if (q == 2 && b.direct && hs[q]->r.table) ds[q]->first = d_index + hs[q]->off_first_t;
            ds[q]->second = hs[q]->r.table ? nullptr : d_index + hs[q]->off_second;
            ds[q]->ld = hs[q]->r.F;
            ds[q]->table = hs[q]->r.table ? 1 : 0;
            if (q == 2 && b.direct && hs[q]->r.table) { ds[q]->ld = hs[q]->r.S; ds[q]->table = 2; }
            ds[q]->fast_first = hs[q]->fast_first;
            ds[q]->first_affine = ds[q]->second_affine = 0;
            ds[q]->f0 = ds[q]->f1 = ds[q]->s0 = ds[q]->s1 = 0;
            if (!hs[q]->r.table) {
                int64_t a0, a1;
                if (affine_of(hs[q]->r.first, a0, a1)) { ds[q]->first_affine = 1; ds[q]->f0 = a0; ds[q]->f1 = a1; }
                if (affine_of(hs[q]->r.second, a0, a1)) { ds[q]->second_affine = 1; ds[q]->s0 = a0; ds[q]->s1 = a1; }
            }
        }
    }
    p->d_index = d_index;
    // materialisation jobs; identical gathers of the same list share one job
    {
        std::vector<DevGather> jobs;
        std::vector<const HostOperand *> job_src;
        std::vector<int32_t> job_buf, job_grp;
        std::vector<int64_t> job_base;
        std::unordered_map<uint64_t, std::vector<size_t>> seen;
        const int32_t slot = (int32_t)nb_max;
        int64_t next = 0;
        for (size_t i = 0; i < p->blocks.size(); ++i) {
            auto &b = p->blocks[i];
            DevBlock &d = dblocks[i];
            for (int q = 0; q < 2; ++q) {
                if (!(q == 0 ? b.mat_a : b.mat_b)) continue;
                const HostOperand &ho = q == 0 ? b.A : b.B;
                DevOperand &dop = q == 0 ? d.A : d.B;
                const int32_t buf = q == 0 ? b.buf1 : b.buf2;
                const int64_t base = q == 0 ? b.off1 : b.off2;
                const int64_t K = b.K, X = q == 0 ? b.M : b.N;
                const int kfast = ho.fast_first;
                uint64_t h = 1469598103934665603ull;
                auto mix = [&](uint64_t v) { h = (h ^ v) * 1099511628211ull; };
                mix((uint64_t)buf); mix((uint64_t)base); mix((uint64_t)K); mix((uint64_t)X);
                mix((uint64_t)ho.r.table); mix((uint64_t)kfast);
                for (auto v : ho.r.first) mix((uint64_t)(uint32_t)v);
                for (auto v : ho.r.second) mix((uint64_t)(uint32_t)v);
                size_t j = jobs.size();
                for (size_t c : seen[h]) {
                    const HostOperand *o2 = job_src[c];
                    if (job_buf[c] == buf && job_base[c] == base && job_grp[c] == b.group && jobs[c].K == K &&
                        jobs[c].X == X && jobs[c].kfast == kfast && o2->r.table == ho.r.table &&
                        o2->r.first == ho.r.first && o2->r.second == ho.r.second) {
                        j = c;
                        break;
                    }
                }
                if (j == jobs.size()) {
                    DevGather g;
                    std::memset(&g, 0, sizeof g);
                    g.src = dop;
                    g.src_base = base;
                    g.src_buf = buf;
                    g.dst_base = next;
                    g.K = (int32_t)K;
                    g.X = (int32_t)X;
                    g.kfast = kfast;
                    next += ((K * X + 63) / 64) * 64;
                    jobs.push_back(g);
                    job_src.push_back(&ho);
                    job_buf.push_back(buf);
                    job_grp.push_back(b.group);
                    job_base.push_back(base);
                    seen[h].push_back(j);
                }
                DevOperand m;
                std::memset(&m, 0, sizeof m);
                m.ld = K;
                m.fast_first = kfast;
                m.first_affine = m.second_affine = 1;
                m.f1 = kfast ? 1 : X;
                m.s1 = kfast ? K : 1;
                dop = m;
                (q == 0 ? d.buf_a : d.buf_b) = slot;
                (q == 0 ? d.base_a : d.base_b) = jobs[j].dst_base;
            }
        }
        std::vector<DevGatherTask> gt;
        const int64_t chunk = 2048;   // elements per workgroup (8 per thread)
        double gbytes = 0;
        auto is_const = [&](int32_t slot) {
            return slot >= 0 && slot < (int32_t)p->constant_slot.size() && p->constant_slot[slot];
        };
        // gathers of constant slots first: they run only when needed (and
        // are never skipped: group -1)
        for (int pass = 0; pass < 2; ++pass)
            for (size_t j = 0; j < jobs.size(); ++j) {
                if (is_const(jobs[j].src_buf) != (pass == 0)) continue;
                const int64_t n = (int64_t)jobs[j].K * jobs[j].X;
                for (int64_t e0 = 0; e0 < n; e0 += chunk)
                    gt.push_back({(int32_t)j, (int32_t)e0, (int32_t)std::min(n, e0 + chunk),
                                  pass == 0 ? -1 : job_grp[j]});
                if (pass == 1) gbytes += 2.0 * p->elem * (double)n;
                if (pass == 0) p->ngtasks_static = (int64_t)gt.size();
            }
        p->ngjobs = (int64_t)jobs.size();
        p->ngtasks = (int64_t)gt.size();
        p->scratch_elems = next;
        p->gather_bytes = gbytes;
        if (!jobs.empty()) {
            if (dmalloc(p, (void **)&p->d_scratch, (size_t)next * p->elem, "hipMalloc(scratch)") ||
                dmalloc(p, (void **)&p->d_gjobs, jobs.size() * sizeof(DevGather), "hipMalloc(gather jobs)") ||
                check(hipMemcpy(p->d_gjobs, jobs.data(), jobs.size() * sizeof(DevGather), hipMemcpyHostToDevice), "hipMemcpy(gather jobs)") ||
                dmalloc(p, (void **)&p->d_gtasks, gt.size() * sizeof(DevGatherTask), "hipMalloc(gather tasks)") ||
                check(hipMemcpy(p->d_gtasks, gt.data(), gt.size() * sizeof(DevGatherTask), hipMemcpyHostToDevice), "hipMemcpy(gather tasks)"))
                return g_err;
        }
    }
    // norm screening: one job per distinct operand block (slot, offset,
    // length), summed in chunks of NCHUNK elements
    if (p->screen) {
        constexpr int64_t NCHUNK = 1 << 16;
        std::map<std::tuple<int32_t, int64_t, int64_t>, int32_t> jobs;
        std::vector<DevNormTask> nt;
        std::vector<int64_t> rows(1, 0);
        auto job = [&](int32_t buf, int64_t off, int64_t len) -> int32_t {
            auto key = std::make_tuple(buf, off, len);
            auto it = jobs.find(key);
            if (it != jobs.end()) return it->second;
            const int32_t j = (int32_t)jobs.size();
            jobs[key] = j;
            for (int64_t e0 = 0; e0 < len; e0 += NCHUNK) {
                DevNormTask t;
                t.off = off;
                t.e0 = e0;
                t.e1 = std::min(len, e0 + NCHUNK);
                t.buf = buf;
                t.pad = 0;
                nt.push_back(t);
            }
            rows.push_back((int64_t)nt.size());
            return j;
        };
        for (size_t i = 0; i < p->blocks.size(); ++i) {
            const HostBlock &b = p->blocks[i];
            dblocks[i].zero_a = job(b.buf1, b.off1, b.len1);
            dblocks[i].zero_b = job(b.buf2, b.off2, b.len2);
        }
        p->nnorm_tasks = (int64_t)nt.size();
        p->nnorm_jobs = (int64_t)jobs.size();
        if (p->nnorm_jobs) {
            if (dmalloc(p, (void **)&p->d_ntasks, std::max<size_t>(1, nt.size()) * sizeof(DevNormTask), "hipMalloc(norm tasks)") ||
                dmalloc(p, (void **)&p->d_npartial, std::max<size_t>(1, nt.size()) * sizeof(double), "hipMalloc(norm partials)") ||
                dmalloc(p, (void **)&p->d_nrow, rows.size() * sizeof(int64_t), "hipMalloc(norm rows)") ||
                dmalloc(p, (void **)&p->d_nzero, (size_t)p->nnorm_jobs * sizeof(int32_t), "hipMalloc(norm flags)"))
                return g_err;
            if ((!nt.empty() && check(hipMemcpy(p->d_ntasks, nt.data(), nt.size() * sizeof(DevNormTask), hipMemcpyHostToDevice), "hipMemcpy(norm tasks)")) ||
                check(hipMemcpy(p->d_nrow, rows.data(), rows.size() * sizeof(int64_t), hipMemcpyHostToDevice), "hipMemcpy(norm rows)"))
                return g_err;
        }
    }
    if (dmalloc(p, (void **)&p->d_blocks, std::max<size_t>(1, dblocks.size()) * sizeof(DevBlock), "hipMalloc(blocks)")) return g_err;
    if (!dblocks.empty() && check(hipMemcpy(p->d_blocks, dblocks.data(), dblocks.size() * sizeof(DevBlock), hipMemcpyHostToDevice), "hipMemcpy(blocks)")) return g_err;
    if (dmalloc(p, (void **)&p->d_tasks, std::max<size_t>(1, tasks.size()) * sizeof(DevTask), "hipMalloc(tasks)")) return g_err;
    if (!tasks.empty() && check(hipMemcpy(p->d_tasks, tasks.data(), tasks.size() * sizeof(DevTask), hipMemcpyHostToDevice), "hipMemcpy(tasks)")) return g_err;
    p->nbufs_max = nb_max;
    if (dmalloc(p, (void **)&p->d_bufs, (nb_max + 2) * sizeof(void *), "hipMalloc(bufs)")) return g_err;
    p->ntasks = (int64_t)tasks.size();
    for (auto &sp : p->spins) {
        if (!sp.ngroups) continue;
        if (dmalloc(p, (void **)&sp.d_gptr, sp.gptr.size() * 8, "hipMalloc(spin)") ||
            dmalloc(p, (void **)&sp.d_elem, std::max<size_t>(1, sp.elem.size()) * 8, "hipMalloc(spin)") ||
            dmalloc(p, (void **)&sp.d_gmat, sp.gmat.size() * 8, "hipMalloc(spin)") ||
            dmalloc(p, (void **)&sp.d_mats, std::max<size_t>(1, sp.mats.size()) * 8, "hipMalloc(spin)") ||
            check(hipMemcpy(sp.d_gptr, sp.gptr.data(), sp.gptr.size() * 8, hipMemcpyHostToDevice), "hipMemcpy(spin)") ||
            check(hipMemcpy(sp.d_elem, sp.elem.data(), sp.elem.size() * 8, hipMemcpyHostToDevice), "hipMemcpy(spin)") ||
            check(hipMemcpy(sp.d_gmat, sp.gmat.data(), sp.gmat.size() * 8, hipMemcpyHostToDevice), "hipMemcpy(spin)") ||
            check(hipMemcpy(sp.d_mats, sp.mats.data(), sp.mats.size() * 8, hipMemcpyHostToDevice), "hipMemcpy(spin)"))
            return g_err;
        if (!p->keep_host) {
            std::vector<int64_t>().swap(sp.gptr);
            std::vector<int64_t>().swap(sp.elem);
            std::vector<int64_t>().swap(sp.gmat);
            std::vector<double>().swap(sp.mats);
        }
    }
    for (auto &tr : p->traces) {
        if (!tr.nrows) continue;
        if (dmalloc(p, (void **)&tr.d_rowptr, (tr.nrows + 1) * 8, "hipMalloc(trace)") ||
            dmalloc(p, (void **)&tr.d_rowdst, tr.nrows * 8, "hipMalloc(trace)") ||
            dmalloc(p, (void **)&tr.d_col, tr.nnz * 8, "hipMalloc(trace)") ||
            dmalloc(p, (void **)&tr.d_sgn, tr.nnz, "hipMalloc(trace)") ||
            check(hipMemcpy(tr.d_rowptr, tr.rowptr.data(), (tr.nrows + 1) * 8, hipMemcpyHostToDevice), "hipMemcpy(trace)") ||
            check(hipMemcpy(tr.d_rowdst, tr.rowdst.data(), tr.nrows * 8, hipMemcpyHostToDevice), "hipMemcpy(trace)") ||
            check(hipMemcpy(tr.d_col, tr.col.data(), tr.nnz * 8, hipMemcpyHostToDevice), "hipMemcpy(trace)") ||
            check(hipMemcpy(tr.d_sgn, tr.sgn.data(), tr.nnz, hipMemcpyHostToDevice), "hipMemcpy(trace)"))
            return g_err;
        if (!p->keep_host) {
            std::vector<int64_t>().swap(tr.rowptr);
            std::vector<int64_t>().swap(tr.rowdst);
            std::vector<int64_t>().swap(tr.col);
            std::vector<int8_t>().swap(tr.sgn);
        }
    }
    // list initialisations, and in fixed-order mode the zeroing of d_raw
    // (slot nbufs_max + 1): tasks that are skipped or screened leave their
    // range at zero
    std::vector<DevInit> inits_all(p->inits);
    std::vector<int32_t> igroup_all(p->init_group);
    if (p->deterministic && p->raw_elems) {
        DevInit z;
        std::memset(&z, 0, sizeof z);
        z.dst = (int32_t)nb_max + 1;
        z.src = -1;
        z.len = p->raw_elems;
        inits_all.push_back(z);
        igroup_all.push_back(-1);
    }
    if (!inits_all.empty()) {
        std::vector<DevInitTask> it;
        const int64_t chunk = 2048;   // elements per workgroup (8 per thread)
        for (size_t j = 0; j < inits_all.size(); ++j)
            for (int64_t e0 = 0; e0 < inits_all[j].len; e0 += chunk)
                it.push_back({(int32_t)j, igroup_all[j], e0, std::min(inits_all[j].len, e0 + chunk)});
        p->nitasks = (int64_t)it.size();
        if (p->nitasks &&
            (dmalloc(p, (void **)&p->d_inits, inits_all.size() * sizeof(DevInit), "hipMalloc(inits)") ||
             check(hipMemcpy(p->d_inits, inits_all.data(), inits_all.size() * sizeof(DevInit), hipMemcpyHostToDevice), "hipMemcpy(inits)") ||
             dmalloc(p, (void **)&p->d_itasks, it.size() * sizeof(DevInitTask), "hipMalloc(init tasks)") ||
             check(hipMemcpy(p->d_itasks, it.data(), it.size() * sizeof(DevInitTask), hipMemcpyHostToDevice), "hipMemcpy(init tasks)")))
            return g_err;
    }
    p->ntasks_affine = ncls[0];
    {
        std::vector<DevSymStep> steps;
        std::vector<DevSymTask> stasks;
        std::vector<int32_t> a1, a2;
        std::vector<int8_t> ff;
        const int64_t chunk = 1024;   // pairs per workgroup: small lists still fill the GPU
        for (auto &y : p->syms) {
            y.npair = (int64_t)y.i1.size();
            if (!y.npair) continue;
            int64_t base = INT64_MAX, top = 0;
            for (int64_t k = 0; k < y.npair; ++k) {
                base = std::min(base, std::min(y.i1[k], y.i2[k]));
                top = std::max(top, std::max(y.i1[k], y.i2[k]));
            }
            if (top - base >= ((int64_t)1 << 31))
                return fail(GECCO_HIP_EINVAL, "symmetrised list spans more than 2^31 elements");
            DevSymStep st;
            std::memset(&st, 0, sizeof st);
            st.base = base;
            st.pair0 = (int64_t)a1.size();
            st.prefac = y.prefac;
            st.buf = y.buf;
            for (int64_t k = 0; k < y.npair; ++k) {
                a1.push_back((int32_t)(y.i1[k] - base));
                a2.push_back((int32_t)(y.i2[k] - base));
                ff.push_back(y.f[k]);
            }
            for (int64_t p0 = 0; p0 < y.npair; p0 += chunk)
                stasks.push_back({(int32_t)steps.size(), y.group, st.pair0 + p0, st.pair0 + std::min(y.npair, p0 + chunk)});
            steps.push_back(st);
            if (!p->keep_host) {
                std::vector<int64_t>().swap(y.i1);
                std::vector<int64_t>().swap(y.i2);
                std::vector<int8_t>().swap(y.f);
            }
            p->payload += 2.0 * p->elem * (double)y.npair;
            p->index_bytes += 9.0 * (double)y.npair;
        }
        p->nsymtasks = (int64_t)stasks.size();
        if (p->nsymtasks &&
            (dmalloc(p, (void **)&p->d_symsteps, steps.size() * sizeof(DevSymStep), "hipMalloc(sym)") ||
             check(hipMemcpy(p->d_symsteps, steps.data(), steps.size() * sizeof(DevSymStep), hipMemcpyHostToDevice), "hipMemcpy(sym)") ||
             dmalloc(p, (void **)&p->d_symtasks, stasks.size() * sizeof(DevSymTask), "hipMalloc(sym)") ||
             check(hipMemcpy(p->d_symtasks, stasks.data(), stasks.size() * sizeof(DevSymTask), hipMemcpyHostToDevice), "hipMemcpy(sym)") ||
             dmalloc(p, (void **)&p->d_sym_i1, a1.size() * 4, "hipMalloc(sym)") ||
             dmalloc(p, (void **)&p->d_sym_i2, a2.size() * 4, "hipMalloc(sym)") ||
             dmalloc(p, (void **)&p->d_sym_f, ff.size(), "hipMalloc(sym)") ||
             check(hipMemcpy(p->d_sym_i1, a1.data(), a1.size() * 4, hipMemcpyHostToDevice), "hipMemcpy(sym)") ||
             check(hipMemcpy(p->d_sym_i2, a2.data(), a2.size() * 4, hipMemcpyHostToDevice), "hipMemcpy(sym)") ||
             check(hipMemcpy(p->d_sym_f, ff.data(), ff.size(), hipMemcpyHostToDevice), "hipMemcpy(sym)")))
            return g_err;
    }
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) ncu = prop.multiProcessorCount;
    }
    // one workgroup per task: the dispatcher balances uneven task sizes
    p->grid = (int)std::min<int64_t>(p->ntasks, (int64_t)1 << 30);
    (void)ncu;
    p->flops = flops;
    p->payload += payload;
    p->index_bytes += 4.0 * (double)pool.size();
    p->finalized = true;
    // host copies of the index tables are no longer needed
    if (!p->keep_host) plan_release_host(p);
    return ok();
}

static int plan_execute(gecco_hip_plan *p, double *const *bufs, int64_t nbufs, hipStream_t s) {
    if (!p->finalized) return fail(GECCO_HIP_ESTATE, "plan not finalized");
    if (nbufs < p->nbufs_max) return fail(GECCO_HIP_EINVAL, "plan uses %lld buffer slots, %lld given", (long long)p->nbufs_max, (long long)nbufs);
    for (int64_t i = 0; i < p->nbufs_max; ++i)
        if (!bufs[i]) return fail(GECCO_HIP_EINVAL, "buffer slot %lld is null", (long long)i);
    if (p->d_skip && (!p->traces.empty() || !p->spins.empty()))
        return fail(GECCO_HIP_EINVAL, "skip flags are not supported for plans with trace or spin-projection steps");
    if (p->d_skip && p->deterministic)
        return fail(GECCO_HIP_EINVAL, "skip flags are not supported for fixed-order plans");
    if (p->ntasks == 0 && p->syms.empty() && p->nitasks == 0 && p->traces.empty() && p->spins.empty()) return ok();
    std::vector<void *> want(bufs, bufs + p->nbufs_max);
    if (p->d_scratch || p->d_raw) want.push_back(p->d_scratch);   // slot nbufs_max: scratch
    if (p->d_raw) want.push_back(p->d_raw);                         // slot nbufs_max + 1: fixed-order tiles
    if (want != p->last_bufs) {
        // slot table changes only when the caller passes different buffers;
        // the copy is ordered on the same stream before the launch
        if (check(hipMemcpyAsync(p->d_bufs, want.data(), want.size() * sizeof(void *), hipMemcpyHostToDevice, s), "hipMemcpyAsync(bufs)")) return g_err;
        if (check(hipStreamSynchronize(s), "hipStreamSynchronize")) return g_err;
        p->last_bufs = want;
        p->static_done = false;   // constant slots may hold other data now
    }
    const int64_t na = p->ntasks_affine;
    const int64_t ng = p->ntasks - na;
    // the launches' task lists: the full ones, or the live ones
    const bool cp = p->compacted;
    const gecco::DevTask *ta = cp ? p->d_ltasks : p->d_tasks, *tg = (cp ? p->d_ltasks : p->d_tasks) + na;
    const int64_t nla = cp ? p->lt[0] : na, nlg = cp ? p->lt[1] : ng;
    const gecco::DevInitTask *ti = cp ? p->d_litasks : p->d_itasks;
    const int64_t nli = cp ? p->lt[2] : p->nitasks;
    const gecco::DevSymTask *ts = cp ? p->d_lsymtasks : p->d_symtasks;
    const int64_t nls = cp ? p->lt[3] : p->nsymtasks;
    auto mark = [&](int cls, bool begin) -> int {
        if (!p->timing) return 0;
        hipEvent_t e;
        if (check(hipEventCreate(&e), "hipEventCreate") || check(hipEventRecord(e, s), "hipEventRecord")) return g_err;
        if (begin) p->pending.push_back({cls, {e, nullptr}});
        else p->pending.back().second.second = e;
        return 0;
    };
    if (nli &&
        gecco_launch_init(p->d_inits, ti, (int)nli, p->d_bufs, p->d_skip, p->elem, s))
        return fail(GECCO_HIP_EDEVICE, "init launch failed: %s", hipGetErrorString(hipGetLastError()));
    const int32_t *zero = nullptr;
    if (p->screen && p->nnorm_jobs) {
        if (gecco_launch_norms(p->d_ntasks, p->nnorm_tasks, p->d_nrow, p->nnorm_jobs, p->d_bufs,
                               p->d_npartial, p->d_nzero, p->elem, s))
            return fail(GECCO_HIP_EDEVICE, "norm launch failed: %s", hipGetErrorString(hipGetLastError()));
        zero = p->d_nzero;
    }
    const int64_t g0 = p->static_done ? p->ngtasks_static : 0;
    if (cp) {
        // live constant-slot gathers (first execute only), then the others
        const int64_t ns = p->static_done ? 0 : p->lt[4], nd = p->lt[5];
        if (ns + nd) {
            if (mark(3, true)) return g_err;
            if ((ns && gecco_launch_gather(p->d_gjobs, p->d_lgtasks, (int)ns, p->d_bufs, p->d_scratch,
                                           p->d_skip, p->elem, s)) ||
                (nd && gecco_launch_gather(p->d_gjobs, p->d_lgtasks + p->ngtasks_static, (int)nd, p->d_bufs,
                                           p->d_scratch, p->d_skip, p->elem, s)))
                return fail(GECCO_HIP_EDEVICE, "gather launch failed: %s", hipGetErrorString(hipGetLastError()));
            if (mark(3, false)) return g_err;
        }
    } else if (p->ngtasks > g0) {
        if (mark(3, true)) return g_err;
        if (gecco_launch_gather(p->d_gjobs, p->d_gtasks + g0, (int)(p->ngtasks - g0), p->d_bufs,
                                p->d_scratch, p->d_skip, p->elem, s))
            return fail(GECCO_HIP_EDEVICE, "gather launch failed: %s", hipGetErrorString(hipGetLastError()));
        if (mark(3, false)) return g_err;
    }
    p->static_done = true;
    if (nla) {
        if (mark(0, true)) return g_err;
        if (gecco_launch_grouped_affine(p->d_blocks, ta, (int)nla, p->d_bufs, (int)nla,
                                        p->affine_variant, p->d_skip, zero, p->d_raw, p->elem, s))
            return fail(GECCO_HIP_EDEVICE, "kernel launch failed: %s", hipGetErrorString(hipGetLastError()));
        if (mark(0, false)) return g_err;
    }
    if (nlg) {
        if (mark(1, true)) return g_err;
        if (gecco_launch_grouped(p->d_blocks, tg, (int)nlg, p->d_bufs, (int)nlg, p->d_skip,
                                 zero, p->elem, s))
            return fail(GECCO_HIP_EDEVICE, "kernel launch failed: %s", hipGetErrorString(hipGetLastError()));
        if (mark(1, false)) return g_err;
    }
    // fixed-order mode: each result element's tile contributions, in task order
    for (auto &tr : p->dtraces)
        if (tr.nrows && gecco_launch_trace(p->d_raw, (char *)bufs[tr.buf_dst], tr.d_rowptr, tr.d_rowdst, tr.d_col,
                                           tr.d_sgn, tr.nrows, 1.0, p->elem, s))
            return fail(GECCO_HIP_EDEVICE, "fixed-order sum launch failed");
    for (auto &tr : p->traces)
        if (tr.nrows && gecco_launch_trace((const char *)bufs[tr.buf_src] + tr.off_src * p->elem,
                                           (char *)bufs[tr.buf_dst] + tr.off_dst * p->elem,
                                           tr.d_rowptr, tr.d_rowdst, tr.d_col, tr.d_sgn, tr.nrows,
                                           tr.fact, p->elem, s))
            return fail(GECCO_HIP_EDEVICE, "trace launch failed");
    if (nls) {
        if (mark(2, true)) return g_err;
        if (gecco_launch_sym_batched(p->d_symsteps, ts, (int)nls, p->d_bufs,
                                     p->d_sym_i1, p->d_sym_i2, p->d_sym_f, p->d_skip, p->elem, s))
            return fail(GECCO_HIP_EDEVICE, "sym_ab launch failed");
        if (mark(2, false)) return g_err;
    }
    for (auto &sp : p->spins)
        if (sp.ngroups && gecco_launch_spin_project((char *)bufs[sp.buf] + sp.off * p->elem, sp.d_gptr, sp.d_elem,
                                                    sp.d_gmat, sp.d_mats, sp.ngroups, p->elem, s))
            return fail(GECCO_HIP_EDEVICE, "spin projection launch failed");
    return ok();
}

// ------------------------------------------------------------ staging
struct Stage {
    double *p = nullptr;
    size_t n = 0;
};
static Stage g_stage[3];

static int stage_get(int i, size_t n, double **out) {
    if (g_stage[i].n < n) {
        if (g_stage[i].p) (void)hipFree(g_stage[i].p);
        g_stage[i].p = nullptr;
        g_stage[i].n = 0;
        if (check(hipMalloc(&g_stage[i].p, std::max<size_t>(n, 1) * sizeof(double)), "hipMalloc(stage)")) return g_err;
        g_stage[i].n = std::max<size_t>(n, 1);
    }
    *out = g_stage[i].p;
    return ok();
}

static bool is_device_ptr(const void *p) {
    hipPointerAttribute_t attr;
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice;
}

// ------------------------------------------------ cached single-call plans
// GeCCo makes the same kernel calls (maps, lengths, factor) every iteration.
// A one-call plan is kept per distinct call -- keyed by the resolved index
// data of its operands, its dimensions, its factor and the device -- so a
// repeated call does no device allocation and no upload: it only runs.
static uint64_t mix64(uint64_t h, uint64_t x) {
    h ^= x + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
    h ^= h >> 31;
    h *= 0xbf58476d1ce4e5b9ull;
    h ^= h >> 29;
    return h;
}

struct CallKey {
    uint64_t h1 = 0x243f6a8885a308d3ull, h2 = 0x13198a2e03707344ull, n = 0;
    void add(uint64_t x) {
        h1 = mix64(h1, x);
        h2 = mix64(h2 ^ 0x5bd1e995u, x * 0xff51afd7ed558ccdull + n);
        ++n;
    }
    void add(const std::vector<int32_t> &v) {
        add((uint64_t)v.size());
        size_t i = 0;
        for (; i + 1 < v.size(); i += 2) add((uint64_t)(uint32_t)v[i] | ((uint64_t)(uint32_t)v[i + 1] << 32));
        if (i < v.size()) add((uint64_t)(uint32_t)v[i]);
    }
    bool operator==(const CallKey &o) const { return h1 == o.h1 && h2 == o.h2 && n == o.n; }
};

struct CallKeyHash {
    size_t operator()(const CallKey &k) const { return (size_t)(k.h1 ^ (k.h2 * 31)); }
};

static CallKey call_key(const HostBlock &hb, int dev) {
    CallKey k;
    uint64_t xf;
    std::memcpy(&xf, &hb.xfac, 8);
    k.add(xf);
    k.add((uint64_t)dev);
    const int64_t dims[5] = {hb.M, hb.N, hb.K, hb.len12, (int64_t)hb.swap | ((int64_t)hb.injective << 1)};
    for (int64_t d : dims) k.add((uint64_t)d);
    for (const HostOperand *o : {&hb.A, &hb.B, &hb.C}) {
        k.add((uint64_t)o->r.table | ((uint64_t)o->fast_first << 1));
        k.add((uint64_t)o->r.F);
        k.add((uint64_t)o->r.S);
        k.add(o->r.first);
        k.add(o->r.second);
    }
    return k;
}

static std::mutex g_calls_mu;
static std::unordered_map<CallKey, gecco_hip_plan *, CallKeyHash> g_calls;
static double g_calls_bytes = 0;
static const size_t CALLS_MAX = 16384;           // cached plans
static const double CALLS_MAX_BYTES = 1.0e9;     // their device index bytes

static void calls_clear() {
    for (auto &kv : g_calls) {
        plan_free_device(kv.second);
        delete kv.second;
    }
    g_calls.clear();
    g_calls_bytes = 0;
}

// One kernel call through the Fortran signature: stage host operands, run
// the call's cached plan (built and finalized on its first occurrence),
// copy a host result back.  The launch is on the null stream; the copy back
// waits for it, device-resident results are ordered before later work on
// that stream, and nothing else synchronises.
static int single_call(const gecco_hip_block_args &a0, double *xop1op2, const double *xop1,
                       const double *xop2) {
    gecco_hip_block_args a = a0;
    a.buf_op1 = 0;
    a.buf_op2 = 1;
    a.buf_op1op2 = 2;
    a.off_op1 = a.off_op2 = a.off_op1op2 = 0;
    const int64_t len1 = prod(a.lstrop1, a.ncblk_op1 + a.nablk_op1);
    const int64_t len2 = prod(a.lstrop2, a.ncblk_op2 + a.nablk_op2);
    const int64_t len12 = prod(a.lstrop1op2, a.ncblk_op1op2 + a.nablk_op1op2);
    HostBlock hb;
    if (build_block(&a, hb)) return g_err;
    int dev = 0;
    if (check(hipGetDevice(&dev), "hipGetDevice")) return g_err;
    const CallKey key = call_key(hb, dev);
    std::lock_guard<std::mutex> lk(g_calls_mu);
    gecco_hip_plan *plan = nullptr;
    auto it = g_calls.find(key);
    if (it != g_calls.end()) {
        plan = it->second;
    } else {
        if (g_calls.size() >= CALLS_MAX || g_calls_bytes >= CALLS_MAX_BYTES) calls_clear();
        plan = new gecco_hip_plan();
        plan->blocks.push_back(std::move(hb));
        if (plan_finalize(plan)) {
            plan_free_device(plan);
            delete plan;
            return g_err;
        }
        g_calls.emplace(key, plan);
        g_calls_bytes += plan->index_bytes + plan->scratch_elems * 8.0 + 1024.0;
    }
    double *d1 = const_cast<double *>(xop1), *d2 = const_cast<double *>(xop2), *d12 = xop1op2;
    const bool h1 = !is_device_ptr(xop1), h2 = !is_device_ptr(xop2), h12 = !is_device_ptr(xop1op2);
    int rc = 0;
    if (h1 && !rc) rc = stage_get(0, len1, &d1) || check(hipMemcpy(d1, xop1, len1 * 8, hipMemcpyHostToDevice), "stage op1");
    if (h2 && !rc) rc = stage_get(1, len2, &d2) || check(hipMemcpy(d2, xop2, len2 * 8, hipMemcpyHostToDevice), "stage op2");
    if (h12 && !rc) rc = stage_get(2, len12, &d12) || check(hipMemcpy(d12, xop1op2, len12 * 8, hipMemcpyHostToDevice), "stage res");
    double *bufs[3] = {d1, d2, d12};
    if (!rc) rc = plan_execute(plan, bufs, 3, 0);
    if (!rc && h12) rc = check(hipMemcpy(xop1op2, d12, len12 * 8, hipMemcpyDeviceToHost), "unstage res");
    return rc ? g_err : ok();
}

// ------------------------------------------ the allocation-free entry point
// Device memory of the calls of contr_blk1blk2_blocked_mm_noalloc_ (index
// tables, task lists, materialised operands, the buffer-slot table).  Each
// device holds its own copy for the life of the process; its address on a
// device comes from hipGetSymbolAddress on the first call there.
#define NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
static __device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static std::unordered_map<int, char *> g_scratch_ptr;   // device -> its copy of scratch

static int scratch_address(char **out) {
    int dev = 0;
    if (check(hipGetDevice(&dev), "hipGetDevice")) return g_err;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it == g_scratch_ptr.end()) {
        void *ptr = nullptr;
        if (check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)")) return g_err;
        it = g_scratch_ptr.emplace(dev, (char *)ptr).first;
    }
    *out = it->second;
    return ok();
}

static int noalloc_call(const gecco_hip_block_args &a0, double *xop1op2, const double *xop1,
                        const double *xop2) {
    if (!is_device_ptr(xop1) || !is_device_ptr(xop2) || !is_device_ptr(xop1op2))
        return fail(GECCO_HIP_EINVAL, "contr_blk1blk2_blocked_mm_noalloc_: operands must be device memory");
    gecco_hip_block_args a = a0;
    a.buf_op1 = 0;
    a.buf_op2 = 1;
    a.buf_op1op2 = 2;
    a.off_op1 = a.off_op2 = a.off_op1op2 = 0;
    char *base = nullptr;
    if (scratch_address(&base)) return g_err;
    gecco_hip_plan plan;
    plan.arena = base;
    plan.arena_bytes = NOALLOC_SCRATCH_BYTES;
    plan.blocks.emplace_back();
    if (build_block(&a, plan.blocks.back())) return g_err;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    // the previous call's kernels may still read the scratch
    if (check(hipStreamSynchronize(0), "hipStreamSynchronize")) return g_err;
    if (plan_finalize(&plan)) return g_err;
    double *bufs[3] = {const_cast<double *>(xop1), const_cast<double *>(xop2), xop1op2};
    return plan_execute(&plan, bufs, 3, 0);
}

int plan_add_block_internal(gecco_hip_plan *plan, const gecco_hip_block_args *args) {
    if (plan->finalized) return fail(GECCO_HIP_ESTATE, "plan already finalized");
    HostBlock hb;
    if (build_block(args, hb)) return g_err;
    hb.group = plan->cur_group;
    plan->blocks.push_back(std::move(hb));
    return ok();
}

int plan_add_sym_pairs(gecco_hip_plan *plan, int32_t buf, std::vector<int64_t> &&i1,
                       std::vector<int64_t> &&i2, std::vector<int8_t> &&f, double prefac) {
    if (plan->finalized) return fail(GECCO_HIP_ESTATE, "plan already finalized");
    if (buf < 0) return fail(GECCO_HIP_EINVAL, "negative buffer slot");
    SymStep y;
    y.buf = buf;
    y.prefac = prefac;
    y.i1 = std::move(i1);
    y.i2 = std::move(i2);
    y.f = std::move(f);
    y.group = plan->cur_group;
    plan->syms.push_back(std::move(y));
    return ok();
}

int plan_add_trace_rows(gecco_hip_plan *plan, int32_t buf_src, int64_t off_src, int32_t buf_dst,
                        int64_t off_dst, double fact, std::vector<int64_t> &&dst,
                        std::vector<int64_t> &&src, std::vector<int8_t> &&sg) {
    if (plan->finalized) return fail(GECCO_HIP_ESTATE, "plan already finalized");
    if (buf_src < 0 || buf_dst < 0 || off_src < 0 || off_dst < 0)
        return fail(GECCO_HIP_EINVAL, "negative buffer slot or offset");
    TraceStep tr;
    tr.buf_src = buf_src;
    tr.buf_dst = buf_dst;
    tr.off_src = off_src;
    tr.off_dst = off_dst;
    tr.fact = fact;
    // CSR by destination; entries of one row keep the enumeration order
    std::vector<size_t> ord(dst.size());
    for (size_t i = 0; i < ord.size(); ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return dst[a] < dst[b]; });
    tr.rowptr.push_back(0);
    for (size_t j = 0; j < ord.size(); ++j) {
        const size_t i = ord[j];
        if (tr.rowdst.empty() || tr.rowdst.back() != dst[i]) {
            if (!tr.rowdst.empty()) tr.rowptr.push_back((int64_t)j);
            tr.rowdst.push_back(dst[i]);
        }
        tr.col.push_back(src[i]);
        tr.sgn.push_back(sg[i]);
    }
    if (!tr.rowdst.empty()) tr.rowptr.push_back((int64_t)ord.size());
    tr.nrows = (int64_t)tr.rowdst.size();
    tr.nnz = (int64_t)tr.col.size();
    plan->traces.push_back(std::move(tr));
    return ok();
}

int plan_add_spin_groups(gecco_hip_plan *plan, int32_t buf, int64_t off, std::vector<int64_t> &&gptr,
                         std::vector<int64_t> &&elem, std::vector<int64_t> &&gmat, std::vector<double> &&mats) {
    if (plan->finalized) return fail(GECCO_HIP_ESTATE, "plan already finalized");
    if (buf < 0 || off < 0) return fail(GECCO_HIP_EINVAL, "negative buffer slot or offset");
    SpinStep sp;
    sp.buf = buf;
    sp.off = off;
    sp.group = plan->cur_group;
    sp.ngroups = gptr.empty() ? 0 : (int64_t)gptr.size() - 1;
    sp.gptr = std::move(gptr);
    sp.elem = std::move(elem);
    sp.gmat = std::move(gmat);
    sp.mats = std::move(mats);
    plan->spins.push_back(std::move(sp));
    return ok();
}

int set_error(int code, const char *msg) { return fail(code, "%s", msg); }
void clear_error() { g_err = 0; g_msg.clear(); }

}  // namespace gecco

using namespace gecco;

extern "C" {

int gecco_hip_last_error(char *msg, int64_t msglen) {
    if (msg && msglen > 0) {
        snprintf(msg, (size_t)msglen, "%s", g_msg.c_str());
    }
    return g_err;
}

#define FILL_ARGS(A)                                                                  \
    A.xfac = *xfac;                                                                   \
    A.tra_op1 = *tra_op1; A.tra_op2 = *tra_op2; A.tra_op1op2 = *tra_op1op2;            \
    A.ncblk_op1 = *ncblk_op1; A.nablk_op1 = *nablk_op1;                               \
    A.ncblk_ex1 = *ncblk_ex1; A.nablk_ex1 = *nablk_ex1;                               \
    A.ncblk_op2 = *ncblk_op2; A.nablk_op2 = *nablk_op2;                               \
    A.ncblk_ex2 = *ncblk_ex2; A.nablk_ex2 = *nablk_ex2;                               \
    A.ncblk_cnt = *ncblk_cnt; A.nablk_cnt = *nablk_cnt;                               \
    A.ncblk_op1op2 = *ncblk_op1op2; A.nablk_op1op2 = *nablk_op1op2;                   \
    A.hpvxop1c = hpvxop1c; A.hpvxop1a = hpvxop1a; A.hpvxop2c = hpvxop2c;              \
    A.hpvxop2a = hpvxop2a; A.hpvxop1op2c = hpvxop1op2c; A.hpvxop1op2a = hpvxop1op2a;  \
    A.lstrop1 = lstrop1; A.lstrop2 = lstrop2; A.lstrop1op2 = lstrop1op2;              \
    A.lstr_ex1 = lstr_ex1; A.lstr_ex2 = lstr_ex2; A.lstr_cnt = lstr_cnt;              \
    A.map_info_12c = map_info_12c; A.map_info_12a = map_info_12a;                     \
    A.map_info_1c = map_info_1c; A.map_info_1a = map_info_1a;                         \
    A.map_info_2c = map_info_2c; A.map_info_2a = map_info_2a;                         \
    A.map_ex1ex2c = map_ex1ex2c; A.map_ex1ex2a = map_ex1ex2a;                         \
    A.map_ex1cntc = map_ex1cntc; A.map_ex1cnta = map_ex1cnta;                         \
    A.map_ex2cntc = map_ex2cntc; A.map_ex2cnta = map_ex2cnta;

void contr_blk1blk2_blocked_mm_(
    const double *xfac, double *xop1op2, const double *xop1, const double *xop2,
    const int64_t *tra_op1, const int64_t *tra_op2, const int64_t *tra_op1op2, double *xscr,
    const int64_t *lenscr, const int64_t *lenblock, const int64_t *lenblock_cnt,
    const int64_t *ncblk_op1, const int64_t *nablk_op1, const int64_t *ncblk_ex1,
    const int64_t *nablk_ex1, const int64_t *ncblk_op2, const int64_t *nablk_op2,
    const int64_t *ncblk_ex2, const int64_t *nablk_ex2, const int64_t *ncblk_cnt,
    const int64_t *nablk_cnt, const int64_t *ncblk_op1op2, const int64_t *nablk_op1op2,
    const int64_t *hpvxop1c, const int64_t *hpvxop1a, const int64_t *hpvxop2c,
    const int64_t *hpvxop2a, const int64_t *hpvxop1op2c, const int64_t *hpvxop1op2a,
    const int64_t *lstrop1, const int64_t *lstrop2, const int64_t *lstrop1op2,
    const int64_t *lstr_ex1, const int64_t *lstr_ex2, const int64_t *lstr_cnt,
    const int64_t *map_info_12c, const int64_t *map_info_12a, const int64_t *map_info_1c,
    const int64_t *map_info_1a, const int64_t *map_info_2c, const int64_t *map_info_2a,
    const int64_t *map_ex1ex2c, const int64_t *map_ex1ex2a, const int64_t *map_ex1cntc,
    const int64_t *map_ex1cnta, const int64_t *map_ex2cntc, const int64_t *map_ex2cnta) {
    (void)xscr; (void)lenscr; (void)lenblock; (void)lenblock_cnt;
    g_err = 0;
    g_msg.clear();
    gecco_hip_block_args a;
    std::memset(&a, 0, sizeof a);
    FILL_ARGS(a)
    single_call(a, xop1op2, xop1, xop2);
}

void contr_blk1blk2_wmaps_c_(
    const double *xfac, double *xop1op2, const double *xop1, const double *xop2,
    const int64_t *tra_op1, const int64_t *tra_op2, const int64_t *tra_op1op2,
    const int64_t *ncblk_op1, const int64_t *nablk_op1, const int64_t *ncblk_ex1,
    const int64_t *nablk_ex1, const int64_t *ncblk_op2, const int64_t *nablk_op2,
    const int64_t *ncblk_ex2, const int64_t *nablk_ex2, const int64_t *ncblk_cnt,
    const int64_t *nablk_cnt, const int64_t *ncblk_op1op2, const int64_t *nablk_op1op2,
    const int64_t *hpvxop1c, const int64_t *hpvxop1a, const int64_t *hpvxop2c,
    const int64_t *hpvxop2a, const int64_t *hpvxop1op2c, const int64_t *hpvxop1op2a,
    const int64_t *lstrop1, const int64_t *lstrop2, const int64_t *lstrop1op2,
    const int64_t *lstr_ex1, const int64_t *lstr_ex2, const int64_t *lstr_cnt,
    const int64_t *map_info_12c, const int64_t *map_info_12a, const int64_t *map_info_1c,
    const int64_t *map_info_1a, const int64_t *map_info_2c, const int64_t *map_info_2a,
    const int64_t *map_ex1ex2c, const int64_t *map_ex1ex2a, const int64_t *map_ex1cntc,
    const int64_t *map_ex1cnta, const int64_t *map_ex2cntc, const int64_t *map_ex2cnta) {
    g_err = 0;
    g_msg.clear();
    gecco_hip_block_args a;
    std::memset(&a, 0, sizeof a);
    FILL_ARGS(a)
    single_call(a, xop1op2, xop1, xop2);
}

void contr_blk1blk2_blocked_mm_noalloc_(
    const double *xfac, double *xop1op2, const double *xop1, const double *xop2,
    const int64_t *tra_op1, const int64_t *tra_op2, const int64_t *tra_op1op2, double *xscr,
    const int64_t *lenscr, const int64_t *lenblock, const int64_t *lenblock_cnt,
    const int64_t *ncblk_op1, const int64_t *nablk_op1, const int64_t *ncblk_ex1,
    const int64_t *nablk_ex1, const int64_t *ncblk_op2, const int64_t *nablk_op2,
    const int64_t *ncblk_ex2, const int64_t *nablk_ex2, const int64_t *ncblk_cnt,
    const int64_t *nablk_cnt, const int64_t *ncblk_op1op2, const int64_t *nablk_op1op2,
    const int64_t *hpvxop1c, const int64_t *hpvxop1a, const int64_t *hpvxop2c,
    const int64_t *hpvxop2a, const int64_t *hpvxop1op2c, const int64_t *hpvxop1op2a,
    const int64_t *lstrop1, const int64_t *lstrop2, const int64_t *lstrop1op2,
    const int64_t *lstr_ex1, const int64_t *lstr_ex2, const int64_t *lstr_cnt,
    const int64_t *map_info_12c, const int64_t *map_info_12a, const int64_t *map_info_1c,
    const int64_t *map_info_1a, const int64_t *map_info_2c, const int64_t *map_info_2a,
    const int64_t *map_ex1ex2c, const int64_t *map_ex1ex2a, const int64_t *map_ex1cntc,
    const int64_t *map_ex1cnta, const int64_t *map_ex2cntc, const int64_t *map_ex2cnta) {
    (void)xscr; (void)lenscr; (void)lenblock; (void)lenblock_cnt;
    g_err = 0;
    g_msg.clear();
    gecco_hip_block_args a;
    std::memset(&a, 0, sizeof a);
    FILL_ARGS(a)
    noalloc_call(a, xop1op2, xop1, xop2);
}

int64_t gecco_hip_noalloc_scratch_bytes(void) { return (int64_t)NOALLOC_SCRATCH_BYTES; }

int gecco_hip_call_cache(int64_t *entries, int64_t clear) {
    g_err = 0;
    std::lock_guard<std::mutex> lk(g_calls_mu);
    if (clear) calls_clear();
    if (entries) *entries = (int64_t)g_calls.size();
    return ok();
}

int gecco_hip_plan_create(gecco_hip_plan **plan) {
    g_err = 0;
    if (!plan) return fail(GECCO_HIP_EINVAL, "null plan pointer");
    *plan = new gecco_hip_plan();
    if (const char *e = getenv("GECCO_HIP_AFFINE_TILE")) (*plan)->affine_variant = atoi(e);
    if (const char *e = getenv("GECCO_HIP_TARGET_TASKS")) (*plan)->target_tasks = std::max(1L, atol(e));
    if (const char *e = getenv("GECCO_HIP_XCD_ORDER")) (*plan)->xcd_order = atoi(e) != 0;
    if (const char *e = getenv("GECCO_HIP_MIN_KCHUNK")) (*plan)->min_kchunk = std::max(32L, atol(e));
    if (const char *e = getenv("GECCO_HIP_MAX_KCHUNK")) (*plan)->max_kchunk = std::max(64L, atol(e));
    if (const char *e = getenv("GECCO_HIP_MATERIALIZE_MIN")) (*plan)->materialize_min = std::max(0L, atol(e));
    if (const char *e = getenv("GECCO_HIP_CHAIN")) (*plan)->chain = atoi(e) != 0;
    if (const char *e = getenv("GECCO_HIP_SKINNY_MAX")) (*plan)->skinny_max = std::max(0L, atol(e));
    if (const char *e = getenv("GECCO_HIP_SKINNY_KCHUNK")) (*plan)->skinny_kchunk = std::max(128L, atol(e));
    if (const char *e = getenv("GECCO_HIP_DIRECT_KMAX")) (*plan)->direct_kmax = std::min(32L, std::max(0L, atol(e)));
    return 0;
}

int gecco_hip_plan_destroy(gecco_hip_plan *plan) {
    if (!plan) return 0;
    plan_free_device(plan);
    delete plan;
    return 0;
}

int gecco_hip_plan_add_block(gecco_hip_plan *plan, const gecco_hip_block_args *args) {
    g_err = 0;
    if (!plan || !args) return fail(GECCO_HIP_EINVAL, "null argument");
    if (plan->finalized) return fail(GECCO_HIP_ESTATE, "plan already finalized");
    if (args->buf_op1 < 0 || args->buf_op2 < 0 || args->buf_op1op2 < 0 ||
        args->off_op1 < 0 || args->off_op2 < 0 || args->off_op1op2 < 0)
        return fail(GECCO_HIP_EINVAL, "negative buffer slot or offset");
    HostBlock hb;
    if (build_block(args, hb)) return g_err;
    hb.group = plan->cur_group;
    plan->blocks.push_back(std::move(hb));
    return 0;
}

int gecco_hip_plan_finalize(gecco_hip_plan *plan) {
    g_err = 0;
    if (!plan) return fail(GECCO_HIP_EINVAL, "null plan");
    int rc = plan_finalize(plan);
    if (rc) plan_free_device(plan);
    return rc;
}

int gecco_hip_plan_execute(gecco_hip_plan *plan, double *const *bufs, int64_t nbufs,
                           void *stream) {
    g_err = 0;
    if (!plan || (!bufs && nbufs)) return fail(GECCO_HIP_EINVAL, "null argument");
    return plan_execute(plan, bufs, nbufs, (hipStream_t)stream);
}

int gecco_hip_plan_stats(const gecco_hip_plan *plan, double *flops, int64_t *nblocks,
                         int64_t *ntasks, double *payload_bytes, double *index_bytes) {
    if (!plan) return fail(GECCO_HIP_EINVAL, "null plan");
    double f = 0;
    for (auto &b : plan->blocks) f += 2.0 * (double)b.M * (double)b.N * (double)b.K;
    if (flops) *flops = f;
    if (nblocks) *nblocks = (int64_t)plan->blocks.size();
    if (ntasks) *ntasks = plan->ntasks;
    if (payload_bytes) *payload_bytes = plan->payload;
    if (index_bytes) *index_bytes = plan->index_bytes;
    return 0;
}

int gecco_hip_plan_block_info(const gecco_hip_plan *plan, int64_t i, int64_t *info, int64_t ninfo) {
    // info = [M, N, K, swap, injective, A.table, B.table, C.table,
    //         A.fast_first, B.fast_first, A affine (first,second), B affine, C affine]
    if (!plan || i < 0 || i >= (int64_t)plan->blocks.size() || ninfo < 16)
        return fail(GECCO_HIP_EINVAL, "plan_block_info: bad arguments");
    const HostBlock &b = plan->blocks[i];
    info[0] = b.M; info[1] = b.N; info[2] = b.K; info[3] = b.swap; info[4] = b.injective;
    info[5] = b.A.r.table; info[6] = b.B.r.table; info[7] = b.C.r.table;
    info[8] = b.A.fast_first; info[9] = b.B.fast_first;
    int64_t a0, a1;
    info[10] = !b.A.r.table && affine_of(b.A.r.first, a0, a1);
    info[11] = !b.A.r.table && affine_of(b.A.r.second, a0, a1);
    info[12] = !b.B.r.table && affine_of(b.B.r.first, a0, a1);
    info[13] = !b.B.r.table && affine_of(b.B.r.second, a0, a1);
    info[14] = !b.C.r.table && affine_of(b.C.r.first, a0, a1);
    info[15] = !b.C.r.table && affine_of(b.C.r.second, a0, a1);
    if (ninfo >= 19) { info[16] = b.strided; info[17] = b.mat_a; info[18] = b.mat_b; }
    if (ninfo >= 20) info[19] = b.direct;
    if (ninfo >= 23) { info[20] = b.buf12; info[21] = b.off12; info[22] = b.len12; }
    if (ninfo >= 24) info[23] = b.next;
    if (ninfo >= 25) info[24] = b.skinny;
    if (ninfo >= 30) {
        // operands in kernel orientation (buf1/off1 hold A: op2 when
        // swapped) and the factor's bit pattern
        info[25] = b.buf1;
        info[26] = b.off1;
        info[27] = b.buf2;
        info[28] = b.off2;
        std::memcpy(&info[29], &b.xfac, sizeof(double));
    }
    return 0;
}

int gecco_hip_plan_block_index(const gecco_hip_plan *plan, int64_t i, int64_t which,
                               int32_t *first, int64_t nfirst, int32_t *second, int64_t nsecond) {
    if (!plan || i < 0 || i >= (int64_t)plan->blocks.size() || which < 0 || which > 2)
        return fail(GECCO_HIP_EINVAL, "plan_block_index: bad arguments");
    if (plan->finalized) return fail(GECCO_HIP_ESTATE, "index tables are released by finalize");
    const HostBlock &b = plan->blocks[i];
    const Resolved &r = which == 0 ? b.A.r : which == 1 ? b.B.r : b.C.r;
    if ((int64_t)r.first.size() > nfirst || (int64_t)r.second.size() > nsecond)
        return fail(GECCO_HIP_EINVAL, "plan_block_index: output too small");
    if (first) std::copy(r.first.begin(), r.first.end(), first);
    if (second) std::copy(r.second.begin(), r.second.end(), second);
    return 0;
}

int gecco_hip_plan_add_init(gecco_hip_plan *plan, int32_t buf_dst, int64_t off_dst, int64_t len,
                            int32_t buf_src, int64_t off_src, double factor) {
    g_err = 0;
    if (!plan) return fail(GECCO_HIP_EINVAL, "null plan");
    if (plan->finalized) return fail(GECCO_HIP_ESTATE, "plan already finalized");
    if (buf_dst < 0 || off_dst < 0 || len < 0 || (buf_src >= 0 && off_src < 0))
        return fail(GECCO_HIP_EINVAL, "add_init: bad slot, offset or length");
    if (buf_src == buf_dst && buf_src >= 0 && off_src < off_dst + len && off_dst < off_src + len &&
        off_src != off_dst)
        return fail(GECCO_HIP_EINVAL, "add_init: overlapping source and destination");
    DevInit j;
    std::memset(&j, 0, sizeof j);
    j.dst = buf_dst;
    j.src = buf_src < 0 ? -1 : buf_src;
    j.off_dst = off_dst;
    j.off_src = buf_src < 0 ? 0 : off_src;
    j.len = len;
    j.factor = factor;
    plan->inits.push_back(j);
    plan->init_group.push_back(plan->cur_group);
    return 0;
}

int gecco_hip_plan_add_sum(gecco_hip_plan *plan, int32_t buf_dst, int64_t off_dst, int64_t len,
                           int32_t buf_src, int64_t off_src, double factor) {
    g_err = 0;
    if (!plan) return fail(GECCO_HIP_EINVAL, "null plan");
    if (plan->finalized) return fail(GECCO_HIP_ESTATE, "plan already finalized");
    if (buf_dst < 0 || off_dst < 0 || len < 0 || buf_src < 0 || off_src < 0)
        return fail(GECCO_HIP_EINVAL, "add_sum: bad slot, offset or length");
    if (buf_src == buf_dst && off_src < off_dst + len && off_dst < off_src + len)
        return fail(GECCO_HIP_EINVAL, "add_sum: source and destination overlap");
    DevInit j;
    std::memset(&j, 0, sizeof j);
    j.dst = buf_dst;
    j.src = buf_src;
    j.off_dst = off_dst;
    j.off_src = off_src;
    j.len = len;
    j.factor = factor;
    j.accumulate = 1;
    plan->inits.push_back(j);
    plan->init_group.push_back(plan->cur_group);
    return 0;
}

int gecco_hip_plan_set_constant(gecco_hip_plan *plan, int32_t slot) {
    if (!plan) return fail(GECCO_HIP_EINVAL, "null plan");
    if (plan->finalized) return fail(GECCO_HIP_ESTATE, "plan already finalized");
    if (slot < 0) return fail(GECCO_HIP_EINVAL, "negative slot");
    if ((int32_t)plan->constant_slot.size() <= slot) plan->constant_slot.resize(slot + 1, 0);
    plan->constant_slot[slot] = 1;
    return 0;
}

int gecco_hip_plan_set_tuning(gecco_hip_plan *plan, int64_t variant, int64_t target_tasks,
                              int64_t min_kchunk, int64_t max_kchunk) {
    if (!plan) return fail(GECCO_HIP_EINVAL, "null plan");
    if (plan->finalized) return fail(GECCO_HIP_ESTATE, "plan already finalized");
    int bm, bn, bk;
    if (variant >= 0 && gecco_affine_variant_shape((int)variant, &bm, &bn, &bk))
        return fail(GECCO_HIP_EINVAL, "unknown tile variant %lld", (long long)variant);
    if (variant >= 0) plan->affine_variant = (int)variant;
    if (target_tasks > 0) plan->target_tasks = target_tasks;
    if (min_kchunk > 0) plan->min_kchunk = std::max<int64_t>(32, min_kchunk);
    if (max_kchunk > 0) plan->max_kchunk = std::max<int64_t>(64, max_kchunk);
    return 0;
}

int gecco_hip_plan_set_precision(gecco_hip_plan *plan, int64_t elem_bytes) {
    if (!plan) return fail(GECCO_HIP_EINVAL, "null plan");
    if (plan->finalized) return fail(GECCO_HIP_ESTATE, "plan already finalized");
    if (elem_bytes != 4 && elem_bytes != 8)
        return fail(GECCO_HIP_EINVAL, "precision: 8 (float64) or 4 (float32) bytes per element");
    plan->elem = (int)elem_bytes;
    return 0;
}

int gecco_hip_plan_keep_host(gecco_hip_plan *plan, int64_t on) {
    if (!plan) return fail(GECCO_HIP_EINVAL, "null plan");
    if (!plan->finalized) { plan->keep_host = on != 0; return 0; }
    if (on) return fail(GECCO_HIP_ESTATE, "keep_host after finalize: the host data is released");
    if (plan->keep_host) plan_release_host(plan);
    return 0;
}

int gecco_hip_plan_retune(gecco_hip_plan *plan, int64_t variant, int64_t target_tasks, int64_t min_kchunk,
                          int64_t max_kchunk) {
    if (!plan) return fail(GECCO_HIP_EINVAL, "null plan");
    if (!plan->finalized || !plan->keep_host)
        return fail(GECCO_HIP_ESTATE, "retune needs a finalized plan that keeps its host data");
    int bm, bn, bk;
    if (variant >= 0 && gecco_affine_variant_shape((int)variant, &bm, &bn, &bk))
        return fail(GECCO_HIP_EINVAL, "unknown tile variant %lld", (long long)variant);
    plan_free_device(plan);
    plan->finalized = false;
    plan->static_done = false;
    plan->last_bufs.clear();
    plan->payload = plan->index_bytes = 0;
    plan->nsymtasks = plan->nitasks = plan->ngtasks = plan->ngtasks_static = plan->ngjobs = 0;
    if (variant >= 0) plan->affine_variant = (int)variant;
    if (target_tasks > 0) plan->target_tasks = target_tasks;
    if (min_kchunk > 0) plan->min_kchunk = std::max<int64_t>(32, min_kchunk);
    if (max_kchunk > 0) plan->max_kchunk = std::max<int64_t>(64, max_kchunk);
    if (plan_finalize(plan)) { plan_free_device(plan); return g_err; }
    return 0;
}

int gecco_hip_plan_set_group(gecco_hip_plan *plan, int32_t group) {
    if (!plan) return fail(GECCO_HIP_EINVAL, "null plan");
    if (plan->finalized) return fail(GECCO_HIP_ESTATE, "plan already finalized");
    if (group < 0) return fail(GECCO_HIP_EINVAL, "negative instance group");
    plan->cur_group = group;
    plan->max_group = std::max(plan->max_group, group);
    return 0;
}

int gecco_hip_plan_set_skip(gecco_hip_plan *plan, const int32_t *flags, int64_t nflags) {
    if (!plan) return fail(GECCO_HIP_EINVAL, "null plan");
    if (flags && nflags <= plan->max_group)
        return fail(GECCO_HIP_EINVAL, "skip flags: %lld given, the plan records groups up to %d",
                    (long long)nflags, plan->max_group);
    if (flags && !is_device_ptr(flags))
        return fail(GECCO_HIP_EINVAL, "skip flags must be device memory");
    plan->d_skip = flags;
    plan->nskip = flags ? nflags : 0;
    plan->compacted = false;   // live lists belong to the flags they were made from
    return 0;
}

int gecco_hip_plan_compact(gecco_hip_plan *plan, int64_t on, void *stream) {
    clear_error();
    if (!plan) return fail(GECCO_HIP_EINVAL, "null plan");
    if (!plan->finalized) return fail(GECCO_HIP_ESTATE, "compact: plan not finalized");
    if (!on) {
        plan->compacted = false;
        return 0;
    }
    if (!plan->d_skip) return fail(GECCO_HIP_ESTATE, "compact: the plan has no skip flags");
    if (plan->deterministic) return fail(GECCO_HIP_EINVAL, "compact: fixed-order plans keep every task");
    if (plan->arena) return fail(GECCO_HIP_EINVAL, "compact: plan without device memory of its own");
    gecco_hip_plan *p = plan;
    hipStream_t s = (hipStream_t)stream;
    const int64_t na = p->ntasks_affine;
    struct Part { const void *src; void *dst; int64_t n; int rec, grp; };
    if (!p->d_ltasks) {
        const int64_t nc = [&] {
            auto ch = [](int64_t n) { return (n + 1023) / 1024; };
            return ch(na) + ch(p->ntasks - na) + ch(p->nitasks) + ch(p->nsymtasks) + ch(p->ngtasks_static) +
                   ch(p->ngtasks - p->ngtasks_static);
        }();
        if (dmalloc(p, (void **)&p->d_ltasks, std::max<int64_t>(1, p->ntasks) * sizeof(gecco::DevTask), "hipMalloc(live tasks)") ||
            dmalloc(p, (void **)&p->d_litasks, std::max<int64_t>(1, p->nitasks) * sizeof(gecco::DevInitTask), "hipMalloc(live tasks)") ||
            dmalloc(p, (void **)&p->d_lsymtasks, std::max<int64_t>(1, p->nsymtasks) * sizeof(gecco::DevSymTask), "hipMalloc(live tasks)") ||
            dmalloc(p, (void **)&p->d_lgtasks, std::max<int64_t>(1, p->ngtasks) * sizeof(gecco::DevGatherTask), "hipMalloc(live tasks)") ||
            dmalloc(p, (void **)&p->d_lchunk, std::max<int64_t>(1, nc) * sizeof(int32_t), "hipMalloc(live tasks)") ||
            dmalloc(p, (void **)&p->d_lcount, 6 * sizeof(int64_t), "hipMalloc(live tasks)"))
            return g_err;
    }
    const Part parts[6] = {
        {p->d_tasks, p->d_ltasks, na, (int)sizeof(gecco::DevTask), (int)offsetof(gecco::DevTask, grp)},
        {p->d_tasks + na, p->d_ltasks + na, p->ntasks - na, (int)sizeof(gecco::DevTask), (int)offsetof(gecco::DevTask, grp)},
        {p->d_itasks, p->d_litasks, p->nitasks, (int)sizeof(gecco::DevInitTask), (int)offsetof(gecco::DevInitTask, grp)},
        {p->d_symtasks, p->d_lsymtasks, p->nsymtasks, (int)sizeof(gecco::DevSymTask), (int)offsetof(gecco::DevSymTask, grp)},
        {p->d_gtasks, p->d_lgtasks, p->ngtasks_static, (int)sizeof(gecco::DevGatherTask), (int)offsetof(gecco::DevGatherTask, grp)},
        {p->d_gtasks + p->ngtasks_static, p->d_lgtasks + p->ngtasks_static, p->ngtasks - p->ngtasks_static,
         (int)sizeof(gecco::DevGatherTask), (int)offsetof(gecco::DevGatherTask, grp)},
    };
    int64_t chunk0 = 0;
    bool launched[6] = {false, false, false, false, false, false};
    for (int i = 0; i < 6; ++i) {
        const Part &q = parts[i];
        if (q.n <= 0) continue;
        if (gecco_launch_compact(q.src, q.dst, q.n, q.rec, q.grp, p->d_skip, p->d_lchunk + chunk0,
                                 p->d_lcount + i, s))
            return fail(GECCO_HIP_EDEVICE, "live-task launch failed: %s", hipGetErrorString(hipGetLastError()));
        launched[i] = true;
        chunk0 += (q.n + 1023) / 1024;
    }
    int64_t cnt[6] = {0, 0, 0, 0, 0, 0};
    if (check(hipStreamSynchronize(s), "hipStreamSynchronize") ||
        check(hipMemcpy(cnt, p->d_lcount, sizeof cnt, hipMemcpyDeviceToHost), "hipMemcpy(live counts)"))
        return g_err;
    for (int i = 0; i < 6; ++i) p->lt[i] = launched[i] ? cnt[i] : 0;
    p->compacted = true;
    return 0;
}

int gecco_hip_plan_live_tasks(const gecco_hip_plan *plan, int64_t *counts, int64_t ncounts) {
    if (!plan || !counts || ncounts < 1) return fail(GECCO_HIP_EINVAL, "live_tasks: null argument");
    const int64_t na = plan->ntasks_affine;
    const int64_t full[6] = {na, plan->ntasks - na, plan->nitasks, plan->nsymtasks, plan->ngtasks_static,
                             plan->ngtasks - plan->ngtasks_static};
    for (int64_t i = 0; i < std::min<int64_t>(ncounts, 6); ++i) counts[i] = plan->compacted ? plan->lt[i] : full[i];
    return 0;
}

int gecco_hip_plan_set_screen(gecco_hip_plan *plan, int64_t on) {
    if (!plan) return fail(GECCO_HIP_EINVAL, "null plan");
    if (plan->finalized) return fail(GECCO_HIP_ESTATE, "plan already finalized");
    plan->screen = on != 0;
    return 0;
}

int gecco_hip_plan_set_deterministic(gecco_hip_plan *plan, int64_t on) {
    if (!plan) return fail(GECCO_HIP_EINVAL, "null plan");
    if (plan->finalized) return fail(GECCO_HIP_ESTATE, "plan already finalized");
    plan->deterministic = on != 0;
    return 0;
}

int gecco_hip_plan_set_chain(gecco_hip_plan *plan, int64_t on) {
    if (!plan) return fail(GECCO_HIP_EINVAL, "null plan");
    if (plan->finalized) return fail(GECCO_HIP_ESTATE, "plan already finalized");
    plan->chain = on != 0;
    return 0;
}

int gecco_hip_plan_set_materialize(gecco_hip_plan *plan, int64_t min_dim) {
    if (!plan) return fail(GECCO_HIP_EINVAL, "null plan");
    if (plan->finalized) return fail(GECCO_HIP_ESTATE, "plan already finalized");
    if (min_dim < 0) return fail(GECCO_HIP_EINVAL, "negative threshold");
    plan->materialize_min = min_dim;
    return 0;
}

int gecco_hip_plan_workspace(const gecco_hip_plan *plan, int64_t *scratch_bytes,
                             int64_t *gather_jobs, double *gather_bytes, int64_t *staged_tasks) {
    if (!plan) return fail(GECCO_HIP_EINVAL, "null plan");
    if (staged_tasks) *staged_tasks = plan->nstaged;
    if (scratch_bytes) *scratch_bytes = 8 * plan->scratch_elems;
    if (gather_jobs) *gather_jobs = plan->ngjobs;
    if (gather_bytes) *gather_bytes = plan->gather_bytes;
    return 0;
}

int gecco_hip_plan_set_timing(gecco_hip_plan *plan, int64_t on) {
    if (!plan) return fail(GECCO_HIP_EINVAL, "null plan");
    plan->timing = on != 0;
    return 0;
}

int gecco_hip_plan_timing(gecco_hip_plan *plan, double *ms, int64_t *launches, double *flops) {
    // ms[5], launches[5]: accumulated times and counts of the strided-GEMM
    // launches, the gather-kernel launches, the symmetrisation launches, the
    // operand-materialisation launches and the direct-kernel launches since
    // the last reset (waits for the recorded events); flops[3]: per-execute
    // FLOP of the strided-GEMM, gather and direct kernels.  Resets the
    // accumulators.
    if (!plan) return fail(GECCO_HIP_EINVAL, "null plan");
    for (auto &pe : plan->pending) {
        float t = 0.f;
        if (check(hipEventSynchronize(pe.second.second), "hipEventSynchronize")) return g_err;
        if (check(hipEventElapsedTime(&t, pe.second.first, pe.second.second), "hipEventElapsedTime")) return g_err;
        plan->ms[pe.first] += t;
        plan->launches[pe.first] += 1;
        (void)hipEventDestroy(pe.second.first);
        (void)hipEventDestroy(pe.second.second);
    }
    plan->pending.clear();
    for (int i = 0; i < 5; ++i) {
        if (ms) ms[i] = plan->ms[i];
        if (launches) launches[i] = plan->launches[i];
        plan->ms[i] = 0;
        plan->launches[i] = 0;
    }
    if (flops) { flops[0] = plan->flops_affine; flops[1] = plan->flops_general; flops[2] = plan->flops_direct; }
    return 0;
}

int gecco_hip_malloc(void **ptr, int64_t bytes) {
    g_err = 0;
    if (!ptr || bytes < 0) return fail(GECCO_HIP_EINVAL, "bad argument");
    if (check(hipMalloc(ptr, (size_t)std::max<int64_t>(bytes, 1)), "hipMalloc")) return g_err;
    return 0;
}
int gecco_hip_free(void *ptr) { return ptr ? check(hipFree(ptr), "hipFree") : 0; }
int gecco_hip_memcpy_htod(void *dst, const void *src, int64_t bytes) {
    return check(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
}
int gecco_hip_memcpy_dtoh(void *dst, const void *src, int64_t bytes) {
    return check(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
}
int gecco_hip_synchronize(void) { return check(hipDeviceSynchronize(), "hipDeviceSynchronize"); }

int gecco_hip_task_order(int64_t n, const int32_t *patch, const int64_t *cost, int64_t *perm) {
    g_err = 0;
    if (n < 0 || (n > 0 && (!patch || !cost || !perm))) return fail(GECCO_HIP_EINVAL, "bad task order arguments");
    xcd_task_order(n, patch, cost, perm);
    return 0;
}

}  // extern "C"
