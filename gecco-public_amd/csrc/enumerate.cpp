// Plan builder: orbital strings, string maps, ME-list layout and the block
// enumerator of a binary contraction, producing the kernel calls that GeCCo's
// contr_op1op2_wmaps_c would make (contract/contr_op1op2_wmaps_c.f:964-2070),
// recorded into a gecco_hip_plan for one grouped launch.
//
// Formulation (own, checked bit-exactly against the reference restatement
// in the tests):
//  * a string is its sorted list of spin-orbitals so = 2*(orb-1)+spin
//    (spin 0 = alpha, 1 = beta), i.e. ascending orbitals, alpha before beta
//    (strings/iordstr2.f:68-92);
//  * strings of one (Ms, IRREP) block are ordered by their occupation codes
//    (0 empty < 1 alpha < 2 beta < 3 doubly) compared from the highest
//    orbital down -- the order of the 4-slope graph index
//    (strings/weight_4sg.f, strings/idx4sg_sb.f: arcs 0,+1,-1,+2 weighted in
//    that order);
//  * the merge sign of (str1, str2) is the parity of sorting str1 ++ str2
//    (strings/iordstr2.f:61-105), 0 if a spin-orbital repeats;
//  * the spin-flip sign is the product of the alpha-before-beta reordering
//    parities of the string and of its flipped string
//    (strings/set_flipmap_kernel.f, std_spsign);
//  * the list layout is occupation -> Ms(A) descending -> IRREP(A) ->
//    Ms/IRREP distributions in next_msgamdist2 order -> dense C x A block
//    (operators/set_op_dim2.f:146-459, operators/next_msgamdist2.f).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <array>
#include <cstring>
#include <map>
#include <array>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/gecco_hip.h"
#include "../../include/gecco_hip_model.h"

namespace gecco {
int plan_add_block_internal(gecco_hip_plan *plan, const gecco_hip_block_args *args);
int plan_add_sym_pairs(gecco_hip_plan *plan, int32_t buf, std::vector<int64_t> &&i1,
                       std::vector<int64_t> &&i2, std::vector<int8_t> &&f, double prefac);
int set_error(int code, const char *msg);
int plan_add_spin_groups(gecco_hip_plan *plan, int32_t buf, int64_t off, std::vector<int64_t> &&gptr,
                         std::vector<int64_t> &&elem, std::vector<int64_t> &&gmat, std::vector<double> &&mats);
int plan_add_trace_rows(gecco_hip_plan *plan, int32_t buf_src, int64_t off_src, int32_t buf_dst,
                        int64_t off_dst, double fact, std::vector<int64_t> &&dst,
                        std::vector<int64_t> &&src, std::vector<int8_t> &&sg);

static inline int mult(int a, int b) { return ((a - 1) ^ (b - 1)) + 1; }
static const int HPVXBLKSEQ[4] = {4, 2, 3, 1};  // include/hpvxseq.h:4

struct StrKey {
    uint64_t w[2] = {0, 0};
    bool operator==(const StrKey &o) const { return w[0] == o.w[0] && w[1] == o.w[1]; }
};
struct StrKeyHash {
    size_t operator()(const StrKey &k) const {
        return std::hash<uint64_t>()(k.w[0] * 0x9E3779B97F4A7C15ull ^ (k.w[1] + 0x632BE59BD9B4E019ull));
    }
};

static StrKey make_key(const uint16_t *so, int n) {
    StrKey k;
    for (int i = 0; i < n; ++i) k.w[i >> 2] |= (uint64_t)(so[i] + 1) << (16 * (i & 3));
    return k;
}

static void unpack(const StrKey &k, int n, uint16_t *so) {
    for (int i = 0; i < n; ++i) so[i] = (uint16_t)(((k.w[i >> 2] >> (16 * (i & 3))) & 0xFFFF) - 1);
}

struct GraphBlock {
    std::vector<StrKey> strs;
    std::unordered_map<StrKey, int32_t, StrKeyHash> index;
};

struct Graph {
    int space = 0, nel = 0;
    std::vector<GraphBlock> blk;  // [(idxms-1)*nsym + (igam-1)]
};

// compare two strings of equal length by occupation codes, highest orbital first
static bool colex_less(const uint16_t *a, const uint16_t *b, int n) {
    // walk from the end, orbital by orbital
    int i = n - 1, j = n - 1;
    while (i >= 0 && j >= 0) {
        int oa = a[i] >> 1, ob = b[j] >> 1;
        if (oa != ob) return oa < ob;  // the string with the higher orbital is larger
        int ca = 0, cb = 0;
        int ii = i, jj = j;
        while (ii >= 0 && (a[ii] >> 1) == oa) { ca |= 1 << (a[ii] & 1); --ii; }
        while (jj >= 0 && (b[jj] >> 1) == ob) { cb |= 1 << (b[jj] & 1); --jj; }
        // code: alpha only =1, beta only =2, both =3  (bitmask 1,2,3 coincide)
        if (ca != cb) return ca < cb;
        i = ii;
        j = jj;
    }
    return j >= 0 && i < 0;  // the string with lower-lying leftovers is larger
}

}  // namespace gecco

struct gecco_hip_orbs {
    int nsym = 1;
    std::vector<int> irr[5];  // hpvx 1..4: sorted IRREP per orbital
    std::map<std::pair<int, int>, std::unique_ptr<gecco::Graph>> graphs;
    // cached string maps, key = (space, nel1, idxms1, g1, nel2, idxms2, g2)
    std::map<std::array<int, 7>, std::vector<int64_t>> maps;
    std::map<std::array<int, 4>, std::vector<int64_t>> flips;
};

namespace gecco {

static Graph &graph(gecco_hip_orbs *o, int space, int nel) {
    auto key = std::make_pair(space, nel);
    auto it = o->graphs.find(key);
    if (it != o->graphs.end()) return *it->second;
    auto g = std::make_unique<Graph>();
    g->space = space;
    g->nel = nel;
    const int nsym = o->nsym;
    g->blk.resize((size_t)(nel + 1) * nsym);
    const std::vector<int> &irr = o->irr[space];
    const int norb = (int)irr.size();
    std::vector<std::vector<uint16_t>> tmp((size_t)(nel + 1) * nsym);
    // alpha set of size na, beta set of size nb (combinations, ascending)
    for (int na = 0; na <= nel; ++na) {
        const int nb = nel - na;
        if (na > norb || nb > norb) continue;
        std::vector<int> A(na), B(nb);
        for (int i = 0; i < na; ++i) A[i] = i;
        for (;;) {
            int ga = 1;
            for (int x : A) ga = mult(ga, irr[x]);
            for (int i = 0; i < nb; ++i) B[i] = i;
            for (;;) {
                int gg = ga;
                for (int x : B) gg = mult(gg, irr[x]);
                const int ms = na - nb;
                const int idxms = (nel - ms) / 2 + 1;
                std::vector<uint16_t> &dst = tmp[(size_t)(idxms - 1) * nsym + gg - 1];
                // merge to sorted spin-orbital list
                int ia = 0, ib = 0;
                while (ia < na || ib < nb) {
                    if (ib >= nb || (ia < na && A[ia] <= B[ib])) dst.push_back((uint16_t)(2 * A[ia++]));
                    else dst.push_back((uint16_t)(2 * B[ib++] + 1));
                }
                int k = nb - 1;
                while (k >= 0 && B[k] == norb - nb + k) --k;
                if (k < 0) break;
                ++B[k];
                for (int i = k + 1; i < nb; ++i) B[i] = B[i - 1] + 1;
            }
            int k = na - 1;
            while (k >= 0 && A[k] == norb - na + k) --k;
            if (k < 0) break;
            ++A[k];
            for (int i = k + 1; i < na; ++i) A[i] = A[i - 1] + 1;
        }
    }
    for (size_t b = 0; b < tmp.size(); ++b) {
        const size_t ns = nel ? tmp[b].size() / nel : 0;
        std::vector<uint32_t> order(ns);
        for (size_t i = 0; i < ns; ++i) order[i] = (uint32_t)i;
        const uint16_t *base = tmp[b].data();
        std::sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
            return colex_less(base + (size_t)x * nel, base + (size_t)y * nel, nel);
        });
        GraphBlock &gb = g->blk[b];
        gb.strs.reserve(ns);
        for (size_t i = 0; i < ns; ++i) {
            StrKey k = make_key(base + (size_t)order[i] * nel, nel);
            gb.index.emplace(k, (int32_t)i);
            gb.strs.push_back(k);
        }
    }
    if (nel == 0) {  // the empty string: Ms 0, IRREP 1
        g->blk[0].strs.push_back(StrKey());
        g->blk[0].index.emplace(StrKey(), 0);
    }
    Graph &ref = *g;
    o->graphs.emplace(key, std::move(g));
    return ref;
}

static int64_t lenstr(gecco_hip_orbs *o, int space, int nel, int idxms, int igam) {
    if (nel < 0 || idxms < 1 || idxms > nel + 1 || igam < 1 || igam > o->nsym) return 0;
    return (int64_t)graph(o, space, nel).blk[(size_t)(idxms - 1) * o->nsym + igam - 1].strs.size();
}

// (str1, str2) -> str12 map, str1 fastest (strings/set_strmap_kernel.f:55-118)
static const std::vector<int64_t> &strmap(gecco_hip_orbs *o, int space, int n1, int im1, int g1,
                                          int n2, int im2, int g2) {
    std::array<int, 7> key{space, n1, im1, g1, n2, im2, g2};
    auto it = o->maps.find(key);
    if (it != o->maps.end()) return it->second;
    const int nsym = o->nsym;
    const Graph &G1 = graph(o, space, n1), &G2 = graph(o, space, n2);
    const Graph &G12 = graph(o, space, n1 + n2);
    const GraphBlock &B1 = G1.blk[(size_t)(im1 - 1) * nsym + g1 - 1];
    const GraphBlock &B2 = G2.blk[(size_t)(im2 - 1) * nsym + g2 - 1];
    const int ms12 = (n1 - 2 * (im1 - 1)) + (n2 - 2 * (im2 - 1));
    const int im12 = (n1 + n2 - ms12) / 2 + 1;
    const GraphBlock &B12 = G12.blk[(size_t)(im12 - 1) * nsym + mult(g1, g2) - 1];
    std::vector<int64_t> m;
    m.reserve(B1.strs.size() * B2.strs.size());
    uint16_t s1[8], s2[8], s12[16];
    for (const StrKey &k2 : B2.strs) {
        unpack(k2, n2, s2);
        for (const StrKey &k1 : B1.strs) {
            unpack(k1, n1, s1);
            int i = 0, j = 0, t = 0, inv = 0;
            bool clash = false;
            while (i < n1 || j < n2) {
                if (j >= n2 || (i < n1 && s1[i] < s2[j])) s12[t++] = s1[i++];
                else {
                    if (i < n1 && s1[i] == s2[j]) { clash = true; break; }
                    inv += n1 - i;
                    s12[t++] = s2[j++];
                }
            }
            if (clash) { m.push_back(0); continue; }
            auto f = B12.index.find(make_key(s12, n1 + n2));
            int64_t idx = f == B12.index.end() ? -1 : f->second;
            if (idx < 0) { m.push_back(0); continue; }
            m.push_back((inv & 1) ? -(idx + 1) : (idx + 1));
        }
    }
    return o->maps.emplace(key, std::move(m)).first->second;
}

static int spsign(const uint16_t *s, int n) {
    int nb = 0, trp = 0;
    for (int i = 0; i < n; ++i) {
        if (s[i] & 1) ++nb;
        else trp += nb;
    }
    return (trp & 1) ? -1 : 1;
}

// string -> spin-flipped string, signed (strings/set_flipmap_kernel.f)
static const std::vector<int64_t> &flipmap(gecco_hip_orbs *o, int space, int nel, int im, int g) {
    std::array<int, 4> key{space, nel, im, g};
    auto it = o->flips.find(key);
    if (it != o->flips.end()) return it->second;
    const int nsym = o->nsym;
    const Graph &G = graph(o, space, nel);
    const GraphBlock &B = G.blk[(size_t)(im - 1) * nsym + g - 1];
    const GraphBlock &BF = G.blk[(size_t)(nel - (im - 1)) * nsym + g - 1];
    std::vector<int64_t> m;
    uint16_t s[8], f[8];
    for (const StrKey &k : B.strs) {
        unpack(k, nel, s);
        for (int i = 0; i < nel; ++i) f[i] = s[i] ^ 1;
        // restore alpha-before-beta on doubly occupied orbitals
        for (int i = 0; i + 1 < nel; ++i)
            if ((f[i] >> 1) == (f[i + 1] >> 1) && f[i] > f[i + 1]) std::swap(f[i], f[i + 1]);
        const int sg = spsign(s, nel) * spsign(f, nel);
        auto fi = BF.index.find(make_key(f, nel));
        m.push_back(sg * (int64_t)(fi->second + 1));
    }
    return o->flips.emplace(key, std::move(m)).first->second;
}

// ---------------------------------------------------------- distributions
// operators/next_msgamdist2.f: all (Ms, IRREP) distributions of a condensed
// C/A occupation, innermost C IRREP, then C Ms, then A IRREP, then A Ms.
struct DistSet {
    std::vector<std::vector<int>> ms;   // per entry
    std::vector<std::vector<int>> gam;
};

static void ms_dists(const std::vector<int> &occ, int mst, std::vector<std::vector<int>> &out) {
    const int n = (int)occ.size();
    out.clear();
    if (n == 0) {
        if (mst == 0) out.push_back({});
        return;
    }
    // first index fastest, each from occ(i) downwards; last fixed by mst
    std::vector<int> d(n);
    for (int i = 0; i < n - 1; ++i) d[i] = occ[i];
    for (;;) {
        int s = 0;
        for (int i = 0; i < n - 1; ++i) s += d[i];
        d[n - 1] = mst - s;
        if (std::abs(d[n - 1]) <= occ[n - 1]) out.push_back(d);
        int i = 0;
        for (; i < n - 1; ++i) {
            if (d[i] > -occ[i]) { d[i] -= 2; break; }
            d[i] = occ[i];
        }
        if (i == n - 1) break;
    }
}

static void gam_dists(int n, int gamt, int nsym, std::vector<std::vector<int>> &out) {
    out.clear();
    if (n == 0) {
        if (gamt == 1) out.push_back({});
        return;
    }
    std::vector<int> g(n, 1);
    for (;;) {
        int acc = gamt;
        for (int i = 0; i < n - 1; ++i) acc = mult(acc, g[i]);
        g[n - 1] = acc;
        out.push_back(g);
        int i = 0;
        for (; i < n - 1; ++i) {
            if (g[i] < nsym) { ++g[i]; break; }
            g[i] = 1;
        }
        if (i == n - 1) break;
    }
}

struct Dist {
    int64_t off = 0, len = 0;
    std::vector<int> msc, msa, gc, ga, imc, ima;
    std::vector<int64_t> lens, ldc, lda;
};

}  // namespace gecco

struct gecco_hip_list {
    gecco_hip_orbs *orbs = nullptr;
    int occ_c[5] = {0}, occ_a[5] = {0};     // summed over the vertices
    int njoined = 1;
    std::vector<std::array<int, 10>> vocc;  // per vertex: C[1..4] at 1..4, A[1..4] at 6..9
    int absym = 1;
    std::vector<int> oc, hc, oa, ha;  // condensed (hpvxblkseq order, vertex inner)
    std::vector<int> vc, va;          // vertex of every condensed component
    int nc = 0, na = 0;
    std::vector<gecco::Dist> dists;
    std::map<std::vector<int>, int> find;  // (idxmsa, igama, imc, gc, ima, ga) -> dist
    std::map<std::pair<int, int>, int64_t> blen;  // (idxmsa, igama) -> length
    int64_t length = 0;
    int tra = 0;   // adjoint view of a stored list (gecco_hip_list_adjoint)
};

namespace gecco {

static int idxms_of(int occ, int ms) { return (occ - ms) / 2 + 1; }

static int msa2idxms4op(int msa, int occ_a, int occ_c) {  // operators/msa2idxms4op.f:24-27
    int idx = (occ_a - msa) / 2 + 1;
    int off = occ_a - occ_c;
    if (off > 0) idx -= off / 2;
    return idx;
}

static void set_op_ldim(const std::vector<int> &hc, const std::vector<int> &ha,
                        const std::vector<int64_t> &lens, std::vector<int64_t> &ldc,
                        std::vector<int64_t> &lda) {
    const size_t nc = hc.size(), na = ha.size();
    ldc.assign(nc, 0);
    lda.assign(na, 0);
    int64_t ld = 1;
    for (int h : HPVXBLKSEQ) {
        for (size_t i = 0; i < nc; ++i)
            if (hc[i] == h) { ldc[i] = ld; ld *= lens[i]; }
        for (size_t i = 0; i < na; ++i)
            if (ha[i] == h) { lda[i] = ld; ld *= lens[nc + i]; }
    }
}

static std::vector<int> find_key(int idxmsa, int igama, const std::vector<int> &imc,
                                 const std::vector<int> &gc, const std::vector<int> &ima,
                                 const std::vector<int> &ga) {
    std::vector<int> k{idxmsa, igama};
    k.insert(k.end(), imc.begin(), imc.end());
    k.insert(k.end(), gc.begin(), gc.end());
    k.insert(k.end(), ima.begin(), ima.end());
    k.insert(k.end(), ga.begin(), ga.end());
    return k;
}

// operators/condense_occ.f: the vertex loop inside the hpvxblkseq loop
static void condense_n(const std::vector<std::array<int, 10>> &vocc, std::vector<int> &oc,
                       std::vector<int> &hc, std::vector<int> &vc, std::vector<int> &oa,
                       std::vector<int> &ha, std::vector<int> &va) {
    for (int h : HPVXBLKSEQ)
        for (size_t v = 0; v < vocc.size(); ++v) {
            if (vocc[v][h] > 0) { oc.push_back(vocc[v][h]); hc.push_back(h); vc.push_back((int)v); }
            if (vocc[v][5 + h] > 0) { oa.push_back(vocc[v][5 + h]); ha.push_back(h); va.push_back((int)v); }
        }
}

static void build_list(gecco_hip_list *L) {
    gecco_hip_orbs *o = L->orbs;
    const int nsym = o->nsym;
    condense_n(L->vocc, L->oc, L->hc, L->vc, L->oa, L->ha, L->va);
    for (int x : L->oc) L->nc += x;
    for (int x : L->oa) L->na += x;
    int64_t idx = 0;
    int idxmsa = 0;
    std::vector<std::vector<int>> mcs, mas, gcs, gas;
    for (int msa = L->na; msa >= -L->na; msa -= 2) {
        const int msc = msa;
        if (std::abs(msc) > L->nc) continue;
        ++idxmsa;
        for (int ga_t = 1; ga_t <= nsym; ++ga_t) {
            const int gc_t = ga_t;
            const int64_t start = idx;
            ms_dists(L->oa, msa, mas);
            gam_dists((int)L->oa.size(), ga_t, nsym, gas);
            ms_dists(L->oc, msc, mcs);
            gam_dists((int)L->oc.size(), gc_t, nsym, gcs);
            for (auto &ma : mas)
                for (auto &ga : gas)
                    for (auto &mc : mcs)
                        for (auto &gc : gcs) {
                            Dist d;
                            d.msc = mc; d.msa = ma; d.gc = gc; d.ga = ga;
                            int64_t len = 1;
                            for (size_t i = 0; i < L->oc.size(); ++i) {
                                d.imc.push_back(idxms_of(L->oc[i], mc[i]));
                                d.lens.push_back(lenstr(o, L->hc[i], L->oc[i], d.imc.back(), gc[i]));
                                len *= d.lens.back();
                            }
                            for (size_t i = 0; i < L->oa.size(); ++i) {
                                d.ima.push_back(idxms_of(L->oa[i], ma[i]));
                                d.lens.push_back(lenstr(o, L->ha[i], L->oa[i], d.ima.back(), ga[i]));
                                len *= d.lens.back();
                            }
                            if (len <= 0) continue;
                            set_op_ldim(L->hc, L->ha, d.lens, d.ldc, d.lda);
                            d.off = idx;
                            d.len = len;
                            L->find[find_key(idxmsa, ga_t, d.imc, d.gc, d.ima, d.ga)] = (int)L->dists.size();
                            L->dists.push_back(std::move(d));
                            idx += len;
                        }
            L->blen[{idxmsa, ga_t}] = idx - start;
        }
    }
    L->length = idx;
}

static const Dist *list_find(const gecco_hip_list *L, int msa_tot, int igama,
                             const std::vector<int> &imc, const std::vector<int> &gc,
                             const std::vector<int> &ima, const std::vector<int> &ga) {
    int idxmsa = msa2idxms4op(msa_tot, L->na, L->nc);
    auto it = L->find.find(find_key(idxmsa, igama, imc, gc, ima, ga));
    return it == L->find.end() ? nullptr : &L->dists[it->second];
}

static int64_t block_len(const gecco_hip_list *L, int msa_tot, int igama) {
    int idxmsa = msa2idxms4op(msa_tot, L->na, L->nc);
    auto it = L->blen.find({idxmsa, igama});
    return it == L->blen.end() ? 0 : it->second;
}

// --------------------------------------------------------- the enumerator
struct Side {  // condensed occupation of one string part (EX1 C, CNT A, ...)
    std::vector<int> occ, hpvx;
};

static Side side_of(const int *occ) {
    Side s;
    for (int h : HPVXBLKSEQ)
        if (occ[h] > 0) { s.occ.push_back(occ[h]); s.hpvx.push_back(h); }
    return s;
}

static std::vector<int64_t> map_info(const std::vector<int> &tgt, const Side &s1, const Side &s2) {
    std::vector<int64_t> mi;
    for (int h : tgt) {
        std::vector<int64_t> i1, i2;
        for (size_t i = 0; i < s1.hpvx.size(); ++i)
            if (s1.hpvx[i] == h) i1.push_back((int64_t)i + 1);
        for (size_t i = 0; i < s2.hpvx.size(); ++i)
            if (s2.hpvx[i] == h) i2.push_back((int64_t)i + 1);
        mi.push_back((int64_t)i1.size());
        mi.insert(mi.end(), i1.begin(), i1.end());
        mi.push_back((int64_t)i2.size());
        mi.insert(mi.end(), i2.begin(), i2.end());
    }
    return mi;
}

struct DistStr {  // one distribution of a string part
    std::vector<int> ms, gam, idxms;
    std::vector<int64_t> len;
};

static void merge_dis(const std::vector<int64_t> &mi, size_t ntgt, const DistStr &d1,
                      const DistStr &d2, std::vector<int> &ms, std::vector<int> &gam) {
    ms.assign(ntgt, 0);
    gam.assign(ntgt, 1);
    size_t k = 0;
    for (size_t t = 0; t < ntgt; ++t) {
        int64_t n = mi[k++];
        for (int64_t i = 0; i < n; ++i, ++k) {
            ms[t] += d1.ms[mi[k] - 1];
            gam[t] = mult(gam[t], d1.gam[mi[k] - 1]);
        }
        n = mi[k++];
        for (int64_t i = 0; i < n; ++i, ++k) {
            ms[t] += d2.ms[mi[k] - 1];
            gam[t] = mult(gam[t], d2.gam[mi[k] - 1]);
        }
    }
}

// strings/get_strmap_blk_c.f:68-200 for one-vertex operators
static void strmap_blk(gecco_hip_orbs *o, const std::vector<int> &tgt_h, const std::vector<int64_t> &mi,
                       const Side &s1, const DistStr &d1, const Side &s2, const DistStr &d2,
                       std::vector<int64_t> &out) {
    out.clear();
    size_t k = 0;
    for (size_t t = 0; t < tgt_h.size(); ++t) {
        int64_t n1 = mi[k++];
        int64_t i1 = n1 ? mi[k] - 1 : -1;
        k += n1;
        int64_t n2 = mi[k++];
        int64_t i2 = n2 ? mi[k] - 1 : -1;
        k += n2;
        if (n1 && n2) {
            const auto &m = strmap(o, tgt_h[t], s1.occ[i1], d1.idxms[i1], d1.gam[i1], s2.occ[i2],
                                   d2.idxms[i2], d2.gam[i2]);
            out.insert(out.end(), m.begin(), m.end());
        } else {
            int64_t n = n1 ? d1.len[i1] : d2.len[i2];
            for (int64_t i = 1; i <= n; ++i) out.push_back(i);
        }
    }
}

static void dist_strings(gecco_hip_orbs *o, const Side &c, const Side &a, int msc, int msa, int gc,
                         int ga, std::vector<DistStr> &out) {
    out.clear();
    const int nsym = o->nsym;
    std::vector<std::vector<int>> mcs, mas, gcs, gas;
    ms_dists(a.occ, msa, mas);
    gam_dists((int)a.occ.size(), ga, nsym, gas);
    ms_dists(c.occ, msc, mcs);
    gam_dists((int)c.occ.size(), gc, nsym, gcs);
    for (auto &ma : mas)
        for (auto &g_a : gas)
            for (auto &mc : mcs)
                for (auto &g_c : gcs) {
                    DistStr d;
                    bool zero = false;
                    d.ms = mc;
                    d.ms.insert(d.ms.end(), ma.begin(), ma.end());
                    d.gam = g_c;
                    d.gam.insert(d.gam.end(), g_a.begin(), g_a.end());
                    for (size_t i = 0; i < c.occ.size(); ++i) {
                        d.idxms.push_back(idxms_of(c.occ[i], mc[i]));
                        d.len.push_back(lenstr(o, c.hpvx[i], c.occ[i], d.idxms.back(), g_c[i]));
                        zero |= d.len.back() == 0;
                    }
                    for (size_t i = 0; i < a.occ.size(); ++i) {
                        d.idxms.push_back(idxms_of(a.occ[i], ma[i]));
                        d.len.push_back(lenstr(o, a.hpvx[i], a.occ[i], d.idxms.back(), g_a[i]));
                        zero |= d.len.back() == 0;
                    }
                    if (!zero) out.push_back(std::move(d));
                }
}

static DistStr part(const DistStr &d, size_t from, size_t n) {
    DistStr r;
    r.ms.assign(d.ms.begin() + from, d.ms.begin() + from + n);
    r.gam.assign(d.gam.begin() + from, d.gam.begin() + from + n);
    r.idxms.assign(d.idxms.begin() + from, d.idxms.begin() + from + n);
    r.len.assign(d.len.begin() + from, d.len.begin() + from + n);
    return r;
}

struct ContrCtx {
    gecco_hip_plan *plan;
    gecco_hip_orbs *o;
    const gecco_hip_list *L1, *L2, *L12;
    Side ex1c, ex1a, ex2c, ex2a, cntc, cnta;
    std::vector<int64_t> mi12c, mi12a, mi1c, mi1a, mi2c, mi2a;
    double fact;
    int32_t buf1, buf2, buf12;
    int64_t off1, off2, off12;
    int64_t nblocks = 0;
    double flops = 0;
};

static int emit(ContrCtx &c, double xfac, const Dist &d1, const Dist &d2, const Dist &d12,
                const DistStr &e1, const DistStr &e2, const DistStr &cn) {
    std::vector<int64_t> m12c, m12a, m1c, m1a, m2c, m2a;
    const size_t n1c = c.ex1c.occ.size(), n2c = c.ex2c.occ.size(), ncc = c.cntc.occ.size();
    const DistStr e1c = part(e1, 0, n1c), e1a = part(e1, n1c, c.ex1a.occ.size());
    const DistStr e2c = part(e2, 0, n2c), e2a = part(e2, n2c, c.ex2a.occ.size());
    const DistStr cc = part(cn, 0, ncc), ca = part(cn, ncc, c.cnta.occ.size());
    strmap_blk(c.o, c.L12->hc, c.mi12c, c.ex1c, e1c, c.ex2c, e2c, m12c);
    strmap_blk(c.o, c.L12->ha, c.mi12a, c.ex2a, e2a, c.ex1a, e1a, m12a);
    strmap_blk(c.o, c.L1->hc, c.mi1c, c.cntc, cc, c.ex1c, e1c, m1c);
    strmap_blk(c.o, c.L1->ha, c.mi1a, c.cnta, ca, c.ex1a, e1a, m1a);
    strmap_blk(c.o, c.L2->hc, c.mi2c, c.cnta, ca, c.ex2c, e2c, m2c);
    strmap_blk(c.o, c.L2->ha, c.mi2a, c.cntc, cc, c.ex2a, e2a, m2a);
    auto v64 = [](const std::vector<int> &v) { return std::vector<int64_t>(v.begin(), v.end()); };
    std::vector<int64_t> h1c = v64(c.L1->hc), h1a = v64(c.L1->ha), h2c = v64(c.L2->hc),
                         h2a = v64(c.L2->ha), h12c = v64(c.L12->hc), h12a = v64(c.L12->ha);
    std::vector<int64_t> zero{0};
    auto P = [&](const std::vector<int64_t> &v) { return v.empty() ? zero.data() : v.data(); };
    gecco_hip_block_args a;
    std::memset(&a, 0, sizeof a);
    a.xfac = xfac;
    a.tra_op1 = c.L1->tra; a.tra_op2 = c.L2->tra; a.tra_op1op2 = c.L12->tra;
    a.ncblk_op1 = (int64_t)c.L1->hc.size(); a.nablk_op1 = (int64_t)c.L1->ha.size();
    a.ncblk_ex1 = (int64_t)n1c; a.nablk_ex1 = (int64_t)c.ex1a.occ.size();
    a.ncblk_op2 = (int64_t)c.L2->hc.size(); a.nablk_op2 = (int64_t)c.L2->ha.size();
    a.ncblk_ex2 = (int64_t)n2c; a.nablk_ex2 = (int64_t)c.ex2a.occ.size();
    a.ncblk_cnt = (int64_t)ncc; a.nablk_cnt = (int64_t)c.cnta.occ.size();
    a.ncblk_op1op2 = (int64_t)c.L12->hc.size(); a.nablk_op1op2 = (int64_t)c.L12->ha.size();
    a.hpvxop1c = P(h1c); a.hpvxop1a = P(h1a); a.hpvxop2c = P(h2c); a.hpvxop2a = P(h2a);
    a.hpvxop1op2c = P(h12c); a.hpvxop1op2a = P(h12a);
    a.lstrop1 = P(d1.lens); a.lstrop2 = P(d2.lens); a.lstrop1op2 = P(d12.lens);
    a.lstr_ex1 = P(e1.len); a.lstr_ex2 = P(e2.len); a.lstr_cnt = P(cn.len);
    a.map_info_12c = P(c.mi12c); a.map_info_12a = P(c.mi12a);
    a.map_info_1c = P(c.mi1c); a.map_info_1a = P(c.mi1a);
    a.map_info_2c = P(c.mi2c); a.map_info_2a = P(c.mi2a);
    a.map_ex1ex2c = P(m12c); a.map_ex1ex2a = P(m12a);
    a.map_ex1cntc = P(m1c); a.map_ex1cnta = P(m1a);
    a.map_ex2cntc = P(m2c); a.map_ex2cnta = P(m2a);
    a.buf_op1 = c.buf1; a.buf_op2 = c.buf2; a.buf_op1op2 = c.buf12;
    a.off_op1 = c.off1 + d1.off; a.off_op2 = c.off2 + d2.off; a.off_op1op2 = c.off12 + d12.off;
    int rc = plan_add_block_internal(c.plan, &a);
    if (rc) return rc;
    ++c.nblocks;
    double n = 2.0;
    for (auto x : e1.len) n *= (double)x;
    for (auto x : e2.len) n *= (double)x;
    for (auto x : cn.len) n *= (double)x;
    c.flops += n;
    return 0;
}

static int enumerate(ContrCtx &c, bool use_tr) {
    const gecco_hip_list *L1 = c.L1, *L2 = c.L2, *L12 = c.L12;
    gecco_hip_orbs *o = c.o;
    const int nsym = o->nsym;
    const int na[3] = {L1->na, L2->na, L12->na}, nc[3] = {L1->nc, L2->nc, L12->nc};
    int nc_cnt = 0, na_cnt = 0, nc_ex1 = 0, na_ex1 = 0, nc_ex2 = 0, na_ex2 = 0;
    for (int x : c.cntc.occ) nc_cnt += x;
    for (int x : c.cnta.occ) na_cnt += x;
    for (int x : c.ex1c.occ) nc_ex1 += x;
    for (int x : c.ex1a.occ) na_ex1 += x;
    for (int x : c.ex2c.occ) nc_ex2 += x;
    for (int x : c.ex2a.occ) na_ex2 += x;
    const bool tr = use_tr && L1->absym != 0 && L2->absym != 0 && L12->absym != 0 &&
                    L1->absym * L2->absym == L12->absym;
    const int idx_restr = na[1] > na[0] ? 2 : 1;
    std::vector<DistStr> D1, D2, DC;
    std::vector<int> msr, gr, msr_a, gr_a, imr, ima, m1c, g1c, m1a, g1a, m2c, g2c, m2a, g2a;
    for (int ms3 = na[2]; ms3 >= -na[2]; ms3 -= 2)
        for (int ms2 = na[1]; ms2 >= -na[1]; ms2 -= 2)
            for (int ms1 = na[0]; ms1 >= -na[0]; ms1 -= 2) {
                const int msa[3] = {ms1, ms2, ms3};
                if (std::abs(ms1) > nc[0] || std::abs(ms2) > nc[1] || std::abs(ms3) > nc[2]) continue;
                const int msc_ac = ms1 + ms2 - ms3;
                if (na_cnt + nc_cnt < std::abs(msc_ac)) continue;
                const int mr = msa[idx_restr - 1];
                if (tr && (ms3 < 0 || (ms3 == 0 && mr < 0))) continue;
                double fac0 = 1.0;
                if (tr && ms3 > 0) fac0 *= 2.0;
                if (tr && ms3 == 0 && mr > 0) fac0 *= 2.0;
                for (int msc_a = na_cnt; msc_a >= -na_cnt; msc_a -= 2) {
                    const int msc_c = msc_ac - msc_a;
                    if (std::abs(msc_c) > nc_cnt) continue;
                    const int mex1c = ms1 - msc_c, mex1a = ms1 - msc_a;
                    const int mex2a = ms2 - msc_c, mex2c = ms2 - msc_a;
                    if (std::abs(mex1c) > nc_ex1 || std::abs(mex1a) > na_ex1 ||
                        std::abs(mex2c) > nc_ex2 || std::abs(mex2a) > na_ex2) continue;
                    if (tr && mr == 0 && ms3 == 0 && msc_c - msc_a > 0) continue;
                    double fac = fac0;
                    if (tr && mr == 0 && ms3 == 0 && msc_c - msc_a < 0) fac *= 2.0;
                    for (int g3 = 1; g3 <= nsym; ++g3)
                        for (int g2 = 1; g2 <= nsym; ++g2)
                            for (int g1 = 1; g1 <= nsym; ++g1) {
                                if (!block_len(L1, ms1, g1) || !block_len(L2, ms2, g2) ||
                                    !block_len(L12, ms3, g3)) continue;
                                const int gcac = mult(mult(g1, g2), g3);
                                for (int gca = 1; gca <= nsym; ++gca) {
                                    const int gcc = mult(gca, gcac);
                                    if (c.cntc.occ.empty() && gcc > 1) continue;
                                    if (c.cnta.occ.empty() && gca > 1) continue;
                                    const int gx1a = mult(g1, gca), gx1c = mult(g1, gcc);
                                    const int gx2a = mult(g2, gcc), gx2c = mult(g2, gca);
                                    if ((c.ex1c.occ.empty() && gx1c > 1) || (c.ex1a.occ.empty() && gx1a > 1) ||
                                        (c.ex2c.occ.empty() && gx2c > 1) || (c.ex2a.occ.empty() && gx2a > 1))
                                        continue;
                                    dist_strings(o, c.ex1c, c.ex1a, mex1c, mex1a, gx1c, gx1a, D1);
                                    dist_strings(o, c.ex2c, c.ex2a, mex2c, mex2a, gx2c, gx2a, D2);
                                    dist_strings(o, c.cntc, c.cnta, msc_c, msc_a, gcc, gca, DC);
                                    const size_t n1c = c.ex1c.occ.size(), n2c = c.ex2c.occ.size();
                                    const size_t ncc = c.cntc.occ.size();
                                    for (const DistStr &e1 : D1) {
                                        const DistStr e1c = part(e1, 0, n1c), e1a = part(e1, n1c, c.ex1a.occ.size());
                                        for (const DistStr &e2 : D2) {
                                            const DistStr e2c = part(e2, 0, n2c), e2a = part(e2, n2c, c.ex2a.occ.size());
                                            merge_dis(c.mi12c, L12->hc.size(), e1c, e2c, msr, gr);
                                            merge_dis(c.mi12a, L12->ha.size(), e2a, e1a, msr_a, gr_a);
                                            imr.clear(); ima.clear();
                                            for (size_t i = 0; i < L12->oc.size(); ++i) imr.push_back(idxms_of(L12->oc[i], msr[i]));
                                            for (size_t i = 0; i < L12->oa.size(); ++i) ima.push_back(idxms_of(L12->oa[i], msr_a[i]));
                                            const Dist *d12 = list_find(L12, ms3, g3, imr, gr, ima, gr_a);
                                            if (!d12) continue;
                                            for (const DistStr &cn : DC) {
                                                const DistStr cc = part(cn, 0, ncc), ca = part(cn, ncc, c.cnta.occ.size());
                                                merge_dis(c.mi1c, L1->hc.size(), cc, e1c, m1c, g1c);
                                                merge_dis(c.mi1a, L1->ha.size(), ca, e1a, m1a, g1a);
                                                merge_dis(c.mi2c, L2->hc.size(), ca, e2c, m2c, g2c);
                                                merge_dis(c.mi2a, L2->ha.size(), cc, e2a, m2a, g2a);
                                                std::vector<int> i1c, i1a, i2c, i2a;
                                                for (size_t i = 0; i < L1->oc.size(); ++i) i1c.push_back(idxms_of(L1->oc[i], m1c[i]));
                                                for (size_t i = 0; i < L1->oa.size(); ++i) i1a.push_back(idxms_of(L1->oa[i], m1a[i]));
                                                for (size_t i = 0; i < L2->oc.size(); ++i) i2c.push_back(idxms_of(L2->oc[i], m2c[i]));
                                                for (size_t i = 0; i < L2->oa.size(); ++i) i2a.push_back(idxms_of(L2->oa[i], m2a[i]));
                                                const Dist *d1 = list_find(L1, ms1, g1, i1c, g1c, i1a, g1a);
                                                const Dist *d2 = list_find(L2, ms2, g2, i2c, g2c, i2a, g2a);
                                                if (!d1 || !d2) continue;
                                                int rc = emit(c, c.fact * fac, *d1, *d2, *d12, e1, e2, cn);
                                                if (rc) return rc;
                                            }
                                        }
                                    }
                                }
                            }
                }
            }
    return 0;
}

// ------------------------------------------------------ the partial trace
// Sources (J component, OPtr component; -1 = none) of every target
// component of a map_info array (contract/set_strmapdim_c2.f layout).
static std::vector<std::pair<int, int>> map_sources(const std::vector<int64_t> &mi, size_t ntgt) {
    std::vector<std::pair<int, int>> out;
    size_t k = 0;
    for (size_t t = 0; t < ntgt; ++t) {
        int64_t n1 = mi[k++];
        int a = n1 ? (int)mi[k] - 1 : -1;
        k += n1;
        int64_t n2 = mi[k++];
        int b = n2 ? (int)mi[k] - 1 : -1;
        k += n2;
        out.push_back({a, b});
    }
    return out;
}

static void split_index(int64_t idx, const std::vector<int64_t> &lens, std::vector<int64_t> &out) {
    out.resize(lens.size());
    for (size_t i = 0; i < lens.size(); ++i) {
        out[i] = idx % lens[i];
        idx /= lens[i];
    }
}

// (OPtr element, OP element, sign) of OPtr(IC,IA) += sum_J OP(J.IC, J.IA),
// J the same string on the C and A sides: the loops of
// contract/trace_op.f:460-700 (Ms and IRREP pairs of (OP(A), OPtr(A)),
// distributions of OPtr, distributions of J with IRREP innermost as
// operators/next_msgamdist_diag.f) and of contract/trace_op_blk.f:80-150.
static int trace_triples(const gecco_hip_list *Lop, const gecco_hip_list *Ltr, const int *cnt5,
                         std::vector<int64_t> &dst, std::vector<int64_t> &src,
                         std::vector<int8_t> &sg) {
    gecco_hip_orbs *o = Lop->orbs;
    const int nsym = o->nsym;
    const Side cnt = side_of(cnt5);
    Side trc, tra;
    trc.occ = Ltr->oc; trc.hpvx = Ltr->hc;
    tra.occ = Ltr->oa; tra.hpvx = Ltr->ha;
    const std::vector<int64_t> mic = map_info(Lop->hc, cnt, trc), mia = map_info(Lop->ha, cnt, tra);
    const auto srcc = map_sources(mic, Lop->hc.size()), srca = map_sources(mia, Lop->ha.size());
    int na_cnt = 0;
    for (int x : cnt.occ) na_cnt += x;
    const size_t ntc = Ltr->oc.size(), nta = Ltr->oa.size();
    std::vector<std::vector<int>> mcs, gcs;
    std::vector<int> mc, gc, ma, ga, imc, ima;
    std::vector<int64_t> mapc, mapa, sa, sc, sj;
    for (int ms_tr = Ltr->na; ms_tr >= -Ltr->na; ms_tr -= 2)
        for (int ms_op = Lop->na; ms_op >= -Lop->na; ms_op -= 2) {
            if (std::abs(ms_op) > Lop->nc || std::abs(ms_tr) > Ltr->nc) continue;
            const int msc = ms_op - ms_tr;
            if (std::abs(msc) > na_cnt) continue;
            for (int g_tr = 1; g_tr <= nsym; ++g_tr)
                for (int g_op = 1; g_op <= nsym; ++g_op) {
                    const int gcnt = mult(g_op, g_tr);
                    if (!block_len(Lop, ms_op, g_op)) continue;
                    ms_dists(cnt.occ, msc, mcs);
                    gam_dists((int)cnt.occ.size(), gcnt, nsym, gcs);
                    for (const Dist &dt : Ltr->dists) {
                        int mtot = 0, gtot = 1;
                        for (int x : dt.msa) mtot += x;
                        for (int x : dt.ga) gtot = mult(gtot, x);
                        if (mtot != ms_tr || gtot != g_tr) continue;
                        DistStr tC, tA;
                        tC.ms = dt.msc; tC.gam = dt.gc; tC.idxms = dt.imc;
                        tC.len.assign(dt.lens.begin(), dt.lens.begin() + ntc);
                        tA.ms = dt.msa; tA.gam = dt.ga; tA.idxms = dt.ima;
                        tA.len.assign(dt.lens.begin() + ntc, dt.lens.end());
                        for (const auto &mj : mcs)
                            for (const auto &gj : gcs) {
                                DistStr cn;
                                cn.ms = mj;
                                cn.gam = gj;
                                bool zero = false;
                                for (size_t i = 0; i < cnt.occ.size(); ++i) {
                                    cn.idxms.push_back(idxms_of(cnt.occ[i], mj[i]));
                                    cn.len.push_back(lenstr(o, cnt.hpvx[i], cnt.occ[i], cn.idxms.back(), gj[i]));
                                    zero |= cn.len.back() == 0;
                                }
                                if (zero) continue;
                                merge_dis(mic, Lop->hc.size(), cn, tC, mc, gc);
                                merge_dis(mia, Lop->ha.size(), cn, tA, ma, ga);
                                imc.clear(); ima.clear();
                                for (size_t i = 0; i < Lop->oc.size(); ++i) imc.push_back(idxms_of(Lop->oc[i], mc[i]));
                                for (size_t i = 0; i < Lop->oa.size(); ++i) ima.push_back(idxms_of(Lop->oa[i], ma[i]));
                                const Dist *dop = list_find(Lop, ms_op, g_op, imc, gc, ima, ga);
                                if (!dop) continue;
                                strmap_blk(o, Lop->hc, mic, cnt, cn, trc, tC, mapc);
                                strmap_blk(o, Lop->ha, mia, cnt, cn, tra, tA, mapa);
                                // per OP component: map block offset and source lengths
                                std::vector<int64_t> offc, offa, n1c, n1a;
                                int64_t acc = 0;
                                for (auto &pr : srcc) {
                                    offc.push_back(acc);
                                    const int64_t n1 = pr.first >= 0 ? cn.len[pr.first] : 1;
                                    const int64_t n2 = pr.second >= 0 ? tC.len[pr.second] : 1;
                                    n1c.push_back(n1);
                                    acc += n1 * n2;
                                }
                                acc = 0;
                                for (auto &pr : srca) {
                                    offa.push_back(acc);
                                    const int64_t n1 = pr.first >= 0 ? cn.len[pr.first] : 1;
                                    const int64_t n2 = pr.second >= 0 ? tA.len[pr.second] : 1;
                                    n1a.push_back(n1);
                                    acc += n1 * n2;
                                }
                                int64_t n_tra = 1, n_trc = 1, n_cnt = 1;
                                for (auto x : tA.len) n_tra *= x;
                                for (auto x : tC.len) n_trc *= x;
                                for (auto x : cn.len) n_cnt *= x;
                                for (int64_t ita = 0; ita < n_tra; ++ita) {
                                    split_index(ita, tA.len, sa);
                                    for (int64_t icn = 0; icn < n_cnt; ++icn) {
                                        split_index(icn, cn.len, sj);
                                        int sgn_a = 1;
                                        int64_t off_a = 0;
                                        bool ok = true;
                                        for (size_t t = 0; t < srca.size() && ok; ++t) {
                                            const int64_t i1 = srca[t].first >= 0 ? sj[srca[t].first] : 0;
                                            const int64_t i2 = srca[t].second >= 0 ? sa[srca[t].second] : 0;
                                            const int64_t e = mapa[offa[t] + i1 + n1a[t] * i2];
                                            if (!e) { ok = false; break; }
                                            if (e < 0) sgn_a = -sgn_a;
                                            off_a += (std::abs(e) - 1) * dop->lda[t];
                                        }
                                        if (!ok) continue;
                                        int64_t tr_a = dt.off;
                                        for (size_t i = 0; i < nta; ++i) tr_a += sa[i] * dt.lda[i];
                                        for (int64_t itc = 0; itc < n_trc; ++itc) {
                                            split_index(itc, tC.len, sc);
                                            int sgn = sgn_a;
                                            int64_t off = dop->off + off_a;
                                            bool okc = true;
                                            for (size_t t = 0; t < srcc.size(); ++t) {
                                                const int64_t i1 = srcc[t].first >= 0 ? sj[srcc[t].first] : 0;
                                                const int64_t i2 = srcc[t].second >= 0 ? sc[srcc[t].second] : 0;
                                                const int64_t e = mapc[offc[t] + i1 + n1c[t] * i2];
                                                if (!e) { okc = false; break; }
                                                if (e < 0) sgn = -sgn;
                                                off += (std::abs(e) - 1) * dop->ldc[t];
                                            }
                                            if (!okc) continue;
                                            int64_t tr = tr_a;
                                            for (size_t i = 0; i < ntc; ++i) tr += sc[i] * dt.ldc[i];
                                            dst.push_back(tr);
                                            src.push_back(off);
                                            sg.push_back((int8_t)sgn);
                                        }
                                    }
                                }
                            }
                    }
                }
        }
    return 0;
}

// ------------------------------------------------------------ the reorder
// map_info of one side of a reorder: per target component (vertex v,
// space h) the I0 component with the same (v, h) and, on vertex kv, the K
// component of space h; k_first puts K as source 1 (A side:
// contract/reo_blk_wmaps_c.f:517-548 builds the A maps from (K, I0)).
static std::vector<int64_t> reo_map_info(const std::vector<int> &th, const std::vector<int> &tv,
                                         const std::vector<int> &i0h, const std::vector<int> &i0v,
                                         const std::vector<int> &kh, int kv, bool k_first) {
    std::vector<int64_t> mi;
    for (size_t t = 0; t < th.size(); ++t) {
        std::vector<int64_t> i0, k;
        for (size_t j = 0; j < i0h.size(); ++j)
            if (i0h[j] == th[t] && i0v[j] == tv[t]) i0.push_back((int64_t)j + 1);
        if (tv[t] == kv)
            for (size_t j = 0; j < kh.size(); ++j)
                if (kh[j] == th[t]) k.push_back((int64_t)j + 1);
        const auto &s1 = k_first ? k : i0;
        const auto &s2 = k_first ? i0 : k;
        mi.push_back((int64_t)s1.size());
        mi.insert(mi.end(), s1.begin(), s1.end());
        mi.push_back((int64_t)s2.size());
        mi.insert(mi.end(), s2.begin(), s2.end());
    }
    return mi;
}

// contract/split_msgmdis.f for one side: the Ms/IRREP of the unknown (I0)
// components from the target distribution and the known (K) ones
static bool split_dis(const std::vector<int64_t> &mi, size_t ntgt, const std::vector<int> &ms_t,
                      const std::vector<int> &g_t, const DistStr &known, bool known_first,
                      size_t n_unknown, std::vector<int> &ms, std::vector<int> &gam) {
    ms.assign(n_unknown, INT32_MIN);
    gam.assign(n_unknown, 0);
    size_t k = 0;
    for (size_t t = 0; t < ntgt; ++t) {
        std::vector<int64_t> s[2];
        for (int q = 0; q < 2; ++q) {
            const int64_t n = mi[k++];
            for (int64_t i = 0; i < n; ++i) s[q].push_back(mi[k++] - 1);
        }
        const auto &kn = known_first ? s[0] : s[1];
        const auto &un = known_first ? s[1] : s[0];
        int m = ms_t[t], g = g_t[t];
        for (auto j : kn) { m -= known.ms[j]; g = mult(g, known.gam[j]); }
        if (un.empty()) {
            if (m != 0 || g != 1) return false;
            continue;
        }
        const int64_t j = un[0];
        if (ms[j] != INT32_MIN && (ms[j] != m || gam[j] != g)) return false;
        ms[j] = m;
        gam[j] = g;
    }
    for (auto x : ms)
        if (x == INT32_MIN) return false;
    return true;
}

// (REO element, ORI element, sign) of REO(I') += sign_reo s_ori s_reo ORI(I)
// for the shift string K moving from vertex frm of ORI to vertex to of REO:
// the loops of contract/reo_op_wmaps_c.f:380-560 (ORI distributions) and
// contract/reo_blk_wmaps_c.f:210-840 (Ms and IRREP of K, K distributions,
// split/merge, string loops K(A), I0(A), K(C), I0(C)).
static int reorder_triples(const gecco_hip_list *Lo, const gecco_hip_list *Lr, int frm, int to,
                           const int *kc5, const int *ka5, int sign_reo,
                           std::vector<int64_t> &dst, std::vector<int64_t> &src,
                           std::vector<int8_t> &sg) {
    gecco_hip_orbs *o = Lo->orbs;
    const int nsym = o->nsym;
    // I0 = ORI without K on vertex frm; REO must be I0 with K on vertex to
    std::vector<std::array<int, 10>> i0 = Lo->vocc, chk;
    for (int h = 1; h <= 4; ++h) {
        i0[frm][h] -= kc5[h];
        i0[frm][5 + h] -= ka5[h];
        if (i0[frm][h] < 0 || i0[frm][5 + h] < 0)
            return set_error(GECCO_HIP_EINVAL, "add_reorder: K is not part of the from-vertex");
    }
    chk = i0;
    if ((int)chk.size() < Lr->njoined) chk.resize(Lr->njoined, std::array<int, 10>{});
    if (to >= (int)chk.size()) return set_error(GECCO_HIP_EINVAL, "add_reorder: to-vertex out of range");
    for (int h = 1; h <= 4; ++h) { chk[to][h] += kc5[h]; chk[to][5 + h] += ka5[h]; }
    // REO's vertices: those of I0 with K on vertex `to` (numbered as ORI's);
    // when REO has fewer, the vertices left empty are dropped (an empty
    // vertex holds no component, so the layout is unchanged)
    std::vector<int> vmap(chk.size(), -1);
    int nkeep = 0;
    for (size_t v = 0; v < chk.size(); ++v) {
        bool empty = true;
        for (int x : chk[v]) empty = empty && x == 0;
        if ((int)chk.size() == Lr->njoined || !empty) vmap[v] = nkeep++;
    }
    if (nkeep != Lr->njoined) return set_error(GECCO_HIP_EINVAL, "add_reorder: vertex counts differ");
    for (size_t v = 0; v < chk.size(); ++v)
        if (vmap[v] >= 0)
            for (int x = 0; x < 10; ++x)
                if (chk[v][x] != Lr->vocc[vmap[v]][x])
                    return set_error(GECCO_HIP_EINVAL, "add_reorder: REO is not I0 with K on the to-vertex");
    std::vector<int> i0c_o, i0c_h, i0c_v, i0a_o, i0a_h, i0a_v;
    condense_n(i0, i0c_o, i0c_h, i0c_v, i0a_o, i0a_h, i0a_v);
    std::vector<int> i0c_vr(i0c_v), i0a_vr(i0a_v);   // I0's vertices in REO's numbering
    for (int &v : i0c_vr) v = vmap[v];
    for (int &v : i0a_vr) v = vmap[v];
    Side kc = side_of(kc5), ka = side_of(ka5);
    Side i0c, i0a;
    i0c.occ = i0c_o; i0c.hpvx = i0c_h;
    i0a.occ = i0a_o; i0a.hpvx = i0a_h;
    const auto mi_oc = reo_map_info(Lo->hc, Lo->vc, i0c_h, i0c_v, kc.hpvx, frm, false);
    const auto mi_oa = reo_map_info(Lo->ha, Lo->va, i0a_h, i0a_v, ka.hpvx, frm, true);
    const auto mi_rc = reo_map_info(Lr->hc, Lr->vc, i0c_h, i0c_vr, kc.hpvx, vmap[to], false);
    const auto mi_ra = reo_map_info(Lr->ha, Lr->va, i0a_h, i0a_vr, ka.hpvx, vmap[to], true);
    const auto so_c = map_sources(mi_oc, Lo->hc.size()), so_a = map_sources(mi_oa, Lo->ha.size());
    const auto sr_c = map_sources(mi_rc, Lr->hc.size()), sr_a = map_sources(mi_ra, Lr->ha.size());
    int kc_max = 0, ka_max = 0, i0c_max = 0, i0a_max = 0;
    for (int x : kc.occ) kc_max += x;
    for (int x : ka.occ) ka_max += x;
    for (int x : i0c_o) i0c_max += x;
    for (int x : i0a_o) i0a_max += x;
    std::vector<DistStr> DK;
    std::vector<int> mc, gc, ma, ga, rmc, rgc, rma, rga, imc, ima;
    std::vector<int64_t> m_oc, m_oa, m_rc, m_ra, ska, sia, skc, sic;
    // per target component of a map: (source-1 comp, source-2 comp, n1, offset)
    struct Tg { int a, b; int64_t n1, off; };
    auto targets = [](const std::vector<std::pair<int, int>> &srcs, const DistStr &d1,
                      const DistStr &d2) {
        std::vector<Tg> t;
        int64_t acc = 0;
        for (auto &pr : srcs) {
            const int64_t n1 = pr.first >= 0 ? d1.len[pr.first] : 1;
            const int64_t n2 = pr.second >= 0 ? d2.len[pr.second] : 1;
            t.push_back({pr.first, pr.second, n1, acc});
            acc += n1 * n2;
        }
        return t;
    };
    auto addr = [](const std::vector<Tg> &tg, const std::vector<int64_t> &map,
                   const std::vector<int64_t> &ld, const std::vector<int64_t> &i1,
                   const std::vector<int64_t> &i2, int &sgn, int64_t &off) {
        for (size_t t = 0; t < tg.size(); ++t) {
            const int64_t x1 = tg[t].a >= 0 ? i1[tg[t].a] : 0;
            const int64_t x2 = tg[t].b >= 0 ? i2[tg[t].b] : 0;
            const int64_t e = map[tg[t].off + x1 + tg[t].n1 * x2];
            if (!e) return false;
            if (e < 0) sgn = -sgn;
            off += (std::abs(e) - 1) * ld[t];
        }
        return true;
    };
    for (const Dist &d : Lo->dists) {
        int ms_op_c = 0, ms_op_a = 0, g_op = 1;
        for (int x : d.msc) ms_op_c += x;
        for (int x : d.msa) ms_op_a += x;
        for (int x : d.ga) g_op = mult(g_op, x);
        for (int ms_ka = ka_max; ms_ka >= -ka_max; ms_ka -= 2) {
            if (std::abs(ms_op_a - ms_ka) > i0a_max) continue;
            for (int ms_kc = kc_max; ms_kc >= -kc_max; ms_kc -= 2) {
                if (std::abs(ms_op_c - ms_kc) > i0c_max) continue;
                for (int g_ka = 1; g_ka <= nsym; ++g_ka)
                    for (int g_kc = 1; g_kc <= nsym; ++g_kc) {
                        dist_strings(o, kc, ka, ms_kc, ms_ka, g_kc, g_ka, DK);
                        for (const DistStr &dk : DK) {
                            const DistStr dkc = part(dk, 0, kc.occ.size()), dka = part(dk, kc.occ.size(), ka.occ.size());
                            if (!split_dis(mi_oc, Lo->hc.size(), d.msc, d.gc, dkc, false, i0c_o.size(), mc, gc) ||
                                !split_dis(mi_oa, Lo->ha.size(), d.msa, d.ga, dka, true, i0a_o.size(), ma, ga))
                                continue;
                            DistStr di0c, di0a;
                            bool bad = false;
                            for (size_t i = 0; i < i0c_o.size() && !bad; ++i) {
                                if (std::abs(mc[i]) > i0c_o[i] || (i0c_o[i] - mc[i]) % 2) { bad = true; break; }
                                di0c.ms.push_back(mc[i]); di0c.gam.push_back(gc[i]);
                                di0c.idxms.push_back(idxms_of(i0c_o[i], mc[i]));
                                di0c.len.push_back(lenstr(o, i0c_h[i], i0c_o[i], di0c.idxms.back(), gc[i]));
                                bad = di0c.len.back() == 0;
                            }
                            for (size_t i = 0; i < i0a_o.size() && !bad; ++i) {
                                if (std::abs(ma[i]) > i0a_o[i] || (i0a_o[i] - ma[i]) % 2) { bad = true; break; }
                                di0a.ms.push_back(ma[i]); di0a.gam.push_back(ga[i]);
                                di0a.idxms.push_back(idxms_of(i0a_o[i], ma[i]));
                                di0a.len.push_back(lenstr(o, i0a_h[i], i0a_o[i], di0a.idxms.back(), ga[i]));
                                bad = di0a.len.back() == 0;
                            }
                            if (bad) continue;
                            merge_dis(mi_rc, Lr->hc.size(), di0c, dkc, rmc, rgc);
                            merge_dis(mi_ra, Lr->ha.size(), dka, di0a, rma, rga);
                            imc.clear(); ima.clear();
                            int ms_r = 0;
                            for (size_t i = 0; i < Lr->oc.size(); ++i) imc.push_back(idxms_of(Lr->oc[i], rmc[i]));
                            for (size_t i = 0; i < Lr->oa.size(); ++i) { ima.push_back(idxms_of(Lr->oa[i], rma[i])); ms_r += rma[i]; }
                            const Dist *dr = list_find(Lr, ms_r, g_op, imc, rgc, ima, rga);
                            if (!dr) continue;
                            strmap_blk(o, Lo->hc, mi_oc, i0c, di0c, kc, dkc, m_oc);
                            strmap_blk(o, Lo->ha, mi_oa, ka, dka, i0a, di0a, m_oa);
                            strmap_blk(o, Lr->hc, mi_rc, i0c, di0c, kc, dkc, m_rc);
                            strmap_blk(o, Lr->ha, mi_ra, ka, dka, i0a, di0a, m_ra);
                            const auto t_oc = targets(so_c, di0c, dkc), t_oa = targets(so_a, dka, di0a);
                            const auto t_rc = targets(sr_c, di0c, dkc), t_ra = targets(sr_a, dka, di0a);
                            int64_t nka = 1, nia = 1, nkc = 1, nic = 1;
                            for (auto x : dka.len) nka *= x;
                            for (auto x : di0a.len) nia *= x;
                            for (auto x : dkc.len) nkc *= x;
                            for (auto x : di0c.len) nic *= x;
                            for (int64_t xka = 0; xka < nka; ++xka) {
                                split_index(xka, dka.len, ska);
                                for (int64_t xia = 0; xia < nia; ++xia) {
                                    split_index(xia, di0a.len, sia);
                                    int sa = sign_reo;
                                    int64_t oa = d.off, ra = dr->off;
                                    if (!addr(t_oa, m_oa, d.lda, ska, sia, sa, oa) ||
                                        !addr(t_ra, m_ra, dr->lda, ska, sia, sa, ra))
                                        continue;
                                    for (int64_t xkc = 0; xkc < nkc; ++xkc) {
                                        split_index(xkc, dkc.len, skc);
                                        for (int64_t xic = 0; xic < nic; ++xic) {
                                            split_index(xic, di0c.len, sic);
                                            int s2 = sa;
                                            int64_t oo = oa, rr = ra;
                                            if (!addr(t_oc, m_oc, d.ldc, sic, skc, s2, oo) ||
                                                !addr(t_rc, m_rc, dr->ldc, sic, skc, s2, rr))
                                                continue;
                                            dst.push_back(rr);
                                            src.push_back(oo);
                                            sg.push_back((int8_t)s2);
                                        }
                                    }
                                }
                            }
                        }
                    }
            }
        }
    }
    return 0;
}

// sym_ab pairs of a list (contract/sym_ab_blk.f:214-377)
static int sym_pairs(gecco_hip_list *L, int64_t off, std::vector<int64_t> &i1,
                     std::vector<int64_t> &i2, std::vector<int8_t> &fs, double relfac) {
    gecco_hip_orbs *o = L->orbs;
    const int msc_max = L->nc, msa_max = L->na;
    double relfac2 = relfac;
    if (((((msc_max + msa_max) / 2 + 1) * (msc_max - msa_max) / 2) % 2) != 0) relfac2 = -relfac;
    auto stdsign = [](const std::vector<int> &ms, const std::vector<int> &occ) {
        if (ms.size() <= 1) return 1;
        int nb = 0, trp = 0;
        for (size_t b = 0; b < ms.size(); ++b) {
            int nal = (occ[b] + ms[b]) / 2, nbe = (occ[b] - ms[b]) / 2;
            trp += nal * nb;
            nb += nbe;
        }
        return (trp & 1) ? -1 : 1;
    };
    for (const Dist &d : L->dists) {
        int msa_tot = 0;
        for (int x : d.msa) msa_tot += x;
        if (msa_tot < 0) continue;
        std::vector<int> mc2, ma2, imc2, ima2;
        for (int x : d.msc) mc2.push_back(-x);
        for (int x : d.msa) ma2.push_back(-x);
        for (size_t i = 0; i < d.msc.size(); ++i) imc2.push_back(idxms_of(L->oc[i], mc2[i]));
        for (size_t i = 0; i < d.msa.size(); ++i) ima2.push_back(idxms_of(L->oa[i], ma2[i]));
        int igama = 1;
        for (int g : d.ga) igama = mult(igama, g);
        const Dist *d2 = list_find(L, -msa_tot, igama, imc2, d.gc, ima2, d.ga);
        if (!d2) return set_error(GECCO_HIP_EINVAL, "sym_ab: flipped distribution missing");
        if (msa_tot == 0 && d.off > d2->off) continue;
        const int gsign = stdsign(d.msc, L->oc) * stdsign(d.msa, L->oa) * stdsign(mc2, L->oc) *
                          stdsign(ma2, L->oa);
        const size_t ncb = L->oc.size(), nab = L->oa.size();
        std::vector<const std::vector<int64_t> *> fl;
        for (size_t k = 0; k < ncb; ++k) fl.push_back(&flipmap(o, L->hc[k], L->oc[k], d.imc[k], d.gc[k]));
        for (size_t k = 0; k < nab; ++k) fl.push_back(&flipmap(o, L->ha[k], L->oa[k], d.ima[k], d.ga[k]));
        int64_t lenc = 1, lena = 1;
        for (size_t k = 0; k < ncb; ++k) lenc *= d.lens[k];
        for (size_t k = 0; k < nab; ++k) lena *= d.lens[ncb + k];
        for (int64_t ia = 0; ia < lena; ++ia) {
            int64_t st = ia, a1 = 0, a2 = 0;
            int as = gsign;
            for (size_t k = 0; k < nab; ++k) {
                int64_t s = st % d.lens[ncb + k];
                int64_t m = (*fl[ncb + k])[s];
                a1 += s * d.lda[k];
                a2 += ((m < 0 ? -m : m) - 1) * d.lda[k];
                if (m < 0) as = -as;
                st /= d.lens[ncb + k];
            }
            for (int64_t ic = 0; ic < lenc; ++ic) {
                int64_t st2 = ic, c1 = 0, c2 = 0;
                int cs = 1;
                for (size_t k = 0; k < ncb; ++k) {
                    int64_t s = st2 % d.lens[k];
                    int64_t m = (*fl[k])[s];
                    c1 += s * d.ldc[k];
                    c2 += ((m < 0 ? -m : m) - 1) * d.ldc[k];
                    if (m < 0) cs = -cs;
                    st2 /= d.lens[k];
                }
                const int64_t x1 = d.off + a1 + c1, x2 = d2->off + a2 + c2;
                if (x1 > x2) continue;
                const double f = cs * as * relfac2;
                i1.push_back(off + x1);
                i2.push_back(off + x2);
                fs.push_back((int8_t)(f > 0 ? 1 : -1));
            }
        }
    }
    return 0;
}

}  // namespace gecco

using namespace gecco;

extern "C" {

int gecco_hip_orbs_create(int64_t nsym, int64_t nh, const int64_t *irrep_h, int64_t np,
                          const int64_t *irrep_p, gecco_hip_orbs **out) {
    if (!out || (nsym != 1 && nsym != 2 && nsym != 4 && nsym != 8) || nh < 0 || np < 0 ||
        nh > 4000 || np > 4000)
        return set_error(GECCO_HIP_EINVAL, "orbs: bad arguments");
    auto *o = new gecco_hip_orbs();
    o->nsym = (int)nsym;
    for (int64_t i = 0; i < nh; ++i) {
        if (irrep_h[i] < 1 || irrep_h[i] > nsym) { delete o; return set_error(GECCO_HIP_EINVAL, "orbs: bad IRREP"); }
        o->irr[1].push_back((int)irrep_h[i]);
    }
    for (int64_t i = 0; i < np; ++i) {
        if (irrep_p[i] < 1 || irrep_p[i] > nsym) { delete o; return set_error(GECCO_HIP_EINVAL, "orbs: bad IRREP"); }
        o->irr[2].push_back((int)irrep_p[i]);
    }
    std::sort(o->irr[1].begin(), o->irr[1].end());
    std::sort(o->irr[2].begin(), o->irr[2].end());
    *out = o;
    return 0;
}

int gecco_hip_orbs_destroy(gecco_hip_orbs *o) {
    delete o;
    return 0;
}

int64_t gecco_hip_orbs_lenstr(gecco_hip_orbs *o, int64_t hpvx, int64_t nel, int64_t idxms,
                              int64_t igam) {
    if (!o || hpvx < 1 || hpvx > 2 || nel > 8) return -1;
    return lenstr(o, (int)hpvx, (int)nel, (int)idxms, (int)igam);
}

int gecco_hip_orbs_strmap(gecco_hip_orbs *o, int64_t hpvx, int64_t n1, int64_t idxms1, int64_t g1,
                          int64_t n2, int64_t idxms2, int64_t g2, int64_t *out, int64_t maxlen,
                          int64_t *len) {
    if (!o || hpvx < 1 || hpvx > 2 || n1 + n2 > 8 || n1 < 0 || n2 < 0)
        return set_error(GECCO_HIP_EINVAL, "strmap: bad arguments");
    const auto &m = strmap(o, (int)hpvx, (int)n1, (int)idxms1, (int)g1, (int)n2, (int)idxms2, (int)g2);
    *len = (int64_t)m.size();
    if ((int64_t)m.size() > maxlen) return set_error(GECCO_HIP_EINVAL, "strmap: buffer too small");
    std::copy(m.begin(), m.end(), out);
    return 0;
}

int gecco_hip_orbs_flipmap(gecco_hip_orbs *o, int64_t hpvx, int64_t nel, int64_t idxms, int64_t g,
                           int64_t *out, int64_t maxlen, int64_t *len) {
    if (!o || hpvx < 1 || hpvx > 2 || nel > 8 || nel < 0)
        return set_error(GECCO_HIP_EINVAL, "flipmap: bad arguments");
    const auto &m = flipmap(o, (int)hpvx, (int)nel, (int)idxms, (int)g);
    *len = (int64_t)m.size();
    if ((int64_t)m.size() > maxlen) return set_error(GECCO_HIP_EINVAL, "flipmap: buffer too small");
    std::copy(m.begin(), m.end(), out);
    return 0;
}

int gecco_hip_list_create_n(gecco_hip_orbs *o, int64_t njoined, const int64_t *occ, int64_t absym,
                            gecco_hip_list **out) {
    if (!o || !occ || !out || njoined < 1 || njoined > 8)
        return set_error(GECCO_HIP_EINVAL, "list: null argument or bad vertex count");
    auto *L = new gecco_hip_list();
    L->orbs = o;
    L->absym = (int)absym;
    L->njoined = (int)njoined;
    L->vocc.assign(njoined, std::array<int, 10>{});
    int nc = 0, na = 0;
    for (int64_t v = 0; v < njoined; ++v)
        for (int ca = 0; ca < 2; ++ca)
            for (int h = 1; h <= 4; ++h) {
                const int64_t n = occ[v * 8 + ca * 4 + h - 1];
                if (n < 0 || n > 8 || ((h == 3 || h == 4) && n)) {
                    delete L;
                    return set_error(GECCO_HIP_EINVAL, "list: occupations must be H/P, 0..8");
                }
                L->vocc[v][ca * 5 + h] = (int)n;
                (ca ? L->occ_a : L->occ_c)[h] += (int)n;
                (ca ? na : nc) += (int)n;
            }
    for (int h = 1; h <= 4; ++h)
        if (L->occ_c[h] > 8 || L->occ_a[h] > 8) {
            delete L;
            return set_error(GECCO_HIP_EINVAL, "list: more than 8 electrons of one space");
        }
    if ((nc - na) % 2) {
        delete L;
        return set_error(GECCO_HIP_EINVAL, "list: Ms(total) = 0 needs creator and annihilator counts of equal parity");
    }
    build_list(L);
    *out = L;
    return 0;
}

int gecco_hip_list_create(gecco_hip_orbs *o, const int64_t *occ_c, const int64_t *occ_a,
                          int64_t absym, gecco_hip_list **out) {
    if (!o || !occ_c || !occ_a || !out) return set_error(GECCO_HIP_EINVAL, "list: null argument");
    int64_t occ[8];
    for (int h = 0; h < 4; ++h) { occ[h] = occ_c[h]; occ[4 + h] = occ_a[h]; }
    return gecco_hip_list_create_n(o, 1, occ, absym, out);
}

int gecco_hip_list_adjoint(const gecco_hip_list *L, gecco_hip_list **out) {
    // C and A sides swap; every distribution keeps its stored offset and
    // length, its component labels swap sides, and its lengths and leading
    // dimensions follow the swapped component order.  The block labels
    // (Ms(A), IRREP(A)) of the view equal the stored ones: Ms(C) = Ms(A) and
    // IRREP(C) = IRREP(A) in every block of a list with Ms(total) = 0 and
    // IRREP(total) = 1, and msa2idxms4op gives the view's block the stored
    // block's index for either order of (occ_a, occ_c).
    if (!L || !out) return set_error(GECCO_HIP_EINVAL, "list_adjoint: null argument");
    if (L->njoined != 1) return set_error(GECCO_HIP_EINVAL, "list_adjoint: one-vertex lists only");
    auto *A = new gecco_hip_list();
    A->orbs = L->orbs;
    for (int h = 0; h < 5; ++h) { A->occ_c[h] = L->occ_a[h]; A->occ_a[h] = L->occ_c[h]; }
    A->njoined = 1;
    A->vocc = L->vocc;
    for (auto &v : A->vocc)
        for (int h = 0; h < 5; ++h) std::swap(v[h], v[5 + h]);
    A->absym = L->absym;
    A->oc = L->oa; A->hc = L->ha; A->vc = L->va;
    A->oa = L->oc; A->ha = L->hc; A->va = L->vc;
    A->nc = L->na; A->na = L->nc;
    const size_t ncs = L->oc.size();
    for (const Dist &d : L->dists) {
        Dist e;
        e.msc = d.msa; e.msa = d.msc; e.gc = d.ga; e.ga = d.gc;
        e.imc = d.ima; e.ima = d.imc;
        e.lens.assign(d.lens.begin() + ncs, d.lens.end());
        e.lens.insert(e.lens.end(), d.lens.begin(), d.lens.begin() + ncs);
        e.ldc = d.lda; e.lda = d.ldc;
        e.off = d.off; e.len = d.len;
        A->dists.push_back(std::move(e));
    }
    for (const auto &kv : L->find) {
        const std::vector<int> &k = kv.first;    // idxmsa, igama, imc, gc, ima, ga
        const size_t nca = L->oa.size();
        std::vector<int> key{k[0], k[1]};
        key.insert(key.end(), k.begin() + 2 + 2 * ncs, k.begin() + 2 + 2 * ncs + nca);
        key.insert(key.end(), k.begin() + 2 + 2 * ncs + nca, k.end());
        key.insert(key.end(), k.begin() + 2, k.begin() + 2 + ncs);
        key.insert(key.end(), k.begin() + 2 + ncs, k.begin() + 2 + 2 * ncs);
        A->find[key] = kv.second;
    }
    A->blen = L->blen;
    A->length = L->length;
    A->tra = !L->tra;
    *out = A;
    return 0;
}

int gecco_hip_list_destroy(gecco_hip_list *L) {
    delete L;
    return 0;
}

int64_t gecco_hip_list_length(const gecco_hip_list *L) { return L ? L->length : -1; }

int64_t gecco_hip_list_ndist(const gecco_hip_list *L) { return L ? (int64_t)L->dists.size() : -1; }

int gecco_hip_list_dist(const gecco_hip_list *L, int64_t i, int64_t *info, int64_t ninfo) {
    // info = [off, len, msa_tot, igama, ncomp_c, ncomp_a, then per C comp
    //         (ms, idxms, gam, len, ldim), then per A comp (same)]
    if (!L || i < 0 || i >= (int64_t)L->dists.size()) return set_error(GECCO_HIP_EINVAL, "list_dist: bad index");
    const Dist &d = L->dists[i];
    std::vector<int64_t> v;
    int msa = 0, ga = 1;
    for (int x : d.msa) msa += x;
    for (int x : d.ga) ga = mult(ga, x);
    v = {d.off, d.len, msa, ga, (int64_t)d.msc.size(), (int64_t)d.msa.size()};
    for (size_t k = 0; k < d.msc.size(); ++k) {
        int64_t tmp[5] = {d.msc[k], d.imc[k], d.gc[k], d.lens[k], d.ldc[k]};
        v.insert(v.end(), tmp, tmp + 5);
    }
    for (size_t k = 0; k < d.msa.size(); ++k) {
        int64_t tmp[5] = {d.msa[k], d.ima[k], d.ga[k], d.lens[d.msc.size() + k], d.lda[k]};
        v.insert(v.end(), tmp, tmp + 5);
    }
    if ((int64_t)v.size() > ninfo) return set_error(GECCO_HIP_EINVAL, "list_dist: buffer too small");
    std::copy(v.begin(), v.end(), info);
    return 0;
}

int gecco_hip_list_spinorbs(const gecco_hip_list *L, int16_t *out, int64_t nout) {
    // per element: spin-orbital index so = 2*(orb-1)+spin of every electron,
    // C components then A components (component order of the list), each
    // component's string in ascending order
    if (!L || !out) return set_error(GECCO_HIP_EINVAL, "list_spinorbs: null argument");
    const int nel = L->nc + L->na;
    if (nout < L->length * (int64_t)nel) return set_error(GECCO_HIP_EINVAL, "list_spinorbs: buffer too small");
    gecco_hip_orbs *o = L->orbs;
    const int nsym = o->nsym;
    const size_t ncb = L->oc.size();
    for (const Dist &d : L->dists) {
        std::vector<const GraphBlock *> gb;
        std::vector<int> occ, hp;
        for (size_t k = 0; k < ncb; ++k) {
            gb.push_back(&graph(o, L->hc[k], L->oc[k]).blk[(size_t)(d.imc[k] - 1) * nsym + d.gc[k] - 1]);
            occ.push_back(L->oc[k]);
        }
        for (size_t k = 0; k < L->oa.size(); ++k) {
            gb.push_back(&graph(o, L->ha[k], L->oa[k]).blk[(size_t)(d.ima[k] - 1) * nsym + d.ga[k] - 1]);
            occ.push_back(L->oa[k]);
        }
        std::vector<int64_t> ld(d.ldc);
        ld.insert(ld.end(), d.lda.begin(), d.lda.end());
        const size_t nc = gb.size();
        std::vector<int64_t> idx(nc, 0);
        for (int64_t e = 0; e < d.len; ++e) {
            int64_t off = d.off;
            for (size_t c = 0; c < nc; ++c) off += idx[c] * ld[c];
            int16_t *dst = out + off * nel;
            uint16_t so[8];
            for (size_t c = 0; c < nc; ++c) {
                unpack(gb[c]->strs[idx[c]], occ[c], so);
                for (int q = 0; q < occ[c]; ++q) *dst++ = (int16_t)so[q];
            }
            for (size_t c = 0; c < nc; ++c) {   // odometer, first component fastest
                if (++idx[c] < (int64_t)gb[c]->strs.size()) break;
                idx[c] = 0;
            }
        }
    }
    return 0;
}

int gecco_hip_list_diagonal(const gecco_hip_list *L, const double *eps_h, const double *eps_p,
                            double *out, int64_t nout) {
    // out[e] = sum of eps over the C strings - sum over the A strings
    // (contract/dia4op_ev.f:298-365 with fac = +1 for C, -1 for A)
    if (!L || !eps_h || !eps_p || !out) return set_error(GECCO_HIP_EINVAL, "list_diagonal: null argument");
    if (nout < L->length) return set_error(GECCO_HIP_EINVAL, "list_diagonal: buffer too small");
    gecco_hip_orbs *o = L->orbs;
    const int nsym = o->nsym;
    for (const Dist &d : L->dists) {
        std::vector<const GraphBlock *> gb;
        std::vector<int> occ;
        std::vector<const double *> eps;
        std::vector<double> fac;
        for (size_t k = 0; k < L->oc.size(); ++k) {
            gb.push_back(&graph(o, L->hc[k], L->oc[k]).blk[(size_t)(d.imc[k] - 1) * nsym + d.gc[k] - 1]);
            occ.push_back(L->oc[k]);
            eps.push_back(L->hc[k] == 1 ? eps_h : eps_p);
            fac.push_back(1.0);
        }
        for (size_t k = 0; k < L->oa.size(); ++k) {
            gb.push_back(&graph(o, L->ha[k], L->oa[k]).blk[(size_t)(d.ima[k] - 1) * nsym + d.ga[k] - 1]);
            occ.push_back(L->oa[k]);
            eps.push_back(L->ha[k] == 1 ? eps_h : eps_p);
            fac.push_back(-1.0);
        }
        std::vector<int64_t> ld(d.ldc);
        ld.insert(ld.end(), d.lda.begin(), d.lda.end());
        const size_t nc = gb.size();
        // per component: the orbital-energy sum of each string
        std::vector<std::vector<double>> sums(nc);
        for (size_t c = 0; c < nc; ++c) {
            sums[c].resize(gb[c]->strs.size());
            uint16_t so[8];
            for (size_t i = 0; i < gb[c]->strs.size(); ++i) {
                unpack(gb[c]->strs[i], occ[c], so);
                double x = 0.0;
                for (int q = 0; q < occ[c]; ++q) x += eps[c][so[q] >> 1];
                sums[c][i] = fac[c] * x;
            }
        }
        std::vector<int64_t> idx(nc, 0);
        for (int64_t e = 0; e < d.len; ++e) {
            int64_t off = d.off;
            double x = 0.0;
            for (size_t c = 0; c < nc; ++c) {
                off += idx[c] * ld[c];
                x += sums[c][idx[c]];
            }
            out[off] = x;
            for (size_t c = 0; c < nc; ++c) {   // odometer, first component fastest
                if (++idx[c] < (int64_t)gb[c]->strs.size()) break;
                idx[c] = 0;
            }
        }
    }
    return 0;
}

int gecco_hip_plan_add_contraction(gecco_hip_plan *plan, const gecco_hip_list *op1,
                                   const gecco_hip_list *op2, const gecco_hip_list *res,
                                   const int64_t *cnt_c, const int64_t *cnt_a, double fact,
                                   int64_t use_tr, int32_t buf1, int64_t off1, int32_t buf2,
                                   int64_t off2, int32_t buf12, int64_t off12, int64_t *nblocks,
                                   double *flops) {
    if (!plan || !op1 || !op2 || !res || !cnt_c || !cnt_a)
        return set_error(GECCO_HIP_EINVAL, "add_contraction: null argument");
    if (op1->orbs != op2->orbs || op1->orbs != res->orbs)
        return set_error(GECCO_HIP_EINVAL, "add_contraction: lists of different orbital sets");
    if (op1->njoined > 1 || op2->njoined > 1 || res->njoined > 1)
        return set_error(GECCO_HIP_EINVAL, "add_contraction: one-vertex lists only");
    int cc[5] = {0}, ca[5] = {0}, e1c[5] = {0}, e1a[5] = {0}, e2c[5] = {0}, e2a[5] = {0};
    for (int h = 1; h <= 4; ++h) {
        cc[h] = (int)cnt_c[h - 1];
        ca[h] = (int)cnt_a[h - 1];
        e1c[h] = op1->occ_c[h] - cc[h];
        e1a[h] = op1->occ_a[h] - ca[h];
        e2c[h] = op2->occ_c[h] - ca[h];
        e2a[h] = op2->occ_a[h] - cc[h];
        if (cc[h] < 0 || ca[h] < 0 || e1c[h] < 0 || e1a[h] < 0 || e2c[h] < 0 || e2a[h] < 0)
            return set_error(GECCO_HIP_EINVAL, "add_contraction: contraction exceeds an operand");
        if (res->occ_c[h] != e1c[h] + e2c[h] || res->occ_a[h] != e1a[h] + e2a[h])
            return set_error(GECCO_HIP_EINVAL, "add_contraction: result occupation != EX1 + EX2");
    }
    ContrCtx c;
    c.plan = plan;
    c.o = op1->orbs;
    c.L1 = op1; c.L2 = op2; c.L12 = res;
    c.ex1c = side_of(e1c); c.ex1a = side_of(e1a);
    c.ex2c = side_of(e2c); c.ex2a = side_of(e2a);
    c.cntc = side_of(cc); c.cnta = side_of(ca);
    c.mi12c = map_info(res->hc, c.ex1c, c.ex2c);
    c.mi12a = map_info(res->ha, c.ex2a, c.ex1a);
    c.mi1c = map_info(op1->hc, c.cntc, c.ex1c);
    c.mi1a = map_info(op1->ha, c.cnta, c.ex1a);
    c.mi2c = map_info(op2->hc, c.cnta, c.ex2c);
    c.mi2a = map_info(op2->ha, c.cntc, c.ex2a);
    c.fact = fact;
    c.buf1 = buf1; c.buf2 = buf2; c.buf12 = buf12;
    c.off1 = off1; c.off2 = off2; c.off12 = off12;
    int rc = enumerate(c, use_tr != 0);
    if (nblocks) *nblocks = c.nblocks;
    if (flops) *flops = c.flops;
    return rc;
}

namespace gecco {

typedef std::vector<std::array<int, 10>> VOcc;   // per vertex: C[h] at h, A[h] at 5 + h
typedef std::vector<std::pair<std::vector<int>, std::vector<int>>> Merge;

static VOcc vocc_of(const int64_t *iocc, int64_t nj) {
    VOcc v((size_t)nj, std::array<int, 10>{});
    for (int64_t j = 0; j < nj; ++j)
        for (int h = 1; h <= 4; ++h) {
            v[j][h] = (int)iocc[8 * j + h - 1];
            v[j][5 + h] = (int)iocc[8 * j + 4 + h - 1];
        }
    return v;
}

// contract/condense_merge_map.f short form -> per target vertex (source-1
// vertices, source-2 vertices), 1-based; false when malformed
static bool merge_of(const int64_t *mm, size_t ntgt, int n1, int n2, Merge &out) {
    out.clear();
    size_t k = 0;
    for (size_t t = 0; t < ntgt; ++t) {
        std::pair<std::vector<int>, std::vector<int>> p;
        for (int q = 0; q < 2; ++q) {
            const int64_t n = mm[k++];
            if (n < 0 || n > 16) return false;
            for (int64_t i = 0; i < n; ++i) {
                const int64_t v = mm[k++];
                if (v < 1 || v > (q ? n2 : n1)) return false;
                (q ? p.second : p.first).push_back((int)v);
            }
        }
        out.push_back(std::move(p));
    }
    return true;
}

// contract/set_idx_seq.f (sep_ca) and contract/set_mapping_info.f:84-170
// for one side (0: C, 1: A) of a target operator whose vertices take the
// source vertices of merge; dag swaps a source's C and A sides (CNT as seen
// from OP2)
static std::vector<int64_t> mapping_info(const VOcc &tgt, const Merge &merge, const VOcc &s1,
                                         bool dag1, const VOcc &s2, bool dag2, int side) {
    const int c1 = dag1 ? 1 - side : side, c2 = dag2 ? 1 - side : side;
    auto seq = [](const VOcc &v, int s) {
        std::map<std::pair<int, int>, int64_t> q;
        int64_t k = 0;
        for (int h : HPVXBLKSEQ)
            for (size_t j = 0; j < v.size(); ++j)
                if (v[j][5 * s + h] > 0) q[{h, (int)j + 1}] = ++k;
        return q;
    };
    const auto q1 = seq(s1, c1), q2 = seq(s2, c2);
    std::vector<int64_t> mi;
    for (int h : HPVXBLKSEQ)
        for (size_t t = 0; t < tgt.size(); ++t) {
            if (tgt[t][5 * side + h] == 0) continue;
            std::vector<int64_t> i1, i2;
            for (int v : merge[t].first)
                if (s1[v - 1][5 * c1 + h] > 0) i1.push_back(q1.at({h, v}));
            for (int v : merge[t].second)
                if (s2[v - 1][5 * c2 + h] > 0) i2.push_back(q2.at({h, v}));
            mi.push_back((int64_t)i1.size());
            mi.insert(mi.end(), i1.begin(), i1.end());
            mi.push_back((int64_t)i2.size());
            mi.insert(mi.end(), i2.begin(), i2.end());
        }
    return mi;
}

static void sides_of(const VOcc &v, Side &c, Side &a) {
    std::vector<int> oc, hc, vc, oa, ha, va;
    condense_n(v, oc, hc, vc, oa, ha, va);
    c.occ = oc; c.hpvx = hc;
    a.occ = oa; a.hpvx = ha;
}

// every target component takes at most one component of each source list
static bool single_sources(const std::vector<int64_t> &mi, size_t ntgt) {
    size_t k = 0;
    for (size_t t = 0; t < ntgt; ++t)
        for (int q = 0; q < 2; ++q) {
            const int64_t n = mi[k];
            if (n > 1) return false;
            k += 1 + (size_t)n;
        }
    return true;
}

}  // namespace gecco

int gecco_hip_plan_add_contraction_n(gecco_hip_plan *plan, const gecco_hip_list *op1,
                                     const gecco_hip_list *op2, const gecco_hip_list *res,
                                     int64_t njoined_cnt, const int64_t *iocc_ex1,
                                     const int64_t *iocc_ex2, const int64_t *iocc_cnt,
                                     const int64_t *merge_op1, const int64_t *merge_op2,
                                     const int64_t *merge_op1op2, const int64_t *merge_op2op1,
                                     double fact, int64_t use_tr, int32_t buf1, int64_t off1,
                                     int32_t buf2, int64_t off2, int32_t buf12, int64_t off12,
                                     int64_t *nblocks, double *flops) {
    if (!plan || !op1 || !op2 || !res || !iocc_ex1 || !iocc_ex2 || (njoined_cnt > 0 && !iocc_cnt) ||
        !merge_op1 || !merge_op2 || !merge_op1op2 || !merge_op2op1)
        return set_error(GECCO_HIP_EINVAL, "add_contraction_n: null argument");
    if (op1->orbs != op2->orbs || op1->orbs != res->orbs)
        return set_error(GECCO_HIP_EINVAL, "add_contraction_n: lists of different orbital sets");
    if (op1->tra || op2->tra || res->tra)
        return set_error(GECCO_HIP_EINVAL, "add_contraction_n: stored lists only, no adjoint views");
    if (njoined_cnt < 0 || njoined_cnt > 16)
        return set_error(GECCO_HIP_EINVAL, "add_contraction_n: bad number of contraction vertices");
    const VOcc ex1 = vocc_of(iocc_ex1, op1->njoined), ex2 = vocc_of(iocc_ex2, op2->njoined);
    const VOcc cnt = njoined_cnt ? vocc_of(iocc_cnt, njoined_cnt) : VOcc{};
    for (const VOcc *v : {&ex1, &ex2, &cnt})
        for (const auto &a : *v)
            for (int x : a)
                if (x < 0) return set_error(GECCO_HIP_EINVAL, "add_contraction_n: negative occupation");
    Merge m1, m2, m12, m21;
    const int nc = (int)njoined_cnt, n1 = op1->njoined, n2 = op2->njoined;
    if (!merge_of(merge_op1, (size_t)n1, nc, n1, m1) || !merge_of(merge_op2, (size_t)n2, nc, n2, m2) ||
        !merge_of(merge_op1op2, (size_t)res->njoined, n1, n2, m12) ||
        !merge_of(merge_op2op1, (size_t)res->njoined, n2, n1, m21))
        return set_error(GECCO_HIP_EINVAL, "add_contraction_n: malformed merge map");
    // every operand vertex is the sum of its merged parts (OP2 sees CNT
    // daggered; the result's creators follow merge_op1op2, its annihilators
    // merge_op2op1)
    auto sums_to = [](const std::array<int, 10> &tgt, const std::pair<std::vector<int>, std::vector<int>> &p,
                      const VOcc &s1, bool dag1, const VOcc &s2, int side) {
        for (int h = 1; h <= 4; ++h) {
            int n = 0;
            for (int v : p.first) n += s1[v - 1][5 * (dag1 ? 1 - side : side) + h];
            for (int v : p.second) n += s2[v - 1][5 * side + h];
            if (n != tgt[5 * side + h]) return false;
        }
        return true;
    };
    for (int v = 0; v < n1; ++v)
        for (int s = 0; s < 2; ++s)
            if (!sums_to(op1->vocc[v], m1[v], cnt, false, ex1, s))
                return set_error(GECCO_HIP_EINVAL, "add_contraction_n: op1 vertex != CNT + EX1");
    for (int v = 0; v < n2; ++v)
        for (int s = 0; s < 2; ++s)
            if (!sums_to(op2->vocc[v], m2[v], cnt, true, ex2, s))
                return set_error(GECCO_HIP_EINVAL, "add_contraction_n: op2 vertex != CNT^+ + EX2");
    for (int v = 0; v < res->njoined; ++v)
        if (!sums_to(res->vocc[v], m12[v], ex1, false, ex2, 0) || !sums_to(res->vocc[v], m21[v], ex2, false, ex1, 1))
            return set_error(GECCO_HIP_EINVAL, "add_contraction_n: result vertex != EX1 + EX2");
    ContrCtx c;
    c.plan = plan;
    c.o = op1->orbs;
    c.L1 = op1; c.L2 = op2; c.L12 = res;
    sides_of(ex1, c.ex1c, c.ex1a);
    sides_of(ex2, c.ex2c, c.ex2a);
    sides_of(cnt, c.cntc, c.cnta);
    c.mi12c = mapping_info(res->vocc, m12, ex1, false, ex2, false, 0);
    c.mi12a = mapping_info(res->vocc, m21, ex2, false, ex1, false, 1);
    c.mi1c = mapping_info(op1->vocc, m1, cnt, false, ex1, false, 0);
    c.mi1a = mapping_info(op1->vocc, m1, cnt, false, ex1, false, 1);
    c.mi2c = mapping_info(op2->vocc, m2, cnt, true, ex2, false, 0);
    c.mi2a = mapping_info(op2->vocc, m2, cnt, true, ex2, false, 1);
    // route 3 carries one string per target component and source list
    // (contract/collect_block_impr.f with ireo from contract/set_strmapdim_c2.f)
    if (!single_sources(c.mi12c, res->hc.size()) || !single_sources(c.mi12a, res->ha.size()) ||
        !single_sources(c.mi1c, op1->hc.size()) || !single_sources(c.mi1a, op1->ha.size()) ||
        !single_sources(c.mi2c, op2->hc.size()) || !single_sources(c.mi2a, op2->ha.size()))
        return set_error(GECCO_HIP_EINVAL,
                         "add_contraction_n: multi map needed (two components of one source merge into one component)");
    c.fact = fact;
    c.buf1 = buf1; c.buf2 = buf2; c.buf12 = buf12;
    c.off1 = off1; c.off2 = off2; c.off12 = off12;
    int rc = enumerate(c, use_tr != 0);
    if (nblocks) *nblocks = c.nblocks;
    if (flops) *flops = c.flops;
    return rc;
}

int gecco_hip_plan_add_trace(gecco_hip_plan *plan, const gecco_hip_list *op,
                             const gecco_hip_list *trop, const int64_t *cnt, double fact,
                             int32_t buf_op, int64_t off_op, int32_t buf_trop, int64_t off_trop,
                             int64_t *nterms) {
    if (!plan || !op || !trop || !cnt) return set_error(GECCO_HIP_EINVAL, "add_trace: null argument");
    if (op->orbs != trop->orbs) return set_error(GECCO_HIP_EINVAL, "add_trace: lists of different orbital sets");
    if (op->njoined > 1 || trop->njoined > 1) return set_error(GECCO_HIP_EINVAL, "add_trace: one-vertex lists only");
    if (op->tra || trop->tra) return set_error(GECCO_HIP_EINVAL, "add_trace: stored lists only, no adjoint views");
    int c5[5] = {0};
    for (int h = 1; h <= 4; ++h) {
        c5[h] = (int)cnt[h - 1];
        if (c5[h] < 0 || op->occ_c[h] - c5[h] != trop->occ_c[h] || op->occ_a[h] - c5[h] != trop->occ_a[h])
            return set_error(GECCO_HIP_EINVAL, "add_trace: OP occupation minus J does not match OPtr");
    }
    std::vector<int64_t> dst, src;
    std::vector<int8_t> sg;
    int rc = trace_triples(op, trop, c5, dst, src, sg);
    if (rc) return rc;
    if (nterms) *nterms = (int64_t)dst.size();
    return plan_add_trace_rows(plan, buf_op, off_op, buf_trop, off_trop, fact, std::move(dst),
                               std::move(src), std::move(sg));
}

int gecco_hip_plan_add_reorder(gecco_hip_plan *plan, const gecco_hip_list *ori,
                               const gecco_hip_list *reo, int64_t from_vtx, int64_t to_vtx,
                               const int64_t *shift_c, const int64_t *shift_a, double fact,
                               int64_t sign_reo, int32_t buf_ori, int64_t off_ori, int32_t buf_reo,
                               int64_t off_reo, int64_t *nterms) {
    if (!plan || !ori || !reo || !shift_c || !shift_a)
        return set_error(GECCO_HIP_EINVAL, "add_reorder: null argument");
    if (ori->orbs != reo->orbs) return set_error(GECCO_HIP_EINVAL, "add_reorder: lists of different orbital sets");
    if (ori->tra || reo->tra) return set_error(GECCO_HIP_EINVAL, "add_reorder: stored lists only, no adjoint views");
    if (from_vtx < 0 || from_vtx >= ori->njoined || to_vtx < 0 ||
        to_vtx >= std::max(ori->njoined, reo->njoined))
        return set_error(GECCO_HIP_EINVAL, "add_reorder: vertex out of range");
    if (sign_reo != 1 && sign_reo != -1) return set_error(GECCO_HIP_EINVAL, "add_reorder: sign must be +1 or -1");
    int kc[5] = {0}, ka[5] = {0};
    for (int h = 1; h <= 4; ++h) {
        kc[h] = (int)shift_c[h - 1];
        ka[h] = (int)shift_a[h - 1];
        if (kc[h] < 0 || ka[h] < 0) return set_error(GECCO_HIP_EINVAL, "add_reorder: negative shift occupation");
    }
    std::vector<int64_t> dst, src;
    std::vector<int8_t> sg;
    int rc = reorder_triples(ori, reo, (int)from_vtx, (int)to_vtx, kc, ka, (int)sign_reo, dst, src, sg);
    if (rc) return rc;
    if (nterms) *nterms = (int64_t)dst.size();
    return plan_add_trace_rows(plan, buf_ori, off_ori, buf_reo, off_reo, fact, std::move(dst),
                               std::move(src), std::move(sg));
}

// Singlet spin projection of a list (contract/spin_prj_list.f,
// contract/spin_prj_blk.f with s2 = 0): the elements of a particle-conserving
// one-vertex list fall into groups with the same spatial orbitals on each
// side and different spins.  A group's element operators tau_e (a+ of the C
// spin orbitals, a of the A ones in reverse order) are orthonormal, and the
// singlet (rank-0 spin tensor) operators of the group are those X with
// [S+, X] = 0 (S+ = sum_p a+_p,alpha a_p,beta; total Ms is zero already).
// The commutator maps tau_e to operators with one beta creator turned alpha
// (+1) or one alpha annihilator turned beta (-1), each brought to a sorted
// order with its permutation sign; the projector onto the null space of that
// map (eigenvectors of A^T A with eigenvalue 0, cyclic Jacobi) replaces the
// reference's genealogical spin couplings -- it is the same orthogonal
// projector onto the singlet subspace.  Distinct projectors are stored once.
static int sort_parity(std::vector<int> &v) {
    int sg = 1;
    for (size_t i = 1; i < v.size(); ++i)
        for (size_t j = i; j > 0 && v[j - 1] > v[j]; --j) {
            std::swap(v[j - 1], v[j]);
            sg = -sg;
        }
    return sg;
}

static void jacobi_eigen(std::vector<double> a, int n, std::vector<double> &vec, std::vector<double> &val) {
    vec.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) vec[(size_t)i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                tot += a[(size_t)i * n + j] * a[(size_t)i * n + j];
                if (i != j) off += a[(size_t)i * n + j] * a[(size_t)i * n + j];
            }
        if (off <= 1e-30 * (tot + 1e-300)) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = a[(size_t)p * n + q];
                if (std::fabs(apq) < 1e-300) continue;
                const double app = a[(size_t)p * n + p], aqq = a[(size_t)q * n + q];
                const double th = 0.5 * (aqq - app) / apq;
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
                for (int k = 0; k < n; ++k) {     // a <- J^T a J, columns then rows
                    const double akp = a[(size_t)k * n + p], akq = a[(size_t)k * n + q];
                    a[(size_t)k * n + p] = c * akp - sn * akq;
                    a[(size_t)k * n + q] = sn * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = a[(size_t)p * n + k], aqk = a[(size_t)q * n + k];
                    a[(size_t)p * n + k] = c * apk - sn * aqk;
                    a[(size_t)q * n + k] = sn * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = vec[(size_t)k * n + p], vkq = vec[(size_t)k * n + q];
                    vec[(size_t)k * n + p] = c * vkp - sn * vkq;
                    vec[(size_t)k * n + q] = sn * vkp + c * vkq;
                }
            }
    }
    val.resize((size_t)n);
    for (int i = 0; i < n; ++i) val[(size_t)i] = a[(size_t)i * n + i];
}

int gecco_hip_plan_add_spin_project(gecco_hip_plan *plan, gecco_hip_list *L, int32_t buf, int64_t off) {
    if (!plan || !L) return set_error(GECCO_HIP_EINVAL, "add_spin_project: null argument");
    if (L->tra) return set_error(GECCO_HIP_EINVAL, "add_spin_project: give the stored list, not its adjoint view");
    if (L->njoined != 1) return set_error(GECCO_HIP_EINVAL, "add_spin_project: one-vertex lists only");
    int ncel = 0, nael = 0;
    for (int o : L->oc) ncel += o;
    for (int o : L->oa) nael += o;
    if (ncel != nael) return set_error(GECCO_HIP_EINVAL, "add_spin_project: only particle-conserving lists");
    const int nel = ncel + nael;
    if (nel > 6) return set_error(GECCO_HIP_EINVAL, "add_spin_project: at most three creators and annihilators");
    std::vector<int> space;
    for (size_t k = 0; k < L->oc.size(); ++k) space.insert(space.end(), (size_t)L->oc[k], L->hc[k]);
    for (size_t k = 0; k < L->oa.size(); ++k) space.insert(space.end(), (size_t)L->oa[k], L->ha[k]);
    for (int sp : space)
        if (sp != 1 && sp != 2) return set_error(GECCO_HIP_EINVAL, "add_spin_project: hole and particle spaces only");
    const int nh2 = 2 * (int)L->orbs->irr[1].size();
    std::vector<int16_t> so((size_t)std::max<int64_t>(1, L->length * nel));
    if (nel > 0 && gecco_hip_list_spinorbs(L, so.data(), (int64_t)so.size())) return 1;
    auto mode = [&](int64_t e, int q) {
        const int m = so[(size_t)e * nel + q];
        return space[(size_t)q] == 2 ? nh2 + m : m;
    };
    // groups: same spatial orbitals (as multisets) on each side
    std::map<std::vector<int>, std::vector<int64_t>> groups;
    for (int64_t e = 0; e < L->length; ++e) {
        std::vector<int> c, a;
        for (int q = 0; q < ncel; ++q) c.push_back(mode(e, q) >> 1);
        for (int q = ncel; q < nel; ++q) a.push_back(mode(e, q) >> 1);
        std::sort(c.begin(), c.end());
        std::sort(a.begin(), a.end());
        c.push_back(-1);
        c.insert(c.end(), a.begin(), a.end());
        groups[c].push_back(e);
    }
    std::vector<int64_t> gptr(1, 0), elem, gmat;
    std::vector<double> mats;
    std::map<std::vector<int64_t>, int64_t> known;
    for (auto &kv : groups) {
        const std::vector<int64_t> &es = kv.second;
        const int n = (int)es.size();
        if (n > 32) return set_error(GECCO_HIP_EINVAL, "add_spin_project: group larger than 32 elements");
        std::map<std::vector<int>, int> rows;
        std::vector<std::vector<std::pair<int, double>>> col((size_t)n);
        auto term = [&](int j, std::vector<int> c, std::vector<int> a, double coef) {
            std::vector<int> sc = c, sa = a;
            const int pc = sort_parity(sc), pa = sort_parity(sa);
            for (size_t i = 1; i < sc.size(); ++i)
                if (sc[i] == sc[i - 1]) return;            // a+ a+ of one spin orbital
            for (size_t i = 1; i < sa.size(); ++i)
                if (sa[i] == sa[i - 1]) return;
            sc.push_back(-1);
            sc.insert(sc.end(), sa.begin(), sa.end());
            const int r = rows.emplace(sc, (int)rows.size()).first->second;
            col[(size_t)j].push_back({r, coef * pc * pa});
        };
        for (int j = 0; j < n; ++j) {
            std::vector<int> c, a;
            for (int q = 0; q < ncel; ++q) c.push_back(mode(es[(size_t)j], q));
            for (int q = ncel; q < nel; ++q) a.push_back(mode(es[(size_t)j], q));
            for (int i = 0; i < ncel; ++i)
                if (c[(size_t)i] & 1) {          // beta creator -> alpha
                    std::vector<int> c2 = c;
                    c2[(size_t)i] -= 1;
                    term(j, c2, a, 1.0);
                }
            for (int i = 0; i < nael; ++i)
                if (!(a[(size_t)i] & 1)) {       // alpha annihilator -> beta, sign -1
                    std::vector<int> a2 = a;
                    a2[(size_t)i] += 1;
                    term(j, c, a2, -1.0);
                }
        }
        // M = A^T A, null space projector
        const int nr = (int)rows.size();
        std::vector<double> A((size_t)nr * n, 0.0), M((size_t)n * n, 0.0), vec, val;
        for (int j = 0; j < n; ++j)
            for (auto &rc : col[(size_t)j]) A[(size_t)rc.first * n + j] += rc.second;
        for (int r = 0; r < nr; ++r)
            for (int j = 0; j < n; ++j)
                for (int k = 0; k < n; ++k) M[(size_t)j * n + k] += A[(size_t)r * n + j] * A[(size_t)r * n + k];
        jacobi_eigen(M, n, vec, val);
        double vmax = 0.0;
        for (double v : val) vmax = std::max(vmax, std::fabs(v));
        std::vector<double> P((size_t)n * n, 0.0);
        for (int t = 0; t < n; ++t) {
            if (std::fabs(val[(size_t)t]) > 1e-9 * (1.0 + vmax)) continue;
            for (int i = 0; i < n; ++i)
                for (int k = 0; k < n; ++k)
                    P[(size_t)i * n + k] += vec[(size_t)i * n + t] * vec[(size_t)k * n + t];
        }
        std::vector<int64_t> key(1, n);
        for (double x : P) key.push_back((int64_t)std::llround(x * 1e12));
        auto it = known.find(key);
        int64_t at;
        if (it != known.end()) {
            at = it->second;
        } else {
            at = (int64_t)mats.size();
            mats.insert(mats.end(), P.begin(), P.end());
            known[key] = at;
        }
        elem.insert(elem.end(), es.begin(), es.end());
        gptr.push_back((int64_t)elem.size());
        gmat.push_back(at);
    }
    return plan_add_spin_groups(plan, buf, off, std::move(gptr), std::move(elem), std::move(gmat), std::move(mats));
}

int gecco_hip_plan_add_sym_ab(gecco_hip_plan *plan, gecco_hip_list *L, int32_t buf, int64_t off,
                              double prefac, double relfac) {
    if (!plan || !L) return set_error(GECCO_HIP_EINVAL, "add_sym_ab: null argument");
    if (L->tra) return set_error(GECCO_HIP_EINVAL, "add_sym_ab: give the stored list, not its adjoint view");
    std::vector<int64_t> i1, i2;
    std::vector<int8_t> f;
    int rc = sym_pairs(L, off, i1, i2, f, relfac);
    if (rc) return rc;
    return plan_add_sym_pairs(plan, buf, std::move(i1), std::move(i2), std::move(f), prefac);
}

}  // extern "C"
