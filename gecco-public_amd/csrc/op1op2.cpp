// One binary contraction with contr_op1op2's semantics through a single
// C-ABI call whose arguments are those of contract/contr_op1op2.f:2-17
// flattened: occupations instead of the me_list / operator types, the
// orbital IRREPs instead of orb_info, whole-list buffers instead of the
// file buffers.  Built on the plan API (include/gecco_hip_model.h); every
// distinct contraction keeps its lists and plan for the life of the process,
// so a formula item GeCCo's scheduler repeats every iteration is enumerated
// and uploaded once.
#include <hip/hip_runtime.h>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <vector>
#include "../../include/gecco_hip.h"
#include "../../include/gecco_hip_model.h"

namespace {

struct OrbsEntry {
    gecco_hip_orbs *orbs = nullptr;
    std::map<std::vector<int64_t>, gecco_hip_list *> lists;   // (C[4], A[4]) -> stored list
};

struct CallEntry {
    gecco_hip_plan *plan = nullptr;
    std::vector<gecco_hip_list *> views;    // adjoint views (or reorder lists) owned by the entry
    int64_t len1 = 0, len2 = 0, len12 = 0;
    double *tmp = nullptr;                  // op1op2tmp of a contraction with a reorder
};

std::mutex g_mu;
std::map<std::vector<int64_t>, OrbsEntry> g_orbs;
std::map<std::vector<int64_t>, CallEntry> g_calls;
double *g_stage[3] = {nullptr, nullptr, nullptr};
int64_t g_stage_n[3] = {0, 0, 0};

int stage(int i, int64_t n, double **out) {
    if (g_stage_n[i] < n) {
        if (g_stage[i]) gecco_hip_free(g_stage[i]);
        g_stage[i] = nullptr;
        g_stage_n[i] = 0;
        if (gecco_hip_malloc((void **)&g_stage[i], std::max<int64_t>(n, 1) * 8)) return 1;
        g_stage_n[i] = n;
    }
    *out = g_stage[i];
    return 0;
}

bool on_device(const void *p) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice;
}

// the stored list of occupation (occ_c, occ_a), created once per orbital set
gecco_hip_list *stored(OrbsEntry &oe, const int64_t *occ_c, const int64_t *occ_a) {
    std::vector<int64_t> key(occ_c, occ_c + 4);
    key.insert(key.end(), occ_a, occ_a + 4);
    auto it = oe.lists.find(key);
    if (it != oe.lists.end()) return it->second;
    gecco_hip_list *L = nullptr;
    if (gecco_hip_list_create(oe.orbs, occ_c, occ_a, 1, &L)) return nullptr;
    oe.lists[key] = L;
    return L;
}

// the stored list of vertex-wise occupation iocc(4, 2, nj), created once
// per orbital set
gecco_hip_list *stored_n(OrbsEntry &oe, int64_t nj, const int64_t *iocc) {
    std::vector<int64_t> key{-nj};   // apart from the one-vertex keys
    key.insert(key.end(), iocc, iocc + 8 * nj);
    auto it = oe.lists.find(key);
    if (it != oe.lists.end()) return it->second;
    gecco_hip_list *L = nullptr;
    if (gecco_hip_list_create_n(oe.orbs, nj, iocc, 1, &L)) return nullptr;
    oe.lists[key] = L;
    return L;
}

OrbsEntry *orbs_entry(const int64_t *nsym, const int64_t *norb_h, const int64_t *igam_h,
                      const int64_t *norb_p, const int64_t *igam_p, std::vector<int64_t> &okey) {
    okey = {*nsym, *norb_h, *norb_p};
    okey.insert(okey.end(), igam_h, igam_h + *norb_h);
    okey.insert(okey.end(), igam_p, igam_p + *norb_p);
    OrbsEntry &oe = g_orbs[okey];
    if (!oe.orbs && gecco_hip_orbs_create(*nsym, *norb_h, igam_h, *norb_p, igam_p, &oe.orbs)) return nullptr;
    return &oe;
}

// length of a short-form merge map over ntgt target vertices
int64_t merge_len(const int64_t *mm, int64_t ntgt) {
    int64_t k = 0;
    for (int64_t t = 0; t < 2 * ntgt; ++t) {
        if (mm[k] < 0 || mm[k] > 16) return -1;
        k += 1 + mm[k];
    }
    return k;
}

}  // namespace

extern "C" {

void gecco_hip_contr_op1op2_(const double *xfac, const int64_t *update, const int64_t *iocc_op1,
                             const int64_t *iocc_op2, const int64_t *iocc_op1op2,
                             const int64_t *iocc_cnt, const int64_t *tra_op1, const int64_t *tra_op2,
                             const int64_t *tra_op1op2, const int64_t *use_tr, const int64_t *nsym,
                             const int64_t *norb_h, const int64_t *igam_h, const int64_t *norb_p,
                             const int64_t *igam_p, const double *xop1, const double *xop2,
                             double *xop1op2, int64_t *ierr) {
    *ierr = 0;
    std::lock_guard<std::mutex> lk(g_mu);
    std::vector<int64_t> okey{*nsym, *norb_h, *norb_p};
    okey.insert(okey.end(), igam_h, igam_h + *norb_h);
    okey.insert(okey.end(), igam_p, igam_p + *norb_p);
    OrbsEntry &oe = g_orbs[okey];
    if (!oe.orbs && gecco_hip_orbs_create(*nsym, *norb_h, igam_h, *norb_p, igam_p, &oe.orbs)) {
        *ierr = GECCO_HIP_EINVAL;
        return;
    }
    int dev = 0;
    (void)hipGetDevice(&dev);
    uint64_t xbits;
    std::memcpy(&xbits, xfac, 8);
    std::vector<int64_t> ckey(okey);
    ckey.insert(ckey.end(), {(int64_t)dev, (int64_t)xbits, *update != 0, *use_tr != 0,
                             *tra_op1 != 0, *tra_op2 != 0, *tra_op1op2 != 0});
    for (const int64_t *o : {iocc_op1, iocc_op2, iocc_op1op2, iocc_cnt}) ckey.insert(ckey.end(), o, o + 8);
    CallEntry &ce = g_calls[ckey];
    if (!ce.plan) {
        // the operands as the contraction uses them (iocc: C column, then A
        // column, each over H, P, V, X), stored as the adjoint when transposed
        const int64_t *occ[3] = {iocc_op1, iocc_op2, iocc_op1op2};
        const bool tra[3] = {*tra_op1 != 0, *tra_op2 != 0, *tra_op1op2 != 0};
        gecco_hip_list *used[3];
        int64_t lens[3];
        for (int q = 0; q < 3; ++q) {
            gecco_hip_list *L = tra[q] ? stored(oe, occ[q] + 4, occ[q]) : stored(oe, occ[q], occ[q] + 4);
            if (!L) {
                *ierr = GECCO_HIP_EINVAL;
                g_calls.erase(ckey);
                return;
            }
            lens[q] = gecco_hip_list_length(L);
            used[q] = L;
            if (tra[q]) {
                gecco_hip_list *V = nullptr;
                if (gecco_hip_list_adjoint(L, &V)) {
                    *ierr = GECCO_HIP_EINVAL;
                    g_calls.erase(ckey);
                    return;
                }
                ce.views.push_back(V);
                used[q] = V;
            }
        }
        gecco_hip_plan *p = nullptr;
        int rc = gecco_hip_plan_create(&p);
        // kernel calls with a zero operand block are skipped, as in the
        // reference (contract/contr_op1op2_wmaps_c.f:1652-1655,1767-1770)
        if (!rc) rc = gecco_hip_plan_set_screen(p, 1);
        // update = .false.: the result list starts from zero
        // (contract/contr_op1op2_wmaps_c.f:644,680)
        if (!rc && !*update) rc = gecco_hip_plan_add_init(p, 2, 0, lens[2], -1, 0, 0.0);
        int64_t nb = 0;
        double fl = 0;
        if (!rc)
            rc = gecco_hip_plan_add_contraction(p, used[0], used[1], used[2], iocc_cnt, iocc_cnt + 4, *xfac,
                                                *use_tr, 0, 0, 1, 0, 2, 0, &nb, &fl);
        // use_tr: the spin-flip symmetrisation of the stored result
        // (contract/contr_op1op2_wmaps_c.f:2099-2102)
        if (!rc && *use_tr) {
            gecco_hip_list *R = tra[2] ? stored(oe, occ[2] + 4, occ[2]) : stored(oe, occ[2], occ[2] + 4);
            rc = gecco_hip_plan_add_sym_ab(p, R, 2, 0, 0.5, 1.0);
        }
        if (!rc) rc = gecco_hip_plan_finalize(p);
        if (rc) {
            if (p) gecco_hip_plan_destroy(p);
            for (auto *v : ce.views) gecco_hip_list_destroy(v);
            g_calls.erase(ckey);
            *ierr = rc;
            return;
        }
        ce.plan = p;
        ce.len1 = lens[0];
        ce.len2 = lens[1];
        ce.len12 = lens[2];
    }
    double *d1 = const_cast<double *>(xop1), *d2 = const_cast<double *>(xop2), *d12 = xop1op2;
    const bool h1 = !on_device(xop1), h2 = !on_device(xop2), h12 = !on_device(xop1op2);
    int rc = 0;
    if (h1 && !rc) rc = stage(0, ce.len1, &d1) || gecco_hip_memcpy_htod(d1, xop1, ce.len1 * 8);
    if (h2 && !rc) rc = stage(1, ce.len2, &d2) || gecco_hip_memcpy_htod(d2, xop2, ce.len2 * 8);
    if (h12 && !rc) rc = stage(2, ce.len12, &d12) || gecco_hip_memcpy_htod(d12, xop1op2, ce.len12 * 8);
    double *bufs[3] = {d1, d2, d12};
    if (!rc) rc = gecco_hip_plan_execute(ce.plan, bufs, 3, nullptr);
    if (!rc && h12) rc = gecco_hip_memcpy_dtoh(xop1op2, d12, ce.len12 * 8);
    if (rc) {
        *ierr = gecco_hip_last_error(nullptr, 0);
        if (!*ierr) *ierr = GECCO_HIP_EDEVICE;
    }
}

void gecco_hip_reo_op_(const double *fac, const int64_t *update, const int64_t *sign_reo,
                       const int64_t *njoined_ori, const int64_t *iocc_ori, const int64_t *njoined_reo,
                       const int64_t *iocc_reo, const int64_t *from_vtx, const int64_t *to_vtx,
                       const int64_t *iocc_shift, const int64_t *nsym, const int64_t *norb_h,
                       const int64_t *igam_h, const int64_t *norb_p, const int64_t *igam_p,
                       const double *xori, double *xreo, int64_t *ierr) {
    *ierr = 0;
    std::lock_guard<std::mutex> lk(g_mu);
    std::vector<int64_t> okey{*nsym, *norb_h, *norb_p};
    okey.insert(okey.end(), igam_h, igam_h + *norb_h);
    okey.insert(okey.end(), igam_p, igam_p + *norb_p);
    OrbsEntry &oe = g_orbs[okey];
    if (!oe.orbs && gecco_hip_orbs_create(*nsym, *norb_h, igam_h, *norb_p, igam_p, &oe.orbs)) {
        *ierr = GECCO_HIP_EINVAL;
        return;
    }
    const int64_t nj[2] = {*njoined_ori, *njoined_reo};
    if (nj[0] < 1 || nj[0] > 8 || nj[1] < 1 || nj[1] > 8) {
        *ierr = GECCO_HIP_EINVAL;
        return;
    }
    int dev = 0;
    (void)hipGetDevice(&dev);
    uint64_t xbits;
    std::memcpy(&xbits, fac, 8);
    std::vector<int64_t> ckey(okey);
    ckey.insert(ckey.end(), {-1, (int64_t)dev, (int64_t)xbits, *update != 0, *sign_reo, *from_vtx,
                             *to_vtx, nj[0], nj[1]});
    ckey.insert(ckey.end(), iocc_ori, iocc_ori + 8 * nj[0]);
    ckey.insert(ckey.end(), iocc_reo, iocc_reo + 8 * nj[1]);
    ckey.insert(ckey.end(), iocc_shift, iocc_shift + 8);
    CallEntry &ce = g_calls[ckey];
    if (!ce.plan) {
        // lists of several vertices: iocc(4, 2, njoined) is the vertex-wise
        // (C, A) occupation, the layout gecco_hip_list_create_n takes
        gecco_hip_list *L[2] = {nullptr, nullptr};
        const int64_t *occ[2] = {iocc_ori, iocc_reo};
        int rc = 0;
        for (int q = 0; q < 2 && !rc; ++q) rc = gecco_hip_list_create_n(oe.orbs, nj[q], occ[q], 1, &L[q]);
        gecco_hip_plan *p = nullptr;
        if (!rc) rc = gecco_hip_plan_create(&p);
        const int64_t lr = rc ? 0 : gecco_hip_list_length(L[1]);
        if (!rc && !*update) rc = gecco_hip_plan_add_init(p, 1, 0, lr, -1, 0, 0.0);
        int64_t nt = 0;
        if (!rc)
            rc = gecco_hip_plan_add_reorder(p, L[0], L[1], *from_vtx - 1, *to_vtx - 1, iocc_shift,
                                            iocc_shift + 4, *fac, *sign_reo, 0, 0, 1, 0, &nt);
        if (!rc) rc = gecco_hip_plan_finalize(p);
        if (rc) {
            if (p) gecco_hip_plan_destroy(p);
            for (auto *l : L) if (l) gecco_hip_list_destroy(l);
            g_calls.erase(ckey);
            *ierr = rc;
            return;
        }
        ce.plan = p;
        ce.len1 = gecco_hip_list_length(L[0]);
        ce.len12 = lr;
        ce.views.assign(L, L + 2);
    }
    double *d1 = const_cast<double *>(xori), *d2 = xreo;
    const bool h1 = !on_device(xori), h2 = !on_device(xreo);
    int rc = 0;
    if (h1 && !rc) rc = stage(0, ce.len1, &d1) || gecco_hip_memcpy_htod(d1, xori, ce.len1 * 8);
    if (h2 && !rc) rc = stage(2, ce.len12, &d2) || gecco_hip_memcpy_htod(d2, xreo, ce.len12 * 8);
    double *bufs[2] = {d1, d2};
    if (!rc) rc = gecco_hip_plan_execute(ce.plan, bufs, 2, nullptr);
    if (!rc && h2) rc = gecco_hip_memcpy_dtoh(xreo, d2, ce.len12 * 8);
    if (rc) {
        *ierr = gecco_hip_last_error(nullptr, 0);
        if (!*ierr) *ierr = GECCO_HIP_EDEVICE;
    }
}

void gecco_hip_contr_op1op2_n_(const double *xfac, const int64_t *update, const int64_t *njoined_op1,
                               const int64_t *iocc_op1, const int64_t *njoined_op2, const int64_t *iocc_op2,
                               const int64_t *njoined_op1op2, const int64_t *iocc_op1op2,
                               const int64_t *iocc_op1op2tmp, const int64_t *njoined_cnt,
                               const int64_t *iocc_ex1, const int64_t *iocc_ex2, const int64_t *iocc_cnt,
                               const int64_t *merge_op1, const int64_t *merge_op2,
                               const int64_t *merge_op1op2, const int64_t *merge_op2op1, const int64_t *nreo,
                               const int64_t *from_vtx, const int64_t *to_vtx, const int64_t *iocc_shift,
                               const int64_t *sign_reo, const int64_t *use_tr, const int64_t *nsym,
                               const int64_t *norb_h, const int64_t *igam_h, const int64_t *norb_p,
                               const int64_t *igam_p, const double *xop1, const double *xop2,
                               double *xop1op2, int64_t *ierr) {
    *ierr = 0;
    std::lock_guard<std::mutex> lk(g_mu);
    const int64_t n1 = *njoined_op1, n2 = *njoined_op2, n12 = *njoined_op1op2, nc = *njoined_cnt;
    if (n1 < 1 || n1 > 8 || n2 < 1 || n2 > 8 || n12 < 1 || n12 > 8 || nc < 0 || nc > 16 ||
        (*nreo != 0 && *nreo != 1)) {
        *ierr = GECCO_HIP_EINVAL;
        return;
    }
    std::vector<int64_t> okey;
    OrbsEntry *oe = orbs_entry(nsym, norb_h, igam_h, norb_p, igam_p, okey);
    if (!oe) {
        *ierr = GECCO_HIP_EINVAL;
        return;
    }
    const int64_t l1 = merge_len(merge_op1, n1), l2 = merge_len(merge_op2, n2),
                  l12 = merge_len(merge_op1op2, n12), l21 = merge_len(merge_op2op1, n12);
    if (l1 < 0 || l2 < 0 || l12 < 0 || l21 < 0) {
        *ierr = GECCO_HIP_EINVAL;
        return;
    }
    int dev = 0;
    (void)hipGetDevice(&dev);
    uint64_t xbits;
    std::memcpy(&xbits, xfac, 8);
    std::vector<int64_t> ckey(okey);
    ckey.insert(ckey.end(), {-2, (int64_t)dev, (int64_t)xbits, *update != 0, *use_tr != 0, n1, n2, n12, nc,
                             *nreo});
    ckey.insert(ckey.end(), iocc_op1, iocc_op1 + 8 * n1);
    ckey.insert(ckey.end(), iocc_op2, iocc_op2 + 8 * n2);
    ckey.insert(ckey.end(), iocc_op1op2, iocc_op1op2 + 8 * n12);
    ckey.insert(ckey.end(), iocc_op1op2tmp, iocc_op1op2tmp + 8 * n12);
    ckey.insert(ckey.end(), iocc_ex1, iocc_ex1 + 8 * n1);
    ckey.insert(ckey.end(), iocc_ex2, iocc_ex2 + 8 * n2);
    if (nc) ckey.insert(ckey.end(), iocc_cnt, iocc_cnt + 8 * nc);
    ckey.insert(ckey.end(), merge_op1, merge_op1 + l1);
    ckey.insert(ckey.end(), merge_op2, merge_op2 + l2);
    ckey.insert(ckey.end(), merge_op1op2, merge_op1op2 + l12);
    ckey.insert(ckey.end(), merge_op2op1, merge_op2op1 + l21);
    if (*nreo) {
        ckey.insert(ckey.end(), {*from_vtx, *to_vtx, *sign_reo});
        ckey.insert(ckey.end(), iocc_shift, iocc_shift + 8);
    }
    CallEntry &ce = g_calls[ckey];
    if (!ce.plan) {
        // the kernel calls write op1op2tmp: the target itself, or (with a
        // reorder) a device list zeroed per call and reordered into the
        // target afterwards (contract/contr_op1op2_wmaps_c.f:1962-1995)
        gecco_hip_list *L1 = stored_n(*oe, n1, iocc_op1), *L2 = stored_n(*oe, n2, iocc_op2);
        gecco_hip_list *L12 = stored_n(*oe, n12, iocc_op1op2);
        gecco_hip_list *Lt = *nreo ? stored_n(*oe, n12, iocc_op1op2tmp) : L12;
        int rc = (L1 && L2 && L12 && Lt) ? 0 : GECCO_HIP_EINVAL;
        gecco_hip_plan *p = nullptr;
        if (!rc) rc = gecco_hip_plan_create(&p);
        if (!rc) rc = gecco_hip_plan_set_screen(p, 1);
        const int64_t lt = rc ? 0 : gecco_hip_list_length(Lt), lr = rc ? 0 : gecco_hip_list_length(L12);
        if (!rc && !*update) rc = gecco_hip_plan_add_init(p, 2, 0, lr, -1, 0, 0.0);
        if (!rc && *nreo) rc = gecco_hip_plan_add_init(p, 3, 0, lt, -1, 0, 0.0);
        int64_t nb = 0, nt = 0;
        double fl = 0;
        if (!rc)
            rc = gecco_hip_plan_add_contraction_n(p, L1, L2, Lt, nc, iocc_ex1, iocc_ex2, iocc_cnt, merge_op1,
                                                  merge_op2, merge_op1op2, merge_op2op1, *xfac, *use_tr, 0, 0,
                                                  1, 0, *nreo ? 3 : 2, 0, &nb, &fl);
        if (!rc && *use_tr) rc = gecco_hip_plan_add_sym_ab(p, Lt, *nreo ? 3 : 2, 0, 0.5, 1.0);
        if (!rc && *nreo)
            rc = gecco_hip_plan_add_reorder(p, Lt, L12, *from_vtx - 1, *to_vtx - 1, iocc_shift, iocc_shift + 4,
                                            1.0, *sign_reo, 3, 0, 2, 0, &nt);
        if (!rc) rc = gecco_hip_plan_finalize(p);
        double *tmp = nullptr;
        if (!rc && *nreo && gecco_hip_malloc((void **)&tmp, std::max<int64_t>(lt, 1) * 8)) rc = GECCO_HIP_EDEVICE;
        if (rc) {
            if (p) gecco_hip_plan_destroy(p);
            g_calls.erase(ckey);
            *ierr = rc;
            return;
        }
        ce.plan = p;
        ce.len1 = gecco_hip_list_length(L1);
        ce.len2 = gecco_hip_list_length(L2);
        ce.len12 = lr;
        ce.tmp = tmp;
    }
    double *d1 = const_cast<double *>(xop1), *d2 = const_cast<double *>(xop2), *d12 = xop1op2;
    const bool h1 = !on_device(xop1), h2 = !on_device(xop2), h12 = !on_device(xop1op2);
    int rc = 0;
    if (h1 && !rc) rc = stage(0, ce.len1, &d1) || gecco_hip_memcpy_htod(d1, xop1, ce.len1 * 8);
    if (h2 && !rc) rc = stage(1, ce.len2, &d2) || gecco_hip_memcpy_htod(d2, xop2, ce.len2 * 8);
    if (h12 && !rc) rc = stage(2, ce.len12, &d12) || gecco_hip_memcpy_htod(d12, xop1op2, ce.len12 * 8);
    double *bufs[4] = {d1, d2, d12, ce.tmp ? ce.tmp : d12};
    if (!rc) rc = gecco_hip_plan_execute(ce.plan, bufs, 4, nullptr);
    if (!rc && h12) rc = gecco_hip_memcpy_dtoh(xop1op2, d12, ce.len12 * 8);
    if (rc) {
        *ierr = gecco_hip_last_error(nullptr, 0);
        if (!*ierr) *ierr = GECCO_HIP_EDEVICE;
    }
}

}  // extern "C"
