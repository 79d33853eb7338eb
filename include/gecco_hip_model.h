/*
 * gecco_hip_model.h -- plan builder of the MI355X block-contraction engine.
 *
 * Builds GeCCo's data model on the host -- orbital strings, string maps,
 * ME-list layout -- and enumerates the kernel calls of a binary contraction
 * as contr_op1op2_wmaps_c does (contract/contr_op1op2_wmaps_c.f:964-2070),
 * recording them into a gecco_hip_plan (include/gecco_hip.h).  This is the
 * "option N" boundary: a driver that holds device-resident ME-lists calls
 * gecco_hip_plan_add_contraction once per [CONTR] formula item instead of
 * contr_op1op2 (contract/contr_op1op2.f:2-184, reached from
 * contract/fs_contr_drv.f:273).
 *
 * Scope: operators over hole (H, hpvx 1) and particle (P, hpvx 2) spaces
 * with one orbital subspace each, Ms(total) 0, totally symmetric (gamt 1),
 * as in closed-shell CC; one-vertex operands and results may be adjoint
 * views (gecco_hip_list_adjoint); operators with several vertices go
 * through gecco_hip_plan_add_contraction_n and gecco_hip_plan_add_reorder.
 */
#ifndef GECCO_HIP_MODEL_H
#define GECCO_HIP_MODEL_H

#include <stdint.h>
#include "gecco_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gecco_hip_orbs gecco_hip_orbs;
typedef struct gecco_hip_list gecco_hip_list;

/* Orbital spaces: IRREPs (1..nsym, D2h subgroup labels) of the hole and
 * particle orbitals; orbitals are ordered by IRREP inside each space, as
 * GeCCo's mostnd arrays assume (strings/lexlstr.f:60,80). */
int gecco_hip_orbs_create(int64_t nsym, int64_t nh, const int64_t *irrep_h, int64_t np,
                          const int64_t *irrep_p, gecco_hip_orbs **out);
int gecco_hip_orbs_destroy(gecco_hip_orbs *orbs);
/* Number of strings of (space hpvx, nel, idxms, igam): graph%lenstr_gm. */
int64_t gecco_hip_orbs_lenstr(gecco_hip_orbs *orbs, int64_t hpvx, int64_t nel, int64_t idxms,
                              int64_t igam);
/* (str1,str2)->str12 map block, str1 fastest (strings/set_strmap_kernel.f). */
int gecco_hip_orbs_strmap(gecco_hip_orbs *orbs, int64_t hpvx, int64_t n1, int64_t idxms1,
                          int64_t g1, int64_t n2, int64_t idxms2, int64_t g2, int64_t *out,
                          int64_t maxlen, int64_t *len);
/* Spin-flip map block (strings/set_flipmap_kernel.f). */
int gecco_hip_orbs_flipmap(gecco_hip_orbs *orbs, int64_t hpvx, int64_t nel, int64_t idxms,
                           int64_t igam, int64_t *out, int64_t maxlen, int64_t *len);

/* ME-list of a one-vertex operator; occ_c/occ_a = occupation per space type
 * (H, P, V, X); absym = alpha/beta (spin-flip) symmetry, +1 for singlets.
 * Layout: operators/set_op_dim2.f:146-459. */
int gecco_hip_list_create(gecco_hip_orbs *orbs, const int64_t *occ_c, const int64_t *occ_a,
                          int64_t absym, gecco_hip_list **out);
/* ME-list of an operator with njoined vertices (intermediates with several
 * vertices): occ[v*8 + 0..3] = C occupation of vertex v (H, P, V, X),
 * occ[v*8 + 4..7] = its A occupation.  Components are condensed as
 * operators/condense_occ.f does (X,P,V,H outer, vertex inner); layout as
 * gecco_hip_list_create.  Contractions take them through
 * gecco_hip_plan_add_contraction_n, reorders (gecco_hip_plan_add_reorder)
 * directly; traces and spin projections take one-vertex lists. */
int gecco_hip_list_create_n(gecco_hip_orbs *orbs, int64_t njoined, const int64_t *occ,
                            int64_t absym, gecco_hip_list **out);
/* Adjoint view X^+ of a one-vertex list X, for a contraction operand or
 * result used with the dagger flag (contract/get_bc_info3.f:173,191,391:
 * tra_op1 / tra_op2 / tra_op1op2 = vertex%dagger, passed to the kernels by
 * contract/contr_op1op2_wmaps_c.f:844-855,1903,1931).  The view's C side is
 * X's A side and vice versa; element (c; a) of the view is X's stored
 * element (a; c), addressed through the transposed leading dimensions of
 * contract/set_op_ldim_c.f:51-70 with no sign.  The view shares X's layout
 * (offsets, length): give gecco_hip_plan_add_contraction the view and the
 * buffer slot of X; spin-flip symmetrisation, traces and reorders take the
 * stored list.  Destroy the view with gecco_hip_list_destroy. */
int gecco_hip_list_adjoint(const gecco_hip_list *list, gecco_hip_list **out);
int gecco_hip_list_destroy(gecco_hip_list *list);
int64_t gecco_hip_list_length(const gecco_hip_list *list);
int64_t gecco_hip_list_ndist(const gecco_hip_list *list);
/* info = [off, len, Ms(A), IRREP(A), ncomp_c, ncomp_a,
 *         per C component (ms, idxms, irrep, nstr, ldim), then per A comp]. */
int gecco_hip_list_dist(const gecco_hip_list *list, int64_t idist, int64_t *info, int64_t ninfo);

/* Spin-orbitals of every element (integral import, denominators):
 * out[elem*(nc+na) + q] = 2*(orb-1)+spin (spin 0 alpha, 1 beta) of electron q,
 * C components then A components, each string ascending; orbitals numbered
 * inside their space.  nout >= length*(nc+na). */
int gecco_hip_list_spinorbs(const gecco_hip_list *list, int16_t *out, int64_t nout);

/* Diagonal preconditioner of a list from orbital energies (Fock diagonal),
 * contract/dia4op_ev.f:298-365 (one-electron part): out[e] = sum of eps
 * over the electrons of the C strings - sum over the A strings; eps_h[nh]
 * and eps_p[np] indexed by the orbital inside its space.  Host arrays,
 * nout >= length. */
int gecco_hip_list_diagonal(const gecco_hip_list *list, const double *eps_h, const double *eps_p,
                            double *out, int64_t nout);

/* Record every kernel call of  Res += fact * Op1 . Op2  (contracted part of
 * Op1's C side cnt_c[4], of Op1's A side cnt_a[4], per space type) into the
 * plan, with spin-flip skipping and factors when use_tr (the use_tr keyword,
 * contract/contr_op1op2_wmaps_c.f:454-458,1007-1101).  Lists live in buffer
 * slots buf* at element offsets off*.  Returns the number of kernel calls and
 * their algorithmic FLOP.  With use_tr the caller adds the result's
 * spin-flip symmetrisation with gecco_hip_plan_add_sym_ab. */
int gecco_hip_plan_add_contraction(gecco_hip_plan *plan, const gecco_hip_list *op1,
                                   const gecco_hip_list *op2, const gecco_hip_list *res,
                                   const int64_t *cnt_c, const int64_t *cnt_a, double fact,
                                   int64_t use_tr, int32_t buf1, int64_t off1, int32_t buf2,
                                   int64_t off2, int32_t buf12, int64_t off12, int64_t *nblocks,
                                   double *flops);
/* The same for operators with several vertices (lists from
 * gecco_hip_list_create_n), with contr_op1op2's description of the
 * contraction (contract/contr_op1op2.f:2-17, the arguments iocc_ex1,
 * iocc_ex2, iocc_cnt, merge_op1, merge_op2, merge_op1op2, merge_op2op1,
 * njoined_cnt): iocc_ex1 (4, 2, op1 vertices), iocc_ex2 (4, 2, op2
 * vertices), iocc_cnt (4, 2, njoined_cnt) with the C column before the A
 * column (CNT C meets op1's C side), and the short-form merge maps of
 * contract/get_bc_info3.f:271-330 (condense_merge_map: per target vertex a
 * count and the 1-based source-1 vertices, then the same for source 2):
 * op1 <- (CNT, EX1), op2 <- (CNT^+, EX2), res C <- (EX1, EX2), res A <-
 * (EX2, EX1).  res is the list the kernel calls write (me_op1op2tmp); a
 * target grouped differently is reached by gecco_hip_plan_add_reorder of
 * res afterwards, as contract/contr_op1op2_wmaps_c.f:1962-1995 reorders
 * each result block.  Stored lists only (no adjoint views); a component
 * that would merge two components of one source list is refused (route 3
 * keeps one string per target component and source). */
int gecco_hip_plan_add_contraction_n(gecco_hip_plan *plan, const gecco_hip_list *op1,
                                     const gecco_hip_list *op2, const gecco_hip_list *res,
                                     int64_t njoined_cnt, const int64_t *iocc_ex1,
                                     const int64_t *iocc_ex2, const int64_t *iocc_cnt,
                                     const int64_t *merge_op1, const int64_t *merge_op2,
                                     const int64_t *merge_op1op2, const int64_t *merge_op2op1,
                                     double fact, int64_t use_tr, int32_t buf1, int64_t off1,
                                     int32_t buf2, int64_t off2, int32_t buf12, int64_t off12,
                                     int64_t *nblocks, double *flops);
/* Partial trace (self-contraction) of a list, run after the plan's
 * contractions:  OPtr(IC, IA) += fact * sum_J s * OP(J.IC, J.IA), with the
 * same string J (occupation cnt[4] per space type H, P, V, X) removed from
 * the C and the A side of OP.  Replaces contract/trace_op.f:2-830 with its
 * kernel contract/trace_op_blk.f (dispatched from contract/contr_op1op2.f:162
 * for self-contractions); trop may be the scalar list (a full trace, the
 * xret case).  Returns the number of element terms. */
int gecco_hip_plan_add_trace(gecco_hip_plan *plan, const gecco_hip_list *op,
                             const gecco_hip_list *trop, const int64_t *cnt, double fact,
                             int32_t buf_op, int64_t off_op, int32_t buf_trop, int64_t off_trop,
                             int64_t *nterms);

/* Reorder: the shift string K (occupations shift_c[4], shift_a[4] per space
 * H, P, V, X) moves from vertex from_vtx of ori to vertex to_vtx of reo,
 *   REO(I') += fact * sign_reo * s_ori(I0,K) * s_reo(I0,K) * ORI(I),
 * I0 = ORI without K; C strings merge (I0, K), A strings (K, I0).  Replaces
 * contract/reo_op_wmaps_c.f:2-575 with its block kernel
 * contract/reo_blk_wmaps_c.f:2-854 (called from contract/fs_reo_drv.f:166
 * and for reordered contraction results, contract/contr_op1op2_wmaps_c.f:1987).
 * Vertices are numbered as ORI's; REO may have fewer vertices when some of
 * I0-with-K are empty (they are dropped: an empty vertex holds no
 * component, so a two-vertex list whose second vertex empties has the
 * one-vertex list's layout).
 * sign_reo (+1/-1) is the sign the formula layer assigns to moving K past
 * the operators of intervening vertices (reorder_info%sign_reo).  Runs after
 * the plan's contractions, with the partial traces.  Returns the number of
 * element terms. */
int gecco_hip_plan_add_reorder(gecco_hip_plan *plan, const gecco_hip_list *ori,
                               const gecco_hip_list *reo, int64_t from_vtx, int64_t to_vtx,
                               const int64_t *shift_c, const int64_t *shift_a, double fact,
                               int64_t sign_reo, int32_t buf_ori, int64_t off_ori, int32_t buf_reo,
                               int64_t off_reo, int64_t *nterms);

/* Spin-flip symmetrisation of a result list after the plan's contractions
 * (contract/sym_ab_blk.f, called with prefac 1/2, relfac absym at
 * contract/contr_op1op2_wmaps_c.f:2099-2102). */
int gecco_hip_plan_add_sym_ab(gecco_hip_plan *plan, gecco_hip_list *list, int32_t buf,
                              int64_t off, double prefac, double relfac);

/* Singlet spin projection of a particle-conserving one-vertex list after the
 * plan's contractions and spin-flip projections (contract/spin_prj_list.f,
 * contract/spin_prj_blk.f with s2 = 0, as contract/frm_sched2.f:146-150
 * applies it to a result with me_res%s2 >= 0): every group of elements with
 * the same spatial orbitals and different spins is replaced by its
 * projection onto the singlet operators (the null space of [S+, .]), at most
 * three creators and three annihilators, hole and particle spaces. */
int gecco_hip_plan_add_spin_project(gecco_hip_plan *plan, gecco_hip_list *list, int32_t buf, int64_t off);

/* One binary contraction  Res (+)= xfac * Op1 . Op2  with the semantics of
 * contr_op1op2 (contract/contr_op1op2.f:2-17, called from
 * contract/fs_contr_drv.f:273-288), its arguments flattened from GeCCo's
 * derived types so that a thin Fortran shim can pass them (Fortran calling
 * convention: every argument by reference, INTEGER = int64):
 *   xfac                 the factor (fact * bc_sign of the formula item)
 *   update               0: the result list is zeroed first
 *                        (contract/contr_op1op2_wmaps_c.f:644,680), else
 *                        the contraction is added to it
 *   iocc_op1, iocc_op2, iocc_op1op2
 *                        occupations (ngastp = 4, 2) of the operators as the
 *                        contraction uses them -- iocc(:,1) C over H, P, V,
 *                        X, iocc(:,2) A -- one vertex each, H and P only
 *   iocc_cnt             the contracted strings: (:,1) on Op1's C side,
 *                        (:,2) on Op1's A side
 *   tra_op1, tra_op2, tra_op1op2
 *                        the dagger flags: the list is stored as the adjoint
 *                        (its C and A occupations swapped)
 *   use_tr               spin-flip symmetry (the use_tr keyword), with the
 *                        result's symmetrisation
 *   nsym, norb_h, igam_h[norb_h], norb_p, igam_p[norb_p]
 *                        IRREPs of the hole and particle orbitals (ordered by
 *                        IRREP), from orb_info
 *   xop1, xop2, xop1op2  the whole lists in GeCCo's layout, host or device
 *                        memory (host buffers are staged)
 *   ierr                 0, or a GECCO_HIP_E* code (gecco_hip_last_error
 *                        has the message; the shim turns it into quit).
 * Kernel calls whose Op1 or Op2 block has a squared norm below 1e-28 are
 * skipped, as the reference does (gecco_hip_plan_set_screen).  Every
 * distinct contraction keeps its lists and plan for the life of the
 * process.  The kernels run on the null stream; a host result is copied back
 * before the call returns. */
void gecco_hip_contr_op1op2_(const double *xfac, const int64_t *update, const int64_t *iocc_op1,
                             const int64_t *iocc_op2, const int64_t *iocc_op1op2,
                             const int64_t *iocc_cnt, const int64_t *tra_op1, const int64_t *tra_op2,
                             const int64_t *tra_op1op2, const int64_t *use_tr, const int64_t *nsym,
                             const int64_t *norb_h, const int64_t *igam_h, const int64_t *norb_p,
                             const int64_t *igam_p, const double *xop1, const double *xop2,
                             double *xop1op2, int64_t *ierr);

/* The same entry for operators with several vertices, with the rest of
 * contr_op1op2's argument list (contract/contr_op1op2.f:2-17): njoined_op1,
 * njoined_op2, njoined_op1op2, njoined_cnt; iocc_op1 (4, 2, njoined_op1)
 * and the like, vertex-wise; iocc_ex1, iocc_ex2, iocc_cnt and the four
 * short-form merge maps merge_op1, merge_op2, merge_op1op2, merge_op2op1
 * (as gecco_hip_plan_add_contraction_n takes them); iocc_op1op2tmp, the
 * occupation the kernel calls write (me_op1op2tmp), and the reorder after
 * the kernel calls that reo_info describes (contract/contr_op1op2_wmaps_c.f:
 * 1962-1995): nreo = 0 (op1op2tmp is op1op2) or 1 -- the shift string
 * iocc_shift(4, 2) moves from vertex from_vtx of op1op2tmp to vertex
 * to_vtx of op1op2 (numbered from 1) with the sign sign_reo; the result of
 * the kernel calls then lives in a device list of the call's own and is
 * reordered into the target.  Stored lists only (no dagger flags); the
 * other arguments as gecco_hip_contr_op1op2_. */
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
                               double *xop1op2, int64_t *ierr);

/* Reorder of a whole list with the semantics of reo_op_wmaps_c
 * (contract/reo_op_wmaps_c.f:2-9, called from contract/fs_reo_drv.f:166),
 * its derived-type arguments flattened as for gecco_hip_contr_op1op2_:
 *   REO (+)= fac * sign_reo * (ORI with the shift string moved from vertex
 *   from_vtx to vertex to_vtx), through gecco_hip_plan_add_reorder;
 *   iocc_ori(4, 2, njoined_ori), iocc_reo(4, 2, njoined_reo): vertex-wise
 *   C and A occupations over H, P, V, X; iocc_shift(4, 2): the moved string
 *   (reo_info's shift occupation); vertices numbered from 1; update 0 zeroes
 *   REO first; orbital IRREPs, buffers and ierr as gecco_hip_contr_op1op2_. */
void gecco_hip_reo_op_(const double *fac, const int64_t *update, const int64_t *sign_reo,
                       const int64_t *njoined_ori, const int64_t *iocc_ori, const int64_t *njoined_reo,
                       const int64_t *iocc_reo, const int64_t *from_vtx, const int64_t *to_vtx,
                       const int64_t *iocc_shift, const int64_t *nsym, const int64_t *norb_h,
                       const int64_t *igam_h, const int64_t *norb_p, const int64_t *igam_p,
                       const double *xori, double *xreo, int64_t *ierr);

#ifdef __cplusplus
}
#endif
#endif /* GECCO_HIP_MODEL_H */
