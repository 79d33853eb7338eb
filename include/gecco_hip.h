/*
 * gecco_hip.h -- C ABI of the MI355X (gfx950) block-contraction engine.
 *
 * Two layers:
 *
 *  1. Drop-in kernel entry points with GeCCo's own Fortran argument lists
 *     (all arguments by reference, default INTEGER/LOGICAL = 8 bytes, symbol
 *     = lower case + trailing underscore).  A GeCCo build links this library
 *     instead of compiling the Fortran kernel; the block enumerator
 *     contr_op1op2_wmaps_c calls them unchanged.  Operand buffers may be host
 *     memory (they are staged through device memory) or device memory.
 *
 *  2. A plan API for the grouped launch: every kernel call of a formula step
 *     (or of many independent problems) is recorded once with
 *     gecco_hip_plan_add_block (the same arguments, operands given as buffer
 *     slot + element offset), resolved into device index tables by
 *     gecco_hip_plan_finalize, and then executed any number of times as one
 *     persistent launch by gecco_hip_plan_execute.
 *
 * Errors: the Fortran-named entry points cannot return a status (the
 * reference's subroutines stop through quit(), baserout/quit.F:36-50); they
 * record it, and gecco_hip_last_error() returns it (0 = success).  All other
 * functions return 0 on success and a GECCO_HIP_E* code otherwise.
 */
#ifndef GECCO_HIP_H
#define GECCO_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GECCO_HIP_OK 0
#define GECCO_HIP_EINVAL 1    /* inconsistent dimensions, maps or arguments */
#define GECCO_HIP_ENOMEM 2    /* device allocation failed                    */
#define GECCO_HIP_EDEVICE 3   /* HIP runtime error / no device               */
#define GECCO_HIP_ESTATE 4    /* plan used in the wrong state                */

/* Last error of the calling thread (0 if none) and its message. */
int gecco_hip_last_error(char *msg, int64_t msglen);

/* Route 3 kernel: R(EX1,EX2) += xfac * sum_CNT Op1(CNT,EX1) Op2(CNT^+,EX2)
 * with signed gathers and signed scatter through the string maps.
 * Replaces contract/contr_blk1blk2_blocked_mm.f:2-20 (called from
 * contract/contr_op1op2_wmaps_c.f:1927).  xscr/lenscr/lenblock/lenblock_cnt
 * are accepted for signature compatibility; the GPU kernel needs no host
 * scratch and its tiling does not depend on them.  The first call with a
 * given set of maps, lengths and factor builds a one-call plan (device index
 * tables) that later identical calls reuse, so a call GeCCo repeats every
 * iteration does no device allocation and no upload; the kernels run on the
 * null stream and only a host-resident result is copied back (which waits
 * for them). */
void contr_blk1blk2_blocked_mm_(
    const double *xfac, double *xop1op2, const double *xop1, const double *xop2,
    const int64_t *tra_op1, const int64_t *tra_op2, const int64_t *tra_op1op2,
    double *xscr, const int64_t *lenscr, const int64_t *lenblock,
    const int64_t *lenblock_cnt,
    const int64_t *ncblk_op1, const int64_t *nablk_op1,
    const int64_t *ncblk_ex1, const int64_t *nablk_ex1,
    const int64_t *ncblk_op2, const int64_t *nablk_op2,
    const int64_t *ncblk_ex2, const int64_t *nablk_ex2,
    const int64_t *ncblk_cnt, const int64_t *nablk_cnt,
    const int64_t *ncblk_op1op2, const int64_t *nablk_op1op2,
    const int64_t *hpvxop1c, const int64_t *hpvxop1a,
    const int64_t *hpvxop2c, const int64_t *hpvxop2a,
    const int64_t *hpvxop1op2c, const int64_t *hpvxop1op2a,
    const int64_t *lstrop1, const int64_t *lstrop2, const int64_t *lstrop1op2,
    const int64_t *lstr_ex1, const int64_t *lstr_ex2, const int64_t *lstr_cnt,
    const int64_t *map_info_12c, const int64_t *map_info_12a,
    const int64_t *map_info_1c, const int64_t *map_info_1a,
    const int64_t *map_info_2c, const int64_t *map_info_2a,
    const int64_t *map_ex1ex2c, const int64_t *map_ex1ex2a,
    const int64_t *map_ex1cntc, const int64_t *map_ex1cnta,
    const int64_t *map_ex2cntc, const int64_t *map_ex2cnta);

/* Route 2 kernel (map-driven scalar contraction in the reference), same
 * result as route 3.  Replaces contract/contr_blk1blk2_wmaps_c.f:2-19
 * (called from contract/contr_op1op2_wmaps_c.f:1899).  On the GPU both
 * routes run the same fused gather/MFMA/scatter kernel. */
void contr_blk1blk2_wmaps_c_(
    const double *xfac, double *xop1op2, const double *xop1, const double *xop2,
    const int64_t *tra_op1, const int64_t *tra_op2, const int64_t *tra_op1op2,
    const int64_t *ncblk_op1, const int64_t *nablk_op1,
    const int64_t *ncblk_ex1, const int64_t *nablk_ex1,
    const int64_t *ncblk_op2, const int64_t *nablk_op2,
    const int64_t *ncblk_ex2, const int64_t *nablk_ex2,
    const int64_t *ncblk_cnt, const int64_t *nablk_cnt,
    const int64_t *ncblk_op1op2, const int64_t *nablk_op1op2,
    const int64_t *hpvxop1c, const int64_t *hpvxop1a,
    const int64_t *hpvxop2c, const int64_t *hpvxop2a,
    const int64_t *hpvxop1op2c, const int64_t *hpvxop1op2a,
    const int64_t *lstrop1, const int64_t *lstrop2, const int64_t *lstrop1op2,
    const int64_t *lstr_ex1, const int64_t *lstr_ex2, const int64_t *lstr_cnt,
    const int64_t *map_info_12c, const int64_t *map_info_12a,
    const int64_t *map_info_1c, const int64_t *map_info_1a,
    const int64_t *map_info_2c, const int64_t *map_info_2a,
    const int64_t *map_ex1ex2c, const int64_t *map_ex1ex2a,
    const int64_t *map_ex1cntc, const int64_t *map_ex1cnta,
    const int64_t *map_ex2cntc, const int64_t *map_ex2cnta);

/* The collect step of route 3 (contract/collect_block_impr.f:2-17, called
 * from contract/contr_blk1blk2_blocked_mm.f for each operand batch): the
 * batch block xblk(k_len, l_len) = sign * xop(element) of the K strings
 * [ka_st0/kc_st0 .. ka_nd0/kc_nd0] and L strings [la_st0/lc_st0 ..
 * la_nd0/lc_nd0] through the string maps map_kl_c / map_kl_a, zero where a
 * map entry is 0; kca_trp != 0 runs K's A strings fastest (the operand
 * stored as Op(CNT^+)).  Same arguments and meaning as the reference, all
 * by reference (int64).  The host computes each block element's source
 * and sign; the GPU moves the data.  xblk and xop may be host arrays
 * (staged) or device memory; errors are recorded for gecco_hip_last_error. */
void collect_block_impr_(
    double *xblk, const double *xop, const int64_t *kc_st0, const int64_t *kc_nd0,
    const int64_t *kc_len, const int64_t *ka_st0, const int64_t *ka_nd0, const int64_t *ka_len,
    const int64_t *k_len, const int64_t *kca_trp, const int64_t *lc_st0, const int64_t *lc_nd0,
    const int64_t *lc_len, const int64_t *la_st0, const int64_t *la_nd0, const int64_t *la_len,
    const int64_t *l_len, const int64_t *map_kl_c, const int64_t *map_kl_a,
    const int64_t *ncblk_op, const int64_t *nablk_op, const int64_t *ldim_opc,
    const int64_t *ldim_opa, const int64_t *ncblk_k, const int64_t *nablk_k,
    const int64_t *ncblk_l, const int64_t *nablk_l, const int64_t *lstr_kc,
    const int64_t *lstr_ka, const int64_t *lstr_lc, const int64_t *lstr_la,
    const int64_t *nstr_kc, const int64_t *nstr_ka, const int64_t *nstr_lc,
    const int64_t *nstr_la, const int64_t *ireo_kc, const int64_t *ireo_ka,
    const int64_t *ireo_lc, const int64_t *ireo_la);

/* The scatter step of route 3 (contract/scatter_block.f:2-20): xop(element)
 * += sign * xblk(k, l) over the batch, through the same kind of maps as
 * collect_block_impr_ (atomic adds on the GPU, so elements that several
 * block entries reach sum in an unspecified order).  istr_kc, istr_ka,
 * istr_lc, istr_la are the reference's work arrays and are not touched.
 * Host arrays are staged, device pointers used in place. */
void scatter_block_(
    double *xop, const double *xblk, const int64_t *kc_st0, const int64_t *kc_nd0,
    const int64_t *kc_len, const int64_t *ka_st0, const int64_t *ka_nd0, const int64_t *ka_len,
    const int64_t *k_len, const int64_t *lc_st0, const int64_t *lc_nd0, const int64_t *lc_len,
    const int64_t *la_st0, const int64_t *la_nd0, const int64_t *la_len, const int64_t *l_len,
    const int64_t *map_kl_c, const int64_t *map_kl_a, const int64_t *ncblk_op,
    const int64_t *nablk_op, const int64_t *ldim_opc, const int64_t *ldim_opa,
    const int64_t *ncblk_k, const int64_t *nablk_k, const int64_t *ncblk_l,
    const int64_t *nablk_l, const int64_t *lstr_kc, const int64_t *lstr_ka,
    const int64_t *lstr_lc, const int64_t *lstr_la, const int64_t *nstr_kc,
    const int64_t *nstr_ka, const int64_t *nstr_lc, const int64_t *nstr_la,
    const int64_t *ireo_kc, const int64_t *ireo_ka, const int64_t *ireo_lc,
    const int64_t *ireo_la, int64_t *istr_kc, int64_t *istr_ka, int64_t *istr_lc,
    int64_t *istr_la);

/* The one-call plans kept by contr_blk1blk2_blocked_mm_ and
 * contr_blk1blk2_wmaps_c_ (at most 16384, or about 1 GB of device index
 * data; beyond that the cache is emptied): *entries = their number after
 * the call; clear != 0 frees them first.  No counterpart in the reference. */
int gecco_hip_call_cache(int64_t *entries, int64_t clear);

/* Allocation-free variant of contr_blk1blk2_blocked_mm_ (same arguments,
 * same result).  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  It calls no hipMalloc/hipFree: the
 * call's index tables, task lists, materialised operands and buffer-slot
 * table are placed in a fixed device array of
 * gecco_hip_noalloc_scratch_bytes() bytes (64 MiB) that every device holds
 * for the life of the process; a call that needs more fails with
 * GECCO_HIP_EINVAL (gecco_hip_last_error).  xop1, xop2 and xop1op2 must be
 * device memory.  The array is shared by every call on a device and keeps
 * the last call's contents (nothing in it has to start at zero: each call
 * writes all it reads); a call first waits for the null stream, so calls
 * from one thread follow each other safely, but calls that overlap on one
 * device from several threads are serialised by a lock only within this
 * process.  The kernels run on the null stream. */
void contr_blk1blk2_blocked_mm_noalloc_(
    const double *xfac, double *xop1op2, const double *xop1, const double *xop2,
    const int64_t *tra_op1, const int64_t *tra_op2, const int64_t *tra_op1op2,
    double *xscr, const int64_t *lenscr, const int64_t *lenblock,
    const int64_t *lenblock_cnt,
    const int64_t *ncblk_op1, const int64_t *nablk_op1,
    const int64_t *ncblk_ex1, const int64_t *nablk_ex1,
    const int64_t *ncblk_op2, const int64_t *nablk_op2,
    const int64_t *ncblk_ex2, const int64_t *nablk_ex2,
    const int64_t *ncblk_cnt, const int64_t *nablk_cnt,
    const int64_t *ncblk_op1op2, const int64_t *nablk_op1op2,
    const int64_t *hpvxop1c, const int64_t *hpvxop1a,
    const int64_t *hpvxop2c, const int64_t *hpvxop2a,
    const int64_t *hpvxop1op2c, const int64_t *hpvxop1op2a,
    const int64_t *lstrop1, const int64_t *lstrop2, const int64_t *lstrop1op2,
    const int64_t *lstr_ex1, const int64_t *lstr_ex2, const int64_t *lstr_cnt,
    const int64_t *map_info_12c, const int64_t *map_info_12a,
    const int64_t *map_info_1c, const int64_t *map_info_1a,
    const int64_t *map_info_2c, const int64_t *map_info_2a,
    const int64_t *map_ex1ex2c, const int64_t *map_ex1ex2a,
    const int64_t *map_ex1cntc, const int64_t *map_ex1cnta,
    const int64_t *map_ex2cntc, const int64_t *map_ex2cnta);
/* Size in bytes of the device array used by the allocation-free entry. */
int64_t gecco_hip_noalloc_scratch_bytes(void);

/* ------------------------------------------------------------ plan API */
typedef struct gecco_hip_plan gecco_hip_plan;

/* Description of one kernel call with the arguments of
 * contract/contr_blk1blk2_blocked_mm.f:2-20 (pointers to int64 arrays, held
 * by the caller only for the duration of gecco_hip_plan_add_block). */
typedef struct gecco_hip_block_args {
    double xfac;
    int64_t tra_op1, tra_op2, tra_op1op2;
    int64_t ncblk_op1, nablk_op1, ncblk_ex1, nablk_ex1;
    int64_t ncblk_op2, nablk_op2, ncblk_ex2, nablk_ex2;
    int64_t ncblk_cnt, nablk_cnt, ncblk_op1op2, nablk_op1op2;
    const int64_t *hpvxop1c, *hpvxop1a, *hpvxop2c, *hpvxop2a;
    const int64_t *hpvxop1op2c, *hpvxop1op2a;
    const int64_t *lstrop1, *lstrop2, *lstrop1op2;
    const int64_t *lstr_ex1, *lstr_ex2, *lstr_cnt;
    const int64_t *map_info_12c, *map_info_12a, *map_info_1c, *map_info_1a;
    const int64_t *map_info_2c, *map_info_2a;
    const int64_t *map_ex1ex2c, *map_ex1ex2a, *map_ex1cntc, *map_ex1cnta;
    const int64_t *map_ex2cntc, *map_ex2cnta;
    /* operand placement: buffer slot (index into the array given to
     * gecco_hip_plan_execute) and element offset of the block start */
    int32_t buf_op1, buf_op2, buf_op1op2;
    int64_t off_op1, off_op2, off_op1op2;
} gecco_hip_block_args;

int gecco_hip_plan_create(gecco_hip_plan **plan);
int gecco_hip_plan_destroy(gecco_hip_plan *plan);
/* Record one kernel call.  No device work happens here. */
int gecco_hip_plan_add_block(gecco_hip_plan *plan, const gecco_hip_block_args *args);
/* Resolve every block's signed gathers/scatter into device index tables,
 * choose tiles and the task order, upload.  Must precede execute. */
int gecco_hip_plan_finalize(gecco_hip_plan *plan);
/* Run all blocks as one grouped launch on `stream` (hipStream_t, may be 0).
 * bufs[i] is the device base address of slot i.  Results accumulate. */
int gecco_hip_plan_execute(gecco_hip_plan *plan, double *const *bufs,
                           int64_t nbufs, void *stream);
/* Algorithmic FLOP (sum of 2*n_ex1*n_ex2*n_cnt), blocks, tasks, bytes of
 * operand payload touched once per execute, and device index bytes. */
int gecco_hip_plan_stats(const gecco_hip_plan *plan, double *flops,
                         int64_t *nblocks, int64_t *ntasks,
                         double *payload_bytes, double *index_bytes);

/* Per-launch HIP-event timing of execute (off by default).  plan_timing
 * returns the accumulated ms and launch counts of the five launch classes
 * (ms[5], launches[5]: strided-GEMM contraction launch, gather contraction
 * kernel, spin-flip symmetrisation, operand materialisation, and a fifth
 * class that stays 0 -- the direct short-contraction tasks run in the
 * strided-GEMM launch) since the previous call, and the FLOP per execute of
 * the contraction classes (flops[3]: strided-GEMM tiles, gather kernel,
 * direct tasks); it waits for the recorded events. */
int gecco_hip_plan_set_timing(gecco_hip_plan *plan, int64_t on);
int gecco_hip_plan_timing(gecco_hip_plan *plan, double *ms, int64_t *launches, double *flops);
/* Kernel-call details for diagnostics: [M, N, K, swap, injective, A/B/C
 * table, A/B fast_first, affine flags of A/B/C first/second vectors] and,
 * with ninfo >= 19, [A and B strided, A materialised, B materialised] and,
 * with ninfo >= 20, [direct short-contraction kernel] and, with
 * ninfo >= 23, [result buffer, result block offset, result block length]
 * and, with ninfo >= 24, [next kernel call of its chain, -1 if none; see
 * gecco_hip_plan_set_chain] and, with ninfo >= 25, [16 x 16 wave-split-K
 * tiles: a result with a short side of at most 16] and, with ninfo >= 30,
 * [A buffer slot, A block offset, B buffer slot, B block offset (kernel
 * orientation), bit pattern of the factor xfac as a double].  Index flags are
 * valid before finalize, the kernel-class and chain entries after it. */
int gecco_hip_plan_block_info(const gecco_hip_plan *plan, int64_t iblock, int64_t *info,
                              int64_t ninfo);
/* Resolved signed index data of operand `which` (0 = A, 1 = B, 2 = result C,
 * kernel orientation) of a recorded kernel call, before finalize: entries
 * sign*(1 + element offset), 0 = forbidden.  Separable: first[F] (K for A/B,
 * M for C) and second[S]; table (block_info flags): first[F*S], F fastest,
 * second unused. */
int gecco_hip_plan_block_index(const gecco_hip_plan *plan, int64_t iblock, int64_t which,
                               int32_t *first, int64_t nfirst, int32_t *second, int64_t nsecond);
/* List initialisation run as the first launch of every execute, before
 * the contractions: dst[off_dst + i] = factor * src[off_src + i] for i < len,
 * or 0 when buf_src < 0 -- the zeroing of a result that is not updated
 * (contract/contr_op1op2_wmaps_c.f:644,680) and the copies that start an
 * intermediate from another list (contract/frm_sched2.f, [ADD]/copy).
 * Before finalize only. */
int gecco_hip_plan_add_init(gecco_hip_plan *plan, int32_t buf_dst, int64_t off_dst, int64_t len,
                            int32_t buf_src, int64_t off_src, double factor);
/* Sum of one list into another of the same layout, in the same launch as
 * the initialisations: dst[off_dst + i] += factor * src[off_src + i] for
 * i < len (contract/fs_add_drv.f, command_add_intm of
 * contract/frm_sched2.f:277-286, for lists of equal occupation).  The adds
 * are atomic, so several sums into one list may share a plan; an
 * initialisation of the same destination must be in an earlier plan (the
 * jobs of one launch run concurrently).  Before finalize only. */
int gecco_hip_plan_add_sum(gecco_hip_plan *plan, int32_t buf_dst, int64_t off_dst, int64_t len,
                           int32_t buf_src, int64_t off_src, double factor);
/* Declare that buffer slot `slot` holds the same data at every execute
 * (e.g. integral lists): the materialised gathers of that slot are built on
 * the first execute and again only when the slot's pointer changes.  The
 * caller must not modify the slot's contents between executes.  Before
 * finalize only. */
int gecco_hip_plan_set_constant(gecco_hip_plan *plan, int32_t slot);
/* Operand materialisation: a kernel call whose Op1/Op2 gathers are not
 * strided views and whose result tile counts satisfy min(n_ex1, n_ex2) >=
 * min_dim has those operands written densely (signed, zeros for forbidden
 * strings) into plan-owned scratch once per execute -- the reference's
 * collect step, contract/collect_block_impr.f -- and then runs on the
 * strided-GEMM or direct kernel.  0 disables; default 1.  Before finalize
 * only. */
int gecco_hip_plan_set_materialize(gecco_hip_plan *plan, int64_t min_dim);
/* Chains (default on): kernel calls of the strided-GEMM kernel that add
 * into the same result block through the same result map -- in the
 * reference, the calls of contract/contr_op1op2_wmaps_c.f for the different
 * Ms/IRREP distributions of the contracted strings that land in one result
 * block -- are run as one GEMM over their concatenated contraction ranges,
 * so each result tile is written once.  Before finalize only. */
int gecco_hip_plan_set_chain(gecco_hip_plan *plan, int64_t on);
/* Norm screening (off by default), as contract/contr_op1op2_wmaps_c.f:
 * 1652-1655,1767-1770: at every execute, a kernel call whose Op1 or Op2
 * block (the whole Ms/IRREP distribution block of the operand) has a
 * squared norm below 1e-28 is skipped.  The norms are computed on the
 * device by a launch before the contractions; screened plans run every
 * kernel call as its own GEMM (no chains).  Before finalize only. */
int gecco_hip_plan_set_screen(gecco_hip_plan *plan, int64_t on);
/* Fixed-order summation (off by default; no counterpart in the reference,
 * whose serial loop adds the kernel calls in a fixed order): the results of
 * a plan are bit-reproducible from execute to execute.  Every task of the
 * strided-GEMM launch writes its unsigned tile into plan-owned device memory
 * instead of adding atomically, and one CSR pass per result list then adds
 * each element's contributions -- every kernel call and K chunk that
 * reaches it -- in a fixed (task) order.  Costs one write and one read of
 * every tile element; all kernel calls must run on the strided-GEMM launch
 * (materialised operands, the default).  Before finalize only. */
int gecco_hip_plan_set_deterministic(gecco_hip_plan *plan, int64_t on);
/* Instance groups, for many independent problems (e.g. CC instances of
 * uneven size) recorded into one plan: every kernel call, list
 * initialisation and symmetrisation recorded after set_group(g) belongs to
 * group g (default 0; before finalize only).  set_skip gives device memory
 * flags[nflags] (nflags above the largest group; NULL turns skipping off):
 * every launch of execute then skips the work of group g while flags[g] != 0,
 * read on the device when each task starts -- so a captured graph of the
 * plan skips the instances that have converged since (see
 * gecco_hip_diis_set_flag).  Gathers of constant slots are never skipped;
 * plans with trace steps refuse skip flags.  No counterpart in the
 * reference, which runs one problem per process. */
int gecco_hip_plan_set_group(gecco_hip_plan *plan, int32_t group);
/* Launch shape of the strided-GEMM kernel (before finalize; a negative or
 * zero argument keeps the current value): tile variant (see
 * gecco_hip_plan_block_info's tile shapes: 0 64x64x32, 1 64x64x16 (default
 * for float64 plans; float32 plans default to 9),
 * 2 64x128x16, 3 128x64x16, 4 128x128x16, 5 128x128x8, 6 64x64x16 and
 * 7 64x64x32 with the loads of two K steps in flight; 8, float32 plans
 * only: 64x64x16 GEMM tiles as split-bfloat16 three-term products x = hi + lo,
 * lo.hi + hi.lo + hi.hi on BF16 MFMA with float32 accumulation; 9 64x64x16
 * with each wave 16 rows across the tile, so that waves beyond the block's
 * rows skip their MFMAs), the number of tasks
 * the contraction ranges are split towards (default 2048), and the shortest
 * and longest contraction range of one task (defaults 256, 1024).  The
 * autotuner (gecco-public_amd/autotune.py) picks them per plan by timing.
 * No counterpart in the reference (its batching constants str_block and
 * cnt_block are input keywords, contract/contr_blk1blk2_blocked_mm.f:184-227). */
int gecco_hip_plan_set_tuning(gecco_hip_plan *plan, int64_t variant, int64_t target_tasks,
                              int64_t min_kchunk, int64_t max_kchunk);
/* The launch order the plans give the strided-GEMM launch's tasks (host
 * only, no device work): n tasks in cost order (longest first), patch[t] the
 * patch of task t (up to 16 tiles of one chain and contraction range, which
 * share operand panels), cost[t] its contraction length.  Workgroup w of
 * the launch runs on XCD w mod 8.  Patches go, longest task first, to the XCD
 * with the least work so far; an XCD's patches keep that order with their
 * tasks together; perm[w] is the task at launch position w.  No counterpart
 * in the reference (a serial loop over blocks); exported so that the order
 * can be checked without a GPU.  GECCO_HIP_XCD_ORDER=0 turns the order off
 * for plans created afterwards. */
int gecco_hip_task_order(int64_t n, const int32_t *patch, const int64_t *cost, int64_t *perm);
/* keep_host(1) before finalize: the plan keeps its host index data, so that
 * gecco_hip_plan_retune can finalize it again with other launch settings
 * (arguments as set_tuning) without recording it again; keep_host(0) after
 * finalize releases that data.  retune frees the plan's device data and
 * finalizes anew (the next execute uploads the buffer table again). */
int gecco_hip_plan_keep_host(gecco_hip_plan *plan, int64_t on);
/* Element type of every list the plan touches: 8 = float64 (default), 4 =
 * float32 (before finalize).  A float32 plan reads and writes float lists
 * (the pointers passed to execute then point at float data) and runs its
 * MFMAs as f32 16x16x4 with float32 accumulation; factors stay float64
 * arguments.  No counterpart in the reference, which is float64 only. */
int gecco_hip_plan_set_precision(gecco_hip_plan *plan, int64_t elem_bytes);
int gecco_hip_plan_retune(gecco_hip_plan *plan, int64_t variant, int64_t target_tasks,
                          int64_t min_kchunk, int64_t max_kchunk);
int gecco_hip_plan_set_skip(gecco_hip_plan *plan, const int32_t *flags, int64_t nflags);
/* Live-task lists of a plan with skip flags: compact(plan, 1, stream) copies,
 * on `stream`, every task list of the plan's launches keeping only the tasks
 * of groups whose flag is 0 now (and those of negative groups, e.g. the
 * gathers of constant slots), in their order, then synchronizes `stream`; the
 * launches of execute then have one workgroup per live task instead of one
 * per task, and still read the flags.  Valid while no flag that was set goes
 * back to 0: a caller that clears flags (a restart) calls compact(plan, 0,
 * NULL) first, which returns to the full lists.  The live lists are
 * rewritten in place by every compaction: a graph captured while the plan
 * was compacted must not be replayed after a later compact call (capture
 * again); a graph captured with the full lists stays valid.  Refused for
 * fixed-order plans.  set_skip and retune drop the live
 * lists.  live_tasks reports the task counts the launches use now (strided-
 * GEMM, gather kernel, initialisation, symmetrisation, constant-slot
 * gathers, other gathers; up to ncounts of them).  No counterpart in the
 * reference, which runs one problem per process. */
int gecco_hip_plan_compact(gecco_hip_plan *plan, int64_t on, void *stream);
int gecco_hip_plan_live_tasks(const gecco_hip_plan *plan, int64_t *counts, int64_t ncounts);
/* Scratch bytes held by a finalized plan, number of gather jobs, the HBM
 * bytes they move per execute (one read and one write per element), and the
 * number of tasks whose result tile is reduced in LDS and written back as
 * one contiguous range (result maps that are tables or not injective). */
int gecco_hip_plan_workspace(const gecco_hip_plan *plan, int64_t *scratch_bytes,
                             int64_t *gather_jobs, double *gather_bytes, int64_t *staged_tasks);

/* Device memory helpers for callers without another allocator. */
int gecco_hip_malloc(void **ptr, int64_t bytes);
int gecco_hip_free(void *ptr);
int gecco_hip_memcpy_htod(void *dst, const void *src, int64_t bytes);
int gecco_hip_memcpy_dtoh(void *dst, const void *src, int64_t bytes);
int gecco_hip_synchronize(void);

#ifdef __cplusplus
}
#endif
#endif /* GECCO_HIP_H */
