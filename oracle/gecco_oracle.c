/*
 * Oracle: CPU restatement of GeCCo's block-contraction kernels in plain C.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker or
 * as the CPU baseline.  The product path never links or calls it.
 *
 * Conventions: every integer argument is int64_t (GeCCo builds with 8-byte
 * default INTEGER and LOGICAL, config/x86_64-linux.conf:7,17,28), arrays keep
 * the reference's 1-based meaning (element k of a Fortran array is p[k-1]),
 * LOGICAL is nonzero = true.
 *
 * Restated routines (reference file:line):
 *   orc_idxstnd_batch              contract/idxstnd_batch.f:2-30
 *   orc_set_strmapdim_c2           contract/set_strmapdim_c2.f:2-61
 *   orc_set_op_ldim_c              contract/set_op_ldim_c.f:2-74
 *   orc_collect_block_impr         contract/collect_block_impr.f:2-415
 *   orc_scatter_block              contract/scatter_block.f:2-174
 *   orc_contr_blk1blk2_blocked_mm  contract/contr_blk1blk2_blocked_mm.f:2-605
 *                                  (dgemm('t','n') from an ILP64 BLAS set with
 *                                  orc_set_dgemm, else a plain loop)
 *   orc_contr_blk1blk2_wmaps_c     contract/contr_blk1blk2_wmaps_c.f:2-413
 *
 * Parity status: pinned.  tests/test_oracle_contract.py checks route 3 against
 * route 2, against a blocking-free string-level brute force and against dense
 * antisymmetrised einsum contractions; blocking invariance (str_block and
 * cnt_block sweeps, as test/batch-*.inp do) is checked as well.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef int64_t I;

static const I HPVXSEQ[4] = {4, 2, 3, 1};   /* include/hpvxseq.h:2-4 */

static I ielprd(const I *a, I n) {
    I p = 1;
    for (I i = 0; i < n; ++i) p *= a[i];
    return p;
}

static I isign1(I x) { return x < 0 ? -1 : 1; }
static I iabs1(I x) { return x < 0 ? -x : x; }

void orc_idxstnd_batch(I *len_batch, I *ibstc, I *ibsta, I *ibndc, I *ibnda,
                       I idx_batch, I maxc, I maxa, I maxlen_batch) {
    I st = (idx_batch - 1) * maxlen_batch + 1;
    *ibstc = (st - 1) % maxc + 1;
    *ibsta = (st - 1) / maxc + 1;
    I nd = st + maxlen_batch - 1;
    if (nd > maxc * maxa) nd = maxc * maxa;
    *ibndc = (nd - 1) % maxc + 1;
    *ibnda = (nd - 1) / maxc + 1;
    *len_batch = nd - st + 1;
}

void orc_set_strmapdim_c2(I *nstr1str2, I *nstr1, I *nstr2, I *ireo1, I *ireo2,
                          I nblk12, const I *nstr1r, const I *nstr2r,
                          const I *map12) {
    I idxmap = 0;
    for (I iblk12 = 1; iblk12 <= nblk12; ++iblk12) {
        idxmap++;
        I nblk = map12[idxmap - 1];
        I len1 = 1;
        for (I ib = 1; ib <= nblk; ++ib) {
            idxmap++;
            len1 *= nstr1r[map12[idxmap - 1] - 1];
            ireo1[map12[idxmap - 1] - 1] = iblk12;
        }
        idxmap++;
        I len2 = 1;
        nblk = map12[idxmap - 1];
        for (I ib = 1; ib <= nblk; ++ib) {
            idxmap++;
            len2 *= nstr2r[map12[idxmap - 1] - 1];
            ireo2[map12[idxmap - 1] - 1] = iblk12;
        }
        nstr1str2[iblk12 - 1] = len1 * len2;
        nstr1[iblk12 - 1] = len1;
        nstr2[iblk12 - 1] = len2;
    }
}

void orc_set_op_ldim_c(I *ldimc, I *ldima, const I *hpvxc, const I *hpvxa,
                       const I *nstr, I nc, I na, I transpose) {
    I ld = 1;
    for (int ih = 0; ih < 4; ++ih) {
        I h = HPVXSEQ[ih];
        if (!transpose) {
            for (I i = 0; i < nc; ++i)
                if (hpvxc[i] == h) { ldimc[i] = ld; ld *= nstr[i]; }
            for (I i = 0; i < na; ++i)
                if (hpvxa[i] == h) { ldima[i] = ld; ld *= nstr[nc + i]; }
        } else {
            for (I i = na - 1; i >= 0; --i)
                if (hpvxa[i] == h) { ldima[i] = ld; ld *= nstr[nc + i]; }
            for (I i = nc - 1; i >= 0; --i)
                if (hpvxc[i] == h) { ldimc[i] = ld; ld *= nstr[i]; }
        }
    }
}

/* contract/collect_block_impr.f:2-415 */
void orc_collect_block_impr(
    double *xblk, const double *xop,
    I kc_st0, I kc_nd0, I kc_len, I ka_st0, I ka_nd0, I ka_len, I k_len,
    I kca_trp,
    I lc_st0, I lc_nd0, I lc_len, I la_st0, I la_nd0, I la_len, I l_len,
    const I *map_kl_c, const I *map_kl_a, I ncblk_op, I nablk_op,
    const I *ldim_opc, const I *ldim_opa, I ncblk_k, I nablk_k, I ncblk_l,
    I nablk_l, const I *lstr_kc, const I *lstr_ka, const I *lstr_lc,
    const I *lstr_la, const I *nstr_kc, const I *nstr_ka, const I *nstr_lc,
    const I *nstr_la, const I *ireo_kc, const I *ireo_ka, const I *ireo_lc,
    const I *ireo_la) {
    (void)nstr_lc;
    (void)nstr_la;
    I istr_kc[16], istr_ka[16], istr_lc[16], istr_la[16], istr_kc0[16],
        istr_lc0[16], istr_la0[16];
    for (I i = 0; i < k_len * l_len; ++i) xblk[i] = 0.0;
    for (I i = 0; i < 16; ++i) {
        istr_ka[i] = 1; istr_kc0[i] = 1; istr_la0[i] = 1; istr_lc0[i] = 1;
        istr_kc[i] = 1; istr_lc[i] = 1; istr_la[i] = 1;
    }
    I ireo_kc1 = 1, lstr_kc1 = 1, ireo_ka1 = 1, lstr_ka1 = 1;
    I ireo_lc1 = 1, lstr_lc1 = 1, ireo_la1 = 1, lstr_la1 = 1;
    if (ncblk_k > 0) { ireo_kc1 = ireo_kc[0]; lstr_kc1 = lstr_kc[0]; }
    if (ncblk_l > 0) { ireo_lc1 = ireo_lc[0]; lstr_lc1 = lstr_lc[0]; }
    if (nablk_l > 0) { ireo_la1 = ireo_la[0]; lstr_la1 = lstr_la[0]; }
    if (nablk_k > 0) { ireo_ka1 = ireo_ka[0]; lstr_ka1 = lstr_ka[0]; }

    I istr1 = la_st0 - 1;
    istr_la0[ireo_la1 - 1] = (la_st0 - 1) % lstr_la1;
    for (I icmp = 2; icmp <= nablk_l; ++icmp) {
        istr1 = istr1 / lstr_la[icmp - 2];
        istr_la0[ireo_la[icmp - 1] - 1] = istr1 % lstr_la[icmp - 1] + 1;
    }
    I ka_st, ka_nd;
    if (!kca_trp) { ka_st = ka_st0; ka_nd = ka_nd0; }
    else { ka_st = 1; ka_nd = ka_len; }
    istr1 = ka_st - 1;
    istr_ka[ireo_ka1 - 1] = (ka_st - 1) % lstr_ka1;
    for (I icmp = 2; icmp <= nablk_k; ++icmp) {
        istr1 = istr1 / lstr_ka[icmp - 2];
        istr_ka[ireo_ka[icmp - 1] - 1] = istr1 % lstr_ka[icmp - 1] + 1;
    }
    I kdx0 = 0;
    if (kca_trp) kdx0 = (ka_len - ka_st0 + 1) % ka_len;
    for (I ka = ka_st; ka <= ka_nd; ++ka) {
        I kc_st, kc_nd;
        if (!kca_trp) {
            kc_st = 1; kc_nd = kc_len;
            if (ka == ka_st0) kc_st = kc_st0;
            if (ka == ka_nd0) kc_nd = kc_nd0;
        } else {
            kc_st = kc_st0; kc_nd = kc_nd0;
            if (ka < ka_st0) kc_st = kc_st0 + 1;
            if (ka > ka_nd0) kc_nd = kc_nd0 - 1;
        }
        istr1 = kc_st - 1;
        istr_kc0[ireo_kc1 - 1] = (kc_st - 1) % lstr_kc1;
        for (I icmp = 2; icmp <= ncblk_k; ++icmp) {
            istr1 = istr1 / lstr_kc[icmp - 2];
            istr_kc0[ireo_kc[icmp - 1] - 1] = istr1 % lstr_kc[icmp - 1] + 1;
        }
        if (kca_trp && ka == ka_st0) kdx0 = 0;
        I kdx_st = kdx0;
        if (!kca_trp) kdx0 = kdx0 + kc_nd - kc_st + 1;
        if (kca_trp) kdx0 = kdx0 + 1;
        istr_ka[ireo_ka1 - 1] += 1;
        if (istr_ka[ireo_ka1 - 1] > lstr_ka1) {
            istr_ka[ireo_ka1 - 1] = 1;
            for (I icmp = 2; icmp <= nablk_k; ++icmp) {
                istr_ka[ireo_ka[icmp - 1] - 1] += 1;
                if (istr_ka[ireo_ka[icmp - 1] - 1] <= lstr_ka[icmp - 1]) break;
                istr_ka[ireo_ka[icmp - 1] - 1] = 1;
            }
        }
        for (I i = 0; i < nablk_op; ++i) istr_la[i] = istr_la0[i];
        I ldx0 = 0;
        for (I la = la_st0; la <= la_nd0; ++la) {
            I lc_st = 1, lc_nd = lc_len;
            if (la == la_st0) lc_st = lc_st0;
            if (la == la_nd0) lc_nd = lc_nd0;
            I ldx = ldx0;
            ldx0 = ldx0 + lc_nd - lc_st + 1;
            istr1 = lc_st - 1;
            istr_lc0[ireo_lc1 - 1] = (lc_st - 1) % lstr_lc1;
            for (I icmp = 2; icmp <= ncblk_l; ++icmp) {
                istr1 = istr1 / lstr_lc[icmp - 2];
                istr_lc0[ireo_lc[icmp - 1] - 1] = istr1 % lstr_lc[icmp - 1] + 1;
            }
            istr_la[ireo_la1 - 1] += 1;
            if (istr_la[ireo_la1 - 1] > lstr_la1) {
                istr_la[ireo_la1 - 1] = 1;
                for (I icmp = 2; icmp <= nablk_l; ++icmp) {
                    istr_la[ireo_la[icmp - 1] - 1] += 1;
                    if (istr_la[ireo_la[icmp - 1] - 1] <= lstr_la[icmp - 1]) break;
                    istr_la[ireo_la[icmp - 1] - 1] = 1;
                }
            }
            I ioff = 0, isgna = 1, idx0op = 1;
            int skip = 0;
            for (I icmp = 1; icmp <= nablk_op; ++icmp) {
                I idx1 = istr_la[icmp - 1];
                I idx2 = istr_ka[icmp - 1];
                I idx = (idx1 - 1) * nstr_ka[icmp - 1] + idx2;
                I m = map_kl_a[ioff + idx - 1];
                if (m == 0) { skip = 1; break; }
                isgna *= isign1(m);
                idx0op += (iabs1(m) - 1) * ldim_opa[icmp - 1];
                ioff += nstr_ka[icmp - 1] * nstr_la[icmp - 1];
            }
            if (skip) continue;
            for (I i = 0; i < ncblk_op; ++i) istr_lc[i] = istr_lc0[i];
            for (I lc = lc_st; lc <= lc_nd; ++lc) {
                ldx++;
                istr_lc[ireo_lc1 - 1] += 1;
                if (istr_lc[ireo_lc1 - 1] > lstr_lc1) {
                    istr_lc[ireo_lc1 - 1] = 1;
                    for (I icmp = 2; icmp <= ncblk_l; ++icmp) {
                        istr_lc[ireo_lc[icmp - 1] - 1] += 1;
                        if (istr_lc[ireo_lc[icmp - 1] - 1] <= lstr_lc[icmp - 1]) break;
                        istr_lc[ireo_lc[icmp - 1] - 1] = 1;
                    }
                }
                I kldx = (ldx - 1) * k_len + kdx_st;
                I inc = 1;
                if (kca_trp) kldx = (ldx - 1) * k_len + kdx_st + 1 - ka_len;
                if (kca_trp) inc = ka_len;
                if (ncblk_op == 0) {
                    double xel = (double)isgna * xop[idx0op - 1];
                    for (I kc = kc_st; kc <= kc_nd; ++kc) {
                        kldx += inc;
                        xblk[kldx - 1] = xel;
                    }
                } else if (ncblk_op == 1) {
                    for (I kc = kc_st; kc <= kc_nd; ++kc) {
                        kldx += inc;
                        I idx = (istr_lc[0] - 1) * nstr_kc[0] + kc;
                        I m = map_kl_c[idx - 1];
                        if (m == 0) continue;
                        I isgn = isgna * isign1(m);
                        I idxop = idx0op + (iabs1(m) - 1) * ldim_opc[0];
                        xblk[kldx - 1] = (double)isgn * xop[idxop - 1];
                    }
                } else {
                    for (I i = 0; i < ncblk_op; ++i) istr_kc[i] = istr_kc0[i];
                    for (I kc = kc_st; kc <= kc_nd; ++kc) {
                        kldx += inc;
                        istr_kc[ireo_kc1 - 1] += 1;
                        if (istr_kc[ireo_kc1 - 1] > lstr_kc1) {
                            istr_kc[ireo_kc1 - 1] = 1;
                            for (I icmp = 2; icmp <= ncblk_k; ++icmp) {
                                istr_kc[ireo_kc[icmp - 1] - 1] += 1;
                                if (istr_kc[ireo_kc[icmp - 1] - 1] <= lstr_kc[icmp - 1]) break;
                                istr_kc[ireo_kc[icmp - 1] - 1] = 1;
                            }
                        }
                        I ioff2 = 0, isgn = isgna, idxop = idx0op;
                        int skip2 = 0;
                        for (I icmp = 1; icmp <= ncblk_op; ++icmp) {
                            I idx1 = istr_lc[icmp - 1];
                            I idx2 = istr_kc[icmp - 1];
                            I idx = (idx1 - 1) * nstr_kc[icmp - 1] + idx2;
                            I m = map_kl_c[ioff2 + idx - 1];
                            if (m == 0) { skip2 = 1; break; }
                            isgn *= isign1(m);
                            idxop += (iabs1(m) - 1) * ldim_opc[icmp - 1];
                            ioff2 += nstr_kc[icmp - 1] * nstr_lc[icmp - 1];
                        }
                        if (skip2) continue;
                        xblk[kldx - 1] = (double)isgn * xop[idxop - 1];
                    }
                }
            }
        }
    }
}

/* contract/scatter_block.f:2-174 */
void orc_scatter_block(
    double *xop, const double *xblk,
    I kc_st0, I kc_nd0, I kc_len, I ka_st0, I ka_nd0, I ka_len, I k_len,
    I lc_st0, I lc_nd0, I lc_len, I la_st0, I la_nd0, I la_len, I l_len,
    const I *map_kl_c, const I *map_kl_a, I ncblk_op, I nablk_op,
    const I *ldim_opc, const I *ldim_opa, I ncblk_k, I nablk_k, I ncblk_l,
    I nablk_l, const I *lstr_kc, const I *lstr_ka, const I *lstr_lc,
    const I *lstr_la, const I *nstr_kc, const I *nstr_ka, const I *nstr_lc,
    const I *nstr_la, const I *ireo_kc, const I *ireo_ka, const I *ireo_lc,
    const I *ireo_la, I *istr_kc, I *istr_ka, I *istr_lc, I *istr_la) {
    (void)ka_len; (void)l_len; (void)la_len; (void)nstr_lc;
    I kdx0 = 0;
    for (I ka = ka_st0; ka <= ka_nd0; ++ka) {
        I kc_st = 1, kc_nd = kc_len;
        if (ka == ka_st0) kc_st = kc_st0;
        if (ka == ka_nd0) kc_nd = kc_nd0;
        I kdx_st = kdx0;
        kdx0 = kdx0 + kc_nd - kc_st + 1;
        for (I i = 0; i < nablk_op; ++i) istr_ka[i] = 1;
        I istr1 = ka - 1;
        for (I icmp = 1; icmp <= nablk_k; ++icmp) {
            istr_ka[ireo_ka[icmp - 1] - 1] = istr1 % lstr_ka[icmp - 1] + 1;
            istr1 /= lstr_ka[icmp - 1];
        }
        I ldx0 = 0;
        for (I la = la_st0; la <= la_nd0; ++la) {
            I lc_st = 1, lc_nd = lc_len;
            if (la == la_st0) lc_st = lc_st0;
            if (la == la_nd0) lc_nd = lc_nd0;
            I ldx = ldx0;
            ldx0 = ldx0 + lc_nd - lc_st + 1;
            for (I i = 0; i < nablk_op; ++i) istr_la[i] = 1;
            istr1 = la - 1;
            for (I icmp = 1; icmp <= nablk_l; ++icmp) {
                istr_la[ireo_la[icmp - 1] - 1] = istr1 % lstr_la[icmp - 1] + 1;
                istr1 /= lstr_la[icmp - 1];
            }
            I ioff = 0, isgna = 1, idx0op = 1;
            int skip = 0;
            for (I icmp = 1; icmp <= nablk_op; ++icmp) {
                I idx1 = istr_ka[icmp - 1];
                I idx2 = istr_la[icmp - 1];
                I idx = (idx1 - 1) * nstr_la[icmp - 1] + idx2;
                I m = map_kl_a[ioff + idx - 1];
                if (m == 0) { skip = 1; break; }
                isgna *= isign1(m);
                idx0op += (iabs1(m) - 1) * ldim_opa[icmp - 1];
                ioff += nstr_ka[icmp - 1] * nstr_la[icmp - 1];
            }
            if (skip) continue;
            for (I lc = lc_st; lc <= lc_nd; ++lc) {
                ldx++;
                for (I i = 0; i < ncblk_op; ++i) istr_lc[i] = 1;
                istr1 = lc - 1;
                for (I icmp = 1; icmp <= ncblk_l; ++icmp) {
                    istr_lc[ireo_lc[icmp - 1] - 1] = istr1 % lstr_lc[icmp - 1] + 1;
                    istr1 /= lstr_lc[icmp - 1];
                }
                I kldx = (ldx - 1) * k_len + kdx_st;
                for (I kc = kc_st; kc <= kc_nd; ++kc) {
                    kldx++;
                    for (I i = 0; i < ncblk_op; ++i) istr_kc[i] = 1;
                    istr1 = kc - 1;
                    for (I icmp = 1; icmp <= ncblk_k; ++icmp) {
                        istr_kc[ireo_kc[icmp - 1] - 1] = istr1 % lstr_kc[icmp - 1] + 1;
                        istr1 /= lstr_kc[icmp - 1];
                    }
                    I ioff2 = 0, isgn = isgna, idxop = idx0op;
                    int skip2 = 0;
                    for (I icmp = 1; icmp <= ncblk_op; ++icmp) {
                        I idx1 = istr_lc[icmp - 1];
                        I idx2 = istr_kc[icmp - 1];
                        I idx = (idx1 - 1) * nstr_kc[icmp - 1] + idx2;
                        I m = map_kl_c[ioff2 + idx - 1];
                        if (m == 0) { skip2 = 1; break; }
                        isgn *= isign1(m);
                        idxop += (iabs1(m) - 1) * ldim_opc[icmp - 1];
                        ioff2 += nstr_kc[icmp - 1] * nstr_lc[icmp - 1];
                    }
                    if (skip2) continue;
                    xop[idxop - 1] += (double)isgn * xblk[kldx - 1];
                }
            }
        }
    }
}

/* A BLAS dgemm with 8-byte integers (the reference links an ILP64 BLAS,
 * config/x86_64-linux.conf; contr_blk1blk2_blocked_mm.f:465,523 call
 * dgemm('t','n',...)), given by the caller as a function pointer; NULL keeps
 * the plain loop below. */
typedef void (*dgemm64_fn)(const char *, const char *, const I *, const I *, const I *,
                           const double *, const double *, const I *, const double *,
                           const I *, const double *, double *, const I *, size_t, size_t);
static dgemm64_fn blas_dgemm = 0;

void orc_set_dgemm(void *fn) { blas_dgemm = (dgemm64_fn)fn; }

/* Test hook: route 3's collect and scatter steps through external functions
 * with the reference's Fortran argument lists (every argument by reference;
 * the product's collect_block_impr_ / scatter_block_ entries), so that a
 * whole route-3 contraction checks them call by call; NULL restores the
 * restatements above. */
typedef void (*collect_fn)(double *, const double *, const I *, const I *, const I *, const I *,
                           const I *, const I *, const I *, const I *, const I *, const I *,
                           const I *, const I *, const I *, const I *, const I *, const I *,
                           const I *, const I *, const I *, const I *, const I *, const I *,
                           const I *, const I *, const I *, const I *, const I *, const I *,
                           const I *, const I *, const I *, const I *, const I *, const I *,
                           const I *, const I *, const I *);
typedef void (*scatter_fn)(double *, const double *, const I *, const I *, const I *, const I *,
                           const I *, const I *, const I *, const I *, const I *, const I *,
                           const I *, const I *, const I *, const I *, const I *, const I *,
                           const I *, const I *, const I *, const I *, const I *, const I *,
                           const I *, const I *, const I *, const I *, const I *, const I *,
                           const I *, const I *, const I *, const I *, const I *, const I *,
                           const I *, const I *, I *, I *, I *, I *);
static collect_fn ext_collect = 0;
static scatter_fn ext_scatter = 0;

void orc_set_collect_scatter(void *collect, void *scatter) {
    ext_collect = (collect_fn)collect;
    ext_scatter = (scatter_fn)scatter;
}

static void collect(
    double *xblk, const double *xop, I kc_st0, I kc_nd0, I kc_len, I ka_st0, I ka_nd0,
    I ka_len, I k_len, I kca_trp, I lc_st0, I lc_nd0, I lc_len, I la_st0, I la_nd0, I la_len,
    I l_len, const I *map_kl_c, const I *map_kl_a, I ncblk_op, I nablk_op, const I *ldim_opc,
    const I *ldim_opa, I ncblk_k, I nablk_k, I ncblk_l, I nablk_l, const I *lstr_kc,
    const I *lstr_ka, const I *lstr_lc, const I *lstr_la, const I *nstr_kc, const I *nstr_ka,
    const I *nstr_lc, const I *nstr_la, const I *ireo_kc, const I *ireo_ka, const I *ireo_lc,
    const I *ireo_la) {
    if (!ext_collect) {
        orc_collect_block_impr(xblk, xop, kc_st0, kc_nd0, kc_len, ka_st0, ka_nd0, ka_len, k_len,
                               kca_trp, lc_st0, lc_nd0, lc_len, la_st0, la_nd0, la_len, l_len,
                               map_kl_c, map_kl_a, ncblk_op, nablk_op, ldim_opc, ldim_opa, ncblk_k,
                               nablk_k, ncblk_l, nablk_l, lstr_kc, lstr_ka, lstr_lc, lstr_la,
                               nstr_kc, nstr_ka, nstr_lc, nstr_la, ireo_kc, ireo_ka, ireo_lc,
                               ireo_la);
        return;
    }
    ext_collect(xblk, xop, &kc_st0, &kc_nd0, &kc_len, &ka_st0, &ka_nd0, &ka_len, &k_len, &kca_trp,
                &lc_st0, &lc_nd0, &lc_len, &la_st0, &la_nd0, &la_len, &l_len, map_kl_c, map_kl_a,
                &ncblk_op, &nablk_op, ldim_opc, ldim_opa, &ncblk_k, &nablk_k, &ncblk_l, &nablk_l,
                lstr_kc, lstr_ka, lstr_lc, lstr_la, nstr_kc, nstr_ka, nstr_lc, nstr_la, ireo_kc,
                ireo_ka, ireo_lc, ireo_la);
}

static void scatter(
    double *xop, const double *xblk, I kc_st0, I kc_nd0, I kc_len, I ka_st0, I ka_nd0, I ka_len,
    I k_len, I lc_st0, I lc_nd0, I lc_len, I la_st0, I la_nd0, I la_len, I l_len,
    const I *map_kl_c, const I *map_kl_a, I ncblk_op, I nablk_op, const I *ldim_opc,
    const I *ldim_opa, I ncblk_k, I nablk_k, I ncblk_l, I nablk_l, const I *lstr_kc,
    const I *lstr_ka, const I *lstr_lc, const I *lstr_la, const I *nstr_kc, const I *nstr_ka,
    const I *nstr_lc, const I *nstr_la, const I *ireo_kc, const I *ireo_ka, const I *ireo_lc,
    const I *ireo_la, I *istr_kc, I *istr_ka, I *istr_lc, I *istr_la) {
    if (!ext_scatter) {
        orc_scatter_block(xop, xblk, kc_st0, kc_nd0, kc_len, ka_st0, ka_nd0, ka_len, k_len, lc_st0,
                          lc_nd0, lc_len, la_st0, la_nd0, la_len, l_len, map_kl_c, map_kl_a,
                          ncblk_op, nablk_op, ldim_opc, ldim_opa, ncblk_k, nablk_k, ncblk_l,
                          nablk_l, lstr_kc, lstr_ka, lstr_lc, lstr_la, nstr_kc, nstr_ka, nstr_lc,
                          nstr_la, ireo_kc, ireo_ka, ireo_lc, ireo_la, istr_kc, istr_ka, istr_lc,
                          istr_la);
        return;
    }
    ext_scatter(xop, xblk, &kc_st0, &kc_nd0, &kc_len, &ka_st0, &ka_nd0, &ka_len, &k_len, &lc_st0,
                &lc_nd0, &lc_len, &la_st0, &la_nd0, &la_len, &l_len, map_kl_c, map_kl_a, &ncblk_op,
                &nablk_op, ldim_opc, ldim_opa, &ncblk_k, &nablk_k, &ncblk_l, &nablk_l, lstr_kc,
                lstr_ka, lstr_lc, lstr_la, nstr_kc, nstr_ka, nstr_lc, nstr_la, ireo_kc, ireo_ka,
                ireo_lc, ireo_la, istr_kc, istr_ka, istr_lc, istr_la);
}

/* dgemm('t','n',m,n,k,alpha,A,lda,B,ldb,0,C,ldc): the BLAS routine when one
 * is set, else a plain loop */
static void ref_dgemm_tn(I m, I n, I k, double alpha, const double *A, I lda,
                         const double *B, I ldb, double *C, I ldc) {
    if (blas_dgemm) {
        const double beta = 0.0;
        blas_dgemm("T", "N", &m, &n, &k, &alpha, A, &lda, B, &ldb, &beta, C, &ldc, 1, 1);
        return;
    }
    for (I j = 0; j < n; ++j)
        for (I i = 0; i < m; ++i) {
            double s = 0.0;
            for (I l = 0; l < k; ++l) s += A[l + i * lda] * B[l + j * ldb];
            C[i + j * ldc] = alpha * s;
        }
}

#define MXB 16

/* contract/contr_blk1blk2_blocked_mm.f:2-605.  Returns 0, or 1 where the
 * reference calls quit() (scratch too small, contr_blk1blk2_blocked_mm.f:218-233). */
int orc_contr_blk1blk2_blocked_mm(
    double xfac, double *xop1op2, const double *xop1, const double *xop2,
    I tra_op1, I tra_op2, I tra_op1op2, double *xscr, I lenscr, I lenblock,
    I lenblock_cnt, I ncblk_op1, I nablk_op1, I ncblk_ex1, I nablk_ex1,
    I ncblk_op2, I nablk_op2, I ncblk_ex2, I nablk_ex2, I ncblk_cnt,
    I nablk_cnt, I ncblk_op1op2, I nablk_op1op2, const I *hpvxop1c,
    const I *hpvxop1a, const I *hpvxop2c, const I *hpvxop2a,
    const I *hpvxop1op2c, const I *hpvxop1op2a, const I *lstrop1,
    const I *lstrop2, const I *lstrop1op2, const I *lstr_ex1,
    const I *lstr_ex2, const I *lstr_cnt, const I *map_info_12c,
    const I *map_info_12a, const I *map_info_1c, const I *map_info_1a,
    const I *map_info_2c, const I *map_info_2a, const I *map_ex1ex2c,
    const I *map_ex1ex2a, const I *map_ex1cntc, const I *map_ex1cnta,
    const I *map_ex2cntc, const I *map_ex2cnta) {
    I ireo_ex1a1[MXB], ireo_ex1c1[MXB], ireo_ex1a12[MXB], ireo_ex1c12[MXB];
    I ireo_ex2a2[MXB], ireo_ex2c2[MXB], ireo_ex2a12[MXB], ireo_ex2c12[MXB];
    I ireo_cnta1[MXB], ireo_cntc1[MXB], ireo_cnta2[MXB], ireo_cntc2[MXB];
    I nstr_ex1a1[MXB], nstr_ex1c1[MXB], nstr_ex1a12[MXB], nstr_ex1c12[MXB];
    I nstr_ex2a2[MXB], nstr_ex2c2[MXB], nstr_ex2a12[MXB], nstr_ex2c12[MXB];
    I nstr_cnta1[MXB], nstr_cntc1[MXB], nstr_cnta2[MXB], nstr_cntc2[MXB];
    I nstr_ex1ex2a[MXB], nstr_ex1ex2c[MXB], nstr_ex1cnta[MXB], nstr_ex1cntc[MXB];
    I nstr_ex2cnta[MXB], nstr_ex2cntc[MXB];
    I ldim_op1a[MXB], ldim_op1c[MXB], ldim_op2a[MXB], ldim_op2c[MXB];
    I ldim_op1op2a[MXB], ldim_op1op2c[MXB];
    I istrscr1_c[MXB], istrscr1_a[MXB], istrscr2_c[MXB], istrscr2_a[MXB];
    for (int i = 0; i < MXB; ++i) {
        ireo_ex1a1[i] = ireo_ex1c1[i] = ireo_ex1a12[i] = ireo_ex1c12[i] = 1;
        ireo_ex2a2[i] = ireo_ex2c2[i] = ireo_ex2a12[i] = ireo_ex2c12[i] = 1;
        ireo_cnta1[i] = ireo_cntc1[i] = ireo_cnta2[i] = ireo_cntc2[i] = 1;
    }

    I nstr_ex1c_tot = ielprd(lstr_ex1, ncblk_ex1);
    I nstr_ex1a_tot = ielprd(lstr_ex1 + ncblk_ex1, nablk_ex1);
    I nstr_ex2c_tot = ielprd(lstr_ex2, ncblk_ex2);
    I nstr_ex2a_tot = ielprd(lstr_ex2 + ncblk_ex2, nablk_ex2);
    I nstr_cntc_tot = ielprd(lstr_cnt, ncblk_cnt);
    I nstr_cnta_tot = ielprd(lstr_cnt + ncblk_cnt, nablk_cnt);

    orc_set_strmapdim_c2(nstr_ex1ex2c, nstr_ex1c12, nstr_ex2c12, ireo_ex1c12,
                         ireo_ex2c12, ncblk_op1op2, lstr_ex1, lstr_ex2, map_info_12c);
    orc_set_strmapdim_c2(nstr_ex1ex2a, nstr_ex2a12, nstr_ex1a12, ireo_ex2a12,
                         ireo_ex1a12, nablk_op1op2, lstr_ex2 + ncblk_ex2,
                         lstr_ex1 + ncblk_ex1, map_info_12a);
    orc_set_strmapdim_c2(nstr_ex1cntc, nstr_cntc1, nstr_ex1c1, ireo_cntc1,
                         ireo_ex1c1, ncblk_op1, lstr_cnt, lstr_ex1, map_info_1c);
    orc_set_strmapdim_c2(nstr_ex1cnta, nstr_cnta1, nstr_ex1a1, ireo_cnta1,
                         ireo_ex1a1, nablk_op1, lstr_cnt + ncblk_cnt,
                         lstr_ex1 + ncblk_ex1, map_info_1a);
    orc_set_strmapdim_c2(nstr_ex2cntc, nstr_cnta2, nstr_ex2c2, ireo_cnta2,
                         ireo_ex2c2, ncblk_op2, lstr_cnt + ncblk_cnt, lstr_ex2,
                         map_info_2c);
    orc_set_strmapdim_c2(nstr_ex2cnta, nstr_cntc2, nstr_ex2a2, ireo_cntc2,
                         ireo_ex2a2, nablk_op2, lstr_cnt, lstr_ex2 + ncblk_ex2,
                         map_info_2a);
    orc_set_op_ldim_c(ldim_op1c, ldim_op1a, hpvxop1c, hpvxop1a, lstrop1,
                      ncblk_op1, nablk_op1, tra_op1);
    orc_set_op_ldim_c(ldim_op2c, ldim_op2a, hpvxop2c, hpvxop2a, lstrop2,
                      ncblk_op2, nablk_op2, tra_op2);
    orc_set_op_ldim_c(ldim_op1op2c, ldim_op1op2a, hpvxop1op2c, hpvxop1op2a,
                      lstrop1op2, ncblk_op1op2, nablk_op1op2, tra_op1op2);

    I lscr_op1op2 = nstr_ex1c_tot * nstr_ex1a_tot * nstr_ex2c_tot * nstr_ex2a_tot;
    if (lscr_op1op2 > lenblock * lenblock) lscr_op1op2 = lenblock * lenblock;
    int op1shorter = nstr_ex1c_tot * nstr_ex1a_tot <= nstr_ex2c_tot * nstr_ex2a_tot;
    I maxlen_exl_batch, maxlen_exs_batch, maxlen_cnt_batch;
    I nstr_exsc_tot, nstr_exsa_tot, nstr_exlc_tot, nstr_exla_tot;
    I ncnt_tot = nstr_cntc_tot * nstr_cnta_tot;
    if (op1shorter) {
        maxlen_exl_batch = nstr_ex2c_tot * nstr_ex2a_tot;
        if (maxlen_exl_batch > lenblock) maxlen_exl_batch = lenblock;
        maxlen_exs_batch = nstr_ex1c_tot * nstr_ex1a_tot;
        if (maxlen_exs_batch > lenblock) maxlen_exs_batch = lenblock;
        maxlen_cnt_batch = (lenscr - lscr_op1op2 - 1) /
                               (nstr_ex1c_tot * nstr_ex1a_tot + maxlen_exl_batch) + 1;
        nstr_exsc_tot = nstr_ex1c_tot; nstr_exsa_tot = nstr_ex1a_tot;
        nstr_exlc_tot = nstr_ex2c_tot; nstr_exla_tot = nstr_ex2a_tot;
    } else {
        maxlen_exl_batch = nstr_ex1c_tot * nstr_ex1a_tot;
        if (maxlen_exl_batch > lenblock) maxlen_exl_batch = lenblock;
        maxlen_exs_batch = nstr_ex2c_tot * nstr_ex2a_tot;
        if (maxlen_exs_batch > lenblock) maxlen_exs_batch = lenblock;
        maxlen_cnt_batch = (lenscr - lscr_op1op2 - 1) /
                               (nstr_ex2c_tot * nstr_ex2a_tot + maxlen_exl_batch) + 1;
        nstr_exsc_tot = nstr_ex2c_tot; nstr_exsa_tot = nstr_ex2a_tot;
        nstr_exlc_tot = nstr_ex1c_tot; nstr_exla_tot = nstr_ex1a_tot;
    }
    if (maxlen_cnt_batch > ncnt_tot) maxlen_cnt_batch = ncnt_tot;
    if (lenblock_cnt > 0 && maxlen_cnt_batch > lenblock_cnt)
        maxlen_cnt_batch = lenblock_cnt;
    if (maxlen_cnt_batch < 1) return 1;
    I idxopsscr = 1;
    I idxoplscr = idxopsscr + nstr_exsc_tot * nstr_exsa_tot * maxlen_cnt_batch;
    I idxop1op2scr = idxoplscr + maxlen_cnt_batch * maxlen_exl_batch;
    I idxend = idxop1op2scr + maxlen_exl_batch * maxlen_exs_batch;
    if (idxend > lenscr) return 1;
    I n_cnt_batch = (ncnt_tot - 1) / maxlen_cnt_batch + 1;
    I n_exl_batch = (nstr_exlc_tot * nstr_exla_tot - 1) / maxlen_exl_batch + 1;
    I n_exs_batch = (nstr_exsc_tot * nstr_exsa_tot - 1) / maxlen_exs_batch + 1;

    for (I cnt_batch = 1; cnt_batch <= n_cnt_batch; ++cnt_batch) {
        I ncnt, cc_bst, ca_bst, cc_bnd, ca_bnd;
        orc_idxstnd_batch(&ncnt, &cc_bst, &ca_bst, &cc_bnd, &ca_bnd, cnt_batch,
                          nstr_cntc_tot, nstr_cnta_tot, maxlen_cnt_batch);
        if (op1shorter) {
            collect(
                xscr + idxopsscr - 1, xop1, cc_bst, cc_bnd, nstr_cntc_tot, ca_bst,
                ca_bnd, nstr_cnta_tot, ncnt, 0, 1, nstr_ex1c_tot, nstr_ex1c_tot, 1,
                nstr_ex1a_tot, nstr_ex1a_tot, nstr_ex1c_tot * nstr_ex1a_tot,
                map_ex1cntc, map_ex1cnta, ncblk_op1, nablk_op1, ldim_op1c,
                ldim_op1a, ncblk_cnt, nablk_cnt, ncblk_ex1, nablk_ex1, lstr_cnt,
                lstr_cnt + ncblk_cnt, lstr_ex1, lstr_ex1 + ncblk_ex1, nstr_cntc1,
                nstr_cnta1, nstr_ex1c1, nstr_ex1a1, ireo_cntc1, ireo_cnta1,
                ireo_ex1c1, ireo_ex1a1);
        } else {
            collect(
                xscr + idxopsscr - 1, xop2, ca_bst, ca_bnd, nstr_cnta_tot, cc_bst,
                cc_bnd, nstr_cntc_tot, ncnt, 1, 1, nstr_ex2c_tot, nstr_ex2c_tot, 1,
                nstr_ex2a_tot, nstr_ex2a_tot, nstr_ex2c_tot * nstr_ex2a_tot,
                map_ex2cntc, map_ex2cnta, ncblk_op2, nablk_op2, ldim_op2c,
                ldim_op2a, nablk_cnt, ncblk_cnt, ncblk_ex2, nablk_ex2,
                lstr_cnt + ncblk_cnt, lstr_cnt, lstr_ex2, lstr_ex2 + ncblk_ex2,
                nstr_cnta2, nstr_cntc2, nstr_ex2c2, nstr_ex2a2, ireo_cnta2,
                ireo_cntc2, ireo_ex2c2, ireo_ex2a2);
        }
        for (I exl_batch = 1; exl_batch <= n_exl_batch; ++exl_batch) {
            I nexl, lc_bst, la_bst, lc_bnd, la_bnd;
            orc_idxstnd_batch(&nexl, &lc_bst, &la_bst, &lc_bnd, &la_bnd, exl_batch,
                              nstr_exlc_tot, nstr_exla_tot, maxlen_exl_batch);
            if (op1shorter) {
                collect(
                    xscr + idxoplscr - 1, xop2, ca_bst, ca_bnd, nstr_cnta_tot,
                    cc_bst, cc_bnd, nstr_cntc_tot, ncnt, 1, lc_bst, lc_bnd,
                    nstr_exlc_tot, la_bst, la_bnd, nstr_exla_tot, nexl, map_ex2cntc,
                    map_ex2cnta, ncblk_op2, nablk_op2, ldim_op2c, ldim_op2a,
                    nablk_cnt, ncblk_cnt, ncblk_ex2, nablk_ex2, lstr_cnt + ncblk_cnt,
                    lstr_cnt, lstr_ex2, lstr_ex2 + ncblk_ex2, nstr_cnta2, nstr_cntc2,
                    nstr_ex2c2, nstr_ex2a2, ireo_cnta2, ireo_cntc2, ireo_ex2c2,
                    ireo_ex2a2);
            } else {
                collect(
                    xscr + idxoplscr - 1, xop1, cc_bst, cc_bnd, nstr_cntc_tot,
                    ca_bst, ca_bnd, nstr_cnta_tot, ncnt, 0, lc_bst, lc_bnd,
                    nstr_exlc_tot, la_bst, la_bnd, nstr_exla_tot, nexl, map_ex1cntc,
                    map_ex1cnta, ncblk_op1, nablk_op1, ldim_op1c, ldim_op1a,
                    ncblk_cnt, nablk_cnt, ncblk_ex1, nablk_ex1, lstr_cnt,
                    lstr_cnt + ncblk_cnt, lstr_ex1, lstr_ex1 + ncblk_ex1, nstr_cntc1,
                    nstr_cnta1, nstr_ex1c1, nstr_ex1a1, ireo_cntc1, ireo_cnta1,
                    ireo_ex1c1, ireo_ex1a1);
            }
            for (I exs_batch = 1; exs_batch <= n_exs_batch; ++exs_batch) {
                I nexs, sc_bst, sa_bst, sc_bnd, sa_bnd;
                orc_idxstnd_batch(&nexs, &sc_bst, &sa_bst, &sc_bnd, &sa_bnd,
                                  exs_batch, nstr_exsc_tot, nstr_exsa_tot,
                                  maxlen_exs_batch);
                I idx = idxopsscr + (exs_batch - 1) * maxlen_exs_batch * ncnt;
                if (op1shorter) {
                    ref_dgemm_tn(nexs, nexl, ncnt, xfac, xscr + idx - 1, ncnt,
                                 xscr + idxoplscr - 1, ncnt, xscr + idxop1op2scr - 1,
                                 nexs);
                    scatter(
                        xop1op2, xscr + idxop1op2scr - 1, sc_bst, sc_bnd,
                        nstr_exsc_tot, sa_bst, sa_bnd, nstr_exsa_tot, nexs, lc_bst,
                        lc_bnd, nstr_exlc_tot, la_bst, la_bnd, nstr_exla_tot, nexl,
                        map_ex1ex2c, map_ex1ex2a, ncblk_op1op2, nablk_op1op2,
                        ldim_op1op2c, ldim_op1op2a, ncblk_ex1, nablk_ex1, ncblk_ex2,
                        nablk_ex2, lstr_ex1, lstr_ex1 + ncblk_ex1, lstr_ex2,
                        lstr_ex2 + ncblk_ex2, nstr_ex1c12, nstr_ex1a12, nstr_ex2c12,
                        nstr_ex2a12, ireo_ex1c12, ireo_ex1a12, ireo_ex2c12,
                        ireo_ex2a12, istrscr1_c, istrscr1_a, istrscr2_c, istrscr2_a);
                } else {
                    ref_dgemm_tn(nexl, nexs, ncnt, xfac, xscr + idxoplscr - 1, ncnt,
                                 xscr + idx - 1, ncnt, xscr + idxop1op2scr - 1, nexl);
                    scatter(
                        xop1op2, xscr + idxop1op2scr - 1, lc_bst, lc_bnd,
                        nstr_exlc_tot, la_bst, la_bnd, nstr_exla_tot, nexl, sc_bst,
                        sc_bnd, nstr_exsc_tot, sa_bst, sa_bnd, nstr_exsa_tot, nexs,
                        map_ex1ex2c, map_ex1ex2a, ncblk_op1op2, nablk_op1op2,
                        ldim_op1op2c, ldim_op1op2a, ncblk_ex1, nablk_ex1, ncblk_ex2,
                        nablk_ex2, lstr_ex1, lstr_ex1 + ncblk_ex1, lstr_ex2,
                        lstr_ex2 + ncblk_ex2, nstr_ex1c12, nstr_ex1a12, nstr_ex2c12,
                        nstr_ex2a12, ireo_ex1c12, ireo_ex1a12, ireo_ex2c12,
                        ireo_ex2a12, istrscr1_c, istrscr1_a, istrscr2_c, istrscr2_a);
                }
            }
        }
    }
    return 0;
}

/* contract/contr_blk1blk2_wmaps_c.f:2-413 (route 2, scalar FMA per element) */
void orc_contr_blk1blk2_wmaps_c(
    double xfac, double *xop1op2, const double *xop1, const double *xop2,
    I tra_op1, I tra_op2, I tra_op1op2, I ncblk_op1, I nablk_op1,
    I ncblk_ex1, I nablk_ex1, I ncblk_op2, I nablk_op2, I ncblk_ex2,
    I nablk_ex2, I ncblk_cnt, I nablk_cnt, I ncblk_op1op2, I nablk_op1op2,
    const I *hpvxop1c, const I *hpvxop1a, const I *hpvxop2c,
    const I *hpvxop2a, const I *hpvxop1op2c, const I *hpvxop1op2a,
    const I *lstrop1, const I *lstrop2, const I *lstrop1op2,
    const I *lstr_ex1, const I *lstr_ex2, const I *lstr_cnt,
    const I *map_info_12c, const I *map_info_12a, const I *map_info_1c,
    const I *map_info_1a, const I *map_info_2c, const I *map_info_2a,
    const I *map_ex1ex2c, const I *map_ex1ex2a, const I *map_ex1cntc,
    const I *map_ex1cnta, const I *map_ex2cntc, const I *map_ex2cnta) {
    I ireo_ex1a1[MXB], ireo_ex1c1[MXB], ireo_ex1a12[MXB], ireo_ex1c12[MXB];
    I ireo_ex2a2[MXB], ireo_ex2c2[MXB], ireo_ex2a12[MXB], ireo_ex2c12[MXB];
    I ireo_cnta1[MXB], ireo_cntc1[MXB], ireo_cnta2[MXB], ireo_cntc2[MXB];
    I nstr_ex1a1[MXB], nstr_ex1c1[MXB], nstr_ex1a12[MXB], nstr_ex1c12[MXB];
    I nstr_ex2a2[MXB], nstr_ex2c2[MXB], nstr_ex2a12[MXB], nstr_ex2c12[MXB];
    I nstr_cnta1[MXB], nstr_cntc1[MXB], nstr_cnta2[MXB], nstr_cntc2[MXB];
    I nstr_ex1ex2a[MXB], nstr_ex1ex2c[MXB], nstr_ex1cnta[MXB], nstr_ex1cntc[MXB];
    I nstr_ex2cnta[MXB], nstr_ex2cntc[MXB];
    I ldim_op1a[MXB], ldim_op1c[MXB], ldim_op2a[MXB], ldim_op2c[MXB];
    I ldim_op1op2a[MXB], ldim_op1op2c[MXB];
    I istr_ex1a12[MXB], istr_ex1a1[MXB], istr_ex2a12[MXB], istr_ex2a2[MXB];
    I istr_ex1c12[MXB], istr_ex1c1[MXB], istr_ex2c12[MXB], istr_ex2c2[MXB];
    I istr_cnta1[MXB], istr_cnta2[MXB], istr_cntc1[MXB], istr_cntc2[MXB];
    I idxop1op2a[MXB], idxop1op2c[MXB];

    I nstr_ex1c_tot = ielprd(lstr_ex1, ncblk_ex1);
    I nstr_ex1a_tot = ielprd(lstr_ex1 + ncblk_ex1, nablk_ex1);
    I nstr_ex2c_tot = ielprd(lstr_ex2, ncblk_ex2);
    I nstr_ex2a_tot = ielprd(lstr_ex2 + ncblk_ex2, nablk_ex2);
    I nstr_cntc_tot = ielprd(lstr_cnt, ncblk_cnt);
    I nstr_cnta_tot = ielprd(lstr_cnt + ncblk_cnt, nablk_cnt);
    orc_set_strmapdim_c2(nstr_ex1ex2c, nstr_ex1c12, nstr_ex2c12, ireo_ex1c12,
                         ireo_ex2c12, ncblk_op1op2, lstr_ex1, lstr_ex2, map_info_12c);
    orc_set_strmapdim_c2(nstr_ex1ex2a, nstr_ex2a12, nstr_ex1a12, ireo_ex2a12,
                         ireo_ex1a12, nablk_op1op2, lstr_ex2 + ncblk_ex2,
                         lstr_ex1 + ncblk_ex1, map_info_12a);
    orc_set_strmapdim_c2(nstr_ex1cntc, nstr_cntc1, nstr_ex1c1, ireo_cntc1,
                         ireo_ex1c1, ncblk_op1, lstr_cnt, lstr_ex1, map_info_1c);
    orc_set_strmapdim_c2(nstr_ex1cnta, nstr_cnta1, nstr_ex1a1, ireo_cnta1,
                         ireo_ex1a1, nablk_op1, lstr_cnt + ncblk_cnt,
                         lstr_ex1 + ncblk_ex1, map_info_1a);
    orc_set_strmapdim_c2(nstr_ex2cntc, nstr_cnta2, nstr_ex2c2, ireo_cnta2,
                         ireo_ex2c2, ncblk_op2, lstr_cnt + ncblk_cnt, lstr_ex2,
                         map_info_2c);
    orc_set_strmapdim_c2(nstr_ex2cnta, nstr_cntc2, nstr_ex2a2, ireo_cntc2,
                         ireo_ex2a2, nablk_op2, lstr_cnt, lstr_ex2 + ncblk_ex2,
                         map_info_2a);
    orc_set_op_ldim_c(ldim_op1c, ldim_op1a, hpvxop1c, hpvxop1a, lstrop1,
                      ncblk_op1, nablk_op1, tra_op1);
    orc_set_op_ldim_c(ldim_op2c, ldim_op2a, hpvxop2c, hpvxop2a, lstrop2,
                      ncblk_op2, nablk_op2, tra_op2);
    orc_set_op_ldim_c(ldim_op1op2c, ldim_op1op2a, hpvxop1op2c, hpvxop1op2a,
                      lstrop1op2, ncblk_op1op2, nablk_op1op2, tra_op1op2);

    for (I s_ex1a = 1; s_ex1a <= nstr_ex1a_tot; ++s_ex1a) {
        for (I i = 0; i < nablk_op1op2; ++i) istr_ex1a12[i] = 1;
        for (I i = 0; i < nablk_op1; ++i) istr_ex1a1[i] = 1;
        I istr1 = s_ex1a - 1;
        for (I icmp = 1; icmp <= nablk_ex1; ++icmp) {
            I v = istr1 % lstr_ex1[ncblk_ex1 + icmp - 1] + 1;
            istr_ex1a12[ireo_ex1a12[icmp - 1] - 1] = v;
            istr_ex1a1[ireo_ex1a1[icmp - 1] - 1] = v;
            istr1 /= lstr_ex1[ncblk_ex1 + icmp - 1];
        }
        for (I s_ex2a = 1; s_ex2a <= nstr_ex2a_tot; ++s_ex2a) {
            for (I i = 0; i < nablk_op1op2; ++i) istr_ex2a12[i] = 1;
            for (I i = 0; i < nablk_op2; ++i) istr_ex2a2[i] = 1;
            istr1 = s_ex2a - 1;
            for (I icmp = 1; icmp <= nablk_ex2; ++icmp) {
                I v = istr1 % lstr_ex2[ncblk_ex2 + icmp - 1] + 1;
                istr_ex2a12[ireo_ex2a12[icmp - 1] - 1] = v;
                istr_ex2a2[ireo_ex2a2[icmp - 1] - 1] = v;
                istr1 /= lstr_ex2[ncblk_ex2 + icmp - 1];
            }
            I ioff = 0, isgnra = 1;
            int skip = 0;
            for (I icmp = 1; icmp <= nablk_op1op2; ++icmp) {
                I idx = (istr_ex1a12[icmp - 1] - 1) * nstr_ex2a12[icmp - 1] +
                        istr_ex2a12[icmp - 1];
                I m = map_ex1ex2a[ioff + idx - 1];
                if (m == 0) { skip = 1; break; }
                isgnra *= isign1(m);
                idxop1op2a[icmp - 1] = iabs1(m) - 1;
                ioff += nstr_ex1ex2a[icmp - 1];
            }
            if (skip) continue;
            for (I s_ex1c = 1; s_ex1c <= nstr_ex1c_tot; ++s_ex1c) {
                for (I i = 0; i < ncblk_op1op2; ++i) istr_ex1c12[i] = 1;
                for (I i = 0; i < ncblk_op1; ++i) istr_ex1c1[i] = 1;
                istr1 = s_ex1c - 1;
                for (I icmp = 1; icmp <= ncblk_ex1; ++icmp) {
                    I v = istr1 % lstr_ex1[icmp - 1] + 1;
                    istr_ex1c12[ireo_ex1c12[icmp - 1] - 1] = v;
                    istr_ex1c1[ireo_ex1c1[icmp - 1] - 1] = v;
                    istr1 /= lstr_ex1[icmp - 1];
                }
                for (I s_ex2c = 1; s_ex2c <= nstr_ex2c_tot; ++s_ex2c) {
                    for (I i = 0; i < ncblk_op1op2; ++i) istr_ex2c12[i] = 1;
                    for (I i = 0; i < ncblk_op2; ++i) istr_ex2c2[i] = 1;
                    istr1 = s_ex2c - 1;
                    for (I icmp = 1; icmp <= ncblk_ex2; ++icmp) {
                        I v = istr1 % lstr_ex2[icmp - 1] + 1;
                        istr_ex2c12[ireo_ex2c12[icmp - 1] - 1] = v;
                        istr_ex2c2[ireo_ex2c2[icmp - 1] - 1] = v;
                        istr1 /= lstr_ex2[icmp - 1];
                    }
                    I ioffc = 0, isgnrc = 1;
                    int skipc = 0;
                    for (I icmp = 1; icmp <= ncblk_op1op2; ++icmp) {
                        I idx = (istr_ex2c12[icmp - 1] - 1) * nstr_ex1c12[icmp - 1] +
                                istr_ex1c12[icmp - 1];
                        I m = map_ex1ex2c[ioffc + idx - 1];
                        if (m == 0) { skipc = 1; break; }
                        isgnrc *= isign1(m);
                        idxop1op2c[icmp - 1] = iabs1(m) - 1;
                        ioffc += nstr_ex1ex2c[icmp - 1];
                    }
                    if (skipc) continue;
                    I isgnr = isgnrc * isgnra;
                    I idxop1op2 = 1;
                    for (I i = 0; i < ncblk_op1op2; ++i)
                        idxop1op2 += idxop1op2c[i] * ldim_op1op2c[i];
                    for (I i = 0; i < nablk_op1op2; ++i)
                        idxop1op2 += idxop1op2a[i] * ldim_op1op2a[i];
                    for (I s_cnta = 1; s_cnta <= nstr_cnta_tot; ++s_cnta) {
                        for (I i = 0; i < nablk_op1; ++i) istr_cnta1[i] = 1;
                        for (I i = 0; i < ncblk_op2; ++i) istr_cnta2[i] = 1;
                        istr1 = s_cnta - 1;
                        for (I icmp = 1; icmp <= nablk_cnt; ++icmp) {
                            I v = istr1 % lstr_cnt[ncblk_cnt + icmp - 1] + 1;
                            istr_cnta1[ireo_cnta1[icmp - 1] - 1] = v;
                            istr_cnta2[ireo_cnta2[icmp - 1] - 1] = v;
                            istr1 /= lstr_cnt[ncblk_cnt + icmp - 1];
                        }
                        I io = 0, s1a = 1, idx0op1 = 1;
                        int sk = 0;
                        for (I icmp = 1; icmp <= nablk_op1; ++icmp) {
                            I idx = (istr_ex1a1[icmp - 1] - 1) * nstr_cnta1[icmp - 1] +
                                    istr_cnta1[icmp - 1];
                            I m = map_ex1cnta[io + idx - 1];
                            if (m == 0) { sk = 1; break; }
                            s1a *= isign1(m);
                            idx0op1 += (iabs1(m) - 1) * ldim_op1a[icmp - 1];
                            io += nstr_ex1cnta[icmp - 1];
                        }
                        if (sk) continue;
                        io = 0;
                        I s2c = 1, idx0op2 = 1;
                        for (I icmp = 1; icmp <= ncblk_op2; ++icmp) {
                            I idx = (istr_ex2c2[icmp - 1] - 1) * nstr_cnta2[icmp - 1] +
                                    istr_cnta2[icmp - 1];
                            I m = map_ex2cntc[io + idx - 1];
                            if (m == 0) { sk = 1; break; }
                            s2c *= isign1(m);
                            idx0op2 += (iabs1(m) - 1) * ldim_op2c[icmp - 1];
                            io += nstr_ex2cntc[icmp - 1];
                        }
                        if (sk) continue;
                        I isgn0 = isgnr * s1a * s2c;
                        for (I s_cntc = 1; s_cntc <= nstr_cntc_tot; ++s_cntc) {
                            for (I i = 0; i < ncblk_op1; ++i) istr_cntc1[i] = 1;
                            for (I i = 0; i < nablk_op2; ++i) istr_cntc2[i] = 1;
                            istr1 = s_cntc - 1;
                            for (I icmp = 1; icmp <= ncblk_cnt; ++icmp) {
                                I v = istr1 % lstr_cnt[icmp - 1] + 1;
                                istr_cntc1[ireo_cntc1[icmp - 1] - 1] = v;
                                istr_cntc2[ireo_cntc2[icmp - 1] - 1] = v;
                                istr1 /= lstr_cnt[icmp - 1];
                            }
                            I io2 = 0, s1c = 1, idxop1 = idx0op1;
                            int sk2 = 0;
                            for (I icmp = 1; icmp <= ncblk_op1; ++icmp) {
                                I idx = (istr_ex1c1[icmp - 1] - 1) * nstr_cntc1[icmp - 1] +
                                        istr_cntc1[icmp - 1];
                                I m = map_ex1cntc[io2 + idx - 1];
                                if (m == 0) { sk2 = 1; break; }
                                s1c *= isign1(m);
                                idxop1 += (iabs1(m) - 1) * ldim_op1c[icmp - 1];
                                io2 += nstr_ex1cntc[icmp - 1];
                            }
                            if (sk2) continue;
                            io2 = 0;
                            I s2a = 1, idxop2 = idx0op2;
                            for (I icmp = 1; icmp <= nablk_op2; ++icmp) {
                                I idx = (istr_ex2a2[icmp - 1] - 1) * nstr_cntc2[icmp - 1] +
                                        istr_cntc2[icmp - 1];
                                I m = map_ex2cnta[io2 + idx - 1];
                                if (m == 0) { sk2 = 1; break; }
                                s2a *= isign1(m);
                                idxop2 += (iabs1(m) - 1) * ldim_op2a[icmp - 1];
                                io2 += nstr_ex2cnta[icmp - 1];
                            }
                            if (sk2) continue;
                            double sgn = xfac * (double)(isgn0 * s1c * s2a);
                            xop1op2[idxop1op2 - 1] += sgn * xop1[idxop1 - 1] * xop2[idxop2 - 1];
                        }
                    }
                }
            }
        }
    }
}
