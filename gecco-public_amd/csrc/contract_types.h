// Device-side work description of the grouped block-contraction launch.
//
// One DevBlock is one kernel call of GeCCo's block enumerator
// (contract/contr_op1op2_wmaps_c.f:1892-1949): a GEMM
//     R(m,n) += xfac * sum_k A(k,m) * B(k,n)
// whose operands are read through signed index tables and whose result is
// written through a signed index table (the string maps of
// contract/collect_block_impr.f and contract/scatter_block.f, resolved once
// per plan).  Depending on which side of the result is contiguous in the
// ME-list, A/B are (Op1,Op2) gathers or (Op2,Op1) gathers.
#pragma once
#include <stdint.h>

namespace gecco {

// Signed element index: 0 = Pauli-forbidden (element is zero), otherwise
// sign * (1 + element offset).  Same convention as GeCCo's string maps
// (strings/set_strmap_kernel.f:100-109).
struct DevOperand {
    const int32_t *first;   // separable: per first index; table: full table
    const int32_t *second;  // separable: per second index; table: nullptr
    int64_t ld;             // table: extent of the first index
    int32_t table;          // 1 = table mode
    int32_t fast_first;     // gathers: 1 if consecutive k are adjacent in memory
    // separable vectors that are affine (entry = 1 + a0 + a1*i, all positive)
    // are evaluated arithmetically instead of being loaded
    int32_t first_affine, second_affine;
    int64_t f0, f1, s0, s1;
};

struct DevBlock {
    double xfac;
    int64_t base_a, base_b, base_c;   // element offsets of the block starts
    int32_t buf_a, buf_b, buf_c;      // buffer slots
    int32_t M, N, K;
    int32_t atomic;                   // result tiles may collide
    // next block of the chain (-1: none): blocks of the strided-GEMM kernel
    // that add into the same result block through the same result map; a
    // task's contraction range continues into the next block of its chain
    int32_t next;
    // norm-screening jobs of the A and B operand blocks (-1: not screened)
    int32_t zero_a, zero_b;
    DevOperand A, B, C;               // A: (k,m)  B: (k,n)  C: (m,n)
};

// One chunk of the squared norm of an operand block (norm screening):
// sum of x^2 over elements [e0, e1) of buffer slot buf from element off.
struct DevNormTask {
    int64_t off, e0, e1;
    int32_t buf, pad;
};

struct DevTask {
    // contraction range [k0, k1) of the tile (m0, n0): counted from the start
    // of block blk, continuing along its chain (strided-GEMM kernel) when k1
    // exceeds blk's K
    int32_t blk, m0, n0, k0, k1;
    // span > 0 (direct short-contraction kernel only): the tile's results are
    // first reduced in LDS over the block-relative element range [lo, lo + span)
    // and then written back as one contiguous range -- read-modify-write if
    // the task owns the range (owned = 1), else atomically
    int32_t span, owned, pad;
    // instance group: the task is skipped while skip[grp] != 0 (skip flags
    // given to the plan); grp < 0 never skips
    int32_t grp, pad2;
    int64_t lo;
};

// Materialisation of a gathered operand (the reference's collect step,
// contract/collect_block_impr.f): dense[e] = sign * src[addr(k,x)], or 0 for a
// forbidden element, with e = k + K*x (kfast) or x + X*k, written into
// plan-owned scratch before the strided-GEMM launch.  K*X < 2^31.
struct DevGather {
    DevOperand src;        // index data of the source operand (k, x)
    int64_t src_base;      // element offset of the source block
    int64_t dst_base;      // element offset in the scratch buffer
    int32_t src_buf;       // buffer slot of the source list
    int32_t K, X;
    int32_t kfast;
};

// List initialisation before a plan's contractions: dst[i] = factor *
// src[i] (src < 0: 0) for i < len.
struct DevInit {
    int64_t off_dst, off_src, len;
    double factor;
    int32_t dst, src;
    // 0: dst = factor * src (or 0 without a source); 1: dst += factor * src
    // (atomic adds, so that several sums into one list may share a launch)
    int32_t accumulate, pad;
};

struct DevInitTask {
    int32_t job, grp;      // grp: instance group (see DevTask)
    int64_t e0, e1;
};

// Spin-flip symmetrisation steps of a plan, run as one launch: pairs
// (base + i1[k], base + i2[k], f[k]) of step s occupy [pair0, pair0 + n).
struct DevSymStep {
    int64_t base, pair0;
    double prefac;
    int32_t buf, pad;
};

struct DevSymTask {
    int32_t step, grp;     // grp: instance group (see DevTask)
    int64_t p0, p1;        // pair range (global pair indices)
};

struct DevGatherTask {
    int32_t job;
    int32_t e0, e1;        // element range [e0, e1) of the job's dense array
    int32_t grp;           // instance group (see DevTask)
};

}  // namespace gecco
