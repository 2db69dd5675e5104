// Job descriptors and LDS layout shared by the host planner and the kernels.
#pragma once

#include <stdint.h>

namespace mm {

// How a triple is executed by the LDS-resident sweep.
enum JobKind : int32_t {
    KIND_BLOCK = 0,  // one workgroup per triple, workgroup barriers between slabs
    KIND_WAVE = 1,   // one wave per triple, wave-level ordering between slabs
    KIND_LANE = 2,   // one lane per triple: a one-index chain kept in registers
};

enum JobFlags : int32_t {
    JOB_WRITE_PARTIAL = 1,  // fused: store raw VJP sums to the partial buffer, do not finalise
};

// One (A, b, c) -> G job.  Offsets are in complex elements from the base
// pointers handed to the kernel.  `top` is the shallowest level swept in LDS:
// the job covers the sub-lattice {k : k_0 = ... = k_{top-1} = 0}, whose
// level-`top` slabs have S = stride[top] points.
struct Job {
    int32_t D, kind, top, S;
    int32_t flags, nacc, smax;
    int32_t nf;  // set by the team that swept the job when it stored a non-finite value
    int32_t shape[8];
    int64_t stride[8];
    int32_t kbit[8];   // bit offset of digit k_j in the packed per-point index word
    int32_t kmask[8];  // mask of digit k_j (after the shift)
    int64_t a_off, b_off, c_off, g_off, p_off;
    int32_t lda;  // row stride of A (0: D); a sub-triple A[top:, top:] of a larger one has the larger D
};
static_assert(sizeof(Job) <= 256, "Job must fit its LDS header");

// Number of complex VJP accumulators: upper triangle of dA, db, dc.
__host__ __device__ inline int nacc_of(int D) { return D * (D + 1) / 2 + D + 1; }

// Row index of the coefficient table entry (i, j), j >= i, counted from `top`.
__host__ __device__ inline int coef_row(int D, int top, int i, int j) {
    int tri_i = i * D - i * (i - 1) / 2;
    int tri_t = top * D - top * (top - 1) / 2;
    return tri_i - tri_t + (j - i);
}

__host__ __device__ inline int coef_rows(int D, int top) { return coef_row(D, top, D - 1, D - 1) + 1; }

__host__ __device__ inline int64_t align16(int64_t x) { return (x + 15) & ~int64_t(15); }

// Byte offsets of the LDS regions of one team.
struct LdsLayout {
    int64_t job, ring, kv, coef, sq, rsq, bvec, amat, red, total;
};

// Bytes reserved at the start of a team's LDS region for its Job descriptor.
constexpr int64_t JOB_LDS_BYTES = 256;

// Largest number of level-`top` slab points one thread keeps in registers,
// per kernel instantiation (kernels for D <= 3, or for D <= 8).
__host__ __device__ constexpr int rmax_for(int dmax, bool fused) { return dmax <= 3 ? (fused ? 2 : 4) : (fused ? 1 : 2); }

// The register path (whole triples only, top = 0) holds every level-0 slab
// point of a team in registers; it needs two LDS slab buffers even when
// fused.  Otherwise the fused sweep keeps three.
__host__ __device__ inline bool reg_path(int S, int nwaves, int rmax) { return int64_t(S) <= int64_t(rmax) * nwaves * 64; }

__host__ __device__ inline LdsLayout lds_layout(int D, int top, int S, int smax, int fused, int nwaves, int rmax) {
    LdsLayout L;
    int nr = (fused && !(top == 0 && reg_path(S, nwaves, rmax))) ? 3 : 2;
    int64_t off = 0;
    L.job = off;
    off += JOB_LDS_BYTES;
    L.ring = off;
    off = align16(off + int64_t(nr) * S * 16);
    L.kv = off;
    off = align16(off + int64_t(S) * 4);
    L.coef = off;
    off = align16(off + int64_t(coef_rows(D, top)) * smax * 16);
    L.sq = off;
    off = align16(off + int64_t(smax) * 8);
    L.rsq = off;
    off = align16(off + int64_t(smax) * 8);
    L.bvec = off;
    off = align16(off + int64_t(D + 1) * 16);
    L.amat = off;
    off = align16(off + int64_t(D) * D * 16);
    L.red = off;
    off = align16(off + (fused ? int64_t(nwaves) * nacc_of(D) * 16 : 0));
    L.total = off;
    return L;
}

}  // namespace mm
