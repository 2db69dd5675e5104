// The one-leftover-mode strategy: the conditional density matrices
// out[m, n, k] = G[(m, k), (n, k)] of mode 0 when modes 1.. are PNR-detected
// with k photons (mrmustard/math/lattice/strategies/compactFock/
// singleLeftoverMode_amps.py:288-421; the numpy backend's fast_diagonal.py
// computes the same amplitudes).
//
// It is the diagonal strategy over the tail modes 1.. with every params
// carrying a (c0, c0) block of mode 0: each pivot output at block entry
// (m, n) is the diagonal strategy's formula plus the two lowering terms of
// mode 0 taken from the pivot's own block (write_block, :20-72):
//   out_i(m, n) = (piv(m, n) B_i + A_i0 sqrt(m) piv(m-1, n) + A_i1 sqrt(n) piv(m, n-1)
//                  + sum_j A_i,j+2 G_in_j(m, n)) / K_i.
// Block entries are independent given the pivot block, so the launch
// structure of diag_group_kernel applies with the block entry as the
// innermost index (where the diagonal strategy has its batch column).  The
// off-diagonal pivots' own blocks are the diagonal pivot's outputs at other
// block entries, so each weight runs two launches: the diagonal pivots (which
// also store arr1[2d, p] blocks for the second), then the off-diagonal ones.
#pragma once

#include "diagonal.hpp"

namespace mm {

struct LeftArgs {
    DiagArgs t;  // tail modes: geometry, list / fwd, inboxes (block entry innermost), d >= 1 sub-lattices
    const double2 *A, *B, *c;  // full 2L x 2L, 2L, 1 (L = tail modes + 1)
    int c0, C;                 // leftover cutoff, c0 * c0
    double2 *out;              // (c0, c0, *tail cutoffs): entry (pos, flat p) at pos * P + f
    double2 *piv;              // tail modes x P x C: arr1[2d, p] blocks of the current weight
};

// The leftover block at zero detections: the two-index lattice of
// (A[:2, :2], B[:2], c) (singleLeftoverMode_amps.py:330-341).  Column 0 is a
// sequential recurrence (thread 0, in registers); column n + 1 needs columns n
// and n - 1 only, so each later column is one step of c0 threads over three
// rotating columns in LDS (dynamic, 3 c0 entries).  c0 <= blockDim.x.
template <int MT, bool REORDER>
__global__ __launch_bounds__(1024) void lo_init_kernel(LeftArgs a) {
    extern __shared__ double2 cols[];  // 3 x c0
    constexpr int L = MT + 1;
    const int c0 = a.c0, tid = threadIdx.x;
    const int64_t P = a.t.P;
    const double2 A00 = a.A[src_index<L, REORDER>(0) * 2 * L + src_index<L, REORDER>(0)];
    const double2 A10 = a.A[src_index<L, REORDER>(1) * 2 * L + src_index<L, REORDER>(0)];
    const double2 A11 = a.A[src_index<L, REORDER>(1) * 2 * L + src_index<L, REORDER>(1)];
    const double2 B0 = a.B[src_index<L, REORDER>(0)], B1 = a.B[src_index<L, REORDER>(1)];
    double2 *blk = a.t.inO;  // inbox slot 0 of position 0: entry (m, n) at m * c0 + n
    if (tid == 0) {
        double2 prev = make_double2(0.0, 0.0), cur = a.c[0];
        cols[0] = cur;
        for (int m = 0; m + 1 < c0; ++m) {
            double2 t = cmul(cur, B0);
            if (m > 0) t = cadd(t, cmul(cscale_promoted(A00, a.t.sq[m]), prev));
            prev = cur;
            cur = cdiv_by(t, a.t.sq[m + 1], a.t.rsq[m + 1]);
            cols[m + 1] = cur;
        }
    }
    __syncthreads();
    if (tid < c0) {
        blk[tid * c0] = cols[tid];
        a.out[int64_t(tid * c0) * P] = cols[tid];
    }
    for (int n = 0; n + 1 < c0; ++n) {
        const double2 *cur = cols + (n % 3) * c0, *prv = cols + ((n + 2) % 3) * c0;
        double2 *nxt = cols + ((n + 1) % 3) * c0;
        if (tid < c0) {
            const int m = tid;
            double2 t = cmul(cur[m], B1);
            if (m > 0) t = cadd(t, cmul(cscale_promoted(A10, a.t.sq[m]), cur[m - 1]));
            if (n > 0) t = cadd(t, cmul(cscale_promoted(A11, a.t.sq[n]), prv[m]));
            const double2 v = cdiv_by(t, a.t.sq[n + 1], a.t.rsq[n + 1]);
            nxt[m] = v;
            blk[m * c0 + n + 1] = v;
            a.out[int64_t(m * c0 + n + 1) * P] = v;
        }
        __syncthreads();
    }
}

// Shared set-up of both launches: lane (slot, i), group = (params, block entry).
template <int MT>
struct LoLane {
    int slot, i, m, pos, bm, bn, pm, cm;
    bool lane_ok, act;
    int64_t f, gs, fw, sm;
    int p[MT];
};

template <int MT, int PB>
__device__ __forceinline__ LoLane<MT> lo_lane(const LeftArgs &a, int64_t count, int64_t start, int64_t blk) {
    LoLane<MT> l;
    const int t = threadIdx.x;
    l.slot = t % PB;
    l.i = t / PB;
    l.m = l.i >> 1;
    l.lane_ok = l.i < 2 * MT;
    const int64_t grp = blk * PB + l.slot;
    l.act = l.lane_ok && grp < count * a.C;
    l.f = l.gs = l.sm = 0;
    l.fw = -1;
    l.pos = l.bm = l.bn = 0;
#pragma unroll
    for (int j = 0; j < MT; ++j) l.p[j] = 0;
    if (l.act) {
        const int64_t q = grp / a.C;
        l.pos = int(grp - q * a.C);
        l.bm = l.pos / a.c0;
        l.bn = l.pos - l.bm * a.c0;
        l.gs = start + q;
        l.f = a.t.list[l.gs];
        l.fw = a.t.fwd[l.m * a.t.P + l.gs];
        uint32_t r = uint32_t(l.f);
#pragma unroll
        for (int j = 0; j < MT; ++j) {
            const uint32_t sj = uint32_t(a.t.stride[j]);
            l.p[j] = int(r / sj);
            r -= uint32_t(l.p[j]) * sj;
        }
    }
    l.pm = 0;
    l.cm = 1;
#pragma unroll
    for (int j = 0; j < MT; ++j)
        if (j == l.m) {
            l.pm = l.p[j];
            l.cm = a.t.cut[j];
            l.sm = a.t.stride[j];
        }
    return l;
}

template <int MT>
__device__ __forceinline__ int lo_runmask(const LeftArgs &a, const LoLane<MT> &l) {
    int mask = 0;
    if (!l.act) return 0;
    bool zp = true;
#pragma unroll
    for (int d = 0; d < MT; ++d) {
        if (zp && l.p[d] < a.t.cut[d] - 1) mask |= 1 << d;
        zp = zp && l.p[d] == 0;
    }
    return mask;
}

// (piv B_r + A_r0 sqrt(m) piv(m-1, n) + A_r1 sqrt(n) piv(m, n-1)) for output row r of the full A;
// Br = B_r, sbm = sqrt(m), sbn = sqrt(n), loaded by the caller with its other inputs
template <int L>
__device__ __forceinline__ double2 lo_self_terms(const double2 (*sA)[2 * L], int r, double2 Br, double sbm, double sbn, double2 pv,
                                                 double2 pvm, double2 pvn, int bm, int bn) {
    double2 t = cmul(pv, Br);
    if (bm > 0) t = cadd(t, cmul(sA[r][0], cscale_promoted(pvm, sbm)));
    if (bn > 0) t = cadd(t, cmul(sA[r][1], cscale_promoted(pvn, sbn)));
    return t;
}

// Diagonal pivots of weight w (use_diag_pivot, :224-285).
template <int MT, bool REORDER>
__global__ __launch_bounds__(DG_THREADS) void lo_diag_kernel(LeftArgs a, int w, int64_t count, int64_t start) {
    constexpr int G = 2 * MT, L = MT + 1;
    constexpr int PB = DG_THREADS / G;
    __shared__ double2 sgin[G][PB + 1];
    __shared__ double2 sA[2 * L][2 * L];
    for (int e = threadIdx.x; e < 4 * L * L; e += DG_THREADS)
        sA[e / (2 * L)][e % (2 * L)] = a.A[src_index<L, REORDER>(e / (2 * L)) * 2 * L + src_index<L, REORDER>(e % (2 * L))];
    const LoLane<MT> l = lo_lane<MT, PB>(a, count, start, blockIdx.x);
    const int64_t P = a.t.P, C = a.C;
    const int runmask = lo_runmask<MT>(a, l);
    const bool diag = l.act && (a.t.cut[0] == 1 || l.p[0] < a.t.cut[0] - 1);
    // every input of the lane in one round trip: its block of arr0 and the
    // two lowered entries (read at the entry itself where the index is 0),
    // its diagonal input, B_r and the sqrt entries it uses
    const int64_t base = l.act ? l.gs * C + l.pos : 0;  // inbox slot 0: arr0 block of p
    const double2 g0r = a.t.inO[base];
    const double2 gmr = a.t.inO[base - (l.bm > 0 ? a.c0 : 0)];
    const double2 gnr = a.t.inO[base - (l.bn > 0 ? 1 : 0)];
    const bool von = diag && l.pm > 0;
    const double2 vr = a.t.inD[von ? (l.i * P + l.gs) * C + l.pos : 0];
    const int r = l.lane_ok ? l.i + 2 : 2;
    const double2 Br = a.B[src_index<L, REORDER>(r)];
    const double s0 = a.t.sq[l.pm], s1 = a.t.sq[l.pm + 1], r1 = a.t.rsq[l.pm + 1], sbm = a.t.sq[l.bm], sbn = a.t.sq[l.bn];
    const double2 zero = make_double2(0.0, 0.0);
    const double2 g0 = l.act ? g0r : zero, gm = l.act && l.bm > 0 ? gmr : zero, gn = l.act && l.bn > 0 ? gnr : zero;
    const double2 v = von ? cscale_promoted(vr, s0) : zero;
    if (l.lane_ok) sgin[l.i][l.slot] = v;
    __syncthreads();
    if (diag && l.pm + 1 < l.cm && (l.i != 1 || l.p[0] + 2 < a.t.cut[0])) {
        double2 t = lo_self_terms<L>(sA, r, Br, sbm, sbn, g0, gm, gn, l.bm, l.bn);
#pragma unroll
        for (int j = 0; j < G; ++j) t = cadd(t, cmul(sA[r][j + 2], sgin[j][l.slot]));
        const double2 o = cdiv_by(t, s1, r1);
        a.t.inD[((l.i ^ 1) * P + l.fw) * C + l.pos] = o;
        if (!(l.i & 1) && (runmask >> l.m & 1)) a.piv[(l.m * P + l.gs) * C + l.pos] = o;
    }
}

// Off-diagonal pivots of weight w (use_offDiag_pivot, :99-221).
template <int MT, bool REORDER>
__global__ __launch_bounds__(DG_THREADS) void lo_off_kernel(LeftArgs a, int w, int64_t count, int64_t start) {
    constexpr int G = 2 * MT, L = MT + 1;
    constexpr int PB = DG_THREADS / G;
    __shared__ double2 sgin[MT][G][PB + 1];
    __shared__ double2 sA[2 * L][2 * L];
    for (int e = threadIdx.x; e < 4 * L * L; e += DG_THREADS)
        sA[e / (2 * L)][e % (2 * L)] = a.A[src_index<L, REORDER>(e / (2 * L)) * 2 * L + src_index<L, REORDER>(e % (2 * L))];
    const LoLane<MT> l = lo_lane<MT, PB>(a, count, start, blockIdx.x);
    const int64_t P = a.t.P, C = a.C;
    const int runmask = lo_runmask<MT>(a, l);
    // every input of the lane in one round trip (from the lane's own arr0
    // entry where it needs none): the inputs of each pivot d this params runs
    // (slots d >= 1 by flat index), the arr1[2d, p] blocks with their two
    // lowered entries, B_r and the sqrt entries of the lane's mode and block
    // entry (the lanes 2d and 2d + 1 of pivot d have m = d, so p_d = p_m)
    const int64_t own = l.act ? l.gs * C + l.pos : 0;
    const double2 g0 = l.act ? a.t.inO[own] : make_double2(0.0, 0.0);
    double2 xr[MT], pv[MT], pvm[MT], pvn[MT];
#pragma unroll
    for (int d = 0; d < MT; ++d) {
        const double2 *xp = a.t.inO;
        int64_t xe = own;
        if (l.lane_ok && l.i > 2 * d && (runmask >> d & 1) && l.pm > 0) {
            if (d == 0) {
                xe = (l.i * P + l.gs) * C + l.pos;
            } else if (l.i == 2 * d + 1) {
                xp = a.t.arr2;
                xe = (a.t.off2[d] + l.f - l.sm) * C + l.pos;
            } else {
                xp = (l.i & 1) ? a.t.a1010 : a.t.a1001;
                xe = (a.t.off11[d] + (l.m - d - 1) * a.t.len[d] + l.f - l.sm) * C + l.pos;
            }
        }
        xr[d] = xp[xe];
        const int64_t pb = l.act ? (d * P + l.gs) * C + l.pos : 0;  // arr1[2d, p] block
        pv[d] = a.piv[pb];
        pvm[d] = a.piv[pb - (l.bm > 0 ? a.c0 : 0)];
        pvn[d] = a.piv[pb - (l.bn > 0 ? 1 : 0)];
    }
    const int r = l.lane_ok ? l.i + 2 : 2;
    const double2 Br = a.B[src_index<L, REORDER>(r)];
    const double s0 = a.t.sq[l.pm], s1 = a.t.sq[l.pm + 1], r1 = a.t.rsq[l.pm + 1], s2 = a.t.sq[l.pm + 2], r2 = a.t.rsq[l.pm + 2];
    const double sbm = a.t.sq[l.bm], sbn = a.t.sq[l.bn];
#pragma unroll
    for (int d = 0; d < MT; ++d) {
        if (!l.lane_ok || l.i < 2 * d) continue;
        double2 v = make_double2(0.0, 0.0);
        if (runmask >> d & 1) {
            if (l.i == 2 * d)
                v = cscale_promoted(g0, s1);
            else if (l.pm > 0)
                v = cscale_promoted(xr[d], s0);
        }
        sgin[d][l.i][l.slot] = v;
    }
    __syncthreads();
#pragma unroll
    for (int d = 0; d < MT; ++d) {
        if (!(runmask >> d & 1) || l.i < 2 * d) continue;
        const double2 zero = make_double2(0.0, 0.0);
        double2 t = lo_self_terms<L>(sA, r, Br, sbm, sbn, pv[d], l.bm > 0 ? pvm[d] : zero, l.bn > 0 ? pvn[d] : zero, l.bm, l.bn);
#pragma unroll
        for (int j = 2 * d; j < G; ++j) t = cadd(t, cmul(sA[r][j + 2], sgin[d][j][l.slot]));
        if (l.i == 2 * d + 1) {
            // arr0[p + e_d]: the output, and inbox slot 0 of that params
            const double2 o = cdiv_by(t, s1, r1);
            a.out[int64_t(l.pos) * P + l.f + a.t.stride[d]] = o;
            a.t.inO[l.fw * C + l.pos] = o;
        } else if (l.i == 2 * d) {
            if (l.pm + 2 < l.cm) {
                const double2 o = cdiv_by(t, s2, r2);
                if (d == 0)
                    a.t.inO[(P + l.fw) * C + l.pos] = o;
                else
                    a.t.arr2[(a.t.off2[d] + l.f) * C + l.pos] = o;
            }
        } else if (l.pm + 1 < l.cm) {
            const double2 o = cdiv_by(t, s1, r1);
            if (d == 0) {
                a.t.inO[((l.i ^ 1) * P + l.fw) * C + l.pos] = o;
            } else {
                const int64_t e = (a.t.off11[d] + (l.m - d - 1) * a.t.len[d] + l.f) * C + l.pos;
                if (l.i & 1)
                    a.t.a1001[e] = o;
                else
                    a.t.a1010[e] = o;
            }
        }
    }
}

}  // namespace mm
