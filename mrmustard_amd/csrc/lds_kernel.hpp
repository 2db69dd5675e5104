// The LDS-resident kernel in its three launch modes:
//   UNIFORM: job k = template job shifted by k triples (one shape), assigned
//            statically to a workgroup, a wave or a lane;
//   QUEUE:   the persistent work-queue kernel of a plan: workgroups first take
//            the large triples (largest first), then every wave takes small
//            triples on its own, then every lane takes one-index chains.
#pragma once

#include "sweep.hpp"

namespace mm {

enum LaunchMode : int32_t { MODE_UNIFORM = 0, MODE_QUEUE = 1 };

struct LdsArgs {
    Bases P;
    Tables T;
    int32_t mode;
    int32_t wave_lds;  // bytes of LDS owned by each wave in the wave phase
    Job tmpl;
    int64_t ua, ub, uc, ug;  // uniform mode: per-triple strides of A, b/db, c/dc, G/g
    int64_t n;               // uniform mode: number of triples
    const Job *table;        // queue mode: block jobs, then wave jobs, then lane jobs
    int64_t nb, nw, nl;
    int *counters;           // queue mode: 3 zeroed counters
    int clear_nf;            // one-workgroup launch that starts a multi-launch path: clear P.nf first
};

template <int DMAX, bool FUSED, bool WAVE>
__device__ __forceinline__ void run_team(const Job *src, const JobShift &sh, const Bases &P, const Tables &T, char *lds) {
    load_job<WAVE>(lds, src, sh);
    const int D = reinterpret_cast<const Job *>(lds)->D;
    constexpr int RM = rmax_for(DMAX, FUSED);
    switch (D) {
        case 2:
            if constexpr (DMAX >= 2) sweep_lds<2, FUSED, WAVE, RM>(P, T, lds);
            break;
        case 3:
            if constexpr (DMAX >= 3) sweep_lds<3, FUSED, WAVE, RM>(P, T, lds);
            break;
        case 4:
            if constexpr (DMAX >= 4) sweep_lds<4, FUSED, WAVE, RM>(P, T, lds);
            break;
        case 5:
            if constexpr (DMAX >= 5) sweep_lds<5, FUSED, WAVE, RM>(P, T, lds);
            break;
        case 6:
            if constexpr (DMAX >= 6) sweep_lds<6, FUSED, WAVE, RM>(P, T, lds);
            break;
        case 7:
            if constexpr (DMAX >= 7) sweep_lds<7, FUSED, WAVE, RM>(P, T, lds);
            break;
        case 8:
            if constexpr (DMAX >= 8) sweep_lds<8, FUSED, WAVE, RM>(P, T, lds);
            break;
        default:
            break;
    }
}

template <int DMAX, bool FUSED, int MAXT>
__global__ __launch_bounds__(MAXT) void lds_kernel(LdsArgs a) {
    extern __shared__ __align__(16) char lds[];
    const int lane = threadIdx.x & (MM_WAVE - 1);
    const int wave = threadIdx.x / MM_WAVE;
    const int wpb = blockDim.x / MM_WAVE;
    const Bases P = a.P;
    const Tables T = a.T;
    if (a.clear_nf && threadIdx.x == 0) *P.nf = 0;  // the sweep raises it only after barriers

    if (a.mode == MODE_UNIFORM) {
        JobShift sh;
        sh.ua = a.ua;
        sh.ub = a.ub;
        sh.uc = a.uc;
        sh.ug = a.ug;
        const int kind = a.tmpl.kind;
        if (kind == KIND_BLOCK) {
            for (int64_t k = blockIdx.x; k < a.n; k += gridDim.x) {
                sh.k = k;
                run_team<DMAX, FUSED, false>(&a.tmpl, sh, P, T, lds);
            }
        } else if (kind == KIND_WAVE) {
            for (int64_t k = int64_t(blockIdx.x) * wpb + wave; k < a.n; k += int64_t(gridDim.x) * wpb) {
                sh.k = k;
                run_team<DMAX, FUSED, true>(&a.tmpl, sh, P, T, lds + wave * a.wave_lds);
            }
        } else {
            const int s0 = a.tmpl.shape[0];
            for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < a.n; k += int64_t(gridDim.x) * blockDim.x)
                chain_lane<FUSED>(s0, a.tmpl.a_off + k * a.ua, a.tmpl.b_off + k * a.ub, a.tmpl.c_off + k * a.uc, a.tmpl.g_off + k * a.ug, P, T);
        }
        return;
    }

    // ---- queue mode.  Every loop ends once its counter passes its job count.
    JobShift none;
    none.k = 0;
    none.ua = none.ub = none.uc = none.ug = 0;
    __shared__ int s_idx;
    for (;;) {
        if (threadIdx.x == 0) s_idx = atomicAdd(&a.counters[0], 1);
        __syncthreads();
        const int idx = s_idx;
        __syncthreads();
        if (idx >= a.nb) break;
        run_team<DMAX, FUSED, false>(a.table + idx, none, P, T, lds);
    }
    char *wl = lds + wave * a.wave_lds;
    for (;;) {
        int idx = 0;
        if (lane == 0) idx = atomicAdd(&a.counters[1], 1);
        idx = __builtin_amdgcn_readfirstlane(__shfl(idx, 0, MM_WAVE));
        if (idx >= a.nw) break;
        run_team<DMAX, FUSED, true>(a.table + a.nb + idx, none, P, T, wl);
    }
    for (;;) {
        int base = 0;
        if (lane == 0) base = atomicAdd(&a.counters[2], MM_WAVE);
        base = __builtin_amdgcn_readfirstlane(__shfl(base, 0, MM_WAVE));
        if (base >= a.nl) break;
        const int64_t idx = int64_t(base) + lane;
        if (idx < a.nl) {
            const Job *J = a.table + a.nb + a.nw + idx;
            chain_lane<FUSED>(J->shape[0], J->a_off, J->b_off, J->c_off, J->g_off, P, T);
        }
    }
}

// Uniform batches whose triples all take one workgroup each, with D known at
// compile time and RM points per thread: nothing but that one sweep is in
// the kernel, so it needs fewer registers than lds_kernel and a 1024-thread
// workgroup keeps four waves on every SIMD.
template <int D, bool FUSED, int MAXT, int RM>
__global__ __launch_bounds__(MAXT) void lds_block_kernel(LdsArgs a) {
    extern __shared__ __align__(16) char lds[];
    if (a.clear_nf && threadIdx.x == 0) *a.P.nf = 0;  // the sweep raises it only after barriers
    JobShift sh;
    sh.ua = a.ua;
    sh.ub = a.ub;
    sh.uc = a.uc;
    sh.ug = a.ug;
    for (int64_t k = blockIdx.x; k < a.n; k += gridDim.x) {
        sh.k = k;
        load_job<false>(lds, &a.tmpl, sh);
        sweep_lds<D, FUSED, false, RM>(a.P, a.T, lds);
    }
}

}  // namespace mm
