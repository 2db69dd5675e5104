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
};

template <int DMAX, bool FUSED, bool WAVE>
__device__ __forceinline__ void run_team(const Job *J, const JobShift &sh, const LdsArgs &a, char *lds) {
    switch (J->D) {
        case 2:
            if constexpr (DMAX >= 2) sweep_lds<2, FUSED, WAVE>(J, sh, a.P, a.T, lds);
            break;
        case 3:
            if constexpr (DMAX >= 3) sweep_lds<3, FUSED, WAVE>(J, sh, a.P, a.T, lds);
            break;
        case 4:
            if constexpr (DMAX >= 4) sweep_lds<4, FUSED, WAVE>(J, sh, a.P, a.T, lds);
            break;
        case 5:
            if constexpr (DMAX >= 5) sweep_lds<5, FUSED, WAVE>(J, sh, a.P, a.T, lds);
            break;
        case 6:
            if constexpr (DMAX >= 6) sweep_lds<6, FUSED, WAVE>(J, sh, a.P, a.T, lds);
            break;
        case 7:
            if constexpr (DMAX >= 7) sweep_lds<7, FUSED, WAVE>(J, sh, a.P, a.T, lds);
            break;
        case 8:
            if constexpr (DMAX >= 8) sweep_lds<8, FUSED, WAVE>(J, sh, a.P, a.T, lds);
            break;
        default:
            break;
    }
}

__device__ __forceinline__ JobShift shift_of(const LdsArgs &a, int64_t k) {
    JobShift sh;
    sh.k = k;
    sh.ua = a.ua;
    sh.ub = a.ub;
    sh.uc = a.uc;
    sh.ug = a.ug;
    return sh;
}

__device__ __forceinline__ Job shifted(const LdsArgs &a, int64_t k) {
    Job J = a.tmpl;
    J.a_off += k * a.ua;
    J.b_off += k * a.ub;
    J.c_off += k * a.uc;
    J.g_off += k * a.ug;
    return J;
}

template <int DMAX, bool FUSED, int MAXT>
__global__ __launch_bounds__(MAXT) void lds_kernel(LdsArgs a) {
    extern __shared__ __align__(16) char lds[];
    const int lane = threadIdx.x & (MM_WAVE - 1);
    const int wave = threadIdx.x / MM_WAVE;
    const int wpb = blockDim.x / MM_WAVE;

    if (a.mode == MODE_UNIFORM) {
        const int kind = a.tmpl.kind;
        if (kind == KIND_BLOCK) {
            for (int64_t k = blockIdx.x; k < a.n; k += gridDim.x) run_team<DMAX, FUSED, false>(&a.tmpl, shift_of(a, k), a, lds);
        } else if (kind == KIND_WAVE) {
            for (int64_t k = int64_t(blockIdx.x) * wpb + wave; k < a.n; k += int64_t(gridDim.x) * wpb)
                run_team<DMAX, FUSED, true>(&a.tmpl, shift_of(a, k), a, lds + wave * a.wave_lds);
        } else {
            for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < a.n; k += int64_t(gridDim.x) * blockDim.x) {
                const Job J = shifted(a, k);
                chain_lane<FUSED>(J, a.P, a.T);
            }
        }
        return;
    }

    // ---- queue mode.  Every loop ends once its counter passes its job count.
    __shared__ int s_idx;
    for (;;) {
        if (threadIdx.x == 0) s_idx = atomicAdd(&a.counters[0], 1);
        __syncthreads();
        const int idx = s_idx;
        __syncthreads();
        if (idx >= a.nb) break;
        run_team<DMAX, FUSED, false>(a.table + idx, shift_of(a, 0), a, lds);
    }
    char *wl = lds + wave * a.wave_lds;
    for (;;) {
        int idx = 0;
        if (lane == 0) idx = atomicAdd(&a.counters[1], 1);
        idx = __builtin_amdgcn_readfirstlane(__shfl(idx, 0, MM_WAVE));
        if (idx >= a.nw) break;
        run_team<DMAX, FUSED, true>(a.table + a.nb + idx, shift_of(a, 0), a, wl);
    }
    for (;;) {
        int base = 0;
        if (lane == 0) base = atomicAdd(&a.counters[2], MM_WAVE);
        base = __builtin_amdgcn_readfirstlane(__shfl(base, 0, MM_WAVE));
        if (base >= a.nl) break;
        const int64_t idx = int64_t(base) + lane;
        if (idx < a.nl) {
            const Job J = a.table[a.nb + a.nw + idx];
            chain_lane<FUSED>(J, a.P, a.T);
        }
    }
}

}  // namespace mm
