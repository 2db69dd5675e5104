// libmmhip: C ABI over the MI355X Bargmann->Fock kernels.
//
// Host-side planning: validation (the ValueError conditions of
// backend_manager.py:674-689 plus the shape/D checks the reference leaves to
// chance), choice of execution path per triple, kernel launches.  Device
// state is one sqrt / 1/sqrt table per device, created on first use.

#include "host.hpp"
#include "common.hpp"
#include "global_kernels.hpp"
#include "jobs.hpp"
#include "lds_kernel.hpp"
#include "lds_launch.hpp"
#include "binomial.hpp"
#include "rows_dot.hpp"

using namespace mm;

// Device scratch of mm_hermite_batch_vjp_static: the VJP's per-block partial
// sums when the caller cannot allocate a workspace.  One copy per GPU, held
// until the process ends.  Host code never uses this name as an address: the
// address on each device comes from hipGetSymbolAddress (scratch_address).
static __device__ __align__(256) char scratch[MM_STATIC_SCRATCH_BYTES];

// Non-finite flags of the multi-launch forward paths, one per stream: calls
// on one stream run in order, so no two calls in flight share a flag unless
// more than NF_SLOTS streams are in use (then slots are reused).  Each call
// clears its flag with its first launch and reads it with its last.
constexpr int NF_SLOTS = 1024;
static __device__ int nf_ring[NF_SLOTS];

namespace {

std::mutex g_scratch_mu;
std::map<int, void *> g_scratch;  // device -> its copy of `scratch`

// Look up `scratch` on the current device once (mm_init does it).
int scratch_init() {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch.count(dev)) return MM_OK;
    void *p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(MM_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch.emplace(dev, p);
    return MM_OK;
}

std::mutex g_nf_mu;
std::map<int, int *> g_nf_base;                         // device -> its copy of nf_ring
std::map<std::pair<int, hipStream_t>, int> g_nf_slot;  // (device, stream) -> slot
int g_nf_next = 0;

int nf_base(int dev, int **out) {
    auto it = g_nf_base.find(dev);
    if (it != g_nf_base.end()) {
        *out = it->second;
        return MM_OK;
    }
    void *p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(nf_ring)));
    if (!p) return fail(MM_ERR_HIP, "hipGetSymbolAddress(nf_ring) returned NULL");
    g_nf_base.emplace(dev, static_cast<int *>(p));
    *out = static_cast<int *>(p);
    return MM_OK;
}

int nf_init() {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_nf_mu);
    int *base = nullptr;
    return nf_base(dev, &base);
}

// The flag of stream s on the current device.
int nf_flag(hipStream_t s, int **out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_nf_mu);
    int *base = nullptr;
    const int rc = nf_base(dev, &base);
    if (rc) return rc;
    const auto key = std::make_pair(dev, s);
    auto it = g_nf_slot.find(key);
    int slot = 0;
    if (it == g_nf_slot.end()) {
        slot = g_nf_next++ % NF_SLOTS;
        g_nf_slot.emplace(key, slot);
    } else {
        slot = it->second;
    }
    *out = base + slot;
    return MM_OK;
}

// Address of `scratch` on the current device, or an error if mm_init has not
// run there.
int scratch_address(void **out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch.find(dev);
    if (it == g_scratch.end()) return fail(MM_ERR_INVALID, "mm_init(%d) has not been called on this device", dev);
    *out = it->second;
    return MM_OK;
}

// ------------------------------------------------------------------ geometry

struct Shape {
    int D = 0;
    int32_t s[8] = {0};
    int64_t st[8] = {0};
    int64_t N = 0;
};

int make_shape(const int64_t *shape, int D, Shape *out) {
    if (!shape) return fail(MM_ERR_INVALID, "shape pointer is NULL");
    if (D < 1 || D > MM_MAX_DIMS) return fail(MM_ERR_INVALID, "D=%d outside [1, %d]", D, MM_MAX_DIMS);
    Shape S;
    S.D = D;
    int64_t n = 1;
    for (int i = 0; i < D; ++i) {
        if (shape[i] < 1 || shape[i] >= MM_MAX_CUTOFF)
            return fail(MM_ERR_INVALID, "shape[%d]=%lld outside [1, %d)", i, (long long)shape[i], MM_MAX_CUTOFF);
        S.s[i] = int32_t(shape[i]);
        if (n > (int64_t(1) << 40) / shape[i]) return fail(MM_ERR_INVALID, "lattice too large");
        n *= shape[i];
    }
    S.st[D - 1] = 1;
    for (int i = D - 1; i > 0; --i) S.st[i - 1] = S.st[i] * S.s[i];
    S.N = n;
    *out = S;
    return MM_OK;
}

Geom geom_of(const Shape &S) {
    Geom g;
    memset(&g, 0, sizeof(g));
    g.D = S.D;
    for (int i = 0; i < 8; ++i) {
        g.shape[i] = i < S.D ? S.s[i] : 1;
        g.stride[i] = i < S.D ? S.st[i] : 1;
    }
    g.N = S.N;
    return g;
}

// Job template for the sub-lattice swept in LDS from level `top` down.
Job make_job(const Shape &S, int top, int kind) {
    Job J;
    memset(&J, 0, sizeof(J));
    J.D = S.D;
    J.kind = kind;
    J.top = top;
    J.S = int32_t(S.st[top]);
    J.nacc = nacc_of(S.D);
    int smax = 1;
    for (int j = top; j < S.D; ++j) smax = std::max(smax, int(S.s[j]));
    J.smax = smax;
    for (int i = 0; i < 8; ++i) {
        J.shape[i] = i < S.D ? S.s[i] : 1;
        J.stride[i] = i < S.D ? S.st[i] : 1;
    }
    int bit = 0;
    for (int j = S.D - 1; j > top; --j) {
        int w = 0;
        while ((1 << w) < S.s[j]) ++w;
        J.kbit[j] = bit;
        J.kmask[j] = (1 << w) - 1;
        bit += w;
    }
    return J;
}

bool lds_kernel_dmax8(int D) { return D > 3; }

int kernel_dmax(int D) { return lds_kernel_dmax8(D) ? 8 : 3; }

// LDS bytes of one team sweeping S from level `top` in the kernel
// instantiation for `kdmax` (3 or 8) indices.
int64_t lds_bytes_rm(const Shape &S, int top, bool fused, int nwaves, int rmax) {
    int smax = 1;
    for (int j = top; j < S.D; ++j) smax = std::max(smax, int(S.s[j]));
    return lds_layout(S.D, top, int(S.st[top]), smax, fused ? 1 : 0, nwaves, rmax).total;
}

int64_t lds_bytes(const Shape &S, int top, bool fused, int nwaves, int kdmax) {
    return lds_bytes_rm(S, top, fused, nwaves, rmax_for(kdmax, fused));
}

bool packed_bits_ok(const Shape &S, int top) {
    int bits = 0;
    for (int j = S.D - 1; j > top; --j) {
        int w = 0;
        while ((1 << w) < S.s[j]) ++w;
        bits += w;
    }
    return bits <= 32;
}

int team_threads(int64_t slab, int maxt) {
    int64_t t = ((slab + 63) / 64) * 64;
    t = std::max<int64_t>(64, std::min<int64_t>(t, maxt));
    return int(t);
}

int lds_maxt(int D, bool fused) { return fused ? (lds_kernel_dmax8(D) ? 256 : 512) : (lds_kernel_dmax8(D) ? 512 : 768); }

// Threads for a team sweeping slabs of S points with at most `rmax` points
// per thread in registers: fewest thread-slots T * ceil(S/T), counting a
// team of fewer than 8 waves (2 per SIMD) as proportionally slower, ties to
// more threads.  0 if no T <= maxt holds the slab in registers.
int pick_threads(int64_t S, int maxt, int rmax) {
    int best = 0;
    double best_cost = 0;
    for (int T = MM_WAVE; T <= maxt; T += MM_WAVE) {
        const int64_t nr = (S + T - 1) / T;
        if (nr > rmax) continue;
        const double waves = T / MM_WAVE;
        const double cost = double(T * nr) * (waves < 8 ? 8.0 / waves : 1.0);
        if (best == 0 || cost < best_cost || (cost == best_cost && T > best)) {
            best = T;
            best_cost = cost;
        }
    }
    return best;
}

// Smallest top level whose sweep fits in LDS with `threads` threads (D >= 2).
int choose_top(const Shape &S, bool fused, int max_lds, int maxt, int *threads) {
    for (int top = 0; top < S.D; ++top) {
        int T = team_threads(S.st[top], maxt);
        if (top == 0) {
            const int Tr = pick_threads(S.st[0], maxt, rmax_for(kernel_dmax(S.D), fused));
            if (Tr) T = Tr;
        }
        if (S.st[top] > int64_t(8) * 1024 * 1024) continue;
        if (!packed_bits_ok(S, top)) continue;
        if (lds_bytes(S, top, fused, T / MM_WAVE, kernel_dmax(S.D)) <= max_lds) {
            *threads = T;
            return top;
        }
    }
    *threads = 64;
    return S.D - 1;
}

// Forward-only uniform batches of D <= 3 triples whose level-0 slab has
// LEAN_MIN_S .. LEAN_MAX_S points: one point per thread of a workgroup of up
// to 1024 threads, in lds_block_kernel (about 70 registers: four waves per
// SIMD).  The fused sweep keeps lds_kernel: its one-point-per-thread variant
// needs 128 registers with spills and measured slower than two points per
// thread at two waves per SIMD.
constexpr int LEAN_MAX_S = 1024;
constexpr int LEAN_MIN_S = 512;

// Launch lds_block_kernel for S if it applies; returns 1 if it does not.
int launch_lean(const LdsArgs &a0, const Shape &S, bool fused, int64_t B, const DevState *ds, hipStream_t s) {
    if (fused || S.D < 2 || S.D > 3 || S.st[0] < LEAN_MIN_S || S.st[0] > LEAN_MAX_S || !packed_bits_ok(S, 0)) return 1;
    const int T = pick_threads(S.st[0], 1024, 1);
    if (!T) return 1;
    const int64_t shmem = lds_bytes_rm(S, 0, fused, T / MM_WAVE, 1);
    if (shmem > ds->max_lds) return 1;
    LdsArgs a = a0;
    a.tmpl = make_job(S, 0, KIND_BLOCK);
    const int grid = int(std::min<int64_t>(B, int64_t(1) << 30));
    launch_lds_block(S.D, a, dim3(grid), dim3(T), size_t(shmem), s);
    LAUNCH_CHECK("lds_block_kernel");
    return MM_OK;
}

// Launch the LDS kernel instantiation that covers D / fused.
int launch_lds(const LdsArgs &a, int D, bool fused, int grid, int threads, int64_t shmem, hipStream_t s) {
    const dim3 g(grid), t(threads);
    if (lds_kernel_dmax8(D))
        (fused ? launch_lds_8f : launch_lds_8)(a, g, t, size_t(shmem), s);
    else
        (fused ? launch_lds_3f : launch_lds_3)(a, g, t, size_t(shmem), s);
    LAUNCH_CHECK("lds_kernel");
    return MM_OK;
}

// The sub-lattice {k_0..k_{top-1} = 0} of triple t (offsets in the uniform
// strides of S), swept in LDS by one workgroup.  With two or three indices
// left and room for the whole-lattice register path, it is swept as the
// lattice of the sub-triple (A[top:, top:], b[top:], c), the A rows read with
// the stride of the full triple: a point with k_<top = 0 takes its pivot and
// every nonzero term from the axes >= top, the sub-lattice's C-order strides
// are those of S, and so the values are the same, from the straight-line
// D <= 3 sweep.  A non-finite value is flagged through at.P.nf (or rerun
// in the sub-triple's arithmetic when the caller keeps no flag).
int launch_face(LdsArgs at, const Shape &S, int top, int T, int64_t t, const DevState *ds, hipStream_t s) {
    const int Ds = S.D - top;
    at.n = 1;
    if (top >= 1 && Ds >= 2 && Ds <= 3) {
        Shape Sub;
        Sub.D = Ds;
        for (int i = 0; i < Ds; ++i) {
            Sub.s[i] = S.s[top + i];
            Sub.st[i] = S.st[top + i];
        }
        Sub.N = S.st[top - 1];
        int Ts = 64;
        // 1024-2048-point slabs: two points per thread of up to 1024 threads
        // (lds_block_kernel), four waves per SIMD
        const int T2 = pick_threads(Sub.st[0], 1024, 2);
        if (Ds == 3 && Sub.st[0] > 1024 && T2 && packed_bits_ok(Sub, 0) &&
            lds_bytes_rm(Sub, 0, false, T2 / MM_WAVE, 2) <= ds->max_lds) {
            at.tmpl = make_job(Sub, 0, KIND_BLOCK);
            at.tmpl.lda = S.D;
            at.tmpl.a_off = t * S.D * S.D + int64_t(top) * S.D + top;
            at.tmpl.b_off = t * S.D + top;
            at.tmpl.c_off = t;
            at.tmpl.g_off = t * S.N;
            launch_lds_block2(at, dim3(1), dim3(T2), size_t(lds_bytes_rm(Sub, 0, false, T2 / MM_WAVE, 2)), s);
            LAUNCH_CHECK("lds_block_kernel");
            return MM_OK;
        }
        if (choose_top(Sub, false, ds->max_lds, lds_maxt(Ds, false), &Ts) == 0) {
            at.tmpl = make_job(Sub, 0, KIND_BLOCK);
            at.tmpl.lda = S.D;
            at.tmpl.a_off = t * S.D * S.D + int64_t(top) * S.D + top;
            at.tmpl.b_off = t * S.D + top;
            at.tmpl.c_off = t;
            at.tmpl.g_off = t * S.N;
            return launch_lds(at, Ds, false, 1, Ts, lds_bytes(Sub, 0, false, Ts / MM_WAVE, kernel_dmax(Ds)), s);
        }
    }
    at.tmpl = make_job(S, top, KIND_BLOCK);
    at.tmpl.a_off = t * S.D * S.D;
    at.tmpl.b_off = t * S.D;
    at.tmpl.c_off = t;
    at.tmpl.g_off = t * S.N;
    return launch_lds(at, S.D, false, 1, T, lds_bytes(S, top, false, T / MM_WAVE, kernel_dmax(S.D)), s);
}

Bases bases(const double *A, const double *b, const double *c, const double *g, double *G, double *dA, double *db, double *dc,
            double *partial) {
    Bases P;
    P.A = reinterpret_cast<const double2 *>(A);
    P.b = reinterpret_cast<const double2 *>(b);
    P.c = reinterpret_cast<const double2 *>(c);
    P.g = reinterpret_cast<const double2 *>(g);
    P.G = reinterpret_cast<double2 *>(G);
    P.dA = reinterpret_cast<double2 *>(dA);
    P.db = reinterpret_cast<double2 *>(db);
    P.dc = reinterpret_cast<double2 *>(dc);
    P.partial = reinterpret_cast<double2 *>(partial);
    P.nf = nullptr;
    return P;
}

// --------------------------------------------------------------- slab path

int launch_slab(const SlabArgs &a, int D, int64_t Sd, int num_cu, hipStream_t s) {
    int64_t blocks = (Sd + 255) / 256;
    int64_t cap = int64_t(num_cu) * 16;
    int grid = int(std::max<int64_t>(1, std::min(blocks, cap)));
    switch (D) {
#define CASE(DD) \
    case DD: hipLaunchKernelGGL((fwd_slab_kernel<DD>), dim3(grid), dim3(256), 0, s, a); break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
        default:
            return fail(MM_ERR_INVALID, "bad D");
    }
    LAUNCH_CHECK("fwd_slab_kernel");
    return MM_OK;
}

// Slabs of level d of at most 2^17 points, whose launch costs more than
// their work, go through fwd_slab_multi_kernel: the slab cut into one tile
// per CU, two slabs per launch when a block's first region (tile + halo)
// stays within 2048 points and two region buffers within the LDS, else one.
// Measured on the (40,)^4 density matrix (39 level-0 slabs of 64,000 points),
// the whole forward: one slab per launch of fwd_slab_kernel 246 us, of this
// kernel 219 us, two per launch 199 us; three per launch measured no faster
// than two (a block's redundant halo points cost about as much as the
// launches they save).
struct MultiPlan {
    int nphase = 0, tile = 0, cap = 0, nblk = 0, smax = 1;
    int64_t shmem = 0;
};

bool multi_plan(const Shape &S, int d, const DevState *ds, MultiPlan *p) {
    if (d + 1 >= S.D || S.s[d] < 2) return false;
    const int64_t Sd = S.st[d];
    if (Sd > (int64_t(1) << 17)) return false;
    const int64_t h = S.st[d + 1];
    int bits = 0, smax = 1;
    for (int j = d + 1; j < S.D; ++j) {
        int w = 0;
        while ((1 << w) < S.s[j]) ++w;
        bits += w;
        smax = std::max(smax, int(S.s[j]));
    }
    if (bits > 32) return false;
    const int64_t tile = std::max<int64_t>(64, (Sd + ds->num_cu - 1) / ds->num_cu);
    for (int T = 2; T >= 1; --T) {
        const int64_t cap = std::min<int64_t>(Sd, tile + (T - 1) * h);
        const int64_t shmem = 2 * cap * 16 + int64_t(S.D - 1 - d) * smax * 16;
        if (cap > int64_t(MULTI_THREADS) * MULTI_KMAX || shmem > ds->max_lds) continue;
        p->nphase = T;
        p->tile = int(tile);
        p->cap = int(cap);
        p->nblk = int((Sd + tile - 1) / tile);
        p->smax = smax;
        p->shmem = shmem;
        return true;
    }
    return false;
}

int launch_slab_multi(const SlabArgs &sa, const Shape &S, const MultiPlan &p, int d, int m0, int T, hipStream_t s) {
    SlabMultiArgs a;
    memset(&a, 0, sizeof(a));
    a.g = sa.g;
    a.A = sa.A;
    a.b = sa.b;
    a.G = sa.G;
    a.sq = sa.sq;
    a.rsq = sa.rsq;
    a.nf = sa.nf;
    a.d = d;
    a.m0 = m0;
    a.nphase = T;
    a.tile = p.tile;
    a.halo = int(S.st[d + 1]);
    a.cap = p.cap;
    a.smax = p.smax;
    int bit = 0;
    for (int j = S.D - 1; j > d; --j) {
        int w = 0;
        while ((1 << w) < S.s[j]) ++w;
        a.kbit[j] = bit;
        a.kmask[j] = (1 << w) - 1;
        bit += w;
    }
    const dim3 grid(unsigned(p.nblk)), block(MULTI_THREADS);
    switch (S.D) {
#define CASE(DD)                                                                                      \
    case DD:                                                                                          \
        raise_lds_limit(fwd_slab_multi_kernel<DD>, int(p.shmem));                                     \
        hipLaunchKernelGGL((fwd_slab_multi_kernel<DD>), grid, block, size_t(p.shmem), s, a);          \
        break;
        CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
        default:
            return fail(MM_ERR_INVALID, "bad D");
    }
    LAUNCH_CHECK("fwd_slab_multi_kernel");
    return MM_OK;
}

// Slabs m = 1 .. shape[d] - 1 of level d: several per launch where
// multi_plan allows, else one launch each.
int launch_level(SlabArgs sa, const Shape &S, int d, const DevState *ds, hipStream_t s) {
    MultiPlan p;
    sa.d = d;
    if (multi_plan(S, d, ds, &p)) {
        for (int m0 = 1; m0 < S.s[d]; m0 += p.nphase) {
            const int rc = launch_slab_multi(sa, S, p, d, m0, std::min(p.nphase, S.s[d] - m0), s);
            if (rc) return rc;
        }
        return MM_OK;
    }
    for (int m = 1; m < S.s[d]; ++m) {
        sa.m = m;
        const int rc = launch_slab(sa, S.D, S.st[d], ds->num_cu, s);
        if (rc) return rc;
    }
    return MM_OK;
}

// ------------------------------------------------------ non-finite fix-up

Lat lat_of(const Shape &S) {
    Lat g;
    memset(&g, 0, sizeof(g));
    g.D = S.D;
    for (int i = 0; i < 8; ++i) {
        g.shape[i] = i < S.D ? S.s[i] : 1;
        g.stride[i] = i < S.D ? S.st[i] : 1;
    }
    g.N = S.N;
    return g;
}

// Last launch of a multi-launch path over B uniform triples: reruns a triple
// in the reference's arithmetic when the path met a non-finite value (the
// forward's flag nf, or a non-finite dc when the VJP was computed).
int launch_fixup(const Shape &S, int64_t B, const double *A, const double *b, const double *c, const double *g, double *G, double *dA,
                 double *db, double *dc, const int *nf, bool fwd, bool vjp, const DevState *ds, hipStream_t s, bool stable = false) {
    if (B < 1) return MM_OK;
    if (B > 0x7fffffff) return fail(MM_ERR_UNSUPPORTED, "too many triples for one launch");
    FixArgs a;
    memset(&a, 0, sizeof(a));
    a.g = lat_of(S);
    a.A = reinterpret_cast<const double2 *>(A);
    a.b = reinterpret_cast<const double2 *>(b);
    a.c = reinterpret_cast<const double2 *>(c);
    a.gg = reinterpret_cast<const double2 *>(g);
    a.G = reinterpret_cast<double2 *>(G);
    a.dA = reinterpret_cast<double2 *>(dA);
    a.db = reinterpret_cast<double2 *>(db);
    a.dc = reinterpret_cast<double2 *>(dc);
    a.sq = ds->sq;
    a.ua = int64_t(S.D) * S.D;
    a.ub = S.D;
    a.uc = 1;
    a.ug = S.N;
    a.nf = nf;
    a.fwd = fwd;
    a.vjp = vjp;
    a.stable = stable;
    hipLaunchKernelGGL(exact_fixup_kernel, dim3(unsigned(B)), dim3(1024), 0, s, a);
    LAUNCH_CHECK("exact_fixup_kernel");
    return MM_OK;
}

// ------------------------------------------------------------------ VJP path

// One launch per pivot-row group (RowSplit) of the VJP kernels.
template <int D>
void launch_vjp_groups(const VjpArgs &a, dim3 grid, hipStream_t s) {
    using RS = RowSplit<D>;
    hipLaunchKernelGGL((vjp_kernel<D, 0, RS::B1>), grid, dim3(256), 0, s, a);
    if constexpr (RS::NG >= 2) hipLaunchKernelGGL((vjp_kernel<D, RS::B1, RS::B2>), grid, dim3(256), 0, s, a);
    if constexpr (RS::NG >= 3) hipLaunchKernelGGL((vjp_kernel<D, RS::B2, D>), grid, dim3(256), 0, s, a);
}

// Threads of the finalize block of each triple: one wave per sum column for
// few triples, four waves when many triples fill the GPU anyway.
int finalize_threads(int64_t B) { return B >= 64 ? 256 : 1024; }

int vjp_chunks(int64_t N, int64_t B) {
    int64_t per = 256 * 8;
    int64_t want = (N + per - 1) / per;
    int64_t cap = std::max<int64_t>(1, 8192 / std::max<int64_t>(B, 1));
    return int(std::max<int64_t>(1, std::min<int64_t>(want, std::min<int64_t>(cap, 4096))));
}

size_t vjp_ws_bytes(const Shape &S, int64_t B) {
    return size_t(B) * vjp_chunks(S.N, B) * nacc_of(S.D) * 16 + 256;
}

// VJP of one stored lattice slab by slab (all row groups of a slab in one
// launch, so each slab is fetched about once rather than once per group):
// taken for D >= 5 (two or three row groups) when the slabs are large.
bool vjp_by_slabs(const Shape &S, int64_t B) { return B == 1 && S.D >= 5 && S.st[0] >= (int64_t(1) << 16); }

int launch_vjp_slabs(const Shape &S, const double *G, const double *g, const double *c, double *dA, double *db, double *dc, void *ws,
                     const DevState *ds, hipStream_t s);

int launch_vjp_sums(const Shape &S, int64_t B, const double *G, const double *g, const double *c, double *dA, double *db, double *dc,
                    void *ws, size_t ws_bytes, const DevState *ds, hipStream_t s);

// The stored lattice's VJP, then the non-finite fix-up (fix = false when the
// caller launches its own).
int launch_vjp(const Shape &S, int64_t B, const double *G, const double *g, const double *c, double *dA, double *db, double *dc,
               void *ws, size_t ws_bytes, const DevState *ds, hipStream_t s, bool fix = true) {
    int rc = launch_vjp_sums(S, B, G, g, c, dA, db, dc, ws, ws_bytes, ds, s);
    if (rc || !fix) return rc;
    return launch_fixup(S, B, nullptr, nullptr, c, g, const_cast<double *>(G), dA, db, dc, nullptr, false, true, ds, s);
}

int launch_vjp_sums(const Shape &S, int64_t B, const double *G, const double *g, const double *c, double *dA, double *db, double *dc,
                    void *ws, size_t ws_bytes, const DevState *ds, hipStream_t s) {
    if (!ws || ws_bytes < vjp_ws_bytes(S, B))
        return fail(MM_ERR_WORKSPACE, "VJP workspace of %zu bytes needed, %zu given", vjp_ws_bytes(S, B), ws_bytes);
    if (vjp_by_slabs(S, B)) return launch_vjp_slabs(S, G, g, c, dA, db, dc, ws, ds, s);
    VjpArgs a;
    a.g = geom_of(S);
    a.G = reinterpret_cast<const double2 *>(G);
    a.gg = reinterpret_cast<const double2 *>(g);
    a.partial = reinterpret_cast<double2 *>(ws);
    a.sq = ds->sq;
    a.nchunks = vjp_chunks(S.N, B);
    dim3 grid(a.nchunks, unsigned(B));
    const double2 *c2 = reinterpret_cast<const double2 *>(c);
    double2 *dA2 = reinterpret_cast<double2 *>(dA), *db2 = reinterpret_cast<double2 *>(db), *dc2 = reinterpret_cast<double2 *>(dc);
    switch (S.D) {
#define CASE(DD)                                                                                               \
    case DD:                                                                                                   \
        launch_vjp_groups<DD>(a, grid, s);                                                                     \
        LAUNCH_CHECK("vjp_kernel");                                                                            \
        hipLaunchKernelGGL((vjp_finalize<DD>), dim3(unsigned(B)), dim3(finalize_threads(B)), 0, s, a.partial, a.nchunks, c2, dA2, db2, dc2); \
        LAUNCH_CHECK("vjp_finalize");                                                                          \
        break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
        default:
            return fail(MM_ERR_INVALID, "bad D");
    }
    return MM_OK;
}

// One launch of slab_fused_kernel for level-0 slab m: the VJP terms of slab m
// (cur, p1, p2 = slabs m, m-1, m-2) added into nblk partial rows, and, when
// nxt is not NULL, the forward of slab m + 1 into nxt.
template <int D>
void launch_slab_fused_t(const SlabVjpArgs &v, const Slab0Args &f, bool fwd, int nblk, hipStream_t s) {
    constexpr int NG = RowSplit<D>::NG;
    if (fwd)
        hipLaunchKernelGGL((slab_fused_kernel<D, true>), dim3(unsigned(nblk * (NG + 1))), dim3(256), 0, s, v, f);
    else
        hipLaunchKernelGGL((slab_fused_kernel<D, false>), dim3(unsigned(nblk * NG)), dim3(256), 0, s, v, f);
}

int launch_slab_fused(const Shape &S, const double *A, const double *b, const double2 *cur, const double2 *p1, const double2 *p2,
                      const double2 *gg, double2 *nxt, double2 *partial, int nblk, int m, const DevState *ds, hipStream_t s) {
    SlabVjpArgs v;
    v.g = geom_of(S);
    v.cur = cur;
    v.p1 = p1;
    v.p2 = p2;
    v.gg = gg;
    v.partial = partial;
    v.sq = ds->sq;
    v.m = m;
    Slab0Args f;
    f.g = v.g;
    f.A = reinterpret_cast<const double2 *>(A);
    f.b = reinterpret_cast<const double2 *>(b);
    f.p1 = cur;
    f.p2 = p1;
    f.cur = nxt;
    f.sq = ds->sq;
    f.rsq = ds->rsq;
    f.m = m + 1;
    switch (S.D) {
#define CASE(DD) \
    case DD: launch_slab_fused_t<DD>(v, f, nxt != nullptr, nblk, s); break;
        CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
        default:
            return fail(MM_ERR_INVALID, "bad D");
    }
    LAUNCH_CHECK("slab_fused_kernel");
    return MM_OK;
}

// Blocks per row group of a slab_fused_kernel launch over slabs of S0 points.
int slab_blocks(int64_t S0, const DevState *ds) {
    return int(std::max<int64_t>(1, std::min<int64_t>((S0 + 1023) / 1024, int64_t(ds->num_cu) * 8)));
}

// Sum nblk partial rows of one triple and finish (symmetrise, dc = sum / c).
int launch_finalize(int D, const double2 *partial, int nblk, const double *c, double *dA, double *db, double *dc, hipStream_t s) {
    const double2 *c2 = reinterpret_cast<const double2 *>(c);
    double2 *dA2 = reinterpret_cast<double2 *>(dA), *db2 = reinterpret_cast<double2 *>(db), *dc2 = reinterpret_cast<double2 *>(dc);
    switch (D) {
#define CASE(DD) \
    case DD: hipLaunchKernelGGL((vjp_finalize<DD>), dim3(1), dim3(finalize_threads(1)), 0, s, partial, nblk, c2, dA2, db2, dc2); break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
        default:
            return fail(MM_ERR_INVALID, "bad D");
    }
    LAUNCH_CHECK("vjp_finalize");
    return MM_OK;
}

// -------------------------------------------------------------- stable path

// Rows (of the last axis) listed for the one-workgroup sweeps of a large
// lattice's first and last wavefronts: one per thread of a 1024-thread workgroup.
constexpr int64_t STABLE_CORNER_ROWS = 1024;

int64_t stable_list_cap(const Shape &S) { return std::min<int64_t>(S.N / S.s[S.D - 1], STABLE_CORNER_ROWS); }

// The row table (8 digits per row), the counters of two row lists, the lists.
size_t stable_ws_bytes(const Shape &S) {
    return align256(size_t(S.N / S.s[S.D - 1]) * 16) + 256 + 2 * align256(size_t(stable_list_cap(S)) * 4);
}

int stable_check(const Shape &S) {
    for (int j = 0; j + 1 < S.D; ++j)
        if (S.s[j] > 65535) return fail(MM_ERR_UNSUPPORTED, "stable: shape[%d]=%d > 65535", j, S.s[j]);
    return MM_OK;
}

// Number of rows whose digit sum over axes 0..D-2 is p, for p = 0..sum(shape[j] - 1).
std::vector<int64_t> row_sum_counts(const Shape &S) {
    std::vector<int64_t> cnt(1, 1);
    for (int j = 0; j + 1 < S.D; ++j) {
        const int64_t sj = S.s[j], m = int64_t(cnt.size());
        std::vector<int64_t> pre(m + 1, 0);
        for (int64_t q = 0; q < m; ++q) pre[q + 1] = pre[q] + cnt[q];
        std::vector<int64_t> nxt(m + sj - 1, 0);
        for (int64_t p = 0; p < int64_t(nxt.size()); ++p) {
            const int64_t hi = std::min(p, m - 1), lo = std::max<int64_t>(0, p - sj + 1);
            nxt[p] = lo <= hi ? pre[hi + 1] - pre[lo] : 0;
        }
        cnt.swap(nxt);
    }
    return cnt;
}

template <int D>
void launch_stable_corner(const StableArgs &a, const int32_t *rows, int count, int n0, int n1, hipStream_t s) {
    hipLaunchKernelGGL((stable_corner_kernel<D>), dim3(1), dim3(unsigned(STABLE_CORNER_ROWS)), 0, s, a, rows, count, n0, n1);
}

int launch_stable(const Shape &S, int64_t B, const double *A, const double *b, const double *c, double *G, void *ws, size_t ws_bytes,
                  const DevState *ds, hipStream_t s) {
    int rc = stable_check(S);
    if (rc) return rc;
    if (!ws || ws_bytes < stable_ws_bytes(S))
        return fail(MM_ERR_WORKSPACE, "stable workspace of %zu bytes needed, %zu given", stable_ws_bytes(S), ws_bytes);
    StableArgs a;
    memset(&a, 0, sizeof(a));
    a.g = geom_of(S);
    a.nrows = S.N / S.s[S.D - 1];
    char *base = reinterpret_cast<char *>(ws);
    a.rows = reinterpret_cast<uint16_t *>(base);
    base += align256(size_t(a.nrows) * 16);
    hipLaunchKernelGGL(row_table_kernel, dim3(unsigned((a.nrows + 255) / 256)), dim3(256), 0, s, a.g, const_cast<uint16_t *>(a.rows),
                       a.nrows);
    LAUNCH_CHECK("row_table_kernel");
    for (int j = 0; j < S.D; ++j) a.rstride[j] = 1.0 / double(S.st[j]);
    a.sq = ds->sq;
    a.rsq = ds->rsq;
    a.tstride_a = int64_t(S.D) * S.D;
    a.tstride_b = S.D;
    a.tstride_g = S.N;
    int nmax = 0;
    for (int j = 0; j < S.D; ++j) nmax += S.s[j] - 1;
    // Per wavefront, one workgroup per triple loops over ceil(nrows / 1024)
    // rows per thread, ~1.6 us per row and pivot axis beyond the first (one
    // round per num_cu triples); a launch over all CUs costs ~5.5 us per triple.
    const int64_t rounds = (B + ds->num_cu - 1) / ds->num_cu;
    const double t_block = double((a.nrows + 1023) / 1024) * double(rounds) * 1.6 * std::max(1, S.D - 1);
    if (t_block <= 5.5 * double(B)) {
        a.A = reinterpret_cast<const double2 *>(A);
        a.b = reinterpret_cast<const double2 *>(b);
        a.c = reinterpret_cast<const double2 *>(c);
        a.G = reinterpret_cast<double2 *>(G);
        switch (S.D) {
#define CASE(DD) \
    case DD: hipLaunchKernelGGL((stable_block_kernel<DD>), dim3(unsigned(B)), dim3(1024), 0, s, a); break;
            CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
        }
        LAUNCH_CHECK("stable_block_kernel");
        return MM_OK;
    }

    // One launch per wavefront, except, for D = 2, the first wavefronts
    // n = 1..nlo (rows with digit sum p <= plo) and the last ones n = nhi..nmax
    // (rows with p >= phi), each run by one workgroup when their rows fit it.
    // A point reads D (D + 1) neighbours through one CU's vector-memory pipe:
    // measured, a workgroup step (one point per thread) costs ~1.5 us at D = 2,
    // ~4.6 us at D = 3 (a launch in one-wave blocks: ~4.1 us) and ~7.3 us at D = 4.
    int *nf = nullptr;
    rc = nf_flag(s, &nf);
    if (rc) return rc;
    a.nf = nf;
    const int slast = S.s[S.D - 1];
    int nlo = 0, nhi = nmax + 1, plo = -1, phi = INT32_MAX;
    int64_t nlo_rows = 0, nhi_rows = 0;
    if (a.nrows <= INT32_MAX && S.D == 2) {
        const std::vector<int64_t> cnt = row_sum_counts(S);
        const int P = int(cnt.size()) - 1;
        while (plo < P && nlo_rows + cnt[plo + 1] <= STABLE_CORNER_ROWS) nlo_rows += cnt[++plo];
        phi = P + 1;
        while (phi > 0 && nhi_rows + cnt[phi - 1] <= STABLE_CORNER_ROWS) nhi_rows += cnt[--phi];
        if (plo == P) {  // every row listed: the first sweep covers every wavefront
            nlo = nmax;
        } else {
            nlo = plo;
            nhi = std::max(phi + slast - 1, nlo + 1);
        }
    }
    int *counts = reinterpret_cast<int *>(base);
    int32_t *lo_rows = reinterpret_cast<int32_t *>(base + 256);
    int32_t *hi_rows = reinterpret_cast<int32_t *>(base + 256 + align256(size_t(stable_list_cap(S)) * 4));
    if (nlo >= 1 || nhi <= nmax) {
        hipLaunchKernelGGL(zero_c128_kernel, dim3(1), dim3(64), 0, s, reinterpret_cast<double2 *>(counts), int64_t(1));
        LAUNCH_CHECK("zero_c128_kernel");
        switch (S.D) {
#define CASE(DD)                                                                                                                     \
    case DD:                                                                                                                         \
        hipLaunchKernelGGL((stable_corner_rows_kernel<DD>), dim3(unsigned((a.nrows + 255) / 256)), dim3(256), 0, s, a, plo,        \
                           nhi <= nmax ? phi : INT32_MAX, lo_rows, hi_rows, counts, int(stable_list_cap(S)));                        \
        break;
            CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
        }
        LAUNCH_CHECK("stable_corner_rows_kernel");
    }
    // one-wave blocks while the rows fill fewer than 256 threads per CU: measured
    // faster there ((40,)^4, 64,000 rows: 0.86 against 0.98 ms; (70,70,60): 0.83
    // against 0.89 ms), slower for larger row counts ((20,)^5: 1.32 against 1.22 ms)
    const int tpb = a.nrows < int64_t(ds->num_cu) * 256 ? 64 : 256;
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((a.nrows + tpb - 1) / tpb, int64_t(ds->num_cu) * 16 * (256 / tpb)));
    for (int64_t t = 0; t < B; ++t) {
        a.A = reinterpret_cast<const double2 *>(A) + t * a.tstride_a;
        a.b = reinterpret_cast<const double2 *>(b) + t * a.tstride_b;
        a.c = reinterpret_cast<const double2 *>(c) + t;
        a.G = reinterpret_cast<double2 *>(G) + t * a.tstride_g;
        hipLaunchKernelGGL(set_first_kernel, dim3(1), dim3(1), 0, s, a.G, a.c, t == 0 ? nf : nullptr);
        LAUNCH_CHECK("set_first_kernel");
        if (nlo >= 1) {
            switch (S.D) {
#define CASE(DD) \
    case DD: launch_stable_corner<DD>(a, lo_rows, int(nlo_rows), 1, nlo, s); break;
                CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
            }
            LAUNCH_CHECK("stable_corner_kernel");
        }
        for (int n = nlo + 1; n < nhi; ++n) {
            a.n = n;
            switch (S.D) {
#define CASE(DD) \
    case DD: hipLaunchKernelGGL((stable_wavefront_kernel<DD>), dim3(unsigned(blocks)), dim3(unsigned(tpb)), 0, s, a); break;
                CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
            }
            LAUNCH_CHECK("stable_wavefront_kernel");
        }
        if (nhi <= nmax) {
            switch (S.D) {
#define CASE(DD) \
    case DD: launch_stable_corner<DD>(a, hi_rows, int(nhi_rows), nhi, nmax, s); break;
                CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
            }
            LAUNCH_CHECK("stable_corner_kernel");
        }
    }
    // non-finite values: the reference's sequential sweep of each flagged call's triples
    return launch_fixup(S, B, A, b, c, nullptr, G, nullptr, nullptr, nullptr, nf, true, false, ds, s, true);
}

// The stored lattice's VJP, one slab_fused_kernel launch per level-0 slab
// (no forward role).  The workspace holds at least slab_blocks rows.
int launch_vjp_slabs(const Shape &S, const double *G, const double *g, const double *c, double *dA, double *db, double *dc, void *ws,
                     const DevState *ds, hipStream_t s) {
    const int64_t S0 = S.st[0];
    const int nblk = std::min(slab_blocks(S0, ds), vjp_chunks(S.N, 1));
    double2 *partial = reinterpret_cast<double2 *>(ws);
    hipLaunchKernelGGL(zero_c128_kernel, dim3(64), dim3(256), 0, s, partial, int64_t(nblk) * nacc_of(S.D));
    LAUNCH_CHECK("zero_c128_kernel");
    const double2 *G2 = reinterpret_cast<const double2 *>(G), *g2 = reinterpret_cast<const double2 *>(g);
    for (int m = 0; m < S.s[0]; ++m) {
        const double2 *cur = G2 + int64_t(m) * S0;
        const int rc = launch_slab_fused(S, nullptr, nullptr, cur, m >= 1 ? cur - S0 : cur, m >= 2 ? cur - 2 * S0 : cur, g2 + int64_t(m) * S0,
                                         nullptr, partial, nblk, m, ds, s);
        if (rc) return rc;
    }
    return launch_finalize(S.D, partial, nblk, c, dA, db, dc, s);
}

// ------------------------------------------------------------ vanilla paths

// Workspace of one vanilla forward (+ fused VJP when `fused`).
// A single triple (D >= 3) whose level-0 slab fits in LDS but holds more
// than 3072 points (four or more per thread of the workgroup; with the VJP,
// more than 1024) is swept over all CUs: the face {k_0 = 0} in LDS (team size
// *Tf), then the level-0 slabs.  Measured for the forward: a workgroup step
// costs ~1.7 us at three points per thread and ~3.4 us at four, a slab over
// all CUs ~2.3 us (two per launch); a fused workgroup step adds the VJP terms.
bool spread_single(const Shape &S, bool fused, int64_t B, const DevState *ds, int *Tf) {
    if (B != 1 || S.D < 3 || S.st[0] <= (fused ? 1024 : 3072) || !packed_bits_ok(S, 1)) return false;
    int T = 64;
    if (choose_top(S, fused, ds->max_lds, lds_maxt(S.D, fused), &T) != 0) return false;
    *Tf = team_threads(S.st[1], lds_maxt(S.D, false));
    return lds_bytes(S, 1, false, *Tf / MM_WAVE, kernel_dmax(S.D)) <= ds->max_lds;
}

size_t fwd_ws_bytes(const Shape &S, bool fused, const DevState *ds) {
    if (!fused || S.D == 1) return 0;
    int T = 64, Tf = 64;
    const int top = choose_top(S, true, ds->max_lds, lds_maxt(S.D, true), &T);
    if (top == 0 && !spread_single(S, true, 1, ds, &Tf)) return 0;
    return vjp_ws_bytes(S, 1);
}

int window_slab0(const Shape &S, const double *A, const double *b, double2 *cur, double2 *p1, double2 *p2, int m,
                 const DevState *ds, hipStream_t s);

// Forward of B triples sharing shape S (uniform strides); fused adds the VJP.
int vanilla_uniform(const Shape &S, int64_t B, const double *A, const double *b, const double *c, const double *g, double *G,
                    double *dA, double *db, double *dc, void *ws, size_t ws_bytes, const DevState *ds, hipStream_t s) {
    const bool fused = g != nullptr;
    LdsArgs a;
    memset(&a, 0, sizeof(a));
    a.P = bases(A, b, c, g, G, dA, db, dc, nullptr);
    a.T.sq = ds->sq;
    a.T.rsq = ds->rsq;
    a.mode = MODE_UNIFORM;
    a.ua = int64_t(S.D) * S.D;
    a.ub = S.D;
    a.uc = 1;
    a.ug = S.N;
    a.n = B;
    if (S.D == 1) {
        a.tmpl = make_job(S, 0, KIND_LANE);
        int grid = int((B + 63) / 64);
        return launch_lds(a, 1, fused, grid, 64, 0, s);
    }
    {
        const int rc = launch_lean(a, S, fused, B, ds, s);
        if (rc <= 0) return rc;
    }
    int T = 64;
    const int maxt = lds_maxt(S.D, fused);
    int top = choose_top(S, fused, ds->max_lds, maxt, &T);
    // a single triple with large level-0 slabs over all CUs (spread_single);
    // with the VJP only when the caller gave the VJP workspace for it
    int Tf = 64;
    const bool spread = top == 0 && spread_single(S, fused, B, ds, &Tf) && (!fused || (ws && ws_bytes >= vjp_ws_bytes(S, 1)));
    if (top == 0 && !spread) {
        a.tmpl = make_job(S, 0, KIND_BLOCK);
        int64_t shmem = lds_bytes(S, 0, fused, T / MM_WAVE, kernel_dmax(S.D));
        int grid = int(std::min<int64_t>(B, int64_t(1) << 30));
        return launch_lds(a, S.D, fused, grid, T, shmem, s);
    }
    // Large lattice: LDS sweep of the sub-lattice {k_<top = 0}, then one launch
    // per slab for the levels above, then the VJP reduction.
    if (spread) {
        top = 1;
        T = Tf;
    } else {
        top = choose_top(S, false, ds->max_lds, lds_maxt(S.D, false), &T);
    }
    if (fused && (!ws || ws_bytes < vjp_ws_bytes(S, B)))
        return fail(MM_ERR_WORKSPACE, "fused workspace of %zu bytes needed, %zu given", vjp_ws_bytes(S, B), ws_bytes);
    int *nf = nullptr;
    {
        const int rc = nf_flag(s, &nf);
        if (rc) return rc;
    }
    for (int64_t t = 0; t < B; ++t) {
        LdsArgs at = a;
        at.P.g = nullptr;
        at.P.nf = nf;
        at.clear_nf = t == 0;  // the first launch of the path (one workgroup)
        int rc = launch_face(at, S, top, T, t, ds, s);
        if (rc) return rc;
        SlabArgs sa;
        sa.g = geom_of(S);
        sa.A = reinterpret_cast<const double2 *>(A) + t * a.ua;
        sa.b = reinterpret_cast<const double2 *>(b) + t * a.ub;
        sa.G = reinterpret_cast<double2 *>(G) + t * a.ug;
        sa.sq = ds->sq;
        sa.rsq = ds->rsq;
        sa.nf = nf;
        // with the VJP of a single triple above 512 MB, level 0 alternates the
        // VJP terms of slab m (every row group in one launch) with the forward
        // of slab m + 1, so that each slab is read while it is still in the
        // caches; below that, the forward (two slabs per launch) then one VJP
        // pass take fewer launches and measured faster ((60,)^4, 207 MB: 0.82
        // against 1.12 ms; (10,)^8, 1.6 GB: 6.76 ms interleaved against 6.9)
        const bool fuse_level0 = fused && B == 1 && S.N * 16 > (int64_t(512) << 20);
        for (int d = top - 1; d >= (fuse_level0 ? 1 : 0); --d) {
            rc = launch_level(sa, S, d, ds, s);
            if (rc) return rc;
        }
        if (fuse_level0) {
            const int64_t S0 = S.st[0];
            const int nblk = std::min(slab_blocks(S0, ds), vjp_chunks(S.N, 1));
            double2 *partial = reinterpret_cast<double2 *>(ws);
            hipLaunchKernelGGL(zero_c128_kernel, dim3(64), dim3(256), 0, s, partial, int64_t(nblk) * nacc_of(S.D));
            LAUNCH_CHECK("zero_c128_kernel");
            double2 *G2 = sa.G;
            const double2 *g2 = reinterpret_cast<const double2 *>(g) + t * a.ug;
            for (int m = 0; m < S.s[0]; ++m) {
                double2 *cur = G2 + int64_t(m) * S0;
                rc = launch_slab_fused(S, A + 2 * t * a.ua, b + 2 * t * a.ub, cur, m >= 1 ? cur - S0 : cur, m >= 2 ? cur - 2 * S0 : cur,
                                       g2 + int64_t(m) * S0, nullptr, partial, nblk, m, ds, s);
                if (rc) return rc;
                if (m + 1 < S.s[0]) {
                    rc = window_slab0(S, A + 2 * t * a.ua, b + 2 * t * a.ub, cur + S0, cur, m >= 1 ? cur - S0 : cur, m + 1, ds, s);
                    if (rc) return rc;
                }
            }
            rc = launch_finalize(S.D, partial, nblk, c + 2 * t, dA + 2 * t * a.ua, db + 2 * t * a.ub, dc + 2 * t, s);
            if (rc) return rc;
            return launch_fixup(S, 1, A + 2 * t * a.ua, b + 2 * t * a.ub, c + 2 * t, g + 2 * t * a.ug, G + 2 * t * a.ug, dA + 2 * t * a.ua,
                                db + 2 * t * a.ub, dc + 2 * t, nf, true, true, ds, s);
        }
    }
    if (fused) {
        const int rc = launch_vjp(S, B, G, g, c, dA, db, dc, ws, ws_bytes, ds, s, false);
        if (rc) return rc;
    }
    return launch_fixup(S, B, A, b, c, g, G, dA, db, dc, nf, true, fused, ds, s);
}

// ------------------------------------------------------- recompute VJP path
//
// dL/dA, dL/db, dL/dc from (A, b, c, dL/dG) without the whole of G in
// memory: the face k_0 = 0 is computed into window slot 0, then each level-0
// slab m is recomputed into slot m % 3 from slots (m-1) % 3 and (m-2) % 3 and
// its VJP terms are added to per-block partial sums.  Resident memory is
// three slabs (3 * stride[0] * 16 B) instead of the lattice.  Lattices of at
// most 64 MB, and those whose slabs fit in LDS, take the fused sweep with a
// full G in the workspace (recompute_whole).

int recompute_blocks(const Shape &S, const DevState *ds) { return slab_blocks(S.st[0], ds); }


// Lattices held whole in the workspace instead of the window: those whose
// fused slabs fit in LDS, and any of at most 64 MB (the window's slab-by-slab
// launches cost more than the memory saves there).  *vjp_ws: the fused path
// then also needs the VJP workspace, after G.
bool recompute_whole(const Shape &S, const DevState *ds, bool *vjp_ws) {
    int T = 64, Tf = 64;
    const int top = (S.D == 1) ? 0 : choose_top(S, true, ds->max_lds, lds_maxt(S.D, true), &T);
    const bool whole = S.D == 1 || top == 0 || S.N * 16 <= (int64_t(64) << 20);
    *vjp_ws = whole && S.D > 1 && (top != 0 || spread_single(S, true, 1, ds, &Tf));
    return whole;
}

size_t recompute_ws_bytes(const Shape &S, const DevState *ds) {
    bool vjp_ws = false;
    if (recompute_whole(S, ds, &vjp_ws)) return align256(size_t(S.N) * 16) + (vjp_ws ? vjp_ws_bytes(S, 1) : 0);
    return 3 * align256(size_t(S.st[0]) * 16) + size_t(recompute_blocks(S, ds)) * nacc_of(S.D) * 16 + 256;
}

// The face {k_0 = 0} of the lattice into `face` (stride[0] entries): the LDS
// sweep from a level `top` >= 1 (it writes only flat indices below
// stride[top-1] <= stride[0]), then slab launches for levels top-1 .. 1.
int window_face(const Shape &S, const double *A, const double *b, const double *c, double2 *face, const DevState *ds,
                hipStream_t s) {
    int T = 64;
    int top = choose_top(S, false, ds->max_lds, lds_maxt(S.D, false), &T);
    if (top == 0) {
        top = 1;
        T = team_threads(S.st[1], lds_maxt(S.D, false));
        if (lds_bytes(S, 1, false, T / MM_WAVE, kernel_dmax(S.D)) > ds->max_lds || !packed_bits_ok(S, 1))
            return fail(MM_ERR_UNSUPPORTED, "slab window: face does not fit in LDS");
    }
    LdsArgs a;
    memset(&a, 0, sizeof(a));
    a.P = bases(A, b, c, nullptr, reinterpret_cast<double *>(face), nullptr, nullptr, nullptr, nullptr);
    a.T.sq = ds->sq;
    a.T.rsq = ds->rsq;
    a.mode = MODE_UNIFORM;
    int rc = launch_face(a, S, top, T, 0, ds, s);
    if (rc) return rc;
    SlabArgs sa;
    sa.g = geom_of(S);
    sa.A = reinterpret_cast<const double2 *>(A);
    sa.b = reinterpret_cast<const double2 *>(b);
    sa.G = face;
    sa.sq = ds->sq;
    sa.rsq = ds->rsq;
    sa.nf = nullptr;
    for (int d = top - 1; d >= 1; --d) {
        rc = launch_level(sa, S, d, ds, s);
        if (rc) return rc;
    }
    return MM_OK;
}

// Level-0 slab m (m >= 1) into cur from the two slabs below it.
int window_slab0(const Shape &S, const double *A, const double *b, double2 *cur, double2 *p1, double2 *p2, int m,
                 const DevState *ds, hipStream_t s) {
    const int64_t S0 = S.st[0];
    const int64_t fblocks = std::max<int64_t>(1, std::min<int64_t>((S0 + 255) / 256, int64_t(ds->num_cu) * 16));
    Slab0Args f;
    f.g = geom_of(S);
    f.A = reinterpret_cast<const double2 *>(A);
    f.b = reinterpret_cast<const double2 *>(b);
    f.p1 = p1;
    f.p2 = p2;
    f.cur = cur;
    f.sq = ds->sq;
    f.rsq = ds->rsq;
    f.m = m;
    switch (S.D) {
#define CASE(DD) \
    case DD: hipLaunchKernelGGL((fwd_slab0_kernel<DD>), dim3(unsigned(fblocks)), dim3(256), 0, s, f); break;
        CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
        default:
            return fail(MM_ERR_INVALID, "bad D");
    }
    LAUNCH_CHECK("fwd_slab0_kernel");
    return MM_OK;
}

int vjp_recompute(const Shape &S, const double *A, const double *b, const double *c, const double *g, double *dA, double *db,
                  double *dc, void *ws, size_t ws_bytes, const DevState *ds, hipStream_t s) {
    const size_t need = recompute_ws_bytes(S, ds);
    if (!ws || ws_bytes < need) return fail(MM_ERR_WORKSPACE, "recompute workspace of %zu bytes needed, %zu given", need, ws_bytes);
    bool vjp_ws = false;
    if (recompute_whole(S, ds, &vjp_ws)) {
        const size_t gb = align256(size_t(S.N) * 16);
        return vanilla_uniform(S, 1, A, b, c, g, reinterpret_cast<double *>(ws), dA, db, dc, reinterpret_cast<char *>(ws) + gb,
                               ws_bytes - gb, ds, s);
    }

    // three slabs: slab m + 1 is computed into the slot of slab m - 2 once the
    // VJP terms of slab m (the last to read slab m - 2) are launched
    const size_t slab_bytes = align256(size_t(S.st[0]) * 16);
    char *base = reinterpret_cast<char *>(ws);
    double2 *slot[3];
    for (int i = 0; i < 3; ++i) slot[i] = reinterpret_cast<double2 *>(base + i * slab_bytes);
    double2 *partial = reinterpret_cast<double2 *>(base + 3 * slab_bytes);
    const int nblk = recompute_blocks(S, ds);
    const int NA = nacc_of(S.D);
    hipLaunchKernelGGL(zero_c128_kernel, dim3(64), dim3(256), 0, s, partial, int64_t(nblk) * NA);
    LAUNCH_CHECK("zero_c128_kernel");

    // 1. the face {k_0 = 0} into slot 0
    int rc = window_face(S, A, b, c, slot[0], ds, s);
    if (rc) return rc;
    // 2. level 0: the VJP terms of slab m (every row group in one launch), then slab m + 1
    const int64_t S0 = S.st[0];
    for (int m = 0; m < S.s[0]; ++m) {
        double2 *cur = slot[m % 3];
        rc = launch_slab_fused(S, A, b, cur, m >= 1 ? slot[(m + 2) % 3] : cur, m >= 2 ? slot[(m + 1) % 3] : cur,
                               reinterpret_cast<const double2 *>(g) + int64_t(m) * S0, nullptr, partial, nblk, m, ds, s);
        if (rc) return rc;
        if (m + 1 < S.s[0]) {
            rc = window_slab0(S, A, b, slot[(m + 1) % 3], cur, m >= 1 ? slot[(m + 2) % 3] : cur, m + 1, ds, s);
            if (rc) return rc;
        }
    }
    // 3. sum the block partials and finish (symmetrise, dc = sum / c)
    return launch_finalize(S.D, partial, nblk, c, dA, db, dc, s);
}

// ------------------------------------------------------------------- plans

// ------------------------------------------------------------ binomial path

int binom_W(const Shape &S) {
    int64_t w = 1;
    for (int i = 0; i < S.D; ++i) w += S.s[i] - 1;
    return int(std::min<int64_t>(w, int64_t(1) << 30));
}

int binom_blocks(const Shape &S) {
    const int64_t per = 256 * 16;
    return int(std::max<int64_t>(1, std::min<int64_t>((S.N + per - 1) / per, 1024)));
}

size_t binom_ws_bytes(const Shape &S) { return size_t(binom_blocks(S)) * size_t(binom_W(S)) * 8 + 256; }

template <int D>
void launch_binomial_kernels(const Shape &S, const double2 *c, double max_l2, int64_t global_cutoff, double2 *G, double *norm_out,
                             void *ws, int num_cu, hipStream_t s) {
    const Geom g = geom_of(S);
    const int W = binom_W(S), nblk = binom_blocks(S);
    const int64_t chunk = (S.N + nblk - 1) / nblk;
    double *part = reinterpret_cast<double *>(ws);
    int *wstop = reinterpret_cast<int *>(reinterpret_cast<char *>(ws) + size_t(nblk) * W * 8);
    hipLaunchKernelGGL((weight_norm_kernel<D>), dim3(nblk), dim3(256), size_t(W) * 8, s, G, g, W, chunk, part);
    hipLaunchKernelGGL(binomial_stop_kernel, dim3(1), dim3(256), size_t(W) * 8, s, part, nblk, W, c, max_l2, global_cutoff, wstop, norm_out);
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((S.N + 255) / 256, int64_t(num_cu) * 16));
    hipLaunchKernelGGL((weight_mask_kernel<D>), dim3(unsigned(grid)), dim3(256), 0, s, G, g, wstop);
}

// ------------------------------------------- conversion fused with the contraction

int launch_rows_dot(const double2 *G, const double2 *w, double2 *out, int64_t rows, int64_t K, hipStream_t s) {
    if (rows <= 0) return MM_OK;
    if (K >= 64) {
        const int64_t blocks = (rows + 3) / 4;  // four waves (rows) per 256-thread block
        if (blocks > 0x7fffffff) return fail(MM_ERR_UNSUPPORTED, "too many rows for one launch");
        hipLaunchKernelGGL(rows_dot_wave_kernel, dim3(unsigned(blocks)), dim3(256), 0, s, G, w, out, rows, K);
    } else {
        const int64_t blocks = (rows + 255) / 256;
        if (blocks > 0x7fffffff) return fail(MM_ERR_UNSUPPORTED, "too many rows for one launch");
        hipLaunchKernelGGL(rows_dot_lane_kernel, dim3(unsigned(blocks)), dim3(256), 0, s, G, w, out, rows, K);
    }
    LAUNCH_CHECK("rows_dot kernel");
    return MM_OK;
}

// true when the lattice goes through the three-slab window (else whole, in the workspace)
bool contract_windowed(const Shape &S, const DevState *ds) {
    if (S.D == 1) return false;
    int T = 64;
    return choose_top(S, false, ds->max_lds, lds_maxt(S.D, false), &T) != 0;
}

size_t contract_ws_bytes(const Shape &S, const DevState *ds) {
    if (!contract_windowed(S, ds)) return align256(size_t(S.N) * 16);
    return 3 * align256(size_t(S.st[0]) * 16);
}

struct GlobalJob {
    Shape S;
    int64_t a_off, b_off, c_off, g_off;
};

}  // namespace

struct mm_plan {
    int device = 0;
    bool fused = false;
    int dmax = 3;
    int threads = 256;
    int grid = 1;
    int64_t shmem = 0;
    int wave_lds = 0;
    int64_t nb = 0, nw = 0, nl = 0, n = 0;
    Job *d_jobs = nullptr;
    int *d_counters = nullptr;
    double *d_partial = nullptr;
    size_t partial_bytes = 0;
    std::vector<GlobalJob> global;
    int64_t n_amps = 0, a_elems = 0, b_elems = 0;
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void *gptr[9] = {nullptr};
};

namespace {

int plan_launch(const mm_plan *p, const double *A, const double *b, const double *c, const double *g, double *G, double *dA,
                double *db, double *dc, hipStream_t s) {
    DevState *ds = nullptr;
    int rc = device_state(&ds);
    if (rc) return rc;
    if (p->nb + p->nw + p->nl > 0) {
        hipLaunchKernelGGL(zero_ints_kernel, dim3(1), dim3(64), 0, s, p->d_counters, 3);
        LAUNCH_CHECK("zero_ints_kernel");
        LdsArgs a;
        memset(&a, 0, sizeof(a));
        a.P = bases(A, b, c, g, G, dA, db, dc, nullptr);
        a.T.sq = ds->sq;
        a.T.rsq = ds->rsq;
        a.mode = MODE_QUEUE;
        a.wave_lds = p->wave_lds;
        a.table = p->d_jobs;
        a.nb = p->nb;
        a.nw = p->nw;
        a.nl = p->nl;
        a.counters = p->d_counters;
        rc = launch_lds(a, p->dmax, p->fused, p->grid, p->threads, p->shmem, s);
        if (rc) return rc;
    }
    for (const GlobalJob &gj : p->global) {
        rc = vanilla_uniform(gj.S, 1, A + 2 * gj.a_off, b + 2 * gj.b_off, c + 2 * gj.c_off, g ? g + 2 * gj.g_off : nullptr,
                             G + 2 * gj.g_off, dA ? dA + 2 * gj.a_off : nullptr, db ? db + 2 * gj.b_off : nullptr,
                             dc ? dc + 2 * gj.c_off : nullptr, p->d_partial, p->partial_bytes, ds, s);
        if (rc) return rc;
    }
    return MM_OK;
}

}  // namespace

// ====================================================================== C ABI

extern "C" {

const char *mm_last_error(void) { return g_err.c_str(); }

const char *mm_version(void) { return "mmhip 0.1 (gfx950)"; }

int mm_init(int device) {
    int cur = 0;
    HIP_TRY(hipGetDevice(&cur));
    if (device != cur) HIP_TRY(hipSetDevice(device));
    DevState *ds = nullptr;
    int rc = device_state(&ds);
    if (rc == MM_OK) rc = scratch_init();
    if (rc == MM_OK) rc = nf_init();
    if (device != cur) HIP_TRY(hipSetDevice(cur));
    return rc;
}

size_t mm_hermite_fwd_workspace(const int64_t *shape, int D, int stable) {
    Shape S;
    if (make_shape(shape, D, &S)) return 0;
    return stable ? stable_ws_bytes(S) : 0;
}

int mm_hermite_fwd(const double *A, const double *b, const double *c, const int64_t *shape, int D, int stable, double *G,
                   void *workspace, size_t ws_bytes, void *stream) {
    Shape S;
    int rc = make_shape(shape, D, &S);
    if (rc) return rc;
    if (!A || !b || !c || !G) return fail(MM_ERR_INVALID, "NULL array pointer");
    DevState *ds = nullptr;
    rc = device_state(&ds);
    if (rc) return rc;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (stable) return launch_stable(S, 1, A, b, c, G, workspace, ws_bytes, ds, s);
    return vanilla_uniform(S, 1, A, b, c, nullptr, G, nullptr, nullptr, nullptr, workspace, ws_bytes, ds, s);
}

size_t mm_hermite_batch_fwd_workspace(int64_t B, const int64_t *shape, int D, int stable) {
    (void)B;
    return mm_hermite_fwd_workspace(shape, D, stable);
}

int mm_hermite_batch_fwd(const double *A, const double *b, const double *c, int64_t B, const int64_t *shape, int D, int stable,
                         double *G, void *workspace, size_t ws_bytes, void *stream) {
    Shape S;
    int rc = make_shape(shape, D, &S);
    if (rc) return rc;
    if (B < 0) return fail(MM_ERR_INVALID, "negative batch size");
    if (B == 0) return MM_OK;
    if (!A || !b || !c || !G) return fail(MM_ERR_INVALID, "NULL array pointer");
    DevState *ds = nullptr;
    rc = device_state(&ds);
    if (rc) return rc;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (stable) return launch_stable(S, B, A, b, c, G, workspace, ws_bytes, ds, s);
    return vanilla_uniform(S, B, A, b, c, nullptr, G, nullptr, nullptr, nullptr, workspace, ws_bytes, ds, s);
}

size_t mm_hermite_vjp_workspace(const int64_t *shape, int D) {
    Shape S;
    if (make_shape(shape, D, &S)) return 0;
    return vjp_ws_bytes(S, 1);
}

int mm_hermite_vjp(const double *G, const double *c, const double *dLdG, const int64_t *shape, int D, double *dA, double *db,
                   double *dc, void *workspace, size_t ws_bytes, void *stream) {
    return mm_hermite_batch_vjp(G, c, dLdG, 1, shape, D, dA, db, dc, workspace, ws_bytes, stream);
}

size_t mm_hermite_batch_vjp_workspace(int64_t B, const int64_t *shape, int D) {
    Shape S;
    if (make_shape(shape, D, &S) || B < 1) return 0;
    return vjp_ws_bytes(S, B);
}

int mm_hermite_batch_vjp(const double *G, const double *c, const double *dLdG, int64_t B, const int64_t *shape, int D, double *dA,
                         double *db, double *dc, void *workspace, size_t ws_bytes, void *stream) {
    Shape S;
    int rc = make_shape(shape, D, &S);
    if (rc) return rc;
    if (B < 0) return fail(MM_ERR_INVALID, "negative batch size");
    if (B == 0) return MM_OK;
    if (B > 65535) return fail(MM_ERR_UNSUPPORTED, "batch of %lld > 65535 triples", (long long)B);
    if (!G || !c || !dLdG || !dA || !db || !dc) return fail(MM_ERR_INVALID, "NULL array pointer");
    DevState *ds = nullptr;
    rc = device_state(&ds);
    if (rc) return rc;
    return launch_vjp(S, B, G, dLdG, c, dA, db, dc, workspace, ws_bytes, ds, reinterpret_cast<hipStream_t>(stream));
}

int mm_hermite_batch_vjp_static(const double *G, const double *c, const double *dLdG, int64_t B, const int64_t *shape, int D,
                                double *dA, double *db, double *dc, void *stream) {
    Shape S;
    int rc = make_shape(shape, D, &S);
    if (rc) return rc;
    if (B < 0) return fail(MM_ERR_INVALID, "negative batch size");
    if (B == 0) return MM_OK;
    if (B > 65535) return fail(MM_ERR_UNSUPPORTED, "batch of %lld > 65535 triples", (long long)B);
    if (!G || !c || !dLdG || !dA || !db || !dc) return fail(MM_ERR_INVALID, "NULL array pointer");
    const size_t need = vjp_ws_bytes(S, B);
    if (need > size_t(MM_STATIC_SCRATCH_BYTES))
        return fail(MM_ERR_UNSUPPORTED, "this VJP needs %zu bytes of scratch, above the static buffer's %zu; use mm_hermite_batch_vjp",
                    need, size_t(MM_STATIC_SCRATCH_BYTES));
    void *ws = nullptr;
    rc = scratch_address(&ws);
    if (rc) return rc;
    // the device state exists: mm_init created it with the scratch address
    DevState *ds = nullptr;
    rc = device_state(&ds);
    if (rc) return rc;
    return launch_vjp(S, B, G, dLdG, c, dA, db, dc, ws, size_t(MM_STATIC_SCRATCH_BYTES), ds, reinterpret_cast<hipStream_t>(stream));
}

size_t mm_hermite_fused_workspace(const int64_t *shape, int D) {
    Shape S;
    if (make_shape(shape, D, &S)) return 0;
    DevState *ds = nullptr;
    if (device_state(&ds)) return 0;
    return fwd_ws_bytes(S, true, ds);
}

int mm_hermite_fused(const double *A, const double *b, const double *c, const double *dLdG, const int64_t *shape, int D, double *G,
                     double *dA, double *db, double *dc, void *workspace, size_t ws_bytes, void *stream) {
    return mm_hermite_batch_fused(A, b, c, dLdG, 1, shape, D, G, dA, db, dc, workspace, ws_bytes, stream);
}

size_t mm_hermite_batch_fused_workspace(int64_t B, const int64_t *shape, int D) {
    Shape S;
    if (make_shape(shape, D, &S) || B < 1) return 0;
    DevState *ds = nullptr;
    if (device_state(&ds)) return 0;
    return fwd_ws_bytes(S, true, ds) ? vjp_ws_bytes(S, B) : 0;
}

int mm_hermite_batch_fused(const double *A, const double *b, const double *c, const double *dLdG, int64_t B, const int64_t *shape,
                           int D, double *G, double *dA, double *db, double *dc, void *workspace, size_t ws_bytes, void *stream) {
    Shape S;
    int rc = make_shape(shape, D, &S);
    if (rc) return rc;
    if (B < 0) return fail(MM_ERR_INVALID, "negative batch size");
    if (B == 0) return MM_OK;
    if (!A || !b || !c || !dLdG || !G || !dA || !db || !dc) return fail(MM_ERR_INVALID, "NULL array pointer");
    DevState *ds = nullptr;
    rc = device_state(&ds);
    if (rc) return rc;
    return vanilla_uniform(S, B, A, b, c, dLdG, G, dA, db, dc, workspace, ws_bytes, ds, reinterpret_cast<hipStream_t>(stream));
}

int mm_plan_create(int64_t n, const int32_t *dims, const int64_t *shapes, int fused, const mm_tune *tune, mm_plan **out) {
    if (!out) return fail(MM_ERR_INVALID, "plan output pointer is NULL");
    *out = nullptr;
    if (n < 0) return fail(MM_ERR_INVALID, "negative number of triples");
    if (n > 0 && (!dims || !shapes)) return fail(MM_ERR_INVALID, "NULL dims/shapes");
    DevState *ds = nullptr;
    int rc = device_state(&ds);
    if (rc) return rc;
    std::vector<Shape> S(static_cast<size_t>(n), Shape());
    int dmax = 1;
    for (int64_t k = 0; k < n; ++k) {
        rc = make_shape(shapes + 8 * k, dims[k], &S[size_t(k)]);
        if (rc) return fail(rc, "triple %lld: %s", (long long)k, g_err.c_str());
        dmax = std::max(dmax, int(dims[k]));
    }
    mm_tune tn;
    memset(&tn, 0, sizeof(tn));
    if (tune) tn = *tune;
    const bool fz = fused != 0;
    const int kd = lds_kernel_dmax8(dmax) ? 8 : 3;
    const int maxt = lds_maxt(kd, fz);
    int T = tn.block_threads > 0 ? tn.block_threads : 256;
    if (T % MM_WAVE || T < MM_WAVE || T > maxt) return fail(MM_ERR_INVALID, "block_threads=%d not a multiple of 64 in [64, %d]", T, maxt);
    const int wpb = T / MM_WAVE;
    int wave_max = tn.wave_max_slab == 0 ? 64 : tn.wave_max_slab;

    auto p = new mm_plan();
    HIP_TRY(hipGetDevice(&p->device));
    p->fused = fz;
    p->dmax = kd;
    p->threads = T;
    p->n = n;
    std::vector<Job> blk, wav, lan;
    std::vector<int64_t> cblk, cwav, clan;
    int64_t a_off = 0, b_off = 0, g_off = 0;
    int64_t block_lds = 0, wave_lds = 0;
    for (int64_t k = 0; k < n; ++k) {
        const Shape &Sk = S[size_t(k)];
        const int D = Sk.D;
        int64_t steps = 1;
        for (int j = 0; j < D; ++j) steps += Sk.s[j] - 1;
        const int64_t cost = Sk.N + 64 * steps;
        auto with_offsets = [&](Job J) {
            J.a_off = a_off;
            J.b_off = b_off;
            J.c_off = k;
            J.g_off = g_off;
            return J;
        };
        if (D == 1) {
            lan.push_back(with_offsets(make_job(Sk, 0, KIND_LANE)));
            clan.push_back(cost);
        } else {
            const int64_t S0 = Sk.st[0];
            const int64_t wl = packed_bits_ok(Sk, 0) ? lds_bytes(Sk, 0, fz, 1, kd) : INT64_MAX;
            const int64_t bl = packed_bits_ok(Sk, 0) ? lds_bytes(Sk, 0, fz, wpb, kd) : INT64_MAX;
            if (wave_max > 0 && S0 <= wave_max && wl * wpb <= ds->max_lds) {
                wav.push_back(with_offsets(make_job(Sk, 0, KIND_WAVE)));
                cwav.push_back(cost);
                wave_lds = std::max(wave_lds, wl);
            } else if (bl <= ds->max_lds && S0 <= int64_t(T) * 64) {
                blk.push_back(with_offsets(make_job(Sk, 0, KIND_BLOCK)));
                cblk.push_back(cost);
                block_lds = std::max(block_lds, bl);
            } else {
                GlobalJob gj;
                gj.S = Sk;
                gj.a_off = a_off;
                gj.b_off = b_off;
                gj.c_off = k;
                gj.g_off = g_off;
                p->global.push_back(gj);
                if (fz) p->partial_bytes = std::max(p->partial_bytes, vjp_ws_bytes(Sk, 1));
            }
        }
        a_off += int64_t(D) * D;
        b_off += D;
        g_off += Sk.N;
    }
    p->n_amps = g_off;
    p->a_elems = a_off;
    p->b_elems = b_off;
    // largest first (longest-processing-time order) within each class
    auto sort_by_cost = [](std::vector<Job> &v, std::vector<int64_t> &c) {
        std::vector<size_t> idx(v.size());
        for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
        std::stable_sort(idx.begin(), idx.end(), [&](size_t x, size_t y) { return c[x] > c[y]; });
        std::vector<Job> out(v.size());
        for (size_t i = 0; i < idx.size(); ++i) out[i] = v[idx[i]];
        v.swap(out);
    };
    sort_by_cost(blk, cblk);
    sort_by_cost(wav, cwav);
    sort_by_cost(lan, clan);
    p->nb = int64_t(blk.size());
    p->nw = int64_t(wav.size());
    p->nl = int64_t(lan.size());
    p->wave_lds = int(align16(wave_lds));
    p->shmem = std::max<int64_t>(block_lds, p->nw ? int64_t(p->wave_lds) * wpb : 0);
    int bpc = tn.blocks_per_cu > 0 ? tn.blocks_per_cu
                                   : int(std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(4, 2048 / T),
                                                                                 p->shmem ? ds->max_lds / p->shmem : 4)));
    const int64_t needed = p->nb + (p->nw + wpb - 1) / wpb + (p->nl + T - 1) / T;
    p->grid = int(std::max<int64_t>(1, std::min<int64_t>(int64_t(ds->num_cu) * bpc, needed)));
    std::vector<Job> all;
    all.insert(all.end(), blk.begin(), blk.end());
    all.insert(all.end(), wav.begin(), wav.end());
    all.insert(all.end(), lan.begin(), lan.end());
    if (!all.empty()) {
        HIP_TRY(hipMalloc(&p->d_jobs, sizeof(Job) * all.size()));
        HIP_TRY(hipMemcpy(p->d_jobs, all.data(), sizeof(Job) * all.size(), hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMalloc(&p->d_counters, 64));
    if (p->partial_bytes) HIP_TRY(hipMalloc(&p->d_partial, p->partial_bytes));
    HIP_TRY(hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking));
    *out = p;
    return MM_OK;
}

int mm_plan_destroy(mm_plan *p) {
    if (!p) return MM_OK;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
    if (p->d_jobs) (void)hipFree(p->d_jobs);
    if (p->d_counters) (void)hipFree(p->d_counters);
    if (p->d_partial) (void)hipFree(p->d_partial);
    delete p;
    return MM_OK;
}

int mm_plan_info(const mm_plan *p, int64_t *n_amplitudes, int64_t *a_elems, int64_t *b_elems) {
    if (!p) return fail(MM_ERR_INVALID, "NULL plan");
    if (n_amplitudes) *n_amplitudes = p->n_amps;
    if (a_elems) *a_elems = p->a_elems;
    if (b_elems) *b_elems = p->b_elems;
    return MM_OK;
}

int mm_plan_describe(const mm_plan *p, int64_t *out /* 8 entries */) {
    if (!p || !out) return fail(MM_ERR_INVALID, "NULL plan or output");
    out[0] = p->nb;
    out[1] = p->nw;
    out[2] = p->nl;
    out[3] = int64_t(p->global.size());
    out[4] = p->grid;
    out[5] = p->threads;
    out[6] = p->shmem;
    out[7] = p->wave_lds;
    return MM_OK;
}

static int plan_args_ok(const mm_plan *p, const double *A, const double *b, const double *c, const double *g, double *G, double *dA,
                        double *db, double *dc) {
    if (!p) return fail(MM_ERR_INVALID, "NULL plan");
    if (p->n == 0) return MM_OK;
    if (!A || !b || !c || !G) return fail(MM_ERR_INVALID, "NULL array pointer");
    if (p->fused && (!g || !dA || !db || !dc)) return fail(MM_ERR_INVALID, "fused plan needs dLdG, dA, db, dc");
    if (!p->fused && (g || dA || db || dc)) return fail(MM_ERR_INVALID, "forward plan takes no dLdG/dA/db/dc");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != p->device) return fail(MM_ERR_INVALID, "plan was created on device %d, current device is %d", p->device, dev);
    return MM_OK;
}

int mm_plan_run(const mm_plan *p, const double *A, const double *b, const double *c, const double *dLdG, double *G, double *dA,
                double *db, double *dc, void *stream) {
    int rc = plan_args_ok(p, A, b, c, dLdG, G, dA, db, dc);
    if (rc || p->n == 0) return rc;
    return plan_launch(p, A, b, c, dLdG, G, dA, db, dc, reinterpret_cast<hipStream_t>(stream));
}

int mm_plan_run_graph(mm_plan *p, const double *A, const double *b, const double *c, const double *dLdG, double *G, double *dA,
                      double *db, double *dc, void *stream) {
    int rc = plan_args_ok(p, A, b, c, dLdG, G, dA, db, dc);
    if (rc || p->n == 0) return rc;
    const void *ptrs[9] = {A, b, c, dLdG, G, dA, db, dc, nullptr};
    bool same = p->exec != nullptr;
    for (int i = 0; i < 9 && same; ++i) same = ptrs[i] == p->gptr[i];
    if (!same) {
        if (p->exec) (void)hipGraphExecDestroy(p->exec);
        if (p->graph) (void)hipGraphDestroy(p->graph);
        p->exec = nullptr;
        p->graph = nullptr;
        HIP_TRY(hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal));
        rc = plan_launch(p, A, b, c, dLdG, G, dA, db, dc, p->cap_stream);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(p->cap_stream, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return fail(MM_ERR_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
        p->graph = graph;
        HIP_TRY(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
        for (int i = 0; i < 9; ++i) p->gptr[i] = ptrs[i];
    }
    HIP_TRY(hipGraphLaunch(p->exec, reinterpret_cast<hipStream_t>(stream)));
    return MM_OK;
}

size_t mm_hermite_vjp_recompute_workspace(const int64_t *shape, int D) {
    Shape S;
    if (make_shape(shape, D, &S)) return 0;
    DevState *ds = nullptr;
    if (device_state(&ds)) return 0;
    return recompute_ws_bytes(S, ds);
}

int mm_hermite_vjp_recompute(const double *A, const double *b, const double *c, const double *dLdG, const int64_t *shape, int D,
                             double *dA, double *db, double *dc, void *workspace, size_t ws_bytes, void *stream) {
    Shape S;
    int rc = make_shape(shape, D, &S);
    if (rc) return rc;
    if (!A || !b || !c || !dLdG || !dA || !db || !dc) return fail(MM_ERR_INVALID, "NULL array pointer");
    DevState *ds = nullptr;
    rc = device_state(&ds);
    if (rc) return rc;
    return vjp_recompute(S, A, b, c, dLdG, dA, db, dc, workspace, ws_bytes, ds, reinterpret_cast<hipStream_t>(stream));
}

size_t mm_hermite_binomial_workspace(const int64_t *shape, int D) {
    Shape S;
    if (make_shape(shape, D, &S)) return 0;
    return binom_ws_bytes(S);
}

int mm_hermite_binomial(const double *A, const double *b, const double *c, const int64_t *shape, int D, double max_l2,
                        int64_t global_cutoff, double *G, double *norm, void *workspace, size_t ws_bytes, void *stream) {
    Shape S;
    int rc = make_shape(shape, D, &S);
    if (rc) return rc;
    if (!A || !b || !c || !G) return fail(MM_ERR_INVALID, "NULL array pointer");
    if (global_cutoff < 1) return fail(MM_ERR_INVALID, "global_cutoff must be >= 1, got %lld", (long long)global_cutoff);
    if (binom_W(S) > BINOM_MAX_W) return fail(MM_ERR_UNSUPPORTED, "more than %d weights (sum of shape - D + 1)", BINOM_MAX_W);
    if (!workspace || ws_bytes < binom_ws_bytes(S))
        return fail(MM_ERR_WORKSPACE, "binomial workspace of %zu bytes needed, %zu given", binom_ws_bytes(S), ws_bytes);
    DevState *ds = nullptr;
    rc = device_state(&ds);
    if (rc) return rc;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    rc = vanilla_uniform(S, 1, A, b, c, nullptr, G, nullptr, nullptr, nullptr, nullptr, 0, ds, s);
    if (rc) return rc;
    const double2 *c2 = reinterpret_cast<const double2 *>(c);
    double2 *G2 = reinterpret_cast<double2 *>(G);
    switch (S.D) {
#define CASE(DD) \
    case DD: launch_binomial_kernels<DD>(S, c2, max_l2, global_cutoff, G2, norm, workspace, ds->num_cu, s); break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
        default:
            return fail(MM_ERR_INVALID, "bad D");
    }
    LAUNCH_CHECK("binomial kernels");
    return MM_OK;
}


size_t mm_hermite_fwd_contract_workspace(const int64_t *shape, int D) {
    Shape S;
    if (make_shape(shape, D, &S)) return 0;
    DevState *ds = nullptr;
    if (device_state(&ds)) return 0;
    return contract_ws_bytes(S, ds);
}

int mm_hermite_fwd_contract(const double *A, const double *b, const double *c, const int64_t *shape, int D, int n_head,
                            const double *w, double *out, void *workspace, size_t ws_bytes, void *stream) {
    return mm_hermite_batch_fwd_contract(A, b, c, 1, shape, D, n_head, w, 1, 0, out, workspace, ws_bytes, stream);
}

size_t mm_hermite_batch_fwd_contract_workspace(const int64_t *shape, int D) { return mm_hermite_fwd_contract_workspace(shape, D); }

int mm_hermite_batch_fwd_contract(const double *A, const double *b, const double *c, int64_t B, const int64_t *shape, int D, int n_head,
                                  const double *W, int64_t M, int64_t sW, double *out, void *workspace, size_t ws_bytes, void *stream) {
    Shape S;
    int rc = make_shape(shape, D, &S);
    if (rc) return rc;
    if (n_head < 1 || n_head > D) return fail(MM_ERR_INVALID, "n_head=%d outside [1, %d]", n_head, D);
    if (B < 0 || M < 1 || sW < 0) return fail(MM_ERR_INVALID, "B=%lld, M=%lld, sW=%lld: need B >= 0, M >= 1, sW >= 0", (long long)B,
                                              (long long)M, (long long)sW);
    if (B == 0) return MM_OK;
    if (!A || !b || !c || !W || !out) return fail(MM_ERR_INVALID, "NULL array pointer");
    DevState *ds = nullptr;
    rc = device_state(&ds);
    if (rc) return rc;
    const size_t need = contract_ws_bytes(S, ds);
    if (!workspace || ws_bytes < need)
        return fail(MM_ERR_WORKSPACE, "fused contraction workspace of %zu bytes needed, %zu given", need, ws_bytes);
    int64_t K = 1;
    for (int j = n_head; j < D; ++j) K *= shape[j];
    const int64_t rows = S.N / K;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    // rows x K amplitudes times the K x M matrix of triple t into rows x M outputs
    auto contract_rows = [&](const double2 *Gr, int64_t nrows, const double *Wt, double *ot) -> int {
        if (M == 1) return launch_rows_dot(Gr, reinterpret_cast<const double2 *>(Wt), reinterpret_cast<double2 *>(ot), nrows, K, s);
        return mm_zgemm(reinterpret_cast<const double *>(Gr), K, Wt, M, ot, M, nrows, M, K, s);
    };
    const bool windowed = contract_windowed(S, ds);
    const int64_t S0 = S.st[0], R = S0 / K;  // rows of one level-0 slab: n_head >= 1, so K divides stride[0]
    const size_t slab_bytes = align256(size_t(S0) * 16);
    char *base = reinterpret_cast<char *>(workspace);
    double2 *slot[3] = {reinterpret_cast<double2 *>(base), reinterpret_cast<double2 *>(base + slab_bytes),
                        reinterpret_cast<double2 *>(base + 2 * slab_bytes)};
    for (int64_t t = 0; t < B; ++t) {
        const double *At = A + 2 * t * S.D * S.D, *bt = b + 2 * t * S.D, *ct = c + 2 * t;
        const double *Wt = W + 2 * t * sW;
        double *ot = out + 2 * t * rows * M;
        if (!windowed) {
            double2 *G = reinterpret_cast<double2 *>(workspace);
            rc = vanilla_uniform(S, 1, At, bt, ct, nullptr, reinterpret_cast<double *>(G), nullptr, nullptr, nullptr, nullptr, 0, ds, s);
            if (rc) return rc;
            rc = contract_rows(G, rows, Wt, ot);
            if (rc) return rc;
            continue;
        }
        rc = window_face(S, At, bt, ct, slot[0], ds, s);
        if (rc) return rc;
        for (int m = 0; m < S.s[0]; ++m) {
            if (m >= 1) {
                rc = window_slab0(S, At, bt, slot[m % 3], slot[(m - 1) % 3], m >= 2 ? slot[(m - 2) % 3] : slot[(m - 1) % 3], m, ds, s);
                if (rc) return rc;
            }
            rc = contract_rows(slot[m % 3], R, Wt, ot + 2 * int64_t(m) * R * M);
            if (rc) return rc;
        }
    }
    return MM_OK;
}

}  // extern "C"
