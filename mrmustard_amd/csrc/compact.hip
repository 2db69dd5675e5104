// libmmhip: C ABI of the photon-number-resolved strategies (diagonal and
// one leftover mode, forward and VJP): geometry, workspaces, launches.

#include "host.hpp"
#include "diagonal.hpp"
#include "diagonal_vjp.hpp"
#include "leftover.hpp"
#include "leftover_vjp.hpp"

using namespace mm;

namespace {

// ------------------------------------------------------------ diagonal path

struct DiagGeom {
    DiagArgs a;
    int M = 0;
    int64_t n2 = 0, n11 = 0;  // entries of the d >= 1 arr2 and arr1010 (each), per batch column
    size_t t_bytes = 0, bytes = 0;
};

int diag_geom(const int64_t *cutoffs, int M, int64_t nb, DiagGeom *out) {
    if (!cutoffs) return fail(MM_ERR_INVALID, "cutoffs pointer is NULL");
    if (M < 1 || M > DG_MAXM) return fail(MM_ERR_INVALID, "diagonal needs 1 <= M <= %d modes, got %d", DG_MAXM, M);
    if (nb < 1) return fail(MM_ERR_INVALID, "batch of %lld columns", (long long)nb);
    DiagGeom g;
    memset(&g.a, 0, sizeof(g.a));
    g.M = M;
    int64_t P = 1, S = 0;
    for (int j = 0; j < M; ++j) {
        if (cutoffs[j] < 1 || cutoffs[j] + 2 >= MM_MAX_CUTOFF)
            return fail(MM_ERR_INVALID, "cutoffs[%d]=%lld outside [1, %d)", j, (long long)cutoffs[j], MM_MAX_CUTOFF - 2);
        g.a.cut[j] = int(cutoffs[j]);
        if (P > (int64_t(1) << 40) / (cutoffs[j] * nb)) return fail(MM_ERR_INVALID, "lattice too large");
        P *= cutoffs[j];
        S += cutoffs[j] - 1;
    }
    if (P > 0x7fffffff) return fail(MM_ERR_UNSUPPORTED, "diagonal: prod(cutoffs) = %lld is not below 2^31", (long long)P);
    g.a.P = P;
    g.a.nb = nb;
    g.a.S = int(S);
    int64_t len = 1;
    for (int j = M - 1; j >= 0; --j) {
        g.a.stride[j] = len;
        len *= g.a.cut[j];
        g.a.len[j] = len;
    }
    for (int d = 1; d < M; ++d) {  // sub-lattices of the off-diagonal pivots d >= 1
        g.a.off2[d] = g.n2;
        g.a.off11[d] = g.n11;
        g.n2 += g.a.len[d];
        g.n11 += int64_t(M - d - 1) * g.a.len[d];
    }
    g.t_bytes = align256(size_t(2 * (M + 1)) * size_t(S + 1) * 8) + align256(size_t(P) * 4) +
                align256(size_t(M) * size_t(P) * 4);  // T and Q, list, nbr
    g.bytes = g.t_bytes + size_t(16) * size_t(nb) * size_t(4 * M * P + g.n2 + 2 * g.n11);  // two inboxes + sub-lattices
    *out = g;
    return MM_OK;
}

// T[0][w] for every weight, as the device table holds it (prefix sums per mode)
std::vector<int64_t> diag_weight_counts(const DiagArgs &a, int M) {
    const int W = a.S + 1;
    std::vector<int64_t> row(W, 0), pre(W);
    row[0] = 1;
    for (int j = M - 1; j >= 0; --j) {
        int64_t acc = 0;
        for (int s = 0; s < W; ++s) pre[s] = (acc += row[s]);
        for (int s = 0; s < W; ++s) row[s] = pre[s] - (s >= a.cut[j] ? pre[s - a.cut[j]] : 0);
    }
    return row;
}

// Bytes of LDS for the count and prefix tables in each rank block, or 0 when
// they are too large and come from diag_counts_kernel in global memory.
size_t diag_table_lds(const DiagArgs &a, int M) {
    const size_t b = size_t(2 * (M + 1)) * size_t(a.S + 1) * 8;
    return b <= 48 * 1024 ? b : 0;
}

template <int M, bool REORDER>
int launch_diag_weights(const DiagArgs &a, const std::vector<int64_t> &counts, int num_cu, hipStream_t s) {
    constexpr int PB = DG_THREADS / (2 * M);
    const size_t lds = diag_table_lds(a, M);
    const int64_t rgrid = std::max<int64_t>(1, std::min<int64_t>((a.P + 255) / 256, int64_t(num_cu) * (lds ? 4 : 16)));
    hipLaunchKernelGGL((diag_rank_kernel<M>), dim3(unsigned(rgrid)), dim3(256), lds, s, a, lds ? 1 : 0);
    int64_t start = 0;
    for (int w = 0; w <= a.S; ++w) {
        const int64_t n = counts[w] * a.nb;
        const int64_t blocks = (n + PB - 1) / PB;
        if (blocks > 0x7fffffff) return fail(MM_ERR_UNSUPPORTED, "weight %d has too many params for one launch", w);
        if (blocks > 0)
            hipLaunchKernelGGL((diag_group_kernel<M, REORDER>), dim3(unsigned(blocks)), dim3(DG_THREADS), 0, s, a, w, counts[w], start);
        start += counts[w];
    }
    return MM_OK;
}

// Device pointers of the forward's workspace: T and Q, list, fwd, inboxes, d >= 1 sub-lattices.
void diag_bind(DiagGeom &g, const double *A, const double *B, const double *c, int c_batched, double *G, char *ws,
               const DevState *ds) {
    DiagArgs &a = g.a;
    const int M = g.M;
    const int64_t nb = a.nb;
    a.A = reinterpret_cast<const double2 *>(A);
    a.B = reinterpret_cast<const double2 *>(B);
    a.c = reinterpret_cast<const double2 *>(c);
    a.c_batched = c_batched ? 1 : 0;
    a.sq = ds->sq;
    a.rsq = ds->rsq;
    a.T = reinterpret_cast<int64_t *>(ws);
    a.Q = a.T + int64_t(M + 1) * (a.S + 1);
    a.list = reinterpret_cast<int32_t *>(ws + align256(size_t(2 * (M + 1)) * size_t(a.S + 1) * 8));
    a.fwd = reinterpret_cast<int32_t *>(reinterpret_cast<char *>(a.list) + align256(size_t(a.P) * 4));
    a.arr0 = reinterpret_cast<double2 *>(G);
    a.inD = reinterpret_cast<double2 *>(ws + g.t_bytes);
    a.inO = a.inD + 2 * M * a.P * nb;
    a.arr2 = a.inO + 2 * M * a.P * nb;
    a.a1010 = a.arr2 + g.n2 * nb;
    a.a1001 = a.a1010 + g.n11 * nb;
}

int diag_forward(DiagGeom &g, int reordered, const DevState *ds, hipStream_t s, std::vector<int64_t> *counts_out) {
    DiagArgs &a = g.a;
    const int M = g.M;
    if (diag_table_lds(a, M) == 0) {
        hipLaunchKernelGGL(diag_counts_kernel, dim3(1), dim3(64), 0, s, a, M);
        LAUNCH_CHECK("diag_counts_kernel");
    }
    const std::vector<int64_t> counts = diag_weight_counts(a, M);
    int rc = MM_OK;
    switch (M * 2 + (reordered ? 1 : 0)) {
#define CASE(MM)                                                       \
    case 2 * MM: rc = launch_diag_weights<MM, false>(a, counts, ds->num_cu, s); break; \
    case 2 * MM + 1: rc = launch_diag_weights<MM, true>(a, counts, ds->num_cu, s); break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10)
#undef CASE
        default:
            return fail(MM_ERR_INVALID, "bad M");
    }
    if (rc) return rc;
    LAUNCH_CHECK("diag_group_kernel");
    if (counts_out) *counts_out = counts;
    return MM_OK;
}

// The VJP's blocks: one per PB params of the largest weight, at most 1024.
int diag_vjp_blocks(const DiagGeom &g) {
    const std::vector<int64_t> counts = diag_weight_counts(g.a, g.M);
    const int64_t mx = *std::max_element(counts.begin(), counts.end());
    const int PB = DG_THREADS / (2 * g.M);
    return int(std::max<int64_t>(1, std::min<int64_t>(1024, (mx + PB - 1) / PB)));
}

size_t diag_vjp_bytes(const DiagGeom &g, int nblk) {
    const int G = 2 * g.M;
    const size_t mirror = size_t(16) * size_t(2 * G * g.a.P + g.n2 + 2 * g.n11);
    return align256(g.bytes) + align256(mirror) + size_t(16) * size_t(nblk) * size_t(G) * size_t(G + 1);
}

template <int M, bool REORDER>
void launch_diag_vjp(DiagVjpArgs &v, const std::vector<int64_t> &counts, int nblk, double2 *dA, double2 *dB, hipStream_t s) {
    constexpr int PB = DG_THREADS / (2 * M);
    const int64_t np = int64_t(nblk) * 2 * M * (2 * M + 1);
    hipLaunchKernelGGL(zero_c128_span_kernel, dim3(64), dim3(256), 0, s, v.part, np);
    std::vector<int64_t> starts(counts.size(), 0);
    for (size_t w = 1; w < counts.size(); ++w) starts[w] = starts[w - 1] + counts[w - 1];
    for (int w = v.a.S; w >= 0; --w) {
        // blocks 0 .. nw-1 of the fixed nblk slots: block b always adds to partial row b
        const int nw = int(std::min<int64_t>(nblk, (counts[w] + PB - 1) / PB));
        const int64_t per = int64_t(nw) * PB;
        const int iters = int((counts[w] + per - 1) / per);
        hipLaunchKernelGGL((diag_vjp_kernel<M, REORDER>), dim3(nw), dim3(DG_THREADS), 0, s, v, w, counts[w], starts[w], iters);
    }
    hipLaunchKernelGGL((diag_vjp_finalize<M, REORDER>), dim3(2 * M * (2 * M + 1)), dim3(256), 0, s, v.part, nblk, dA, dB);
}

// ------------------------------------------------------------ one leftover mode

constexpr int LO_MAXT = 7;  // PNR-detected modes

int lo_geom(int64_t output_cutoff, const int64_t *pnr, int Mt, DiagGeom *g, int *c0, size_t *bytes) {
    if (!pnr) return fail(MM_ERR_INVALID, "pnr_cutoffs pointer is NULL");
    if (Mt < 1) return fail(MM_ERR_INVALID, "The number of modes should be greater than 1.");
    if (Mt > LO_MAXT) return fail(MM_ERR_UNSUPPORTED, "1leftoverMode supports up to %d detected modes, got %d", LO_MAXT, Mt);
    if (output_cutoff < 0) return fail(MM_ERR_INVALID, "output_cutoff=%lld is negative", (long long)output_cutoff);
    if (output_cutoff >= 1024) return fail(MM_ERR_UNSUPPORTED, "output_cutoff=%lld: at most 1023", (long long)output_cutoff);
    int64_t tail[DG_MAXM];
    for (int j = 0; j < Mt; ++j) {
        if (pnr[j] < 0) return fail(MM_ERR_INVALID, "pnr_cutoffs[%d]=%lld is negative", j, (long long)pnr[j]);
        tail[j] = pnr[j] + 1;
    }
    *c0 = int(output_cutoff + 1);
    int rc = diag_geom(tail, Mt, int64_t(*c0) * *c0, g);
    if (rc) return rc;
    *bytes = g->bytes + size_t(16) * size_t(Mt) * size_t(g->a.P) * size_t(*c0) * size_t(*c0);  // + piv
    return MM_OK;
}

template <int MT, bool REORDER>
int launch_leftover(LeftArgs &a, const std::vector<int64_t> &counts, int num_cu, hipStream_t s) {
    constexpr int PB = DG_THREADS / (2 * MT);
    const size_t lds = diag_table_lds(a.t, MT);
    const int64_t rgrid = std::max<int64_t>(1, std::min<int64_t>((a.t.P + 255) / 256, int64_t(num_cu) * (lds ? 4 : 16)));
    hipLaunchKernelGGL((diag_rank_kernel<MT>), dim3(unsigned(rgrid)), dim3(256), lds, s, a.t, lds ? 1 : 0);
    const int ithreads = (a.c0 + 63) / 64 * 64;
    hipLaunchKernelGGL((lo_init_kernel<MT, REORDER>), dim3(1), dim3(ithreads), size_t(3) * a.c0 * 16, s, a);
    int64_t start = 0;
    for (int w = 0; w <= a.t.S; ++w) {
        const int64_t blocks = (counts[w] * a.C + PB - 1) / PB;
        if (blocks > 0x7fffffff) return fail(MM_ERR_UNSUPPORTED, "weight %d has too many entries for one launch", w);
        if (blocks > 0) {
            hipLaunchKernelGGL((lo_diag_kernel<MT, REORDER>), dim3(unsigned(blocks)), dim3(DG_THREADS), 0, s, a, w, counts[w], start);
            hipLaunchKernelGGL((lo_off_kernel<MT, REORDER>), dim3(unsigned(blocks)), dim3(DG_THREADS), 0, s, a, w, counts[w], start);
        }
        start += counts[w];
    }
    return MM_OK;
}

// Binds the forward's arrays to the workspace (layout of lo_geom).
void lo_bind(const DiagGeom &g, int c0, const double *A, const double *b, const double *c, double *out, char *ws, DevState *ds,
             LeftArgs &a) {
    const int Mt = g.M;
    a.t = g.a;
    DiagArgs &t = a.t;
    t.sq = ds->sq;
    t.rsq = ds->rsq;
    const int64_t C = int64_t(c0) * c0;
    t.T = reinterpret_cast<int64_t *>(ws);
    t.Q = t.T + int64_t(Mt + 1) * (t.S + 1);
    t.list = reinterpret_cast<int32_t *>(ws + align256(size_t(2 * (Mt + 1)) * size_t(t.S + 1) * 8));
    t.fwd = reinterpret_cast<int32_t *>(reinterpret_cast<char *>(t.list) + align256(size_t(t.P) * 4));
    t.inD = reinterpret_cast<double2 *>(ws + g.t_bytes);
    t.inO = t.inD + 2 * Mt * t.P * C;
    t.arr2 = t.inO + 2 * Mt * t.P * C;
    t.a1010 = t.arr2 + g.n2 * C;
    t.a1001 = t.a1010 + g.n11 * C;
    t.arr0 = nullptr;
    a.piv = reinterpret_cast<double2 *>(ws + g.bytes);
    a.A = reinterpret_cast<const double2 *>(A);
    a.B = reinterpret_cast<const double2 *>(b);
    a.c = reinterpret_cast<const double2 *>(c);
    a.c0 = c0;
    a.C = int(C);
    a.out = reinterpret_cast<double2 *>(out);
}

int lo_forward(LeftArgs &a, int Mt, int reordered, DevState *ds, hipStream_t s, std::vector<int64_t> *counts_out) {
    if (diag_table_lds(a.t, Mt) == 0) {
        hipLaunchKernelGGL(diag_counts_kernel, dim3(1), dim3(64), 0, s, a.t, Mt);
        LAUNCH_CHECK("diag_counts_kernel");
    }
    const std::vector<int64_t> counts = diag_weight_counts(a.t, Mt);
    int rc = MM_OK;
    switch (Mt * 2 + (reordered ? 1 : 0)) {
#define CASE(MM)                                                                   \
    case 2 * MM: rc = launch_leftover<MM, false>(a, counts, ds->num_cu, s); break; \
    case 2 * MM + 1: rc = launch_leftover<MM, true>(a, counts, ds->num_cu, s); break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7)
#undef CASE
        default:
            return fail(MM_ERR_INVALID, "bad number of detected modes");
    }
    if (rc) return rc;
    LAUNCH_CHECK("lo kernels");
    if (counts_out) *counts_out = counts;
    return MM_OK;
}

// The VJP's blocks: one per PB groups (params, block entry) of the largest weight, at most 1024.
int lo_vjp_blocks(const DiagGeom &g, int Mt, int64_t C) {
    const std::vector<int64_t> counts = diag_weight_counts(g.a, Mt);
    const int64_t mx = *std::max_element(counts.begin(), counts.end());
    const int PB = DG_THREADS / (2 * Mt);
    return int(std::max<int64_t>(1, std::min<int64_t>(1024, (mx * C + PB - 1) / PB)));
}

// forward workspace, adjoint mirrors of the inboxes and sub-lattices, adP, partial rows
size_t lo_vjp_bytes(const DiagGeom &g, int Mt, int64_t C, size_t fwd_bytes, int nblk) {
    const int GL = 2 * (Mt + 1);
    const size_t mirror = size_t(16) * size_t(C) * size_t(4 * Mt * g.a.P + g.n2 + 2 * g.n11);
    const size_t adp = size_t(16) * size_t(Mt) * size_t(g.a.P) * size_t(C);
    return align256(fwd_bytes) + align256(mirror) + align256(adp) + size_t(16) * size_t(nblk + 1) * GL * (GL + 1);
}

template <int MT, bool REORDER>
void launch_lo_vjp(LeftVjpArgs &v, const std::vector<int64_t> &counts, int nblk, double2 *dA, double2 *dB, hipStream_t s) {
    constexpr int PB = DG_THREADS / (2 * MT);
    constexpr int GL = 2 * (MT + 1);
    const int64_t C = v.a.C;
    auto zero = [&](double2 *p, int64_t n) {
        const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(2048, (n + 255) / 256));
        hipLaunchKernelGGL(zero_c128_span_kernel, dim3(unsigned(grid)), dim3(256), 0, s, p, n);
    };
    // the mirrors, adP and the partial rows are one contiguous span of the workspace
    zero(v.adD, (v.part + int64_t(nblk + 1) * GL * (GL + 1)) - v.adD);
    std::vector<int64_t> starts(counts.size(), 0);
    for (size_t w = 1; w < counts.size(); ++w) starts[w] = starts[w - 1] + counts[w - 1];
    for (int w = v.a.t.S; w >= 0; --w) {
        const int64_t groups = counts[w] * C;
        if (groups == 0) continue;
        // blocks 0 .. nw-1 of the fixed nblk slots: block b always adds to partial row b
        const int nw = int(std::min<int64_t>(nblk, (groups + PB - 1) / PB));
        const int64_t per = int64_t(nw) * PB;
        const int iters = int((groups + per - 1) / per);
        hipLaunchKernelGGL((lo_off_vjp_kernel<MT, REORDER>), dim3(nw), dim3(DG_THREADS), 0, s, v, counts[w], starts[w], iters);
        hipLaunchKernelGGL((lo_diag_vjp_kernel<MT, REORDER>), dim3(nw), dim3(DG_THREADS), 0, s, v, counts[w], starts[w], iters);
    }
    int nt = 64;
    while (nt < v.a.c0) nt <<= 1;  // c0 <= 1024
    hipLaunchKernelGGL((lo_init_vjp_kernel<MT, REORDER>), dim3(1), dim3(nt), 0, s, v, nblk);
    hipLaunchKernelGGL((diag_vjp_finalize<MT + 1, REORDER>), dim3(GL * (GL + 1)), dim3(256), 0, s, v.part, nblk + 1, dA, dB);
}

}  // namespace

extern "C" {

size_t mm_hermite_1leftover_workspace(int64_t output_cutoff, const int64_t *pnr_cutoffs, int Mt) {
    DiagGeom g;
    int c0 = 0;
    size_t bytes = 0;
    return lo_geom(output_cutoff, pnr_cutoffs, Mt, &g, &c0, &bytes) ? 0 : bytes;
}

int mm_hermite_1leftover(const double *A, const double *b, const double *c, int64_t output_cutoff, const int64_t *pnr_cutoffs,
                         int Mt, int reordered, double *out, void *workspace, size_t ws_bytes, void *stream) {
    DiagGeom g;
    int c0 = 0;
    size_t need = 0;
    int rc = lo_geom(output_cutoff, pnr_cutoffs, Mt, &g, &c0, &need);
    if (rc) return rc;
    if (!A || !b || !c || !out) return fail(MM_ERR_INVALID, "NULL array pointer");
    if (!workspace || ws_bytes < need) return fail(MM_ERR_WORKSPACE, "1leftover workspace of %zu bytes needed, %zu given", need, ws_bytes);
    DevState *ds = nullptr;
    rc = device_state(&ds);
    if (rc) return rc;
    LeftArgs a;
    lo_bind(g, c0, A, b, c, out, reinterpret_cast<char *>(workspace), ds, a);
    std::vector<int64_t> counts;
    return lo_forward(a, Mt, reordered, ds, reinterpret_cast<hipStream_t>(stream), &counts);
}

size_t mm_hermite_1leftover_vjp_workspace(int64_t output_cutoff, const int64_t *pnr_cutoffs, int Mt) {
    DiagGeom g;
    int c0 = 0;
    size_t bytes = 0;
    if (lo_geom(output_cutoff, pnr_cutoffs, Mt, &g, &c0, &bytes)) return 0;
    return lo_vjp_bytes(g, Mt, int64_t(c0) * c0, bytes, lo_vjp_blocks(g, Mt, int64_t(c0) * c0));
}

int mm_hermite_1leftover_vjp(const double *A, const double *b, const double *c, int64_t output_cutoff, const int64_t *pnr_cutoffs,
                             int Mt, int reordered, const double *dLdout, double *out, double *dA, double *db, double *dc,
                             void *workspace, size_t ws_bytes, void *stream) {
    DiagGeom g;
    int c0 = 0;
    size_t fwd_bytes = 0;
    int rc = lo_geom(output_cutoff, pnr_cutoffs, Mt, &g, &c0, &fwd_bytes);
    if (rc) return rc;
    if (!A || !b || !c || !dLdout || !out || !dA || !db || !dc) return fail(MM_ERR_INVALID, "NULL array pointer");
    const int64_t C = int64_t(c0) * c0;
    const int nblk = lo_vjp_blocks(g, Mt, C);
    const size_t need = lo_vjp_bytes(g, Mt, C, fwd_bytes, nblk);
    if (!workspace || ws_bytes < need)
        return fail(MM_ERR_WORKSPACE, "1leftover VJP workspace of %zu bytes needed, %zu given", need, ws_bytes);
    DevState *ds = nullptr;
    rc = device_state(&ds);
    if (rc) return rc;
    char *ws = reinterpret_cast<char *>(workspace);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    LeftVjpArgs v;
    lo_bind(g, c0, A, b, c, out, ws, ds, v.a);
    std::vector<int64_t> counts;
    rc = lo_forward(v.a, Mt, reordered, ds, s, &counts);
    if (rc) return rc;
    const int64_t P = g.a.P;
    const int64_t nmirror = C * (4 * Mt * P + g.n2 + 2 * g.n11);
    v.cot = reinterpret_cast<const double2 *>(dLdout);
    v.adD = reinterpret_cast<double2 *>(ws + align256(fwd_bytes));
    v.adO = v.adD + 2 * Mt * P * C;
    v.ad2 = v.adO + 2 * Mt * P * C;
    v.ad1010 = v.ad2 + g.n2 * C;
    v.ad1001 = v.ad1010 + g.n11 * C;
    v.adP = reinterpret_cast<double2 *>(ws + align256(fwd_bytes) + align256(size_t(16) * size_t(nmirror)));
    v.part = reinterpret_cast<double2 *>(reinterpret_cast<char *>(v.adP) + align256(size_t(16) * size_t(Mt * P * C)));
    v.dc = reinterpret_cast<double2 *>(dc);
    double2 *dA2 = reinterpret_cast<double2 *>(dA), *dB2 = reinterpret_cast<double2 *>(db);
    switch (Mt * 2 + (reordered ? 1 : 0)) {
#define CASE(MM)                                                                        \
    case 2 * MM: launch_lo_vjp<MM, false>(v, counts, nblk, dA2, dB2, s); break; \
    case 2 * MM + 1: launch_lo_vjp<MM, true>(v, counts, nblk, dA2, dB2, s); break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7)
#undef CASE
        default:
            return fail(MM_ERR_INVALID, "bad number of detected modes");
    }
    LAUNCH_CHECK("lo vjp kernels");
    return MM_OK;
}

size_t mm_hermite_diagonal_workspace(const int64_t *cutoffs, int M, int64_t nb) {
    DiagGeom g;
    return diag_geom(cutoffs, M, nb, &g) ? 0 : g.bytes;
}

int mm_hermite_diagonal(const double *A, const double *B, const double *c, int c_batched, const int64_t *cutoffs, int M,
                        int64_t nb, int reordered, double *G, void *workspace, size_t ws_bytes, void *stream) {
    DiagGeom g;
    int rc = diag_geom(cutoffs, M, nb, &g);
    if (rc) return rc;
    if (!A || !B || !c || !G) return fail(MM_ERR_INVALID, "NULL array pointer");
    if (!workspace || ws_bytes < g.bytes)
        return fail(MM_ERR_WORKSPACE, "diagonal workspace of %zu bytes needed, %zu given", g.bytes, ws_bytes);
    DevState *ds = nullptr;
    rc = device_state(&ds);
    if (rc) return rc;
    diag_bind(g, A, B, c, c_batched, G, reinterpret_cast<char *>(workspace), ds);
    return diag_forward(g, reordered, ds, reinterpret_cast<hipStream_t>(stream), nullptr);
}

size_t mm_hermite_diagonal_vjp_workspace(const int64_t *cutoffs, int M) {
    DiagGeom g;
    if (diag_geom(cutoffs, M, 1, &g)) return 0;
    return diag_vjp_bytes(g, diag_vjp_blocks(g));
}

int mm_hermite_diagonal_vjp(const double *A, const double *B, const double *c, const int64_t *cutoffs, int M, int reordered,
                            const double *dLdG, double *G, double *dA, double *dB, double *dc, void *workspace, size_t ws_bytes,
                            void *stream) {
    DiagGeom g;
    int rc = diag_geom(cutoffs, M, 1, &g);
    if (rc) return rc;
    if (!A || !B || !c || !dLdG || !G || !dA || !dB || !dc) return fail(MM_ERR_INVALID, "NULL array pointer");
    const int nblk = diag_vjp_blocks(g);
    const size_t need = diag_vjp_bytes(g, nblk);
    if (!workspace || ws_bytes < need) return fail(MM_ERR_WORKSPACE, "diagonal VJP workspace of %zu bytes needed, %zu given", need, ws_bytes);
    DevState *ds = nullptr;
    rc = device_state(&ds);
    if (rc) return rc;
    char *ws = reinterpret_cast<char *>(workspace);
    diag_bind(g, A, B, c, 0, G, ws, ds);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    std::vector<int64_t> counts;
    rc = diag_forward(g, reordered, ds, s, &counts);
    if (rc) return rc;
    DiagVjpArgs v;
    v.a = g.a;
    v.cot = reinterpret_cast<const double2 *>(dLdG);
    v.adD = reinterpret_cast<double2 *>(ws + align256(g.bytes));
    v.adO = v.adD + 2 * M * g.a.P;
    v.ad2 = v.adO + 2 * M * g.a.P;
    v.ad1010 = v.ad2 + g.n2;
    v.ad1001 = v.ad1010 + g.n11;
    v.part = reinterpret_cast<double2 *>(ws + align256(g.bytes) + align256(size_t(16) * size_t(4 * M * g.a.P + g.n2 + 2 * g.n11)));
    v.dc = reinterpret_cast<double2 *>(dc);
    double2 *dA2 = reinterpret_cast<double2 *>(dA), *dB2 = reinterpret_cast<double2 *>(dB);
    switch (M * 2 + (reordered ? 1 : 0)) {
#define CASE(MM)                                                                  \
    case 2 * MM: launch_diag_vjp<MM, false>(v, counts, nblk, dA2, dB2, s); break; \
    case 2 * MM + 1: launch_diag_vjp<MM, true>(v, counts, nblk, dA2, dB2, s); break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10)
#undef CASE
        default:
            return fail(MM_ERR_INVALID, "bad M");
    }
    LAUNCH_CHECK("diag_vjp_kernel");
    return MM_OK;
}

}  // extern "C"
