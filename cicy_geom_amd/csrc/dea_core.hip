// dea_core.hip -- block-preconditioned persistent CG for the core of the condensed solve.
//
// After the condensed solve's exact eliminations (dea_condensed.hip) the core system is
// small (1-5 % of the rows) and one CG iteration over it moves only a few tens of MB: the
// two-kernel PCG iteration is then bound by launch and reduction latency, not bandwidth.
// Nanoplatelets make it worse numerically: a plate's triangulation edges conduct ~1e4 S
// against ~1e-3 S for its junctions, so each plate contributes a nearly rigid voltage mode
// that Jacobi scaling leaves at an eigenvalue ~1e-7 and CG needs tens of thousands of
// iterations.  Here the core is solved by
//
//   * the Chronopoulos-Gear formulation of PCG (both dot products of a step summed at one
//     reduction, two grid barriers per iteration), run as one cooperative launch per chunk
//     of iterations, every work-group resident on its own CU.  (The pipelined variant with
//     one barrier per iteration loses too much attainable accuracy on these systems: its
//     recursively updated vectors drift and the recursion stalls before 1e-10.);
//   * block-Jacobi preconditioning: row groups given by the caller (cicy_sys_set_blocks;
//     the nodes of one nanoplatelet, which the reference numbers consecutively in
//     LM_matrix_for_gnps, Direct_Electrifying.cpp:307-349) are inverted exactly (dense
//     inverse of the core's diagonal block, computed once per system in FP64), the other
//     rows use the reference's Jacobi preconditioner (Direct_Electrifying.cpp:1662-1676).
//     A plate's inverse is dominated by one constant c (its nearly rigid mode), so it is
//     stored as c 11^T (FP64) plus the remainder R = A_bb^-1 - c 11^T in FP32: R's entries
//     are small, its rounding (~1e-10 of ||R||) stays far below R's smallest eigenvalue, so
//     the preconditioner stays symmetric positive definite at half the bytes.  Each block
//     is solved by one work-group: a work-group-local dense product with the block's
//     residual staged in LDS;
//   * two partitions of the rows, one per phase, each balanced on its own: the core rows
//     are renumbered so that the Jacobi rows come first and the rows of each block follow
//     as one run; the vector updates and the preconditioner (phase u) give each
//     work-group a run of Jacobi rows plus a run of whole blocks, sized by a cost model
//     measured on MI355X (row updates, inverse entries, rounds of block products); the
//     product w = A u (phase w) gives each work-group an equal share of the rows.  Plate
//     rows are few per work-group but dear in phase u, tube rows cheap there but as dear
//     as any row in the product, so one partition for both phases leaves either phase
//     waiting at its barrier for the work-groups the other phase favoured;
//   * vectors in HBM: u and w, which cross between the partitions (and u is gathered by
//     the product), are written and read with device-coherent accesses (dea_grid.h);
//     x, r, p, s are private to the phase-u owner of the row.  The product reads the
//     matrix in sliced ELL (64-row slices, slot-major) so that its column and value loads
//     are coalesced.
//
// Iteration (u = M^-1 r, w = A u, gamma = r.u, delta = w.u; Chronopoulos & Gear 1989):
//   p = u + beta p;  s = w + beta s;  x += alpha p;  r -= alpha s;  u = M^-1 r
//   [barrier: gamma' = r.u, r.r]  stop when r.r <= tol^2 ||ref||^2
//   w = A u
//   [barrier: delta = w.u]
//   beta = gamma' / gamma;  alpha = gamma' / (delta - beta gamma' / alpha)
// If gamma or the denominator of alpha loses positivity (rounding on a badly
// preconditioned core) the solve stops as "stalled" and the condensed solve, which checks
// the full system's true residual after every pass, corrects again from it.
// Sums over work-groups are taken in a fixed order, so solves are bit-reproducible.
#include "dea_grid.h"
#include "dea_internal.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

using namespace cicy;

namespace cicy {

constexpr int kCoreThreads = 1024;
constexpr int kCoreWaves = kCoreThreads / kWave;
constexpr int kCorePairs = kCoreWaves / 2;  // blocks per round: two waves share one block
constexpr int kCoreMaxBlock = 128;     // rows of a dense block (larger groups are split)
// the split size in effect: kCoreMaxBlock, or CICY_CORE_MAX_BLOCK (16..kCoreMaxBlock) when set
// (a tuning knob; host and device layouts read the same value)
inline int core_max_block() {
    if (const char *e = getenv("CICY_CORE_MAX_BLOCK"))
        return std::max(16, std::min(kCoreMaxBlock, atoi(e)));
    return kCoreMaxBlock;
}
constexpr int kCoreChunk = 4096;       // iterations per cooperative launch (at most)
constexpr int kCoreRowsPerGroup = 192; // work-groups: about one per this many rows
constexpr int kCoreUnroll = 12;        // block-inverse columns loaded before use
constexpr int kCoreEnt = 8;            // row slots loaded before their gathers
constexpr int kCoreSums = 4;           // entries per work-group slot of the grid sums
constexpr int kCorePasses = 10;        // corrections from the true residual (stalled solves)
constexpr int kCoreLdsCols = 24 * 1024; // column indices of the product a work-group keeps in LDS
// Phase-u cost per work-group on MI355X, measured (us): kCostRow per row update (p, s, x,
// r, u), kCostEntry per block-inverse entry, kCostRound per round of kCorePairs blocks.
constexpr double kCostRow = 2.0e-3, kCostEntry = 6.3e-5, kCostRound = 3.2;
// Phase-w cost of a row, in units of one stored entry (a row is one lane's chain of
// dependent loads; its entries are the slots of that chain)
constexpr int kCostWRow = 64;
// A Jacobi row's phase-w cost relative to a block row's: its gathers reach columns spread
// over the core, a plate row's stay near its plate (configuration 3, CICY_CORE_TIMING:
// 15.4 us for 1 125 Jacobi rows against 12.3 us for 1 104 block rows per work-group)
constexpr double kCostWJacobi = 1.25;

struct CoreState {
    long long k;         // iterations done in this solve
    long long max_iter;
    double gamma, alpha, beta, rr, stop2;
    int done;            // 0 running, 1 converged, 2 iteration cap, 3 breakdown (not finite),
                         // 4 barrier timeout, 5 stalled (gamma or alpha lost positivity; the
                         // caller corrects from the true residual)
    int started;         // the prologue (u0 = M^-1 r0, w0 = A u0) has run
};

struct CoreDev {
    int m, G;
    // phase u: work-group g updates the Jacobi rows ju[g] .. ju[g+1] and the rows of the
    // blocks bu[g] .. bu[g+1] (rows numbered Jacobi rows first, then block by block)
    const int *ju, *bu;
    // phase w: work-group g forms w for the rows wr[g] .. wr[g+1]
    const int *wr;
    // the off-diagonal entries of both triangles in sliced ELL: work-group g's phase-w rows
    // are cut into slices of 64 rows (wsl[g] .. wsl[g+1]); slice s holds sl_len[s] slots,
    // slot t of its lane-th row at ell_*[sl_off[s] + t * 64 + lane] (column -1: padding),
    // in the row's column order
    const int *wsl;
    const long long *sl_off;
    int lds_cols;         // column indices staged in LDS per work-group (its first slots)
    const int *sl_len;
    const int *ell_col;
    const double *ell_val;
    const double *diag;   // [m]
    const int *blk_row0;  // [nb]
    const int *blk_n;     // [nb]
    const long long *blk_off;  // [nb] offset of the block's remainder (row-major, symmetric)
    const double *blk_c;  // [nb] constant part c of the block's inverse
    const float *binv;    // remainders R = A_bb^-1 - c 11^T
    double *x, *r, *u, *w, *p, *s;
    GridCtl *ctl;
    unsigned long long *part;  // [3][G][kCoreSums] grid-sum slots (grid_sum_slots)
};

}  // namespace cicy

struct cicy_core {
    cicy_ctx *ctx = nullptr;  // its arrays come from the context's pool
    int m = 0, G = 0, nb = 0, blocked_rows = 0, lds_cols = 0;
    long long nnz2 = 0, inv_entries = 0;
    int *ju = nullptr, *bu = nullptr, *wr = nullptr, *perm = nullptr;
    int *blk_row0 = nullptr, *blk_n = nullptr;
    int *wsl = nullptr, *sl_len = nullptr, *ell_col = nullptr;
    long long *blk_off = nullptr, *sl_off = nullptr;
    double *ell_val = nullptr, *diag = nullptr, *blk_c = nullptr;
    float *binv = nullptr;
    double *x = nullptr, *r = nullptr, *u = nullptr, *w = nullptr, *p = nullptr, *s = nullptr;
    unsigned long long *part = nullptr;
    cicy::GridCtl *ctl = nullptr;
    cicy::CoreState *st = nullptr;
    size_t bytes = 0;
};

namespace cicy {
namespace {

// ------------------------------------------------------------------------------------
// host layout
// ------------------------------------------------------------------------------------

struct CoreHost {
    int m = 0, G = 1, mj = 0;    // rows, work-groups, Jacobi rows (numbered first)
    std::vector<int> perm;       // [m] caller's core row of each row
    std::vector<int> rp, col, blk_of, ju, bu, wr, blk_row0, blk_n;
    std::vector<int> wsl, sl_len, ell_col;
    std::vector<long long> sl_off;
    std::vector<double> ell_val;
    std::vector<double> val, diag, blk_c;
    std::vector<float> binv;
    std::vector<long long> blk_off;
};

// In-place Cholesky of the dense SPD matrix a (n x n, row-major) and its inverse into inv.
// Returns false when a pivot is not positive.
bool dense_spd_inverse(int n, std::vector<double> &a, double *inv) {
    for (int j = 0; j < n; j++) {
        double d = a[(size_t)j * n + j];
        for (int k = 0; k < j; k++) d -= a[(size_t)j * n + k] * a[(size_t)j * n + k];
        if (!(d > 0.0) || !std::isfinite(d)) return false;
        const double l = std::sqrt(d);
        a[(size_t)j * n + j] = l;
        for (int i = j + 1; i < n; i++) {
            double v = a[(size_t)i * n + j];
            for (int k = 0; k < j; k++) v -= a[(size_t)i * n + k] * a[(size_t)j * n + k];
            a[(size_t)i * n + j] = v / l;
        }
    }
    // W = L^-1 (lower triangular, row-major), row by row: W_i = -(sum_k<i L_ik W_k) / L_ii
    // with W_ii = 1 / L_ii; every inner loop runs along a row
    std::vector<double> w((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) {
        const double *li = &a[(size_t)i * n];
        double *wi = &w[(size_t)i * n];
        for (int k = 0; k < i; k++) {
            const double lik = li[k];
            const double *wk = &w[(size_t)k * n];
            for (int j = 0; j <= k; j++) wi[j] -= lik * wk[j];
        }
        const double inv_d = 1.0 / li[i];
        for (int j = 0; j < i; j++) wi[j] *= inv_d;
        wi[i] = inv_d;
    }
    // A^-1 = W^T W: lower half as a sum of the rank-1 terms W_k^T W_k, then mirrored, so the
    // inverse is exactly symmetric (the product reads it by columns)
    std::fill(inv, inv + (size_t)n * n, 0.0);
    for (int k = 0; k < n; k++) {
        const double *wk = &w[(size_t)k * n];
        for (int i = 0; i <= k; i++) {
            const double wki = wk[i];
            double *ri = inv + (size_t)i * n;
            for (int j = 0; j <= i; j++) ri[j] += wki * wk[j];
        }
    }
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++) inv[(size_t)i * n + j] = inv[(size_t)j * n + i];
    return true;
}

// The work-groups and their row partitions (phase u: runs of whole blocks, then Jacobi rows
// by the cost model; phase w: runs of equal cost) and the sliced-ELL slices of the phase-w
// rows, for a layout whose rows (rp, renumbered), Jacobi row count mj and blocks are set.
// Returns the ELL size; sl_base / sl_end receive each slice's rows.
size_t core_partitions(CoreHost &H, int max_groups, std::vector<int> &sl_base,
                       std::vector<int> &sl_end) {
    const int m = H.m;
    const int nbk = (int)H.blk_row0.size();
    H.G = std::max(1, std::min(max_groups, (m + kCoreRowsPerGroup - 1) / kCoreRowsPerGroup));
    const int G = H.G;
    // phase u: the blocks cut into G runs of about equal estimated cost, each at most one
    // round of the block product (kCorePairs blocks) where the blocks allow it, then the
    // Jacobi rows shared out so that every work-group's estimate reaches the same level
    std::vector<double> bcost((size_t)nbk);
    double btotal = 0.0;
    for (int b = 0; b < nbk; b++) {
        const double n = H.blk_n[b];
        bcost[b] = kCostEntry * n * n + kCostRow * n + kCostRound / kCorePairs;
        btotal += bcost[b];
    }
    H.bu.assign((size_t)G + 1, nbk);
    {
        const int cap = nbk <= G * kCorePairs ? kCorePairs : (nbk + G - 1) / G;
        int b = 0;
        double acc = 0.0;
        for (int g = 0; g < G; g++) {
            H.bu[g] = b;
            // at least `must` blocks here, so that the rest fits the work-groups left
            const int must = std::max(0, (nbk - b) - (G - g - 1) * cap);
            const double target = btotal * (double)(g + 1) / G;
            for (int cnt = 0; b < nbk && cnt < cap &&
                              (cnt < must || acc + 0.5 * bcost[b] <= target);
                 cnt++)
                acc += bcost[b++];
        }
    }
    std::vector<double> cb((size_t)G, 0.0);
    for (int g = 0; g < G; g++) {
        const int cnt = H.bu[g + 1] - H.bu[g];
        for (int b = H.bu[g]; b < H.bu[g + 1]; b++) {
            const double n = H.blk_n[b];
            cb[g] += kCostEntry * n * n + kCostRow * n;
        }
        cb[g] += kCostRound * (double)((cnt + kCorePairs - 1) / kCorePairs);
    }
    auto rows_at = [&](double T) {
        double t = 0.0;
        for (int g = 0; g < G; g++) t += std::max(0.0, (T - cb[g]) / kCostRow);
        return t;
    };
    double lo = 0.0, hi = *std::max_element(cb.begin(), cb.end()) + kCostRow * (H.mj + 1.0);
    for (int it = 0; it < 100; it++) {
        const double mid = 0.5 * (lo + hi);
        (rows_at(mid) < H.mj ? lo : hi) = mid;
    }
    H.ju.assign((size_t)G + 1, 0);
    {
        long long given = 0;
        for (int g = 0; g < G; g++) {
            long long c = (long long)std::floor(std::max(0.0, (hi - cb[g]) / kCostRow));
            c = std::min<long long>(c, H.mj - given);
            if (g == G - 1) c = H.mj - given;
            given += c;
            H.ju[g + 1] = (int)given;
        }
    }
    // phase w: the rows cut into G runs of equal cost (kCostWRow per row + its entries,
    // times kCostWJacobi for a Jacobi row)
    H.wr.assign((size_t)G + 1, m);
    H.wr[0] = 0;
    {
        auto wcost = [&](int r) {
            const double c = (double)kCostWRow + (double)(H.rp[r + 1] - H.rp[r]);
            return r < H.mj ? kCostWJacobi * c : c;
        };
        double total = 0.0;
        for (int r = 0; r < m; r++) total += wcost(r);
        double acc = 0.0;
        int g = 1;
        for (int r = 0; r < m && g < G; r++) {
            acc += wcost(r);
            while (g < G && acc >= total * (double)g / G) H.wr[g++] = r + 1;
        }
    }
    // sliced ELL of each work-group's phase-w rows: the slices and their offsets first, then
    // the slices filled on host threads
    H.wsl.assign(1, 0);
    H.sl_off.clear();
    H.sl_len.clear();
    size_t ell = 0;
    for (int gg = 0; gg < G; gg++) {
        for (int base = H.wr[gg]; base < H.wr[gg + 1]; base += kWave) {
            const int end = std::min(H.wr[gg + 1], base + kWave);
            int len = 0;
            for (int r = base; r < end; r++) len = std::max(len, H.rp[r + 1] - H.rp[r]);
            H.sl_off.push_back((long long)ell);
            H.sl_len.push_back(len);
            sl_base.push_back(base);
            sl_end.push_back(end);
            ell += (size_t)len * kWave;
        }
        H.wsl.push_back((int)H.sl_len.size());
    }
    return ell;
}

// Core layout from the core's strictly-lower SSS triangle (crp/ccol/cval), its diagonal,
// the full-system row of each core row (ascending) and the caller's row groups
// (ranges[2 b], ranges[2 b + 1]) of the full system.  G work-groups at most.
bool build_core_host(int m, const std::vector<int64_t> &crp, const std::vector<int64_t> &ccol,
                     const std::vector<double> &cval, const std::vector<double> &cdiag,
                     const std::vector<int> &orig, int64_t nblocks, const int64_t *ranges,
                     int max_groups, CoreHost &H, const char **why) {
    H.m = m;
    // full symmetric CSR, columns ascending within each row
    std::vector<int> cnt((size_t)m + 1, 0);
    for (int r = 0; r < m; r++)
        for (int64_t k = crp[r]; k < crp[r + 1]; k++) {
            cnt[r + 1]++;
            cnt[(size_t)ccol[k] + 1]++;
        }
    H.rp.assign((size_t)m + 1, 0);
    for (int r = 0; r < m; r++) H.rp[r + 1] = H.rp[r] + cnt[r + 1];
    H.col.assign((size_t)H.rp[m], 0);
    H.val.assign((size_t)H.rp[m], 0.0);
    std::vector<int> fill(H.rp.begin(), H.rp.end() - 1);
    for (int r = 0; r < m; r++)
        for (int64_t k = crp[r]; k < crp[r + 1]; k++) {
            const int c = (int)ccol[k];
            H.col[fill[r]] = c;
            H.val[fill[r]++] = cval[k];
            H.col[fill[c]] = r;
            H.val[fill[c]++] = cval[k];
        }
    // rows sorted by column on host threads (rows are independent)
    const int nth = host_threads(m, 1 << 14);
    auto sort_rows = [&](int t) {
        std::vector<std::pair<int, double>> row;
        const int r0 = (int)((long long)m * t / nth), r1 = (int)((long long)m * (t + 1) / nth);
        for (int r = r0; r < r1; r++) {
            row.clear();
            for (int k = H.rp[r]; k < H.rp[r + 1]; k++) row.emplace_back(H.col[k], H.val[k]);
            std::sort(row.begin(), row.end());
            for (size_t q = 0; q < row.size(); q++) {
                H.col[(size_t)H.rp[r] + q] = row[q].first;
                H.val[(size_t)H.rp[r] + q] = row[q].second;
            }
        }
    };
    parallel_for(nth, sort_rows);
    H.diag = cdiag;
    // row groups -> core row ranges (split at core_max_block() rows)
    const int maxb = core_max_block();
    std::vector<std::pair<int, int>> groups;
    for (int64_t b = 0; b < nblocks; b++) {
        const int64_t s0 = ranges[2 * b], s1 = ranges[2 * b + 1];
        if (s0 < 0 || s1 < s0) {
            *why = "a row group has a negative start or ends before it starts";
            return false;
        }
        const int q0 = (int)(std::lower_bound(orig.begin(), orig.end(), s0) - orig.begin());
        const int q1 = (int)(std::lower_bound(orig.begin(), orig.end(), s1) - orig.begin());
        for (int q = q0; q < q1; q += maxb) {
            const int e = std::min(q1, q + maxb);
            if (e - q >= 2) groups.emplace_back(q, e);
        }
    }
    // the caller's groups usually come in row order already
    if (!std::is_sorted(groups.begin(), groups.end())) std::sort(groups.begin(), groups.end());
    for (size_t g = 1; g < groups.size(); g++)
        if (groups[g].first < groups[g - 1].second) {
            *why = "row groups overlap";
            return false;
        }
    // dense inverses of the diagonal blocks (in parallel over blocks)
    const size_t ng = groups.size();
    std::vector<long long> off(ng + 1, 0);
    for (size_t g = 0; g < ng; g++) {
        const long long nb = groups[g].second - groups[g].first;
        off[g + 1] = off[g] + nb * nb;
    }
    std::vector<double> inv((size_t)off[ng]);
    std::vector<char> ok(ng, 0);
    auto work = [&](size_t g0, size_t g1) {
        std::vector<double> a;
        for (size_t g = g0; g < g1; g++) {
            const int q0 = groups[g].first, nb = groups[g].second - groups[g].first;
            a.assign((size_t)nb * nb, 0.0);
            for (int i = 0; i < nb; i++) {
                a[(size_t)i * nb + i] = H.diag[q0 + i];
                for (int k = H.rp[q0 + i]; k < H.rp[q0 + i + 1]; k++) {
                    const int c = H.col[k] - q0;
                    if (c >= 0 && c < nb) a[(size_t)i * nb + c] = H.val[k];
                }
            }
            ok[g] = dense_spd_inverse(nb, a, &inv[(size_t)off[g]]) ? 1 : 0;
        }
    };
    {
        const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> th;
        const size_t per = (ng + nt - 1) / nt;
        for (unsigned t = 0; t < nt; t++) {
            const size_t g0 = t * per, g1 = std::min(ng, g0 + per);
            if (g0 < g1) th.emplace_back(work, g0, g1);
        }
        for (auto &t : th) t.join();
    }
    // keep the blocks that factorised (a block that did not stays on Jacobi) and renumber
    // the rows: Jacobi rows first (in order), then the rows of each block, block by block
    std::vector<char> in_blk((size_t)m, 0);
    for (size_t g = 0; g < ng; g++)
        if (ok[g])
            for (int q = groups[g].first; q < groups[g].second; q++) in_blk[q] = 1;
    std::vector<int> &perm = H.perm;
    perm.clear();
    perm.reserve((size_t)m);
    for (int q = 0; q < m; q++)
        if (!in_blk[q]) perm.push_back(q);
    H.mj = (int)perm.size();
    for (size_t g = 0; g < ng; g++)
        if (ok[g])
            for (int q = groups[g].first; q < groups[g].second; q++) perm.push_back(q);
    std::vector<int> iperm((size_t)m);
    for (int i = 0; i < m; i++) iperm[perm[i]] = i;
    {
        std::vector<int> rp2((size_t)m + 1, 0), col2(H.col.size());
        std::vector<double> val2(H.val.size()), diag2((size_t)m);
        for (int i = 0; i < m; i++) rp2[i + 1] = rp2[i] + (H.rp[perm[i] + 1] - H.rp[perm[i]]);
        parallel_for(nth, [&](int t) {
            std::vector<std::pair<int, double>> row;
            const int i0 = (int)((long long)m * t / nth), i1 = (int)((long long)m * (t + 1) / nth);
            for (int i = i0; i < i1; i++) {
                const int o = perm[i];
                diag2[i] = H.diag[o];
                row.clear();
                for (int k = H.rp[o]; k < H.rp[o + 1]; k++)
                    row.emplace_back(iperm[H.col[k]], H.val[k]);
                std::sort(row.begin(), row.end());
                for (size_t q = 0; q < row.size(); q++) {
                    col2[(size_t)rp2[i] + q] = row[q].first;
                    val2[(size_t)rp2[i] + q] = row[q].second;
                }
            }
        });
        H.rp.swap(rp2);
        H.col.swap(col2);
        H.val.swap(val2);
        H.diag.swap(diag2);
    }
    H.blk_of.assign((size_t)m, -1);
    H.blk_row0.clear();
    H.blk_n.clear();
    H.blk_off.clear();
    H.blk_c.clear();
    long long kept = 0;
    for (size_t g = 0; g < ng; g++)
        if (ok[g]) kept += off[g + 1] - off[g];
    H.binv.resize((size_t)kept);
    int row0 = H.mj;
    long long at = 0;
    std::vector<size_t> kept_g;  // the factorised groups, in order
    for (size_t g = 0; g < ng; g++) {
        if (!ok[g]) continue;
        const int nb = groups[g].second - groups[g].first;
        const int id = (int)H.blk_row0.size();
        for (int i = 0; i < nb; i++) H.blk_of[row0 + i] = id;
        H.blk_row0.push_back(row0);
        H.blk_n.push_back(nb);
        H.blk_off.push_back(at);
        kept_g.push_back(g);
        at += (long long)nb * nb;
        row0 += nb;
    }
    // constant part (the mean entry of the inverse) and FP32 remainder of each block, on
    // host threads over blocks
    H.blk_c.assign(kept_g.size(), 0.0);
    {
        const int nbt = host_threads((long long)kept_g.size(), 8);
        parallel_for(nbt, [&](int t) {
            for (size_t b = t; b < kept_g.size(); b += nbt) {
                const size_t g = kept_g[b];
                const double nb = groups[g].second - groups[g].first;
                // the sum of the row sums (each in column order)
                double c = 0.0;
                const int n_ = groups[g].second - groups[g].first;
                for (int i = 0; i < n_; i++) {
                    double t = 0.0;
                    for (int j = 0; j < n_; j++) t += inv[(size_t)off[g] + (size_t)i * n_ + j];
                    c += t;
                }
                c /= nb * nb;
                H.blk_c[b] = c;
                long long o = H.blk_off[b];
                for (long long e = off[g]; e < off[g + 1]; e++)
                    H.binv[(size_t)o++] = (float)(inv[(size_t)e] - c);
            }
        });
    }
    std::vector<int> sl_base, sl_end;
    const size_t ell = core_partitions(H, max_groups, sl_base, sl_end);
    H.ell_col.assign(ell, -1);
    H.ell_val.assign(ell, 0.0);
    const int nsl = (int)sl_base.size();
    parallel_for(host_threads(nsl, 256), [&](int t) {
        const int nt = host_threads(nsl, 256);
        for (int q = (int)((long long)nsl * t / nt); q < (int)((long long)nsl * (t + 1) / nt);
             q++) {
            const int base = sl_base[q], end = sl_end[q];
            const size_t o = (size_t)H.sl_off[q];
            for (int r = base; r < end; r++)
                for (int k = 0; k < H.rp[r + 1] - H.rp[r]; k++) {
                    H.ell_col[o + (size_t)k * kWave + (r - base)] = H.col[(size_t)H.rp[r] + k];
                    H.ell_val[o + (size_t)k * kWave + (r - base)] = H.val[(size_t)H.rp[r] + k];
                }
        }
    });
    return true;
}

// The preconditioner of the layout on the host: out = M^-1 v.
void host_core_precond(const CoreHost &H, const double *v, double *out) {
    for (int i = 0; i < H.m; i++)
        if (H.blk_of[i] < 0) out[i] = (1.0 / H.diag[i]) * v[i];
    for (size_t b = 0; b < H.blk_row0.size(); b++) {
        const int q0 = H.blk_row0[b], nb = H.blk_n[b];
        const float *M = &H.binv[(size_t)H.blk_off[b]];
        double sum = 0.0;
        for (int j = 0; j < nb; j++) sum += v[q0 + j];
        for (int i = 0; i < nb; i++) {
            double z = H.blk_c[b] * sum;
            for (int j = 0; j < nb; j++) z = fma((double)M[(size_t)j * nb + i], v[q0 + j], z);
            out[q0 + i] = z;
        }
    }
}

void host_core_apply(const CoreHost &H, const double *v, double *out) {
    for (int i = 0; i < H.m; i++) {
        double acc = H.diag[i] * v[i];
        for (int k = H.rp[i]; k < H.rp[i + 1]; k++) acc = fma(H.val[k], v[H.col[k]], acc);
        out[i] = acc;
    }
}

constexpr int kStalled = -1;  // host status of a stalled recursion (kernel done == 5)

// The kernel's recurrences on the host (sequential sums): solves A x = rhs from x = 0 and
// returns the iteration count.
long long host_core_cg(const CoreHost &H, const double *rhs, double tol, double ref2,
                       long long max_iter, double *x, int *status) {
    const int m = H.m;
    std::vector<double> r(rhs, rhs + m), u(m), w(m), p(m, 0.0), s(m, 0.0);
    std::fill(x, x + m, 0.0);
    const double stop2 = tol * tol * ref2;
    auto dot = [&](const std::vector<double> &a, const std::vector<double> &b) {
        double t = 0.0;
        for (int i = 0; i < m; i++) t += a[i] * b[i];
        return t;
    };
    host_core_precond(H, r.data(), u.data());
    double gamma = dot(r, u), rr = dot(r, r);
    *status = CICY_OK;
    if (rr <= stop2) return 0;
    host_core_apply(H, u.data(), w.data());
    double delta = dot(w, u);
    double alpha = gamma / delta, beta = 0.0;
    if (!(alpha > 0.0) || !std::isfinite(alpha)) {
        *status = kStalled;
        return 0;
    }
    long long k = 0;
    while (true) {
        for (int i = 0; i < m; i++) {
            p[i] = fma(beta, p[i], u[i]);
            s[i] = fma(beta, s[i], w[i]);
            x[i] = fma(alpha, p[i], x[i]);
            r[i] = fma(-alpha, s[i], r[i]);
        }
        host_core_precond(H, r.data(), u.data());
        const double g2 = dot(r, u);
        rr = dot(r, r);
        k++;
        if (rr <= stop2) return k;
        if (!std::isfinite(rr)) {
            *status = CICY_BREAKDOWN;
            return k;
        }
        if (!(g2 > 0.0)) {
            *status = kStalled;
            return k;
        }
        if (k >= max_iter) {
            *status = CICY_MAXITER;
            return k;
        }
        host_core_apply(H, u.data(), w.data());
        delta = dot(w, u);
        beta = g2 / gamma;
        const double den = delta - beta * g2 / alpha;
        if (!(den > 0.0) || !std::isfinite(den)) {
            *status = kStalled;
            return k;
        }
        alpha = g2 / den;
        gamma = g2;
    }
}

// ------------------------------------------------------------------------------------
// the persistent kernel
// ------------------------------------------------------------------------------------

// dir 0: dst[q] = src[perm[q]] (caller's core order -> the kernel's); dir 1: the inverse
__global__ void k_core_permute(int m, const int *perm, const double *src, double *dst, int dir) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m) return;
    if (dir == 0) dst[q] = src[perm[q]];
    else dst[perm[q]] = src[q];
}

// One chunk of iterations.  All work-groups hold the same scalars (the grid sums are
// bit-identical everywhere), so they take the same branches and leave together.
// TIMED: each work-group adds the wall-clock ticks of phase u, the first grid sum, phase w
// and the second grid sum of its iterations to tim[4 g .. 4 g + 3] (a diagnostic variant,
// launched when CICY_CORE_TIMING is set)
template <bool TIMED>
__global__ __launch_bounds__(kCoreThreads) void k_core_cg(CoreDev a, CoreState *st, int chunk,
                                                          unsigned long long *tim) {
    __shared__ double yb[kCorePairs][kCoreMaxBlock];      // block input in flight
    __shared__ double zp[kCorePairs][2][kCoreMaxBlock];   // its two half products
    __shared__ double lds[kCoreSums * kCoreWaves];
    __shared__ int flag;
    extern __shared__ int scols[];  // [lds_cols] this work-group's first product columns
    const int g = blockIdx.x, tid = threadIdx.x;
    const int wave = tid / kWave, lane = tid & (kWave - 1);
    const int j0 = a.ju[g], j1 = a.ju[g + 1];     // phase u: Jacobi rows
    const int b0 = a.bu[g], b1 = a.bu[g + 1];     // phase u: blocks
    const int w0 = a.wr[g], w1 = a.wr[g + 1];     // phase w: rows
    const int s0 = a.wsl[g], s1 = a.wsl[g + 1];
    // the column indices of the product are the same in every iteration: the first
    // lds_cols of this work-group's slots are read from LDS, so the gathers of u wait for an
    // LDS read instead of a load from HBM
    const long long cbase = s1 > s0 ? a.sl_off[s0] : 0;
    const int ccount = s1 > s0 ? (int)std::min<long long>(a.sl_off[s1 - 1] +
                                                              (long long)a.sl_len[s1 - 1] * kWave -
                                                              cbase,
                                                          a.lds_cols)
                               : 0;
    for (int i = tid; i < ccount; i += kCoreThreads) scols[i] = a.ell_col[cbase + i];
    __syncthreads();
    long long k = st->k;
    const long long max_iter = st->max_iter;
    double gamma = st->gamma, alpha = st->alpha, beta = st->beta, rr = st->rr;
    const double stop2 = st->stop2;
    int done = st->done;
    const bool started = st->started != 0;
    unsigned bar = 0;

    // (A v)_row for this lane's row of slice sl, v gathered from src (coherent); own = v_row
    auto row_product = [&](int sl, int row, bool live, const double *src, double &own) {
        const int len = a.sl_len[sl];
        const long long so = a.sl_off[sl];
        const bool staged = so - cbase + (long long)len * kWave <= ccount;
        const int *cc = staged ? scols + (so - cbase) + lane : a.ell_col + so + lane;
        const double *vv = a.ell_val + so + lane;
        own = live ? ld_coh(&src[row]) : 0.0;
        double acc = live ? a.diag[row] * own : 0.0;
        for (int t = 0; t < len; t += kCoreEnt) {
            int c[kCoreEnt];
            double v[kCoreEnt], gu[kCoreEnt];
#pragma unroll
            for (int e = 0; e < kCoreEnt; e++) {
                const bool ok = t + e < len;
                c[e] = ok ? cc[(t + e) * kWave] : -1;
                v[e] = ok ? vv[(t + e) * kWave] : 0.0;
            }
#pragma unroll
            for (int e = 0; e < kCoreEnt; e++) gu[e] = c[e] >= 0 ? ld_coh(&src[c[e]]) : 0.0;
#pragma unroll
            for (int e = 0; e < kCoreEnt; e++) acc = fma(v[e], gu[e], acc);
        }
        return acc;
    };
    // the update of row i (when update) and its new residual
    auto update_row = [&](int i, bool update) {
        double ri = a.r[i];
        if (update) {
            const double pi = fma(beta, a.p[i], ld_coh(&a.u[i]));
            const double si = fma(beta, a.s[i], ld_coh(&a.w[i]));
            a.p[i] = pi;
            a.s[i] = si;
            a.x[i] = fma(alpha, pi, a.x[i]);
            ri = fma(-alpha, si, ri);
            a.r[i] = ri;
        }
        return ri;
    };
    auto blocks_apply = [&](bool update, double &pg, double &pr) {
        // for the blocks of this work-group: the rows' updates, then u_b = M_b^-1 r_b with
        // coherent stores; pg += r.u, pr += r.r over their rows.  Blocks go kCorePairs at a
        // time; waves 2 q and 2 q + 1 share block base + q: each updates and stages half of
        // its rows, multiplies half of its columns, finalises half of its rows.
        const int slot = wave >> 1, half = wave & 1;
        for (int base = b0; base < b1; base += kCorePairs) {
            const int bi = base + slot;
            const bool act = bi < b1;
            const int nbk = act ? a.blk_n[bi] : 0;
            const int q0 = act ? a.blk_row0[bi] : 0;
            const int jr = half * kWave + lane;  // this lane's row of the block
            if (jr < nbk) yb[slot][jr] = update_row(q0 + jr, update);
            __syncthreads();
            if (act) {
                // partial product over this wave's half of the columns: lane i accumulates
                // rows i and i + 64 from column j of the symmetric remainder (consecutive
                // lanes read consecutive addresses); the loads of kCoreUnroll columns are
                // issued before their products, the last batch predicated.  Lanes past the
                // block read a valid address and discard the result.
                const float *M = a.binv + a.blk_off[bi];
                const int hc = (nbk + 1) >> 1;
                const int jlo = half ? hc : 0, jhi = half ? nbk : hc;
                const int c0 = min(lane, nbk - 1), c1 = min(lane + kWave, nbk - 1);
                double z0 = 0.0, z1 = 0.0;
                for (int j = jlo; j < jhi; j += kCoreUnroll) {
                    float m0[kCoreUnroll], m1[kCoreUnroll];
#pragma unroll
                    for (int t = 0; t < kCoreUnroll; t++) {
                        const int jj = min(j + t, jhi - 1);
                        m0[t] = M[(size_t)jj * nbk + c0];
                        m1[t] = M[(size_t)jj * nbk + c1];
                    }
#pragma unroll
                    for (int t = 0; t < kCoreUnroll; t++) {
                        const double yj = j + t < jhi ? yb[slot][j + t] : 0.0;
                        z0 = fma((double)m0[t], yj, z0);
                        z1 = fma((double)m1[t], yj, z1);
                    }
                }
                zp[slot][half][lane] = z0;
                zp[slot][half][lane + kWave] = z1;
            }
            __syncthreads();
            if (act) {
                double sum = (lane < nbk ? yb[slot][lane] : 0.0) +
                             (lane + kWave < nbk ? yb[slot][lane + kWave] : 0.0);
                sum = wave_sum(sum);
                sum = __shfl(sum, 0, kWave);
                if (jr < nbk) {
                    const double v = (a.blk_c[bi] * sum + zp[slot][0][jr]) + zp[slot][1][jr];
                    const double yv = yb[slot][jr];
                    st_coh(&a.u[q0 + jr], v);
                    pg += yv * v;
                    pr += yv * yv;
                }
            }
            __syncthreads();
        }
    };
    // p = u + beta p, s = w + beta s, x += alpha p, r -= alpha s (when update) and u = M^-1 r
    // for the phase-u rows of this work-group; partial r.u and r.r
    auto phase_u = [&](bool update, double &pg, double &pr) {
        pg = 0.0;
        pr = 0.0;
        for (int i = j0 + tid; i < j1; i += kCoreThreads) {
            const double ri = update_row(i, update);
            const double ui = (1.0 / a.diag[i]) * ri;
            st_coh(&a.u[i], ui);
            pg += ri * ui;
            pr += ri * ri;
        }
        blocks_apply(update, pg, pr);
    };
    // w = A u for the phase-w rows of this work-group, partial w.u
    auto phase_w = [&]() {
        double pd = 0.0;
        for (int sl = s0 + wave; sl < s1; sl += kCoreWaves) {
            const int row = w0 + (sl - s0) * kWave + lane;
            const bool live = row < w1;
            double uo;
            const double wv = row_product(sl, row, live, a.u, uo);
            if (live) {
                st_coh(&a.w[row], wv);
                pd += wv * uo;
            }
        }
        return pd;
    };

    if (!done && !started) {
        double v[2];
        phase_u(false, v[0], v[1]);
        if (!grid_sum_slots<2, kCoreWaves, kCoreSums>(v, a.G, a.part, a.ctl, bar++, lds, &flag)) {
            done = 4;
        } else {
            gamma = v[0];
            rr = v[1];
            if (rr <= stop2) {
                done = 1;
            } else {
                double d[1] = {phase_w()};
                if (!grid_sum_slots<1, kCoreWaves, kCoreSums>(d, a.G, a.part, a.ctl, bar++, lds, &flag)) {
                    done = 4;
                } else {
                    alpha = gamma / d[0];
                    beta = 0.0;
                    if (!(alpha > 0.0) || !isfinite(alpha)) done = 5;
                }
            }
        }
    }
    unsigned long long tacc[4] = {0, 0, 0, 0}, tprev = TIMED ? wall_clock64() : 0;
    auto tick = [&](int j) {
        if (TIMED) {
            const unsigned long long now = wall_clock64();
            tacc[j] += now - tprev;
            tprev = now;
        }
    };
    for (int it = 0; it < chunk && !done; it++) {
        double v[2];
        phase_u(true, v[0], v[1]);
        tick(0);
        if (!grid_sum_slots<2, kCoreWaves, kCoreSums>(v, a.G, a.part, a.ctl, bar++, lds, &flag)) {
            done = 4;
            break;
        }
        tick(1);
        const double g2 = v[0];
        rr = v[1];
        k++;
        if (rr <= stop2) {
            done = 1;
            break;
        }
        if (!isfinite(rr)) {
            done = 3;
            break;
        }
        if (!(g2 > 0.0)) {
            done = 5;
            break;
        }
        if (k >= max_iter) {
            done = 2;
            break;
        }
        double d[1] = {phase_w()};
        tick(2);
        if (!grid_sum_slots<1, kCoreWaves, kCoreSums>(d, a.G, a.part, a.ctl, bar++, lds, &flag)) {
            done = 4;
            break;
        }
        tick(3);
        beta = g2 / gamma;
        const double den = d[0] - beta * g2 / alpha;
        if (!(den > 0.0) || !isfinite(den)) {
            done = 5;
            break;
        }
        alpha = g2 / den;
        gamma = g2;
    }
    if (TIMED && tid == 0)
        for (int j = 0; j < 4; j++) tim[4 * g + j] += tacc[j];
    if (g == 0 && tid == 0) {
        st->k = k;
        st->gamma = gamma;
        st->alpha = alpha;
        st->beta = beta;
        st->rr = rr;
        st->done = done;
        st->started = 1;
    }
}

bool core_alloc(cicy_ctx *ctx, void **dst, size_t bytes) {
    *dst = ctx->pool.get(bytes);
    if (!*dst) set_error(ctx, "hipMalloc of %zu bytes (core) failed", bytes);
    return *dst != nullptr;
}

template <typename T>
bool up_vec(cicy_ctx *ctx, cicy_core *z, T **dst, const std::vector<T> &src) {
    const size_t bytes = std::max<size_t>(1, src.size()) * sizeof(T);
    if (!core_alloc(ctx, (void **)dst, bytes)) return false;
    z->bytes += bytes;
    if (src.empty()) return true;
    return h2d_async(ctx, *dst, src.data(), src.size() * sizeof(T), ctx->stream);
}

bool zeros(cicy_ctx *ctx, cicy_core *z, double **dst, size_t count) {
    const size_t bytes = std::max<size_t>(1, count) * sizeof(double);
    if (!core_alloc(ctx, (void **)dst, bytes)) return false;
    z->bytes += bytes;
    return hip_ok(ctx, hipMemsetAsync(*dst, 0, bytes, ctx->stream), "memset (core)");
}

}  // namespace

void core_free(cicy_core *z) {
    if (!z) return;
    void *ptrs[] = {z->ju, z->bu, z->wr, z->perm, z->blk_row0, z->blk_n, z->wsl, z->sl_len,
                    z->ell_col, z->ell_val, z->diag, z->blk_c, z->x, z->r, z->u, z->w, z->p,
                    z->s, z->binv, z->part, z->blk_off, z->sl_off, z->ctl, z->st};
    for (void *q : ptrs) dev_release(z->ctx, q);
    delete z;
}

int core_info(const cicy_core *z, int what) {
    if (!z) return 0;
    switch (what) {
        case 0: return z->nb;
        case 1: return z->blocked_rows;
        case 2: return z->G;
        case 3: return (int)z->inv_entries;
        default: return 0;
    }
}

cicy_core *core_setup(cicy_ctx *ctx, int m, const std::vector<int64_t> &crp,
                      const std::vector<int64_t> &ccol, const std::vector<double> &cval,
                      const std::vector<double> &cdiag, const std::vector<int> &orig,
                      int64_t nblocks, const int64_t *ranges) {
    int cus = 0, per_cu = 0;
    if (!hip_ok(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                           ctx->device),
                "hipDeviceGetAttribute") ||
        !hip_ok(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                         &per_cu, reinterpret_cast<const void *>(k_core_cg<false>), kCoreThreads, 0),
                "hipOccupancyMaxActiveBlocksPerMultiprocessor"))
        return nullptr;
    if (per_cu < 1) {
        set_error(ctx, "core CG kernel cannot be resident");
        return nullptr;
    }
    // the grid sums read at most kSlotLanes slots per lane of one wave
    cus = std::min(cus, kSlotLanes * kWave);
    StageClock clk;
    CoreHost H;
    const char *why = nullptr;
    if (!build_core_host(m, crp, ccol, cval, cdiag, orig, nblocks, ranges, cus, H, &why)) {
        set_error(ctx, "core blocks: %s", why);
        return nullptr;
    }
    long long most = 0;  // slots of the work-group with the most
    for (int g = 0; g < H.G; g++)
        if (H.wsl[g + 1] > H.wsl[g])
            most = std::max(most, H.sl_off[H.wsl[g + 1] - 1] +
                                      (long long)H.sl_len[H.wsl[g + 1] - 1] * kWave -
                                      H.sl_off[H.wsl[g]]);
    const int lds_cols = (int)std::min<long long>(most, kCoreLdsCols);
    if (!hip_ok(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                         &per_cu, reinterpret_cast<const void *>(k_core_cg<false>), kCoreThreads,
                         (size_t)lds_cols * sizeof(int)),
                "hipOccupancyMaxActiveBlocksPerMultiprocessor"))
        return nullptr;
    if (per_cu < 1) {
        set_error(ctx, "core CG kernel cannot be resident with %d staged columns", lds_cols);
        return nullptr;
    }
    clk.mark("core: layout + block inverses");
    cicy_core *z = new cicy_core();
    z->ctx = ctx;
    z->m = m;
    z->G = H.G;
    z->lds_cols = lds_cols;
    z->nb = (int)H.blk_row0.size();
    z->nnz2 = (long long)H.col.size();
    z->inv_entries = (long long)H.binv.size();
    for (int v : H.blk_n) z->blocked_rows += v;
    bool ok = up_vec(ctx, z, &z->ju, H.ju) && up_vec(ctx, z, &z->bu, H.bu) &&
              up_vec(ctx, z, &z->wr, H.wr) && up_vec(ctx, z, &z->perm, H.perm) &&
              up_vec(ctx, z, &z->wsl, H.wsl) &&
              up_vec(ctx, z, &z->sl_off, H.sl_off) && up_vec(ctx, z, &z->sl_len, H.sl_len) &&
              up_vec(ctx, z, &z->ell_col, H.ell_col) && up_vec(ctx, z, &z->ell_val, H.ell_val) &&
              up_vec(ctx, z, &z->diag, H.diag) && up_vec(ctx, z, &z->blk_row0, H.blk_row0) &&
              up_vec(ctx, z, &z->blk_n, H.blk_n) && up_vec(ctx, z, &z->blk_off, H.blk_off) &&
              up_vec(ctx, z, &z->blk_c, H.blk_c) && up_vec(ctx, z, &z->binv, H.binv) &&
              zeros(ctx, z, &z->x, m) &&
              zeros(ctx, z, &z->r, m) && zeros(ctx, z, &z->u, m) && zeros(ctx, z, &z->w, m) &&
              zeros(ctx, z, &z->p, m) && zeros(ctx, z, &z->s, m) &&
              core_alloc(ctx, (void **)&z->part,
                         3 * kCoreSums * sizeof(unsigned long long) * (size_t)H.G) &&
              core_alloc(ctx, (void **)&z->ctl, sizeof(GridCtl)) &&
              core_alloc(ctx, (void **)&z->st, sizeof(CoreState)) &&
              hip_ok(ctx, hipStreamSynchronize(ctx->stream), "sync");
    clk.mark("core: upload");
    if (!ok) {
        core_free(z);
        return nullptr;
    }
    return z;
}

namespace {

// ---- the layout on the device (core_setup_device) ------------------------------------

constexpr int kLT = 256;  // threads of the layout kernels

#define CORE_LOOP(i, n) \
    for (long long i = (long long)blockIdx.x * kLT + threadIdx.x; i < (long long)(n); \
         i += (long long)gridDim.x * kLT)

inline int core_grid(long long n) {
    return (int)std::max<long long>(1, std::min<long long>((n + kLT - 1) / kLT, 1 << 16));
}

#pragma clang fp contract(off)

// out[t] = first core row whose full-system row is >= vals[t] (lower_bound in orig)
__global__ __launch_bounds__(kLT) void k_group_bounds(int nv, const int64_t *__restrict__ vals,
                                                      const int *__restrict__ orig, int m,
                                                      int *__restrict__ out) {
    CORE_LOOP(t, nv) {
        const int64_t v = vals[t];
        int lo = 0, hi = m;
        while (lo < hi) {
            const int mid = lo + (hi - lo) / 2;
            if ((int64_t)orig[mid] < v) lo = mid + 1;
            else hi = mid;
        }
        out[t] = lo;
    }
}

// (i, j), 0 <= j <= i, of the e-th entry of a lower triangle stored by rows
__device__ __forceinline__ void tri_index(int e, int *i, int *j) {
    int a = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
    while ((a + 1) * (a + 2) / 2 <= e) a++;
    while (a * (a + 1) / 2 > e) a--;
    *i = a;
    *j = e - a * (a + 1) / 2;
}

// Dense inverse of one diagonal block per work-group (work-group b takes group order[b],
// the largest first).  Every entry gets dense_spd_inverse's operations in its order, run
// right-looking so that whole steps go in parallel: Cholesky (at step k the column of L,
// then the trailing update a_ij -= l_ik l_jk: each entry sees k = 0, 1, ... as in the host's
// inner loop), W = L^-1 by forward substitution (acc_ij -= l_ik w_kj at step k, k from j up,
// then times 1 / l_ii), A^-1 = W^T W (each entry summed over k ascending by one thread),
// then c = the sum of the row sums (each in column order) over the block's n^2 entries.
// LDS: the block with an odd row stride (ld = n | 1, no bank conflicts on columns): L in
// the lower triangle, W in the strict upper one (W[i][j] at [j][i]) and W's diagonal after
// the block; the inverse's lower triangle then replaces L.
constexpr int kBigInverseThreads = 1024;  // threads per block inverse beside the layout
template <int T>
__global__ __launch_bounds__(T) void k_block_inverse(const int *__restrict__ order,
                                                       const int *__restrict__ gq0,
                                                       const int *__restrict__ gnb,
                                                       const long long *__restrict__ goff,
                                                       const int *__restrict__ frp,
                                                       const int *__restrict__ fcol,
                                                       const double *__restrict__ fval,
                                                       const double *__restrict__ cdiag,
                                                       double *__restrict__ inv,
                                                       int *__restrict__ ok,
                                                       double *__restrict__ cg) {
    extern __shared__ double A[];
    __shared__ int fail;
    const int g = order[blockIdx.x], q0 = gq0[g], nb = gnb[g], tid = threadIdx.x;
    const int ld = nb | 1;
    double *wd = A + (size_t)nb * ld;
    double *rs = wd + nb;
    double *out = inv + goff[g];
    for (int e = tid; e < nb * ld; e += T) A[e] = 0.0;
    if (tid == 0) fail = 0;
    __syncthreads();
    for (int i = tid; i < nb; i += T) {
        A[i * ld + i] = cdiag[q0 + i];
        for (int k = frp[q0 + i]; k < frp[q0 + i + 1]; k++) {
            const int c = fcol[k] - q0;
            if (c >= 0 && c < nb) A[i * ld + c] = fval[k];
        }
    }
    __syncthreads();
    // Cholesky
    for (int k = 0; k < nb; k++) {
        if (tid == 0) {
            const double d = A[k * ld + k];
            if (!(d > 0.0) || !isfinite(d)) fail = 1;
            else A[k * ld + k] = sqrt(d);
        }
        __syncthreads();
        if (fail) break;
        const double l = A[k * ld + k];
        for (int i = k + 1 + tid; i < nb; i += T) A[i * ld + k] = A[i * ld + k] / l;
        __syncthreads();
        const int r = nb - k - 1, cnt = r * (r + 1) / 2;
        for (int e = tid; e < cnt; e += T) {
            int a, c;
            tri_index(e, &a, &c);
            const int i = k + 1 + a, j = k + 1 + c;
            A[i * ld + j] = A[i * ld + j] - A[i * ld + k] * A[j * ld + k];
        }
        __syncthreads();
    }
    if (fail) {
        if (tid == 0) {
            ok[g] = 0;
            cg[g] = 0.0;
        }
        return;
    }
    // W = L^-1: the strict upper triangle holds the accumulators, then W
    for (int e = tid; e < nb * ld; e += T) {
        const int i = e / ld, j = e % ld;
        if (j > i && j < nb) A[e] = 0.0;
    }
    __syncthreads();
    for (int k = 0; k < nb; k++) {
        const double inv_d = 1.0 / A[k * ld + k];
        for (int j = tid; j < k; j += T) A[j * ld + k] = A[j * ld + k] * inv_d;
        if (tid == 0) wd[k] = inv_d;
        __syncthreads();
        const int rows = nb - k - 1, cols = k + 1;
        for (int e = tid; e < rows * cols; e += T) {
            const int i = k + 1 + e / cols, j = e % cols;
            const double wkj = j == k ? wd[k] : A[j * ld + k];
            A[j * ld + i] = A[j * ld + i] - A[i * ld + k] * wkj;
        }
        __syncthreads();
    }
    // A^-1 = W^T W into the lower triangle (L is no longer needed) and to global memory
    const int ntri = nb * (nb + 1) / 2;
    for (int e = tid; e < ntri; e += T) {
        int i, j;
        tri_index(e, &i, &j);
        double acc = 0.0;
        for (int k = i; k < nb; k++) {
            const double wki = k == i ? wd[i] : A[i * ld + k];
            const double wkj = k == j ? wd[j] : A[j * ld + k];
            acc = acc + wki * wkj;
        }
        A[i * ld + j] = acc;
        out[(size_t)i * nb + j] = acc;
        out[(size_t)j * nb + i] = acc;
    }
    __syncthreads();
    for (int i = tid; i < nb; i += T) {
        double t = 0.0;
        for (int j = 0; j < nb; j++) t = t + (j <= i ? A[i * ld + j] : A[j * ld + i]);
        rs[i] = t;
    }
    __syncthreads();
    if (tid == 0) {
        double c = 0.0;
        for (int i = 0; i < nb; i++) c = c + rs[i];
        const double dn = nb;
        cg[g] = c / (dn * dn);
        ok[g] = 1;
    }
}

// kept block b: its rows' new numbers (STRUCT), or its constant and FP32 remainder
template <bool STRUCT>
__global__ __launch_bounds__(kLT) void k_block_keep(const int *__restrict__ kq0,
                                                    const int *__restrict__ knb,
                                                    const int *__restrict__ krow0,
                                                    const long long *__restrict__ kgoff,
                                                    const long long *__restrict__ kboff,
                                                    const int *__restrict__ kg,
                                                    const double *__restrict__ inv,
                                                    const double *__restrict__ cg,
                                                    int *__restrict__ in_blk,
                                                    int *__restrict__ iperm, int *__restrict__ perm,
                                                    double *__restrict__ blk_c,
                                                    float *__restrict__ binv) {
    const int b = blockIdx.x, q0 = kq0[b], nb = knb[b], row0 = krow0[b];
    if (STRUCT) {
        for (int t = threadIdx.x; t < nb; t += kLT) {
            in_blk[q0 + t] = 1;
            iperm[q0 + t] = row0 + t;
            perm[row0 + t] = q0 + t;
        }
        return;
    }
    const double c = cg[kg[b]];
    const double *src = inv + kgoff[b];
    float *dst = binv + kboff[b];
    for (int e = threadIdx.x; e < nb * nb; e += kLT) dst[e] = (float)(src[e] - c);
    if (threadIdx.x == 0) blk_c[b] = c;
}

__global__ __launch_bounds__(kLT) void k_jacobi_flags(int m, const int *__restrict__ in_blk,
                                                      int *__restrict__ flag) {
    CORE_LOOP(q, m + 1) flag[q] = q < m && !in_blk[q];
}

__global__ __launch_bounds__(kLT) void k_jacobi_perm(int m, const int *__restrict__ in_blk,
                                                     const int *__restrict__ jpos,
                                                     int *__restrict__ iperm,
                                                     int *__restrict__ perm) {
    CORE_LOOP(q, m) {
        if (in_blk[q]) continue;
        iperm[q] = jpos[q];
        perm[jpos[q]] = (int)q;
    }
}

__global__ __launch_bounds__(kLT) void k_perm_len(int m, const int *__restrict__ perm,
                                                  const int *__restrict__ frp,
                                                  int *__restrict__ len) {
    CORE_LOOP(i, m + 1) len[i] = i < m ? frp[perm[i] + 1] - frp[perm[i]] : 0;
}

// Row i of the renumbered layout is row perm[i]: its Jacobi columns (renumbered below mj,
// in the old order, which the renumbering keeps) then its block columns.
__global__ __launch_bounds__(kLT) void k_permute_rows(int m, int mj, const int *__restrict__ perm,
                                                      const int *__restrict__ iperm,
                                                      const int *__restrict__ frp,
                                                      const int *__restrict__ fcol,
                                                      const double *__restrict__ fval,
                                                      const double *__restrict__ cdiag,
                                                      const int *__restrict__ rp2,
                                                      int *__restrict__ col2,
                                                      double *__restrict__ val2,
                                                      double *__restrict__ diag2) {
    CORE_LOOP(i, m) {
        const int o = perm[i];
        int d = rp2[i];
        for (int k = frp[o]; k < frp[o + 1]; k++) {
            const int c = iperm[fcol[k]];
            if (c < mj) {
                col2[d] = c;
                val2[d++] = fval[k];
            }
        }
        for (int k = frp[o]; k < frp[o + 1]; k++) {
            const int c = iperm[fcol[k]];
            if (c >= mj) {
                col2[d] = c;
                val2[d++] = fval[k];
            }
        }
        diag2[i] = cdiag[o];
    }
}

// sliced ELL: slice q, lane l holds row sl_base[q] + l (padding: column -1, value 0)
__global__ __launch_bounds__(kLT) void k_fill_ell(int nsl, const int *__restrict__ sl_base,
                                                  const int *__restrict__ sl_end,
                                                  const long long *__restrict__ sl_off,
                                                  const int *__restrict__ sl_len,
                                                  const int *__restrict__ rp2,
                                                  const int *__restrict__ col2,
                                                  const double *__restrict__ val2,
                                                  int *__restrict__ ell_col,
                                                  double *__restrict__ ell_val) {
    CORE_LOOP(t, (long long)nsl * kWave) {
        const int q = (int)(t / kWave), lane = (int)(t % kWave);
        const int r = sl_base[q] + lane;
        const int len = r < sl_end[q] ? rp2[r + 1] - rp2[r] : 0;
        const long long o = sl_off[q];
        for (int k = 0; k < sl_len[q]; k++) {
            const long long pos = o + (long long)k * kWave + lane;
            ell_col[pos] = k < len ? col2[rp2[r] + k] : -1;
            ell_val[pos] = k < len ? val2[rp2[r] + k] : 0.0;
        }
    }
}

#pragma clang fp contract(on)

template <typename T>
bool core_dev(cicy_ctx *ctx, cicy_core *z, T **dst, size_t count) {
    const size_t bytes = std::max<size_t>(1, count) * sizeof(T);
    if (!core_alloc(ctx, (void **)dst, bytes)) return false;
    z->bytes += bytes;
    return true;
}

// the kernel's residency: work-groups (at most one per CU, at most the grid sums' slots)
bool core_groups(cicy_ctx *ctx, int *cus) {
    int per_cu = 0;
    if (!hip_ok(ctx, hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount,
                                           ctx->device),
                "hipDeviceGetAttribute") ||
        !hip_ok(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                         &per_cu, reinterpret_cast<const void *>(k_core_cg<false>), kCoreThreads, 0),
                "hipOccupancyMaxActiveBlocksPerMultiprocessor"))
        return false;
    if (per_cu < 1) {
        set_error(ctx, "core CG kernel cannot be resident");
        return false;
    }
    *cus = std::min(*cus, kSlotLanes * kWave);
    // CICY_CORE_GROUPS caps the work-groups (a tuning knob; the layout follows it)
    if (const char *e = getenv("CICY_CORE_GROUPS")) *cus = std::max(1, std::min(*cus, atoi(e)));
    return true;
}

// staged LDS columns per work-group (its first slots, at most kCoreLdsCols), checked for
// residency
bool core_lds_cols(cicy_ctx *ctx, const CoreHost &H, int *lds_cols) {
    long long most = 0;  // slots of the work-group with the most
    for (int g = 0; g < H.G; g++)
        if (H.wsl[g + 1] > H.wsl[g])
            most = std::max(most, H.sl_off[H.wsl[g + 1] - 1] +
                                      (long long)H.sl_len[H.wsl[g + 1] - 1] * kWave -
                                      H.sl_off[H.wsl[g]]);
    *lds_cols = (int)std::min<long long>(most, kCoreLdsCols);
    int per_cu = 0;
    if (!hip_ok(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                         &per_cu, reinterpret_cast<const void *>(k_core_cg<false>), kCoreThreads,
                         (size_t)*lds_cols * sizeof(int)),
                "hipOccupancyMaxActiveBlocksPerMultiprocessor"))
        return false;
    if (per_cu < 1) {
        set_error(ctx, "core CG kernel cannot be resident with %d staged columns", *lds_cols);
        return false;
    }
    return true;
}

}  // namespace

// The layout of core_setup built on the device from the core's symmetric CSR (frp, fcol,
// fval: both triangles, columns ascending), its diagonal and the full-system row of each
// core row (ascending), all on the device: equal to build_core_host's bit for bit.  The
// caller's row groups (ranges, host) are cut and checked as there; the dense inverses, the
// renumbering, the renumbered rows and the sliced ELL are built on the device, the
// work-group partitions on the host from the renumbered row lengths.
cicy_core *core_setup_device(cicy_ctx *ctx, int m, const int *frp, const int *fcol,
                             const double *fval, const double *cdiag, const int *orig,
                             int64_t nblocks, const int64_t *ranges) {
    hipStream_t st = ctx->stream;
    int cus = 0;
    if (!core_groups(ctx, &cus)) return nullptr;
    StageClock clk;
    PoolScope ps(ctx);
    // row groups -> core row ranges (split at core_max_block() rows)
    const int maxb = core_max_block();
    for (int64_t b = 0; b < nblocks; b++)
        if (ranges[2 * b] < 0 || ranges[2 * b + 1] < ranges[2 * b]) {
            set_error(ctx, "core blocks: a row group has a negative start or ends before it "
                           "starts");
            return nullptr;
        }
    std::vector<int> qb((size_t)std::max<int64_t>(2 * nblocks, 1));
    if (nblocks > 0) {
        int64_t *dr;
        int *dq;
        if (!ps.get(dr, 2 * (size_t)nblocks) || !ps.get(dq, 2 * (size_t)nblocks) ||
            !h2d_async(ctx, dr, ranges, 2 * (size_t)nblocks * sizeof(int64_t), st))
            return nullptr;
        hipLaunchKernelGGL(k_group_bounds, dim3(core_grid(2 * nblocks)), dim3(kLT), 0, st,
                           (int)(2 * nblocks), dr, orig, m, dq);
        if (!hip_ok(ctx, hipGetLastError(), "k_group_bounds") ||
            !d2h_sync(ctx, qb.data(), dq, 2 * (size_t)nblocks * sizeof(int)))
            return nullptr;
    }
    clk.mark("device core: group bounds");
    std::vector<std::pair<int, int>> groups;
    for (int64_t b = 0; b < nblocks; b++) {
        const int q0 = qb[(size_t)(2 * b)], q1 = qb[(size_t)(2 * b + 1)];
        for (int q = q0; q < q1; q += maxb) {
            const int e = std::min(q1, q + maxb);
            if (e - q >= 2) groups.emplace_back(q, e);
        }
    }
    // the caller's groups usually come in row order already
    if (!std::is_sorted(groups.begin(), groups.end())) std::sort(groups.begin(), groups.end());
    for (size_t g = 1; g < groups.size(); g++)
        if (groups[g].first < groups[g - 1].second) {
            set_error(ctx, "core blocks: row groups overlap");
            return nullptr;
        }
    const int ng = (int)groups.size();
    std::vector<int> gq0((size_t)ng), gnb((size_t)ng);
    std::vector<long long> goff((size_t)ng + 1, 0);
    int nbmax = 0;
    for (int g = 0; g < ng; g++) {
        gq0[(size_t)g] = groups[(size_t)g].first;
        gnb[(size_t)g] = groups[(size_t)g].second - groups[(size_t)g].first;
        nbmax = std::max(nbmax, gnb[(size_t)g]);
        goff[(size_t)g + 1] = goff[(size_t)g] + (long long)gnb[(size_t)g] * gnb[(size_t)g];
    }
    // dense inverses
    int *dgq0, *dgnb, *dok, *dord;
    long long *dgoff;
    double *dinv, *dcg;
    // groups by descending size, equal sizes in group order (a counting sort: sizes are at
    // most nbmax)
    std::vector<int> okh((size_t)ng, 0), gord((size_t)ng);
    {
        std::vector<int> start((size_t)nbmax + 2, 0);
        for (int g = 0; g < ng; g++) start[(size_t)(nbmax - gnb[(size_t)g]) + 1]++;
        for (int v = 0; v <= nbmax; v++) start[(size_t)v + 1] += start[(size_t)v];
        for (int g = 0; g < ng; g++) gord[(size_t)start[(size_t)(nbmax - gnb[(size_t)g])]++] = g;
    }
    if (!ps.get(dgq0, (size_t)ng) || !ps.get(dgnb, (size_t)ng) || !ps.get(dok, (size_t)ng) ||
        !ps.get(dord, (size_t)ng) ||
        !ps.get(dgoff, (size_t)ng + 1) || !ps.get(dinv, (size_t)goff[(size_t)ng]) ||
        !ps.get(dcg, (size_t)ng))
        return nullptr;
    clk.mark("device core: groups + buffers");
    // The inverses run on the second stream while this stream lays the core out on the
    // assumption that every block is kept (a block whose Cholesky fails goes back to Jacobi,
    // which changes the layout): the layout is rebuilt in the rare case that one failed.
    const bool overlap = ng > 0 && ctx->stream2 != nullptr;
    if (ng > 0) {
        const size_t lds = ((size_t)nbmax * (nbmax | 1) + 2 * (size_t)nbmax) * sizeof(double);
        if (!h2d_async(ctx, dgq0, gq0.data(), (size_t)ng * sizeof(int), st) ||
            !h2d_async(ctx, dgnb, gnb.data(), (size_t)ng * sizeof(int), st) ||
            !h2d_async(ctx, dord, gord.data(), (size_t)ng * sizeof(int), st) ||
            !h2d_async(ctx, dgoff, goff.data(), ((size_t)ng + 1) * sizeof(long long), st) ||
            !hip_ok(ctx, hipFuncSetAttribute(reinterpret_cast<const void *>(k_block_inverse<kLT>),
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)lds), "hipFuncSetAttribute") ||
            !hip_ok(ctx, hipFuncSetAttribute(
                             reinterpret_cast<const void *>(k_block_inverse<kBigInverseThreads>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                    "hipFuncSetAttribute"))
            return nullptr;
        clk.mark("device core: group uploads");
        if (overlap) {
            if (!hip_ok(ctx, hipEventRecord(ctx->ev_fork, st), "event") ||
                !hip_ok(ctx, hipStreamWaitEvent(ctx->stream2, ctx->ev_fork, 0), "wait"))
                return nullptr;
            // all the inverses on the second stream, largest first, so this stream goes
            // straight to the layout; a block's inverse is a chain of barrier-separated steps,
            // and more threads per block shorten each step (every entry still gets the same
            // operations), while the small blocks fill the CUs the large ones leave
            hipLaunchKernelGGL(k_block_inverse<kBigInverseThreads>, dim3(ng),
                               dim3(kBigInverseThreads), lds, ctx->stream2, dord, dgq0, dgnb,
                               dgoff, frp, fcol, fval, cdiag, dinv, dok, dcg);
            if (!hip_ok(ctx, hipGetLastError(), "k_block_inverse") ||
                !hip_ok(ctx, hipEventRecord(ctx->ev_join, ctx->stream2), "event"))
                return nullptr;
            std::fill(okh.begin(), okh.end(), 1);
        } else {
            hipLaunchKernelGGL(k_block_inverse<kLT>, dim3(ng), dim3(kLT), lds, st, dord, dgq0, dgnb,
                               dgoff, frp, fcol, fval, cdiag, dinv, dok, dcg);
            if (!hip_ok(ctx, hipGetLastError(), "k_block_inverse") ||
                !d2h_sync(ctx, okh.data(), dok, (size_t)ng * sizeof(int)))
                return nullptr;
        }
    }
    clk.mark("device core: block inverses");
    // the layout for the blocks kept (okh[g] != 0); after_join: the inverses are still running
    // on the second stream, joined before their values are copied
    auto layout = [&](const std::vector<int> &okh, bool after_join) -> cicy_core * {
        // kept blocks, numbered after the Jacobi rows in group order
        CoreHost H;
        H.m = m;
        std::vector<int> kq0, knb, krow0, kg;
        std::vector<long long> kgoff, kboff;
        int blocked = 0;
        for (int g = 0; g < ng; g++)
            if (okh[(size_t)g]) blocked += gnb[(size_t)g];
        H.mj = m - blocked;
        {
            int row0 = H.mj;
            long long at = 0;
            for (int g = 0; g < ng; g++) {
                if (!okh[(size_t)g]) continue;
                const int nb = gnb[(size_t)g];
                kq0.push_back(gq0[(size_t)g]);
                knb.push_back(nb);
                krow0.push_back(row0);
                kg.push_back(g);
                kgoff.push_back(goff[(size_t)g]);
                kboff.push_back(at);
                H.blk_row0.push_back(row0);
                H.blk_n.push_back(nb);
                H.blk_off.push_back(at);
                at += (long long)nb * nb;
                row0 += nb;
            }
        }
        const int nbk = (int)kq0.size();
        cicy_core *z = new cicy_core();
        z->ctx = ctx;
        z->m = m;
        z->nb = nbk;
        z->blocked_rows = blocked;
        z->inv_entries = nbk > 0 ? kboff.back() + (long long)knb.back() * knb.back() : 0;
        auto fail = [&]() -> cicy_core * {
            core_free(z);
            return nullptr;
        };
        int *in_blk, *iperm, *jflag, *jpos, *len2, *rp2, *col2, *dk[4];
        long long *dkl[2];
        double *val2;
        if (!ps.get(in_blk, (size_t)m + 1) || !ps.get(iperm, (size_t)m + 1) ||
            !ps.get(jflag, (size_t)m + 1) || !ps.get(jpos, (size_t)m + 1) ||
            !ps.get(len2, (size_t)m + 1) || !ps.get(rp2, (size_t)m + 1) ||
            !core_dev(ctx, z, &z->perm, (size_t)m) || !core_dev(ctx, z, &z->diag, (size_t)m) ||
            !core_dev(ctx, z, &z->blk_c, (size_t)nbk) ||
            !core_dev(ctx, z, &z->binv, (size_t)z->inv_entries) ||
            !hip_ok(ctx, hipMemsetAsync(in_blk, 0, ((size_t)m + 1) * sizeof(int), st), "memset"))
            return fail();
        for (int k = 0; k < 4; k++)
            if (!ps.get(dk[k], (size_t)nbk)) return fail();
        for (int k = 0; k < 2; k++)
            if (!ps.get(dkl[k], (size_t)nbk)) return fail();
        if (nbk > 0) {
            const std::vector<int> *iv[] = {&kq0, &knb, &krow0, &kg};
            for (int k = 0; k < 4; k++)
                if (!h2d_async(ctx, dk[k], iv[k]->data(), (size_t)nbk * sizeof(int), st))
                    return fail();
            if (!h2d_async(ctx, dkl[0], kgoff.data(), (size_t)nbk * sizeof(long long), st) ||
                !h2d_async(ctx, dkl[1], kboff.data(), (size_t)nbk * sizeof(long long), st))
                return fail();
            hipLaunchKernelGGL(k_block_keep<true>, dim3(nbk), dim3(kLT), 0, st, dk[0], dk[1], dk[2],
                               dkl[0], dkl[1], dk[3], dinv, dcg, in_blk, iperm, z->perm, z->blk_c,
                               z->binv);
        }
        hipLaunchKernelGGL(k_jacobi_flags, dim3(core_grid(m + 1)), dim3(kLT), 0, st, m, in_blk, jflag);
        size_t tmp_bytes = 0;
        unsigned char *tmp = nullptr;
        auto scan = [&](const int *in, int *out, int count) {
            tmp_bytes = 0;
            return hip_ok(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, in, out, count, st),
                          "scan") &&
                   ps.get(tmp, tmp_bytes) &&
                   hip_ok(ctx, hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, in, out, count, st),
                          "scan");
        };
        if (!scan(jflag, jpos, m + 1)) return fail();
        hipLaunchKernelGGL(k_jacobi_perm, dim3(core_grid(m)), dim3(kLT), 0, st, m, in_blk, jpos, iperm,
                           z->perm);
        hipLaunchKernelGGL(k_perm_len, dim3(core_grid(m + 1)), dim3(kLT), 0, st, m, z->perm, frp, len2);
        if (!scan(len2, rp2, m + 1)) return fail();
        H.rp.resize((size_t)m + 1);
        if (!hip_ok(ctx, hipGetLastError(), "core renumbering") ||
            !d2h_sync(ctx, H.rp.data(), rp2, ((size_t)m + 1) * sizeof(int)))
            return fail();
        const int nnz2 = H.rp[(size_t)m];
        if (!ps.get(col2, (size_t)nnz2) || !ps.get(val2, (size_t)nnz2)) return fail();
        hipLaunchKernelGGL(k_permute_rows, dim3(core_grid(m)), dim3(kLT), 0, st, m, H.mj, z->perm,
                           iperm, frp, fcol, fval, cdiag, rp2, col2, val2, z->diag);
        clk.mark("device core: renumbered rows");
        // partitions and slices on the host (from the row lengths), the ELL on the device
        std::vector<int> sl_base, sl_end;
        const size_t ell = core_partitions(H, cus, sl_base, sl_end);
        if (clk.on) {
            int longest = 0;
            for (int r = 0; r < m; r++) longest = std::max(longest, H.rp[(size_t)r + 1] - H.rp[(size_t)r]);
            fprintf(stderr, "[cicy setup]   core: %d rows, %d entries, %zu ELL slots (x%.2f), "
                    "longest row %d, %d Jacobi rows, %zu blocks\n", m, nnz2, ell,
                    nnz2 > 0 ? (double)ell / nnz2 : 0.0, longest, H.mj, H.blk_n.size());
        }
        z->G = H.G;
        z->nnz2 = nnz2;
        int lds_cols = 0;
        if (!core_lds_cols(ctx, H, &lds_cols)) return fail();
        z->lds_cols = lds_cols;
        const int nsl = (int)sl_base.size();
        int *dsb, *dse;
        if (!ps.get(dsb, (size_t)nsl) || !ps.get(dse, (size_t)nsl) ||
            !up_vec(ctx, z, &z->ju, H.ju) || !up_vec(ctx, z, &z->bu, H.bu) ||
            !up_vec(ctx, z, &z->wr, H.wr) || !up_vec(ctx, z, &z->wsl, H.wsl) ||
            !up_vec(ctx, z, &z->sl_off, H.sl_off) || !up_vec(ctx, z, &z->sl_len, H.sl_len) ||
            !up_vec(ctx, z, &z->blk_row0, H.blk_row0) || !up_vec(ctx, z, &z->blk_n, H.blk_n) ||
            !up_vec(ctx, z, &z->blk_off, H.blk_off) || !core_dev(ctx, z, &z->ell_col, ell) ||
            !core_dev(ctx, z, &z->ell_val, ell) ||
            (nsl > 0 && (!h2d_async(ctx, dsb, sl_base.data(), (size_t)nsl * sizeof(int), st) ||
                         !h2d_async(ctx, dse, sl_end.data(), (size_t)nsl * sizeof(int), st))))
            return fail();
        if (nsl > 0)
            hipLaunchKernelGGL(k_fill_ell, dim3(core_grid((long long)nsl * kWave)), dim3(kLT), 0, st,
                               nsl, dsb, dse, z->sl_off, z->sl_len, rp2, col2, val2, z->ell_col,
                               z->ell_val);
        // the inverses' values: after the second stream's inverses when they ran beside this
        if (after_join && !hip_ok(ctx, hipStreamWaitEvent(st, ctx->ev_join, 0), "wait"))
            return fail();
        if (nbk > 0)
            hipLaunchKernelGGL(k_block_keep<false>, dim3(nbk), dim3(kLT), 0, st, dk[0], dk[1], dk[2],
                               dkl[0], dkl[1], dk[3], dinv, dcg, in_blk, iperm, z->perm, z->blk_c,
                               z->binv);
        const bool ok = hip_ok(ctx, hipGetLastError(), "k_fill_ell") &&
                        zeros(ctx, z, &z->x, m) && zeros(ctx, z, &z->r, m) && zeros(ctx, z, &z->u, m) &&
                        zeros(ctx, z, &z->w, m) && zeros(ctx, z, &z->p, m) && zeros(ctx, z, &z->s, m) &&
                        core_alloc(ctx, (void **)&z->part,
                                   3 * kCoreSums * sizeof(unsigned long long) * (size_t)H.G) &&
                        core_alloc(ctx, (void **)&z->ctl, sizeof(GridCtl)) &&
                        core_alloc(ctx, (void **)&z->st, sizeof(CoreState)) &&
                        hip_ok(ctx, hipStreamSynchronize(st), "sync");
        clk.mark("device core: slices + upload");
        if (!ok) return fail();
        return z;
    };
    cicy_core *z = layout(okh, overlap);
    if (!z && overlap) {  // the second stream's inverses still use this call's temporaries
        (void)hipStreamSynchronize(ctx->stream2);
        return nullptr;
    }
    if (z && overlap) {
        std::vector<int> okd((size_t)ng, 0);
        if (!d2h_sync(ctx, okd.data(), dok, (size_t)ng * sizeof(int))) {
            core_free(z);
            return nullptr;
        }
        if (okd != okh) {  // a block went back to Jacobi: lay the core out again
            core_free(z);
            z = layout(okd, false);
        }
    }
    return z;
}

// Solves A x = rhs (core rows, on the device) from x = 0 until ||r|| <= tol sqrt(ref2);
// the solution goes to x_out.  stats: iters and status.
bool solve_core(cicy_ctx *ctx, cicy_core *z, const double *rhs, double tol, long long max_iter,
                double ref2, double *x_out, cicy_solve_stats *stats) {
    hipStream_t stream = ctx->stream;
    const size_t vb = (size_t)z->m * sizeof(double);
    const double max_seconds = [] {
        const char *e = getenv("CICY_MAX_SOLVE_SECONDS");
        return e ? atof(e) : 0.0;
    }();
    auto t0 = std::chrono::steady_clock::now();
    CoreState h;
    memset(&h, 0, sizeof(h));
    h.max_iter = max_iter > 0 ? max_iter : 1;
    h.stop2 = tol * tol * ref2;
    const int pg = (z->m + 255) / 256;
    hipLaunchKernelGGL(k_core_permute, dim3(pg), dim3(256), 0, stream, z->m, z->perm, rhs, z->r,
                       0);
    bool ok = hip_ok(ctx, hipGetLastError(), "k_core_permute") &&
              hip_ok(ctx, hipMemsetAsync(z->x, 0, vb, stream), "memset") &&
              hip_ok(ctx, hipMemsetAsync(z->s, 0, vb, stream), "memset") &&
              hip_ok(ctx, hipMemsetAsync(z->p, 0, vb, stream), "memset") &&
              hip_ok(ctx, hipMemcpyAsync(z->st, &h, sizeof(h), hipMemcpyHostToDevice, stream),
                     "H2D");
    CoreDev a{z->m, z->G, z->ju, z->bu, z->wr, z->wsl, z->sl_off, z->lds_cols, z->sl_len, z->ell_col,
              z->ell_val, z->diag, z->blk_row0, z->blk_n, z->blk_off, z->blk_c, z->binv,
              z->x, z->r, z->u, z->w, z->p, z->s, z->ctl, z->part};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ok = ok && hip_ok(ctx, hipEventCreate(&e0), "hipEventCreate") &&
         hip_ok(ctx, hipEventCreate(&e1), "hipEventCreate") &&
         hip_ok(ctx, hipEventRecord(e0, stream), "event");
    const bool timed = getenv("CICY_CORE_TIMING") != nullptr;
    unsigned long long *tim = nullptr;
    if (timed) {
        tim = static_cast<unsigned long long *>(ctx->pool.get(4 * sizeof(unsigned long long) *
                                                              (size_t)z->G));
        ok = ok && tim &&
             hip_ok(ctx, hipMemsetAsync(tim, 0, 4 * sizeof(unsigned long long) * (size_t)z->G,
                                        stream), "memset");
    }
    const void *kernel = timed ? reinterpret_cast<const void *>(k_core_cg<true>)
                               : reinterpret_cast<const void *>(k_core_cg<false>);
    while (ok && !h.done) {
        int chunk = kCoreChunk;
        void *args[] = {(void *)&a, (void *)&z->st, (void *)&chunk, (void *)&tim};
        ok = hip_ok(ctx, hipMemsetAsync(z->ctl, 0, sizeof(GridCtl), stream), "memset") &&
             hip_ok(ctx, hipMemsetD32Async((hipDeviceptr_t)z->part, (int)kSlotEmptyWord,
                                           3 * kCoreSums * 2 * (size_t)z->G, stream),
                    "memset") &&
             hip_ok(ctx, hipLaunchCooperativeKernel(kernel, dim3(z->G), dim3(kCoreThreads), args,
                                                    (size_t)z->lds_cols * sizeof(int), stream),
                    "hipLaunchCooperativeKernel (core CG)") &&
             d2h_sync(ctx, &h, z->st, sizeof(h));
        if (ok && !h.done && max_seconds > 0 &&
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() >
                max_seconds)
            break;
    }
    float kms = 0.f;
    if (ok && hip_ok(ctx, hipEventRecord(e1, stream), "event") &&
        hip_ok(ctx, hipEventSynchronize(e1), "event"))
        (void)hipEventElapsedTime(&kms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (timed && tim) {
        // per-iteration time of each phase, mean and largest over the work-groups
        std::vector<unsigned long long> t(4 * (size_t)z->G);
        int khz = 0;
        if (ok && d2h_sync(ctx, t.data(), tim, t.size() * sizeof(unsigned long long)) &&
            hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device) ==
                hipSuccess &&
            khz > 0 && h.k > 0) {
            const char *names[4] = {"phase u", "grid sum A", "phase w", "grid sum B"};
            for (int j = 0; j < 4; j++) {
                double mean = 0.0, most = 0.0;
                for (int g = 0; g < z->G; g++) {
                    const double us = (double)t[4 * (size_t)g + j] * 1e3 / khz / (double)h.k;
                    mean += us / z->G;
                    most = std::max(most, us);
                }
                fprintf(stderr, "[cicy core timing] %-10s mean %6.2f us  max %6.2f us per "
                        "iteration (%lld iterations, %d work-groups)\n", names[j], mean, most,
                        h.k, z->G);
            }
        }
        ctx->pool.put(tim);
    }
    if (ok) {
        hipLaunchKernelGGL(k_core_permute, dim3(pg), dim3(256), 0, stream, z->m, z->perm, z->x,
                           x_out, 1);
        ok = hip_ok(ctx, hipGetLastError(), "k_core_permute");
    }
    memset(stats, 0, sizeof(*stats));
    stats->kernel_ms = kms;
    stats->iters = h.k;
    // a stalled recursion is a soft stop (the caller corrects from the true residual)
    stats->status = !ok            ? CICY_DEVICE_ERROR
                    : h.done == 1 ? CICY_OK
                    : h.done == 3 ? CICY_BREAKDOWN
                    : h.done == 4 ? CICY_DEVICE_ERROR
                                  : CICY_MAXITER;
    stats->refinements = h.done == 5 ? 1 : 0;
    if (ok && h.done == 4) {
        set_error(ctx, "core CG: grid barrier timed out (work-groups not co-resident)");
        return false;
    }
    return ok;
}

}  // namespace cicy

// ------------------------------------------------------------------------------------
// C-ABI test hook: the layout and the kernel's recurrences on the host
// ------------------------------------------------------------------------------------

extern "C" int cicy_host_core_block_cg(int64_t m, const int64_t *row_ptr, const int64_t *col_ind,
                                       const double *values, const double *diagonal,
                                       int64_t nnz, int64_t nblocks, const int64_t *ranges,
                                       const double *rhs, double tol, double ref2,
                                       int64_t max_iter, double *x, int64_t *iters,
                                       int64_t *blocked_rows) {
    using namespace cicy;
    if (m <= 0 || !row_ptr || !diagonal || !rhs || !x || row_ptr[m] != nnz) return 0;
    std::vector<int64_t> crp(row_ptr, row_ptr + m + 1), ccol(col_ind, col_ind + nnz);
    std::vector<double> cval(values, values + nnz), cdiag(diagonal, diagonal + m);
    std::vector<int> orig((size_t)m);
    for (int64_t i = 0; i < m; i++) orig[i] = (int)i;
    CoreHost H;
    const char *why = nullptr;
    if (!build_core_host((int)m, crp, ccol, cval, cdiag, orig, nblocks, ranges, 256, H, &why))
        return 0;
    // passes from the true residual, as the condensed solve corrects after a stalled one
    const long long cap = max_iter > 0 ? max_iter : 10 * m + 100;
    const double stop2 = tol * tol * ref2;
    // in the layout's row order (Jacobi rows first, then block by block)
    std::vector<double> b((size_t)m), xh((size_t)m, 0.0), res((size_t)m), dx((size_t)m),
        ax((size_t)m);
    for (int64_t i = 0; i < m; i++) b[i] = res[i] = rhs[H.perm[i]];
    long long k = 0;
    int status = CICY_MAXITER;
    for (int pass = 0; pass < kCorePasses && k < cap; pass++) {
        int st = CICY_OK;
        k += host_core_cg(H, res.data(), tol, ref2, cap - k, dx.data(), &st);
        for (int64_t i = 0; i < m; i++) xh[i] += dx[i];
        host_core_apply(H, xh.data(), ax.data());
        double rr = 0.0;
        for (int64_t i = 0; i < m; i++) {
            res[i] = b[i] - ax[i];
            rr += res[i] * res[i];
        }
        if (rr <= stop2) {
            status = CICY_OK;
            break;
        }
        if (st == CICY_BREAKDOWN) {
            status = CICY_BREAKDOWN;
            break;
        }
    }
    for (int64_t i = 0; i < m; i++) x[H.perm[i]] = xh[i];
    if (iters) *iters = k;
    if (blocked_rows) {
        long long br = 0;
        for (int v : H.blk_n) br += v;
        *blocked_rows = br;
    }
    return status == CICY_OK ? 1 : 0;
}
