// dea_setup.hip -- a system's set-up on the device, from arrays resident in HBM.
//
// The reference hands its solver the SSS arrays of the reduced system (row_ptr, col_ind,
// values, diagonal: From_2d_to_1d_vectors, Direct_Electrifying.cpp:1329-1394) and solves
// each (cluster, direction) system once (Electrical_analysis.cpp:42, :252).  Everything the
// condensed block method builds on top of those arrays is built here, on the device, so a
// new system costs device time only:
//
//   1. the band + remainder layout (BandDev): rows checked, band[r] = a(r, r-1), the other
//      entries as half-entries of both of their rows, each row's own entries first (stored
//      order), then the transposed ones by ascending source row (a stable radix sort on the
//      column);
//   2. the condensation: anchors (rows with a half-entry), paths between them, each path's
//      Thomas factors and Schur terms (one thread per path), the reduced system S;
//   3. the series rounds: in each round the nodes of S with at most kElimDegree neighbours
//      and a usable pivot are candidates, the greedy independent set of the candidates in
//      elim_priority order is eliminated exactly (series reduction / star-mesh), and the
//      adjacency lists of the nodes they touched are rebuilt (fill merged in record order);
//      what is left is the core, handed on as a symmetric CSR.
//
// Every sum has a fixed order and no step depends on thread timing: the result equals the
// host forms (build_band, build_condensed_host, build_series_host) bit for bit, and
// #pragma clang fp contract(off) keeps the compiler from fusing what the host computes as
// separate products and sums.
#include "dea_internal.h"
#include "dea_grid.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

using namespace cicy;

namespace cicy {

// ------------------------------------------------------------------------------------
// device memory pool
// ------------------------------------------------------------------------------------

void *DevPool::get(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    // best fit among the free blocks that are not much larger than the request
    const size_t most = 2 * bytes + (64 << 10);
    int best = -1;
    for (size_t i = 0; i < blocks.size(); i++)
        if (!blocks[i].used && blocks[i].bytes >= bytes && blocks[i].bytes <= most &&
            (best < 0 || blocks[i].bytes < blocks[(size_t)best].bytes))
            best = (int)i;
    if (best >= 0) {
        blocks[(size_t)best].used = true;
        return blocks[(size_t)best].p;
    }
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    blocks.push_back({p, bytes, true});
    return p;
}

bool DevPool::put(void *p) {
    for (auto &b : blocks)
        if (b.p == p) {
            b.used = false;
            return true;
        }
    return false;
}

void dev_release(cicy_ctx *ctx, void *p) {
    if (!p) return;
    if (!ctx || !ctx->pool.put(p)) (void)hipFree(p);
}

void DevPool::clear() {
    for (auto &b : blocks) (void)hipFree(b.p);
    blocks.clear();
}

namespace {

constexpr int kST = 256;  // threads per work-group of the set-up kernels
constexpr int kFloatingPathDev = -2;  // pl of a floating path (the host form's kFloatingPath)
constexpr int kUndecided = 1, kIn = 2, kOut = 3;  // candidate states of a round
constexpr int kMaxMisPasses = 4096;
constexpr int kMisBatch = 16;  // independent-set passes launched per read-back
constexpr int kPathChunkS = 8;  // path rows whose loads the factorisation issues together

inline int grid_of(long long n) {
    return (int)std::max<long long>(1, std::min<long long>((n + kST - 1) / kST, 1 << 16));
}

#define GRID_LOOP(i, n) \
    for (long long i = (long long)blockIdx.x * kST + threadIdx.x; i < (long long)(n); \
         i += (long long)gridDim.x * kST)

#pragma clang fp contract(off)

// ---- 1. band + remainder layout ------------------------------------------------------

// First malformed row: key (row * 4 + kind) << 31 | entry offset in the row; kinds 1 row_ptr
// decreasing or out of [0, nnz], 2 zero or non-finite diagonal, 3 column not strictly below
// the row.  A row's entries are read only after its bounds have been checked.
__device__ __forceinline__ unsigned long long bad_key(long long r, int kind, long long off) {
    return (((unsigned long long)r * 4 + (unsigned long long)kind) << 31) |
           (unsigned long long)off;
}

__global__ __launch_bounds__(kST) void k_rows_check(int n, long long nnz,
                                                    const int64_t *__restrict__ rp,
                                                    const int64_t *__restrict__ col,
                                                    const double *__restrict__ diag,
                                                    unsigned long long *__restrict__ err) {
    GRID_LOOP(r, n) {
        const int64_t a = rp[r], b = rp[r + 1];
        unsigned long long key = ~0ull;
        if (a < 0 || b < a || b > nnz) {
            key = bad_key(r, 1, 0);
        } else if (!(diag[r] != 0.0) || !isfinite(diag[r])) {
            key = bad_key(r, 2, 0);
        } else {
            for (int64_t j = a; j < b; j++) {
                const int64_t c = col[j];
                if (c < 0 || c >= r) {
                    key = bad_key(r, 3, j - a);
                    break;
                }
            }
        }
        if (key != ~0ull) atomicMin(err, key);
    }
}

__global__ __launch_bounds__(64) void k_row_ends(int n, const int64_t *__restrict__ rp,
                                                int64_t *__restrict__ out) {
    if (threadIdx.x == 0) {
        out[0] = rp[0];
        out[1] = rp[n];
    }
}

// band value (the last stored entry of column r-1) and the count of the row's other
// entries; each of those is counted in its column's transposed count
__global__ __launch_bounds__(kST) void k_band_rows(int n, const int64_t *__restrict__ rp,
                                                   const int64_t *__restrict__ col,
                                                   const double *__restrict__ val,
                                                   double *__restrict__ band, int *__restrict__ own,
                                                   int *__restrict__ colcnt) {
    GRID_LOOP(r, n) {
        double bv = 0.0;
        int cnt = 0;
        for (int64_t j = rp[r]; j < rp[r + 1]; j++) {
            const int c = (int)col[j];
            if (c == r - 1) {
                bv = val[j];
            } else {
                cnt++;
                atomicAdd(&colcnt[c], 1);
            }
        }
        band[r] = bv;
        own[r] = cnt;
        if (r == 0) {
            band[n] = 0.0;
            own[n] = 0;
        }
    }
}

__global__ __launch_bounds__(kST) void k_add_counts(int n, const int *__restrict__ a,
                                                    const int *__restrict__ b,
                                                    int *__restrict__ out) {
    GRID_LOOP(i, n + 1) out[i] = i < n ? a[i] + b[i] : 0;
}

// a row's own half-entries at the start of its range, and the list of them (column,
// index) that the transposed placement sorts by column
__global__ __launch_bounds__(kST) void k_band_emit(
    int n, const int64_t *__restrict__ rp, const int64_t *__restrict__ col,
    const double *__restrict__ val, const int *__restrict__ xrp, const int *__restrict__ ostart,
    int *__restrict__ xnbr, double *__restrict__ xval, int *__restrict__ tkey,
    int *__restrict__ tidx, int *__restrict__ trow, double *__restrict__ tval) {
    GRID_LOOP(r, n) {
        const int x0 = xrp[r], o0 = ostart[r];
        int t = 0;
        for (int64_t j = rp[r]; j < rp[r + 1]; j++) {
            const int c = (int)col[j];
            if (c == r - 1) continue;
            xnbr[x0 + t] = c;
            xval[x0 + t] = val[j];
            tkey[o0 + t] = c;
            tidx[o0 + t] = o0 + t;
            trow[o0 + t] = (int)r;
            tval[o0 + t] = val[j];
            t++;
        }
    }
}

// transposed half-entries: the p-th entry of column c (ascending source position) goes after
// c's own entries
__global__ __launch_bounds__(kST) void k_band_transpose(
    int ne, const int *__restrict__ skey, const int *__restrict__ sidx,
    const int *__restrict__ cstart, const int *__restrict__ xrp, const int *__restrict__ own,
    const int *__restrict__ trow, const double *__restrict__ tval, int *__restrict__ xnbr,
    double *__restrict__ xval) {
    GRID_LOOP(p, ne) {
        const int c = skey[p], e = sidx[p];
        const int pos = xrp[c] + own[c] + ((int)p - cstart[c]);
        xnbr[pos] = trow[e];
        xval[pos] = tval[e];
    }
}

__global__ __launch_bounds__(kST) void k_xblk(int nblocks, int n, int rb,
                                              const int *__restrict__ xrp,
                                              int *__restrict__ xblk) {
    GRID_LOOP(k, nblocks + 1) xblk[k] = xrp[std::min<long long>(n, k * (long long)rb)];
}

// The row blocks' launch order, as build_band orders them: the heavy blocks (more than
// `limit` half-entries, four times the mean) first, by descending count with ties in block
// order, then the other blocks in block order.  One work-group walks the blocks in chunks of
// its size, a block scan per chunk placing the light blocks straight into `order` and the
// heavy ones, with their counts, into hv/hc[0, H) in block order (H to *nheavy) for
// k_band_rank.
constexpr int kOrderThreads = 1024;
__global__ __launch_bounds__(kOrderThreads) void k_band_split(int nblocks,
                                                              const int *__restrict__ xblk,
                                                              double limit, int *__restrict__ hv,
                                                              int *__restrict__ hc,
                                                              int *__restrict__ nheavy,
                                                              int *__restrict__ order) {
    using Scan = hipcub::BlockScan<int, kOrderThreads>;
    __shared__ typename Scan::TempStorage tmp;
    int h0 = 0, l0 = 0;  // heavy and light blocks before this chunk
    for (int base = 0; base < nblocks; base += kOrderThreads) {
        const int k = base + (int)threadIdx.x;
        int c = 0;
        bool heavy = false;
        if (k < nblocks) {
            c = xblk[k + 1] - xblk[k];
            heavy = c > 0 && (double)c > limit;
        }
        int off, nh;
        Scan(tmp).ExclusiveSum(heavy ? 1 : 0, off, nh);
        if (k < nblocks) {
            if (heavy) {
                hv[h0 + off] = k;
                hc[h0 + off] = c;
            } else {
                // light blocks go after every heavy one: their final place is fixed once H is
                // known, so keep their rank among the light blocks for now
                order[l0 + ((int)threadIdx.x - off)] = k;
            }
        }
        h0 += nh;
        l0 += std::min(kOrderThreads, nblocks - base) - nh;
        __syncthreads();  // tmp is reused by the next chunk's scan
    }
    if (threadIdx.x == 0) *nheavy = h0;
}

// Each heavy block's rank among the heavy blocks by (descending count, block index), one
// wave per heavy block (its lanes read the heavy list in coalesced strides and the counts
// are summed across the wave); order[0, H) takes the heavy blocks and the light ones
// (listed at light[0, L) by k_band_split) move up by H.
__global__ __launch_bounds__(kST) void k_band_rank(int nblocks, const int *__restrict__ hv,
                                                   const int *__restrict__ hc,
                                                   const int *__restrict__ nheavy,
                                                   const int *__restrict__ light,
                                                   int *__restrict__ order) {
    const int H = *nheavy;
    const int lane = (int)threadIdx.x & (kWave - 1);
    const long long waves = (long long)gridDim.x * (kST / kWave);
    for (long long i = ((long long)blockIdx.x * kST + threadIdx.x) / kWave; i < H; i += waves) {
        const int k = hv[i], c = hc[i];
        int rank = 0;
        for (int j = lane; j < H; j += kWave) {
            const int cj = hc[j], kj = hv[j];
            rank += (cj > c || (cj == c && kj < k)) ? 1 : 0;
        }
        for (int o = kWave / 2; o > 0; o >>= 1) rank += __shfl_xor(rank, o, kWave);
        if (lane == 0) order[rank] = k;
    }
    // the light blocks, in block order after the heavy ones
    for (long long t = (long long)blockIdx.x * kST + threadIdx.x; t < (long long)nblocks - H;
         t += (long long)gridDim.x * kST)
        order[H + t] = light[t];
}

// ---- 2. condensation ---------------------------------------------------------------------

__global__ __launch_bounds__(kST) void k_anchor_flags(int n, const int *__restrict__ xrp,
                                                      int *__restrict__ flag) {
    GRID_LOOP(i, n + 1) flag[i] = i < n ? (xrp[i + 1] > xrp[i] ? 1 : 0) : 0;
}

__global__ __launch_bounds__(kST) void k_anchor_index(int n, const int *__restrict__ flag,
                                                      const int *__restrict__ apos,
                                                      int *__restrict__ anchor_of,
                                                      int *__restrict__ anchors) {
    GRID_LOOP(i, n) {
        if (flag[i]) {
            anchor_of[i] = apos[i];
            anchors[apos[i]] = (int)i;
        } else {
            anchor_of[i] = -1;
        }
    }
}

// a path starts at a non-anchor row whose predecessor is an anchor or not coupled to it
__global__ __launch_bounds__(kST) void k_path_flags(int n, const int *__restrict__ anchor_of,
                                                    const double *__restrict__ band,
                                                    int *__restrict__ pflag) {
    GRID_LOOP(i, n + 1) {
        int f = 0;
        if (i < n)
            f = anchor_of[i] < 0 && (i == 0 || anchor_of[i - 1] >= 0 || band[i] == 0.0);
        pflag[i] = f;
    }
}

// each path's first and last row: a path row whose next row does not continue the path
// ends one; paths are disjoint runs in row order, so the path ending at row i is the last
// one started at or before i
__global__ __launch_bounds__(kST) void k_path_starts(int n, const int *__restrict__ pflag,
                                                     const int *__restrict__ pscan,
                                                     const int *__restrict__ anchor_of,
                                                     const double *__restrict__ band,
                                                     int *__restrict__ ps, int *__restrict__ pe) {
    GRID_LOOP(i, n) {
        if (pflag[i]) ps[pscan[i]] = (int)i;
        if (anchor_of[i] < 0 && (i + 1 == n || anchor_of[i + 1] >= 0 || band[i + 1] == 0.0))
            pe[pscan[i] + pflag[i] - 1] = (int)i;
    }
}

__global__ __launch_bounds__(kST) void k_fill_factors(int n, double *__restrict__ lfac,
                                                      double *__restrict__ mfac) {
    GRID_LOOP(i, n) {
        lfac[i] = 0.0;
        mfac[i] = 1.0;
    }
}

// One thread per path: its neighbouring anchors, Thomas factors, and its Schur terms
// (termL on the left anchor's diagonal, termR on the right one's, fillv the coupling
// between them).  err: lowest path with a non-positive pivot, (path << 2) | kind with kind
// 1 on a path coupled to an anchor, 2 inside a floating path.
__global__ __launch_bounds__(kST) void k_paths(
    int np, int n, const int *__restrict__ ps, const int *__restrict__ anchor_of,
    const double *__restrict__ band, const double *__restrict__ diag, int *__restrict__ pe,
    int *__restrict__ pl, int *__restrict__ pr, double *__restrict__ lfac,
    double *__restrict__ mfac, double *__restrict__ g, double *__restrict__ termL,
    double *__restrict__ termR, double *__restrict__ fillv, unsigned long long *__restrict__ err) {
    GRID_LOOP(p, np) {
        const int s = ps[p], e = pe[p];
        const int L = (s > 0 && band[s] != 0.0) ? s - 1 : -1;
        const int R = (e + 1 < n && band[e + 1] != 0.0) ? e + 1 : -1;
        pl[p] = L >= 0 ? anchor_of[L] : -1;
        pr[p] = R >= 0 ? anchor_of[R] : -1;
        termL[p] = 0.0;
        termR[p] = 0.0;
        fillv[p] = 0.0;
        // the Thomas factors, the loads of kPathChunkS rows issued before their arithmetic;
        // the pivots checked and, for a path with an anchor on the left, g = T^-1 band[s] e_s
        // forward-eliminated as they come
        const double pivot_floor = 1e-12 * diag[e];
        auto bad = [&](double v) { return !(v > pivot_floor) || !isfinite(v); };
        double mprev = diag[s];
        mfac[s] = mprev;
        bool bad_inner = s < e && bad(mprev), bad_last = s == e && bad(mprev);
        const bool left = L >= 0;
        double gprev = 0.0;
        if (left) {
            gprev = band[s];
            g[s] = gprev;
        }
        for (int k0 = s + 1; k0 <= e; k0 += kPathChunkS) {
            double bb[kPathChunkS], dd[kPathChunkS];
#pragma unroll
            for (int t = 0; t < kPathChunkS; t++) {
                const int k = min(k0 + t, e);
                bb[t] = band[k];
                dd[t] = diag[k];
            }
#pragma unroll
            for (int t = 0; t < kPathChunkS; t++) {
                const int k = k0 + t;
                if (k > e) break;
                const double l = bb[t] / mprev;
                mprev = dd[t] - l * bb[t];
                lfac[k] = l;
                mfac[k] = mprev;
                if (k == e) bad_last = bad(mprev);
                else bad_inner = bad_inner || bad(mprev);
                if (left) {
                    gprev = -l * gprev;
                    g[k] = gprev;
                }
            }
        }
        if (bad_inner || bad_last) {
            if (L >= 0 || R >= 0) {
                atomicMin(err, ((unsigned long long)p << 2) | 1ull);
                continue;
            }
            if (bad_inner) {
                atomicMin(err, ((unsigned long long)p << 2) | 2ull);
                continue;
            }
            pl[p] = kFloatingPathDev;
            lfac[e] = 0.0;
            mfac[e] = 1.0;
            continue;
        }
        if (left) {
            // back-substitution of g, the loads of kPathChunkS rows issued together
            double u = gprev / mprev;
            const double ue = u;
            for (int k0 = e - 1; k0 >= s; k0 -= kPathChunkS) {
                double gg[kPathChunkS], bb[kPathChunkS], mm[kPathChunkS];
#pragma unroll
                for (int t = 0; t < kPathChunkS; t++) {
                    const int k = max(k0 - t, s);
                    gg[t] = g[k];
                    bb[t] = band[k + 1];
                    mm[t] = mfac[k];
                }
#pragma unroll
                for (int t = 0; t < kPathChunkS; t++) {
                    const int k = k0 - t;
                    if (k < s) break;
                    u = (gg[t] - bb[t] * u) / mm[t];
                }
            }
            termL[p] = band[s] * u;
            if (R >= 0) fillv[p] = -band[e + 1] * ue;
        }
        if (R >= 0) termR[p] = band[e + 1] * (band[e + 1] / mprev);
    }
}

// S's diagonal: the anchor's own, less the term of the path to its left, then of the path
// to its right (the order the paths come in)
__global__ __launch_bounds__(kST) void k_sdiag(int m, int n, const int *__restrict__ anchors,
                                               const int *__restrict__ anchor_of,
                                               const double *__restrict__ band,
                                               const double *__restrict__ diag,
                                               const int *__restrict__ pscan,
                                               const double *__restrict__ termL,
                                               const double *__restrict__ termR,
                                               double *__restrict__ sdiag) {
    GRID_LOOP(j, m) {
        const int i = anchors[j];
        double v = diag[i];
        if (i > 0 && anchor_of[i - 1] < 0 && band[i] != 0.0) v = v - termR[pscan[i] - 1];
        if (i + 1 < n && anchor_of[i + 1] < 0 && band[i + 1] != 0.0) v = v - termL[pscan[i + 1]];
        sdiag[j] = v;
    }
}

// Entries of S's row j (strictly lower, reduced numbering), in the host form's order: the
// coupling of the path to the left (when it joins two anchors), the band entry to the
// previous row (when that is an anchor), then the row's own half-entries with an anchor
// column, in stored order.  Pass 0 counts them, pass 1 writes them.
template <int PASS>
__global__ __launch_bounds__(kST) void k_s_raw(int m, const int *__restrict__ anchors,
                                               const int *__restrict__ anchor_of,
                                               const double *__restrict__ band,
                                               const int *__restrict__ xrp,
                                               const int *__restrict__ xnbr,
                                               const double *__restrict__ xval,
                                               const int *__restrict__ pscan,
                                               const int *__restrict__ pl,
                                               const double *__restrict__ fillv,
                                               int *__restrict__ cnt, const int *__restrict__ off,
                                               int *__restrict__ rcol, double *__restrict__ rval) {
    GRID_LOOP(j, m) {
        const int i = anchors[j];
        int c = 0;
        int o = PASS ? off[j] : 0;
        auto emit = [&](int col, double v) {
            if (PASS) {
                rcol[o + c] = col;
                rval[o + c] = v;
            }
            c++;
        };
        if (i > 0 && anchor_of[i - 1] < 0 && band[i] != 0.0) {
            const int p = pscan[i] - 1;  // the path ending at i - 1, whose right anchor is i
            if (pl[p] >= 0) emit(pl[p], fillv[p]);
        }
        if (i > 0 && band[i] != 0.0 && anchor_of[i - 1] >= 0) emit(anchor_of[i - 1], band[i]);
        for (int t = xrp[i]; t < xrp[i + 1]; t++) {
            const int cc = xnbr[t];
            if (cc < i && anchor_of[cc] >= 0) emit(anchor_of[cc], xval[t]);
        }
        if (!PASS) cnt[j] = c;
        if (j == 0 && !PASS) cnt[m] = 0;
    }
}

// stable insertion sort of each row's raw entries by column, then the number of distinct
// columns
__global__ __launch_bounds__(kST) void k_rows_sort(int m, const int *__restrict__ off,
                                                   int *__restrict__ col, double *__restrict__ val,
                                                   int *__restrict__ ucnt) {
    GRID_LOOP(j, m) {
        const int b = off[j], e = off[j + 1];
        for (int q = b + 1; q < e; q++) {
            const int cq = col[q];
            const double vq = val[q];
            int w = q;
            while (w > b && col[w - 1] > cq) {
                col[w] = col[w - 1];
                val[w] = val[w - 1];
                w--;
            }
            col[w] = cq;
            val[w] = vq;
        }
        int u = 0;
        for (int q = b; q < e; q++)
            if (q == b || col[q] != col[q - 1]) u++;
        ucnt[j] = u;
        if (j == 0) ucnt[m] = 0;
    }
}

// equal columns of a sorted row summed in order
__global__ __launch_bounds__(kST) void k_rows_sum(int m, const int *__restrict__ off,
                                                  const int *__restrict__ col,
                                                  const double *__restrict__ val,
                                                  const int *__restrict__ rrp,
                                                  int *__restrict__ rcol,
                                                  double *__restrict__ rval) {
    GRID_LOOP(j, m) {
        const int b = off[j], e = off[j + 1];
        int o = rrp[j];
        for (int q = b; q < e;) {
            int w = q;
            double v = 0.0;
            while (w < e && col[w] == col[q]) v += val[w++];
            rcol[o] = col[q];
            rval[o++] = v;
            q = w;
        }
    }
}

// ---- symmetric CSR from a strictly-lower one (columns ascending in every row) ---------

__global__ __launch_bounds__(kST) void k_low_counts(int m, const int *__restrict__ rp,
                                                    const int *__restrict__ col,
                                                    int *__restrict__ lcnt,
                                                    int *__restrict__ tcnt) {
    GRID_LOOP(r, m) {
        lcnt[r] = rp[r + 1] - rp[r];
        for (int k = rp[r]; k < rp[r + 1]; k++) atomicAdd(&tcnt[col[k]], 1);
        if (r == 0) {
            lcnt[m] = 0;
            tcnt[m] = 0;
        }
    }
}

__global__ __launch_bounds__(kST) void k_low_emit(int m, const int *__restrict__ rp,
                                                  const int *__restrict__ col,
                                                  const double *__restrict__ val,
                                                  const int *__restrict__ frp,
                                                  int *__restrict__ fcol, double *__restrict__ fval,
                                                  int *__restrict__ tkey, int *__restrict__ tidx,
                                                  int *__restrict__ trow) {
    GRID_LOOP(r, m) {
        const int o = frp[r];
        for (int k = rp[r]; k < rp[r + 1]; k++) {
            fcol[o + k - rp[r]] = col[k];
            fval[o + k - rp[r]] = val[k];
            tkey[k] = col[k];
            tidx[k] = k;
            trow[k] = (int)r;
        }
    }
}

// the upper entries of row c: rows above it, ascending (the stable sort keeps row order)
__global__ __launch_bounds__(kST) void k_up_place(int ne, const int *__restrict__ skey,
                                                  const int *__restrict__ sidx,
                                                  const int *__restrict__ tstart,
                                                  const int *__restrict__ frp,
                                                  const int *__restrict__ lcnt,
                                                  const int *__restrict__ trow,
                                                  const double *__restrict__ val,
                                                  int *__restrict__ fcol,
                                                  double *__restrict__ fval) {
    GRID_LOOP(p, ne) {
        const int c = skey[p], k = sidx[p];
        const int pos = frp[c] + lcnt[c] + ((int)p - tstart[c]);
        fcol[pos] = trow[k];
        fval[pos] = val[k];
    }
}

// ---- 3. series rounds -------------------------------------------------------------------

__global__ __launch_bounds__(kST) void k_cand(int m, const int *__restrict__ arp,
                                              const double *__restrict__ d,
                                              const double *__restrict__ sd0,
                                              const int *__restrict__ alive,
                                              int *__restrict__ state, int max_degree) {
    GRID_LOOP(k, m) {
        const double dk = d[k];
        const bool c = alive[k] && arp[k + 1] - arp[k] <= max_degree &&
                       dk > 1e-12 * fabs(sd0[k]) && isfinite(dk);
        state[k] = c ? kUndecided : 0;
    }
}

// One pass toward the greedy independent set in elim_priority order: an undecided candidate
// with a selected neighbour is out; one whose earlier candidate neighbours are all out is
// in; the others wait for the next pass (*undecided is set when there are any).  Decisions are final, so
// reading a neighbour's state while it changes only delays a decision.
__global__ __launch_bounds__(kST) void k_mis_pass(int m, const int *__restrict__ arp,
                                                  const int *__restrict__ acol, int *state,
                                                  const unsigned int *__restrict__ before,
                                                  unsigned int *__restrict__ undecided) {
    // the set settled in an earlier pass of this batch (no candidate was left waiting)
    if (before && *before == 0) return;
    bool any_wait = false;
    GRID_LOOP(k, m) {
        bool wait = false;
        if (state[k] == kUndecided) {
            const unsigned long long pk = elim_priority((int)k);
            bool blocked = false;
            for (int t = arp[k]; t < arp[k + 1]; t++) {
                const int j = acol[t];
                const int sj = state[j];
                if (sj == kIn) {
                    blocked = true;
                    break;
                }
                if (sj == kUndecided && elim_priority(j) < pk) wait = true;
            }
            if (blocked) {
                state[k] = kOut;
                wait = false;
            } else if (!wait) {
                state[k] = kIn;
            }
        }
        any_wait = any_wait || wait;
    }
    // one store per work-group with a waiting candidate (only "any left" is read)
    if (__syncthreads_or(any_wait) && threadIdx.x == 0) *undecided = 1u;
}

// The round's independent set in one cooperative launch: each thread sweeps its
// candidates until every one is decided, by k_mis_pass's rule.  Decisions are final and
// depend only on earlier candidates' final decisions, so the set is the one the passes
// give; states are read and written with device-coherent atomics, every work-group is
// resident, and the undecided candidate of lowest priority can always be decided, so
// every thread finishes.  A thread past max_sweeps stops and sets *gave_up (the host then
// finishes the set with passes).
__global__ __launch_bounds__(kST) void k_mis_async(int m, const int *__restrict__ arp,
                                                   const int *__restrict__ acol, int *state,
                                                   int *__restrict__ gave_up,
                                                   unsigned max_sweeps) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long first = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    for (unsigned sweep = 0;; sweep++) {
        bool any_wait = false;
        for (long long k = first; k < m; k += stride) {
            if (ld_coh_i(&state[k]) != kUndecided) continue;
            const unsigned long long pk = elim_priority((int)k);
            bool blocked = false, wait = false;
            for (int t = arp[k]; t < arp[k + 1]; t++) {
                const int j = acol[t];
                const int sj = ld_coh_i(&state[j]);
                if (sj == kIn) {
                    blocked = true;
                    break;
                }
                if (sj == kUndecided && elim_priority(j) < pk) wait = true;
            }
            if (blocked)
                __hip_atomic_store(&state[k], kOut, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            else if (!wait)
                __hip_atomic_store(&state[k], kIn, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            else
                any_wait = true;
        }
        if (!any_wait) return;
        if (sweep >= max_sweeps) {
            __hip_atomic_store(gave_up, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
        __builtin_amdgcn_s_sleep(2);
    }
}
constexpr unsigned kMisMaxSweeps = 1u << 16;

// four counts per node, prefix-summed together in one scan
struct Cnt4 {
    int v[4];
};
struct Cnt4Sum {
    __host__ __device__ Cnt4 operator()(const Cnt4 &a, const Cnt4 &b) const {
        return Cnt4{{a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2], a.v[3] + b.v[3]}};
    }
};

__global__ __launch_bounds__(kST) void k_unpack4(int count, const Cnt4 *__restrict__ in,
                                                 int *__restrict__ o0, int *__restrict__ o1,
                                                 int *__restrict__ o2, int *__restrict__ o3) {
    GRID_LOOP(i, count) {
        const Cnt4 c = in[i];
        o0[i] = c.v[0];
        if (o1) o1[i] = c.v[1];
        if (o2) o2[i] = c.v[2];
        if (o3) o3[i] = c.v[3];
    }
}

__global__ __launch_bounds__(kST) void k_sel_flags(int m, const int *__restrict__ state,
                                                   const int *__restrict__ arp,
                                                   int *__restrict__ sel, Cnt4 *__restrict__ c4) {
    GRID_LOOP(k, m + 1) {
        const int s = k < m && state[k] == kIn;
        const int deg = s ? arp[k + 1] - arp[k] : 0;
        sel[k] = s;
        // selected, record entries, and a bound on the fill the node causes
        c4[k] = Cnt4{{s, deg, deg * (deg - 1), 0}};
    }
}

// records of the round (ascending node), from the state before the round
__global__ __launch_bounds__(kST) void k_records(int m, const int *__restrict__ sel,
                                                 const int *__restrict__ selpos,
                                                 const int *__restrict__ degoff,
                                                 const int *__restrict__ arp,
                                                 const int *__restrict__ acol,
                                                 const double *__restrict__ aval,
                                                 const double *__restrict__ d, int R0, int E0,
                                                 int *__restrict__ e_node, int *__restrict__ e_off,
                                                 double *__restrict__ e_d, int *__restrict__ e_nbr,
                                                 double *__restrict__ e_a,
                                                 int *__restrict__ rec_of) {
    GRID_LOOP(k, m) {
        if (!sel[k]) continue;
        const int q = selpos[k], rec = R0 + q, o = E0 + degoff[k];
        e_node[rec] = (int)k;
        e_off[rec] = o;
        e_d[rec] = d[k];
        rec_of[k] = rec;
        for (int t = arp[k]; t < arp[k + 1]; t++) {
            e_nbr[o + t - arp[k]] = acol[t];
            e_a[o + t - arp[k]] = aval[t];
        }
    }
}

// per surviving node: selected neighbours (contributions), fill entries and an upper bound
// of its new row length
__global__ __launch_bounds__(kST) void k_targets(int m, const int *__restrict__ alive,
                                                 const int *__restrict__ sel,
                                                 const int *__restrict__ arp,
                                                 const int *__restrict__ acol,
                                                 int *__restrict__ tflag, Cnt4 *__restrict__ c4) {
    GRID_LOOP(u, m + 1) {
        int t = 0, s = 0, f = 0, b = 0;
        if (u < m && alive[u] && !sel[u]) {
            for (int q = arp[u]; q < arp[u + 1]; q++) {
                const int j = acol[q];
                if (sel[j]) {
                    s++;
                    f += arp[j + 1] - arp[j] - 1;
                }
            }
            t = s > 0;
            b = arp[u + 1] - arp[u] - s + f;
        }
        tflag[u] = t;
        c4[u] = Cnt4{{t, s, f, b}};
    }
}

// receivers of the round (the forward sweep's gather lists) and the diagonal updates,
// contributions in record order (= ascending neighbour)
__global__ __launch_bounds__(kST) void k_receivers(int m, const int *__restrict__ tflag,
                                                   const int *__restrict__ tpos,
                                                   const int *__restrict__ cpos,
                                                   const int *__restrict__ arp,
                                                   const int *__restrict__ acol,
                                                   const double *__restrict__ aval,
                                                   const int *__restrict__ sel,
                                                   const int *__restrict__ rec_of,
                                                   const double *__restrict__ e_d,
                                                   const int *__restrict__ rs,
                                                   int *__restrict__ rv_node, int *__restrict__ rv_off,
                                                   int *__restrict__ rv_rec, double *__restrict__ rv_a,
                                                   double *__restrict__ d) {
    if (rs[0]) return;  // the round overflowed its arrays (k_round_guard)
    const int RV0 = rs[1], C0 = rs[2];
    GRID_LOOP(u, m) {
        if (!tflag[u]) continue;
        const int t = RV0 + tpos[u];
        int c = C0 + cpos[u];
        rv_node[t] = (int)u;
        rv_off[t] = c;
        double du = d[u];
        for (int q = arp[u]; q < arp[u + 1]; q++) {
            const int j = acol[q];
            if (!sel[j]) continue;
            const int rec = rec_of[j];
            const double a = aval[q];
            rv_rec[c] = rec;
            rv_a[c] = a;
            c++;
            du = du - a * a / e_d[rec];
        }
        d[u] = du;
    }
}

// Round state on the device (rs): rs[0] an overflow flag, rs[1] and rs[2] the receivers
// and contributions written so far, rs[3 + q] the first receiver of round q.  The round's
// arrays are sized from bounds the host reads with the independent set; the exact totals
// come from the targets' scan and stay on the device.
constexpr int kMaxRounds = 1024;
__global__ __launch_bounds__(64) void k_round_guard(int m, const int *__restrict__ tpos,
                                                    const int *__restrict__ cpos,
                                                    const int *__restrict__ fpos,
                                                    const int *__restrict__ upos,
                                                    int *__restrict__ rs, long long ecap,
                                                    long long fill_cap, long long ub_cap) {
    if (threadIdx.x == 0 &&
        ((long long)rs[1] + tpos[m] > ecap || (long long)rs[2] + cpos[m] > ecap ||
         fpos[m] > fill_cap || upos[m] > ub_cap))
        rs[0] = 1;
}
__global__ __launch_bounds__(64) void k_round_end(int m, const int *__restrict__ tpos,
                                                  const int *__restrict__ cpos,
                                                  int *__restrict__ rs, int q) {
    if (threadIdx.x == 0) {
        if (!rs[0]) {
            rs[1] += tpos[m];
            rs[2] += cpos[m];
        }
        rs[3 + q + 1] = rs[1];
    }
}

constexpr int kMergeLists = 8;  // eliminated neighbours merged without staging their fill

// The new row of each surviving node: its old entries without the eliminated neighbours,
// merged with the fill of every eliminated neighbour (record order; equal columns summed
// after the old value, in record order).  Written at ub offsets, true lengths in len.
__global__ __launch_bounds__(kST) void k_merge(
    int m, const int *__restrict__ alive, const int *__restrict__ sel,
    const int *__restrict__ tflag, const int *__restrict__ arp, const int *__restrict__ acol,
    const double *__restrict__ aval, const int *__restrict__ upos, const int *__restrict__ fpos,
    const int *__restrict__ rec_of, const int *__restrict__ e_off, const int *__restrict__ e_nbr,
    const double *__restrict__ e_a, const double *__restrict__ e_d, int *__restrict__ fcol,
    double *__restrict__ fval, int *__restrict__ ocol, double *__restrict__ oval,
    int *__restrict__ len, const int *__restrict__ rs) {
    const bool overflow = rs[0] != 0;  // nothing is written then (k_round_guard)
    GRID_LOOP(u, m + 1) {
        if (overflow || u == m || !alive[u] || sel[u]) {
            len[u] = 0;
            continue;
        }
        int o = upos[u];
        if (!tflag[u]) {
            for (int q = arp[u]; q < arp[u + 1]; q++) {
                ocol[o] = acol[q];
                oval[o++] = aval[q];
            }
            len[u] = arp[u + 1] - arp[u];
            continue;
        }
        // with at most kMergeLists eliminated neighbours: one merge of the old row (without
        // them) and their records' rows (each ascending, the entry for u left out), equal
        // columns summed old value first, then in record order -- the order the sort below
        // gives, without staging the fill
        {
            int lp[kMergeLists], le[kMergeLists], hc[kMergeLists];
            double la[kMergeLists], ld[kMergeLists];
            int nl = 0;
            bool fits = true;
            for (int q = arp[u]; q < arp[u + 1]; q++) {
                const int j = acol[q];
                if (!sel[j]) continue;
                if (nl == kMergeLists) {
                    fits = false;
                    break;
                }
                const int rec = rec_of[j];
                lp[nl] = e_off[rec];
                le[nl] = e_off[rec + 1];
                la[nl] = aval[q];
                ld[nl] = e_d[rec];
                nl++;
            }
            if (fits) {
#pragma unroll
                for (int l = 0; l < kMergeLists; l++) {
                    if (l >= nl) break;
                    while (lp[l] < le[l] && e_nbr[lp[l]] == (int)u) lp[l]++;
                    hc[l] = lp[l] < le[l] ? e_nbr[lp[l]] : INT_MAX;
                }
                const int o0 = o;
                int q = arp[u];
                const int qe = arp[u + 1];
                while (q < qe && sel[acol[q]]) q++;
                for (;;) {
                    const int co = q < qe ? acol[q] : INT_MAX;
                    int c = co;
#pragma unroll
                    for (int l = 0; l < kMergeLists; l++)
                        if (l < nl) c = min(c, hc[l]);
                    if (c == INT_MAX) break;
                    double v = 0.0;
                    bool have = false;
                    if (co == c) {
                        v = aval[q++];
                        have = true;
                        while (q < qe && sel[acol[q]]) q++;
                    }
#pragma unroll
                    for (int l = 0; l < kMergeLists; l++) {
                        if (l >= nl || hc[l] != c) continue;
                        const double fv = -la[l] * e_a[lp[l]] / ld[l];
                        v = have ? v + fv : fv;
                        have = true;
                        lp[l]++;
                        while (lp[l] < le[l] && e_nbr[lp[l]] == (int)u) lp[l]++;
                        hc[l] = lp[l] < le[l] ? e_nbr[lp[l]] : INT_MAX;
                    }
                    ocol[o] = c;
                    oval[o++] = v;
                }
                len[u] = o - o0;
                continue;
            }
        }
        // the fill, in record order, then stably sorted by column
        const int f0 = fpos[u];
        int f = f0;
        for (int q = arp[u]; q < arp[u + 1]; q++) {
            const int j = acol[q];
            if (!sel[j]) continue;
            const int rec = rec_of[j];
            const double au = aval[q], dk = e_d[rec];
            for (int t = e_off[rec]; t < e_off[rec + 1]; t++) {
                const int v = e_nbr[t];
                if (v == u) continue;
                fcol[f] = v;
                fval[f++] = -au * e_a[t] / dk;
            }
        }
        for (int q = f0 + 1; q < f; q++) {
            const int cq = fcol[q];
            const double vq = fval[q];
            int w = q;
            while (w > f0 && fcol[w - 1] > cq) {
                fcol[w] = fcol[w - 1];
                fval[w] = fval[w - 1];
                w--;
            }
            fcol[w] = cq;
            fval[w] = vq;
        }
        // merge
        const int o0 = o;
        int q = arp[u], qe = arp[u + 1], p = f0;
        while (q < qe || p < f) {
            while (q < qe && sel[acol[q]]) q++;
            if (q >= qe && p >= f) break;
            int c;
            double v;
            if (p >= f || (q < qe && acol[q] < fcol[p])) {
                c = acol[q];
                v = aval[q++];
            } else if (q >= qe || fcol[p] < acol[q]) {
                c = fcol[p];
                v = fval[p++];
                while (p < f && fcol[p] == c) v += fval[p++];
            } else {
                c = acol[q];
                v = aval[q++];
                while (p < f && fcol[p] == c) v += fval[p++];
            }
            ocol[o] = c;
            oval[o++] = v;
        }
        len[u] = o - o0;
    }
}

__global__ __launch_bounds__(kST) void k_compact(int m, const int *__restrict__ upos,
                                                 const int *__restrict__ nrp,
                                                 const int *__restrict__ ocol,
                                                 const double *__restrict__ oval,
                                                 int *__restrict__ ncol,
                                                 double *__restrict__ nval,
                                                 const int *__restrict__ rs) {
    if (rs[0]) return;  // the round overflowed its arrays (k_round_guard)
    GRID_LOOP(u, m) {
        const int b = upos[u], o = nrp[u], l = nrp[u + 1] - nrp[u];
        for (int t = 0; t < l; t++) {
            ncol[o + t] = ocol[b + t];
            nval[o + t] = oval[b + t];
        }
    }
}

__global__ __launch_bounds__(kST) void k_kill(int m, const int *__restrict__ sel,
                                              int *__restrict__ alive) {
    GRID_LOOP(k, m) if (sel[k]) alive[k] = 0;
}

__global__ __launch_bounds__(kST) void k_fill_i32(int *__restrict__ p, int n, int v) {
    GRID_LOOP(i, n) p[i] = v;
}
__global__ __launch_bounds__(kST) void k_set_int(int *__restrict__ p, int v) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
}

// the entries at `at` of up to eight arrays (scan totals), gathered for one read-back
struct LastArgs {
    const int *p[8];
};
__global__ __launch_bounds__(64) void k_last(LastArgs a, int k, int at, int *__restrict__ out) {
    if ((int)threadIdx.x < k) out[threadIdx.x] = a.p[threadIdx.x][at];
}
struct PickArgs {
    const int *p[8];
    int at[8];
};
__global__ __launch_bounds__(64) void k_pick_ints(PickArgs a, int k, const int *__restrict__ extra,
                                                 int nextra, int *__restrict__ out) {
    if ((int)threadIdx.x < k) out[threadIdx.x] = a.p[threadIdx.x][a.at[threadIdx.x]];
    for (int t = (int)threadIdx.x; t < nextra; t += 64) out[8 + t] = extra[t];
}

// *bad = 1 when some v[i] is not positive and finite
__global__ __launch_bounds__(kST) void k_nonpositive(int n, const double *__restrict__ v,
                                                     int *__restrict__ bad) {
    bool any = false;
    GRID_LOOP(i, n) any = any || !(v[i] > 0.0) || !isfinite(v[i]);
    if (__syncthreads_or(any) && threadIdx.x == 0) *bad = 1;
}

// the core: the nodes left, in S order, with their rows renumbered (both triangles)
__global__ __launch_bounds__(kST) void k_core_len(int m, const int *__restrict__ alive,
                                                  const int *__restrict__ arp,
                                                  int *__restrict__ flag,
                                                  int *__restrict__ clen) {
    GRID_LOOP(k, m + 1) {
        const int a = k < m && alive[k];
        flag[k] = a;
        clen[k] = a ? arp[k + 1] - arp[k] : 0;
    }
}

__global__ __launch_bounds__(kST) void k_core_fill(
    int m, const int *__restrict__ alive, const int *__restrict__ core_of,
    const int *__restrict__ cpos, const int *__restrict__ arp, const int *__restrict__ acol,
    const double *__restrict__ aval, const double *__restrict__ d,
    const double *__restrict__ sd0, const int *__restrict__ anchors,
    int *__restrict__ core_nodes, int *__restrict__ frp, int *__restrict__ fcol,
    double *__restrict__ fval, double *__restrict__ cdiag, int *__restrict__ orig) {
    GRID_LOOP(k, m) {
        if (!alive[k]) continue;
        const int q = core_of[k], o = cpos[k];
        core_nodes[q] = (int)k;
        frp[q] = o;
        // what the rounds leave of a floating group of tubes: a node without neighbours and
        // without a usable pivot, decoupled with diagonal 1 (as the host form)
        const bool isolated = arp[k] == arp[k + 1];
        const double dk = d[k];
        cdiag[q] = (isolated && isfinite(dk) && !(dk > 1e-12 * fabs(sd0[k]))) ? 1.0 : dk;
        orig[q] = anchors[k];
        for (int t = arp[k]; t < arp[k + 1]; t++) {
            fcol[o + t - arp[k]] = core_of[acol[t]];
            fval[o + t - arp[k]] = aval[t];
        }
    }
}

#pragma clang fp contract(on)

// ------------------------------------------------------------------------------------
// host orchestration
// ------------------------------------------------------------------------------------

int bits_for(long long n) {
    int b = 1;
    while ((1LL << b) <= n) b++;
    return b;
}

struct Cub {
    cicy_ctx *ctx;
    PoolScope &ps;
    // exclusive prefix sums of count int entries (the last entry of `in` is usually a zero
    // placed there so that out[count - 1] is the total)
    bool scan(const int *in, int *out, int count) {
        size_t bytes = 0;
        if (!hip_ok(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, count,
                                                          ctx->stream), "scan"))
            return false;
        void *tmp = nullptr;
        unsigned char *t = nullptr;
        if (!ps.get(t, bytes)) return false;
        tmp = t;
        return hip_ok(ctx, hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, count,
                                                            ctx->stream), "scan");
    }
    // exclusive prefix sums of count Cnt4 entries
    bool scan4(const Cnt4 *in, Cnt4 *out, int count) {
        size_t bytes = 0;
        const Cnt4 zero{{0, 0, 0, 0}};
        if (!hip_ok(ctx, hipcub::DeviceScan::ExclusiveScan(nullptr, bytes, in, out, Cnt4Sum(),
                                                           zero, count, ctx->stream), "scan"))
            return false;
        unsigned char *t = nullptr;
        if (!ps.get(t, bytes)) return false;
        return hip_ok(ctx, hipcub::DeviceScan::ExclusiveScan(t, bytes, in, out, Cnt4Sum(), zero,
                                                             count, ctx->stream), "scan");
    }
    // stable radix sort of (key, value) pairs on the low `bits` bits of the keys
    bool sort(const int *k_in, int *k_out, const int *v_in, int *v_out, int count, int bits) {
        if (count == 0) return true;
        size_t bytes = 0;
        if (!hip_ok(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, k_in, k_out, v_in,
                                                            v_out, count, 0, bits, ctx->stream),
                    "radix sort"))
            return false;
        unsigned char *t = nullptr;
        if (!ps.get(t, bytes)) return false;
        return hip_ok(ctx, hipcub::DeviceRadixSort::SortPairs(t, bytes, k_in, k_out, v_in, v_out,
                                                              count, 0, bits, ctx->stream),
                      "radix sort");
    }
};

// last element of a device int array (a scan's total), waited for
bool last_int(cicy_ctx *ctx, const int *p, long long at, int *out) {
    return d2h_sync(ctx, out, p + at, sizeof(int));
}

// entries of device int arrays (each at its own index) and, optionally, the first nextra
// entries of `extra` (into extra_out), with one wait; `buf` holds 8 + nextra ints
bool pick_ints(cicy_ctx *ctx, int *buf, std::initializer_list<std::pair<const int *, int>> at,
               int *out, const int *extra = nullptr, int nextra = 0, int *extra_out = nullptr) {
    PickArgs a{};
    int k = 0;
    for (const auto &x : at) {
        a.p[k] = x.first;
        a.at[k] = x.second;
        k++;
    }
    hipLaunchKernelGGL(k_pick_ints, dim3(1), dim3(64), 0, ctx->stream, a, k, extra, nextra, buf);
    std::vector<int> h((size_t)8 + nextra);
    if (!hip_ok(ctx, hipGetLastError(), "k_pick_ints") ||
        !d2h_sync(ctx, h.data(), buf, (nextra > 0 ? 8 + (size_t)nextra : (size_t)k) * sizeof(int)))
        return false;
    std::copy(h.begin(), h.begin() + k, out);
    if (nextra > 0) std::copy(h.begin() + 8, h.end(), extra_out);
    return true;
}

// the entries at `at` of several device int arrays, with one wait; `buf` holds 8 ints
bool last_ints(cicy_ctx *ctx, int *buf, std::initializer_list<const int *> arrs, int at,
               int *out) {
    LastArgs a{};
    int k = 0;
    for (const int *p : arrs) a.p[k++] = p;
    hipLaunchKernelGGL(k_last, dim3(1), dim3(64), 0, ctx->stream, a, k, at, buf);
    return hip_ok(ctx, hipGetLastError(), "k_last") &&
           d2h_sync(ctx, out, buf, (size_t)k * sizeof(int));
}

template <typename T>
bool dev_alloc(cicy_ctx *ctx, T **p, size_t count) {
    *p = static_cast<T *>(ctx->pool.get(std::max<size_t>(count, 1) * sizeof(T)));
    if (!*p) set_error(ctx, "hipMalloc of %zu bytes failed", count * sizeof(T));
    return *p != nullptr;
}

#define LAUNCH(kernel, count, ...) \
    hipLaunchKernelGGL(kernel, dim3(grid_of(count)), dim3(kST), 0, st, __VA_ARGS__)

// Symmetric CSR (both triangles, ascending columns) of the strictly-lower CSR (rp, col,
// val) with m rows and nz entries; frp/fcol/fval are allocated by the caller with 2 nz
// entries.
bool lower_to_full(cicy_ctx *ctx, PoolScope &ps, int m, int nz, const int *rp, const int *col,
                   const double *val, int *frp, int *fcol, double *fval) {
    hipStream_t st = ctx->stream;
    Cub cub{ctx, ps};
    int *lcnt, *tcnt, *cnt, *tstart, *tkey, *tidx, *trow, *skey, *sidx;
    if (!ps.get(lcnt, (size_t)m + 1) || !ps.get(tcnt, (size_t)m + 1) ||
        !ps.get(cnt, (size_t)m + 1) || !ps.get(tstart, (size_t)m + 1) ||
        !ps.get(tkey, (size_t)nz) || !ps.get(tidx, (size_t)nz) || !ps.get(trow, (size_t)nz) ||
        !ps.get(skey, (size_t)nz) || !ps.get(sidx, (size_t)nz))
        return false;
    if (!hip_ok(ctx, hipMemsetAsync(tcnt, 0, ((size_t)m + 1) * sizeof(int), st), "memset"))
        return false;
    LAUNCH(k_low_counts, m, m, rp, col, lcnt, tcnt);
    LAUNCH(k_add_counts, m + 1, m, lcnt, tcnt, cnt);
    if (!cub.scan(cnt, frp, m + 1) || !cub.scan(tcnt, tstart, m + 1)) return false;
    LAUNCH(k_low_emit, m, m, rp, col, val, frp, fcol, fval, tkey, tidx, trow);
    if (!cub.sort(tkey, skey, tidx, sidx, nz, bits_for(m))) return false;
    LAUNCH(k_up_place, nz, nz, skey, sidx, tstart, frp, lcnt, trow, val, fcol, fval);
    return hip_ok(ctx, hipGetLastError(), "lower_to_full");
}

}  // namespace

// ------------------------------------------------------------------------------------
// 1. band + remainder layout from the caller's device arrays
// ------------------------------------------------------------------------------------

bool band_from_device(cicy_sys *s, const int64_t *row_ptr, const int64_t *col_ind,
                      const double *values, const double *diag) {
    cicy_ctx *ctx = s->ctx;
    hipStream_t st = ctx->stream;
    const int n = s->n, nnz = s->nnz, rb = s->rb;
    PoolScope ps(ctx);
    Cub cub{ctx, ps};
    // err[0]: the first bad row's key (k_rows_check, which checks every row's bounds before
    // it reads the row); err[1], err[2]: row_ptr[0] and row_ptr[n]; one read-back for all
    unsigned long long *err;
    if (!ps.get(err, 3) ||
        !hip_ok(ctx, hipMemsetAsync(err, 0xff, sizeof(unsigned long long), st), "memset"))
        return false;
    if (n > 0) {
        hipLaunchKernelGGL(k_row_ends, dim3(1), dim3(64), 0, st, n, row_ptr,
                           reinterpret_cast<int64_t *>(err + 1));
        LAUNCH(k_rows_check, n, n, (long long)nnz, row_ptr, col_ind, diag, err);
    }
    unsigned long long got[3] = {~0ull, 0, 0};
    if (!hip_ok(ctx, hipGetLastError(), "k_rows_check") ||
        !d2h_sync(ctx, got, err, (n > 0 ? 3 : 1) * sizeof(unsigned long long)))
        return false;
    const unsigned long long key = got[0];
    if (n > 0 && ((int64_t)got[1] != 0 || (int64_t)got[2] != nnz)) {
        set_error(ctx, "row_ptr must start at 0 and end at nnz (row_ptr[n]=%lld, nnz=%lld)",
                  (long long)(int64_t)got[2], (long long)nnz);
        return false;
    }
    if (key != ~0ull) {
        const long long r = (long long)((key >> 31) >> 2);
        const int kind = (int)((key >> 31) & 3);
        if (kind == 1) {
            set_error(ctx, "row_ptr decreases or exceeds nnz at row %lld", r);
        } else if (kind == 2) {
            set_error(ctx, "zero or non-finite diagonal at row %lld (floating node)", r);
        } else {
            int64_t a = 0, c = 0;
            const long long off = (long long)(key & ((1ull << 31) - 1));
            (void)d2h_sync(ctx, &a, row_ptr + r, 8);
            (void)d2h_sync(ctx, &c, col_ind + a + off, 8);
            set_error(ctx, "col_ind[%lld]=%lld is not strictly below row %lld",
                      (long long)(a + off), (long long)c, r);
        }
        return false;
    }
    // the layout
    const size_t nv = (size_t)n;
    int *own, *colcnt, *cnt, *ostart, *cstart;
    if (!ps.get(own, nv + 1) || !ps.get(colcnt, nv + 1) || !ps.get(cnt, nv + 1) ||
        !ps.get(ostart, nv + 1) || !ps.get(cstart, nv + 1) ||
        !dev_alloc(ctx, &s->band, nv + 1) || !dev_alloc(ctx, &s->xrp, nv + 1) ||
        !hip_ok(ctx, hipMemsetAsync(colcnt, 0, (nv + 1) * sizeof(int), st), "memset"))
        return false;
    s->device_bytes += (nv + 1) * (sizeof(double) + sizeof(int));
    if (n > 0) LAUNCH(k_band_rows, n, n, row_ptr, col_ind, values, s->band, own, colcnt);
    else if (!hip_ok(ctx, hipMemsetAsync(s->band, 0, sizeof(double), st), "memset")) return false;
    if (n == 0 && (!hip_ok(ctx, hipMemsetAsync(own, 0, sizeof(int), st), "memset")))
        return false;
    LAUNCH(k_add_counts, nv + 1, n, own, colcnt, cnt);
    int nhalf = 0, noff = 0;
    int *totbuf, tot[8];
    if (!ps.get(totbuf, 8) || !cub.scan(cnt, s->xrp, n + 1) || !cub.scan(own, ostart, n + 1) ||
        !cub.scan(colcnt, cstart, n + 1) || !last_ints(ctx, totbuf, {s->xrp, ostart}, n, tot))
        return false;
    nhalf = tot[0];
    noff = tot[1];
    s->nx = nhalf;
    int *tkey, *tidx, *trow, *skey, *sidx;
    double *tval;
    if (!dev_alloc(ctx, &s->xnbr, (size_t)nhalf) || !dev_alloc(ctx, &s->xval, (size_t)nhalf) ||
        !ps.get(tkey, (size_t)noff) || !ps.get(tidx, (size_t)noff) || !ps.get(trow, (size_t)noff) ||
        !ps.get(tval, (size_t)noff) || !ps.get(skey, (size_t)noff) || !ps.get(sidx, (size_t)noff))
        return false;
    s->device_bytes += (size_t)nhalf * (sizeof(int) + sizeof(double));
    if (n > 0)
        LAUNCH(k_band_emit, n, n, row_ptr, col_ind, values, s->xrp, ostart, s->xnbr, s->xval,
               tkey, tidx, trow, tval);
    if (!cub.sort(tkey, skey, tidx, sidx, noff, bits_for(n))) return false;
    if (noff > 0)
        LAUNCH(k_band_transpose, noff, noff, skey, sidx, cstart, s->xrp, own, trow, tval, s->xnbr,
               s->xval);
    // row blocks and their launch order (heaviest blocks first, as build_band orders them),
    // found on the device
    int *hv, *hc, *light, *nheavy;
    const size_t nbv = std::max<size_t>((size_t)s->nblocks, 1);
    if (!dev_alloc(ctx, &s->xblk, (size_t)s->nblocks + 1) || !dev_alloc(ctx, &s->xorder, nbv) ||
        !ps.get(hv, nbv) || !ps.get(hc, nbv) || !ps.get(light, nbv) || !ps.get(nheavy, 1))
        return false;
    LAUNCH(k_xblk, s->nblocks + 1, s->nblocks, n, rb, s->xrp, s->xblk);
    if (s->nblocks > 0) {
        const double mean = (double)nhalf / s->nblocks;
        hipLaunchKernelGGL(k_band_split, dim3(1), dim3(kOrderThreads), 0, st, s->nblocks, s->xblk,
                           4.0 * mean, hv, hc, nheavy, light);
        LAUNCH(k_band_rank, s->nblocks, s->nblocks, hv, hc, nheavy, light, s->xorder);
    }
    // the caller's arrays are read only during the call: wait for the last read
    return hip_ok(ctx, hipGetLastError(), "band layout") &&
           hip_ok(ctx, hipMemcpyAsync(s->diag, diag, nv * sizeof(double),
                                      hipMemcpyDeviceToDevice, st), "copy diagonal") &&
           hip_ok(ctx, hipStreamSynchronize(st), "sync");
}

// ------------------------------------------------------------------------------------
// 2 + 3. condensation and series rounds of a system whose band layout is on the device
// ------------------------------------------------------------------------------------

void condensed_free_device(cicy_cond *c) {
    void *ptrs[] = {c->d_frp, c->d_fcol, c->d_orig, c->d_fval, c->d_cdiag};
    for (void *q : ptrs) dev_release(c->ctx, q);
    c->d_frp = c->d_fcol = c->d_orig = nullptr;
    c->d_fval = c->d_cdiag = nullptr;
}

cicy_cond *condensed_setup_device(cicy_sys *s, const char **why, bool *failed) {
    *failed = false;
    cicy_ctx *ctx = s->ctx;
    hipStream_t st = ctx->stream;
    const int n = s->n;
    if (n == 0) {
        *why = "empty system";
        return nullptr;
    }
    StageClock clk;
    PoolScope ps(ctx);
    Cub cub{ctx, ps};
    const size_t nv = (size_t)n;
    cicy_cond *c = new cicy_cond();
    c->ctx = ctx;
    c->on_device = true;
    c->n = n;
    c->nb = std::max(1, std::min(s->nblocks, 1024));
    auto fail = [&]() -> cicy_cond * {
        condensed_free(c);
        *failed = true;
        return nullptr;
    };
    // anchors
    int *flag, *apos;
    if (!ps.get(flag, nv + 1) || !ps.get(apos, nv + 1)) return fail();
    LAUNCH(k_anchor_flags, nv + 1, n, s->xrp, flag);
    int m = 0;
    if (!cub.scan(flag, apos, n + 1) || !last_int(ctx, apos, n, &m)) return fail();
    if ((long long)m * 2 > n) {
        condensed_free(c);
        *why = "more than half of the rows carry junction entries: nothing worth condensing";
        return nullptr;
    }
    c->m = m;
    if (!dev_alloc(ctx, &c->anchor_of, nv) || !dev_alloc(ctx, &c->anchors, (size_t)m))
        return fail();
    LAUNCH(k_anchor_index, n, n, flag, apos, c->anchor_of, c->anchors);
    // paths
    int *pflag, *pscan;
    if (!ps.get(pflag, nv + 1) || !ps.get(pscan, nv + 1)) return fail();
    LAUNCH(k_path_flags, nv + 1, n, c->anchor_of, s->band, pflag);
    int np = 0;
    if (!cub.scan(pflag, pscan, n + 1) || !last_int(ctx, pscan, n, &np)) return fail();
    c->npath = np;
    double *g, *termL, *termR, *fillv;
    unsigned long long *perr;
    if (!dev_alloc(ctx, &c->ps, (size_t)np) || !dev_alloc(ctx, &c->pe, (size_t)np) ||
        !dev_alloc(ctx, &c->pl, (size_t)np) || !dev_alloc(ctx, &c->pr, (size_t)np) ||
        !dev_alloc(ctx, &c->lfac, nv) || !dev_alloc(ctx, &c->mfac, nv) ||
        !dev_alloc(ctx, &c->band, nv + 1) || !dev_alloc(ctx, &c->y, nv) ||
        !dev_alloc(ctx, &c->delta, nv) || !dev_alloc(ctx, &c->rhs, nv) ||
        !dev_alloc(ctx, &c->norm, 2) || !dev_alloc(ctx, &c->counter, 1) ||
        !dev_alloc(ctx, &c->part, 2 * (size_t)c->nb) || !dev_alloc(ctx, &c->gpart, 2 * (size_t)c->nb) ||
        !dev_alloc(ctx, &c->gcount, (size_t)c->nb) || !ps.get(g, nv) ||
        !ps.get(termL, (size_t)np) || !ps.get(termR, (size_t)np) || !ps.get(fillv, (size_t)np) ||
        !ps.get(perr, 1) ||
        !hip_ok(ctx, hipEventCreate(&c->ev0), "hipEventCreate") ||
        !hip_ok(ctx, hipEventCreate(&c->ev1), "hipEventCreate") ||
        !hip_ok(ctx, hipMemsetAsync(c->norm, 0, 2 * sizeof(double), st), "memset") ||
        !hip_ok(ctx, hipMemsetAsync(c->counter, 0, sizeof(unsigned), st), "memset") ||
        !hip_ok(ctx, hipMemsetAsync(c->part, 0, 2 * (size_t)c->nb * sizeof(double), st), "memset") ||
        !hip_ok(ctx, hipMemsetAsync(c->gpart, 0, 2 * (size_t)c->nb * sizeof(double), st), "memset") ||
        !hip_ok(ctx, hipMemsetAsync(c->gcount, 0, (size_t)c->nb * sizeof(unsigned), st), "memset") ||
        !hip_ok(ctx, hipMemsetAsync(perr, 0xff, sizeof(unsigned long long), st), "memset") ||
        !hip_ok(ctx, hipMemcpyAsync(c->band, s->band, (nv + 1) * sizeof(double),
                                    hipMemcpyDeviceToDevice, st), "copy band"))
        return fail();
    LAUNCH(k_path_starts, n, n, pflag, pscan, c->anchor_of, s->band, c->ps, c->pe);
    LAUNCH(k_fill_factors, n, n, c->lfac, c->mfac);
    if (np > 0)
        LAUNCH(k_paths, np, np, n, c->ps, c->anchor_of, s->band, s->diag, c->pe, c->pl, c->pr,
               c->lfac, c->mfac, g, termL, termR, fillv, perr);
    unsigned long long pkey = ~0ull;
    if (!hip_ok(ctx, hipGetLastError(), "paths") || !d2h_sync(ctx, &pkey, perr, 8)) return fail();
    if (pkey != ~0ull) {
        condensed_free(c);
        *why = (pkey & 3) == 1 ? "non-positive pivot on a path coupled to the rest (not an M-matrix)"
                               : "non-positive pivot inside a floating path";
        return nullptr;
    }
    clk.mark("device condense: paths");
    // S: diagonal, rows (raw, sorted, summed)
    double *sdiag;
    int *rcnt, *roff, *rcol_raw, *ucnt, *rrp, *rcol;
    double *rval_raw, *rval;
    if (!ps.get(sdiag, (size_t)m) || !ps.get(rcnt, (size_t)m + 1) || !ps.get(roff, (size_t)m + 1) ||
        !ps.get(ucnt, (size_t)m + 1) || !ps.get(rrp, (size_t)m + 1))
        return fail();
    if (m > 0) {
        LAUNCH(k_sdiag, m, m, n, c->anchors, c->anchor_of, s->band, s->diag, pscan, termL, termR,
               sdiag);
        LAUNCH(k_s_raw<0>, m, m, c->anchors, c->anchor_of, s->band, s->xrp, s->xnbr, s->xval,
               pscan, c->pl, fillv, rcnt, (const int *)nullptr, (int *)nullptr, (double *)nullptr);
    } else if (!hip_ok(ctx, hipMemsetAsync(rcnt, 0, sizeof(int), st), "memset")) {
        return fail();
    }
    int nraw = 0;
    if (!cub.scan(rcnt, roff, m + 1) || !last_int(ctx, roff, m, &nraw) ||
        !ps.get(rcol_raw, (size_t)nraw) || !ps.get(rval_raw, (size_t)nraw))
        return fail();
    if (m > 0) {
        LAUNCH(k_s_raw<1>, m, m, c->anchors, c->anchor_of, s->band, s->xrp, s->xnbr, s->xval,
               pscan, c->pl, fillv, rcnt, roff, rcol_raw, rval_raw);
        LAUNCH(k_rows_sort, m, m, roff, rcol_raw, rval_raw, ucnt);
    } else if (!hip_ok(ctx, hipMemsetAsync(ucnt, 0, sizeof(int), st), "memset")) {
        return fail();
    }
    int nred = 0;
    if (!cub.scan(ucnt, rrp, m + 1) || !last_int(ctx, rrp, m, &nred) ||
        !ps.get(rcol, (size_t)nred) || !ps.get(rval, (size_t)nred))
        return fail();
    if (m > 0) LAUNCH(k_rows_sum, m, m, roff, rcol_raw, rval_raw, rrp, rcol, rval);
    c->nnz_red = nred;
    clk.mark("device condense: reduced rows");

    // ---- series rounds over S --------------------------------------------------------
    const size_t mv = (size_t)m;
    int cap = std::max(2 * nred, 1);
    int *arp, *acol, *brp, *bcol;
    double *aval, *bval;
    int *alive, *state, *sel, *selpos, *degoff, *rec_of, *tflag, *tpos, *cpos, *fpos, *upos, *len,
        *totbuf, *rs;
    Cnt4 *c4, *c4s;
    double *d;
    unsigned int *undecided;
    if (!ps.get(arp, mv + 1) || !ps.get(acol, (size_t)cap) || !ps.get(aval, (size_t)cap) ||
        !ps.get(brp, mv + 1) || !ps.get(alive, mv + 1) || !ps.get(state, mv + 1) ||
        !ps.get(sel, mv + 1) || !ps.get(selpos, mv + 1) || !ps.get(c4, mv + 1) ||
        !ps.get(c4s, mv + 1) ||
        !ps.get(degoff, mv + 1) || !ps.get(rec_of, mv + 1) || !ps.get(tflag, mv + 1) ||
        !ps.get(tpos, mv + 1) || !ps.get(cpos, mv + 1) ||
        !ps.get(fpos, mv + 1) ||
        !ps.get(upos, mv + 1) || !ps.get(len, mv + 1) || !ps.get(d, mv + 1) ||
        !ps.get(undecided, kMisBatch + 1) || !ps.get(totbuf, 8 + 3 + kMaxRounds + 1) ||
        !ps.get(rs, 3 + kMaxRounds + 1) ||
        !hip_ok(ctx, hipMemsetAsync(rs, 0, (3 + kMaxRounds + 1) * sizeof(int), st), "memset") ||
        !hip_ok(ctx, hipMemsetAsync(undecided, 0, (kMisBatch + 1) * sizeof(unsigned), st),
                "memset"))
        return fail();
    // set by k_mis_async when it gives up; cleared again before the passes finish the set
    int *gave_up = reinterpret_cast<int *>(undecided + kMisBatch);
    // the cooperative independent-set grid: every work-group resident, at most one thread
    // per node
    int mis_grid = 0;
    if (mis_async_enabled() && m > 0) {
        int cus = 0, per_cu = 0;
        if (!hip_ok(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                               ctx->device),
                    "hipDeviceGetAttribute") ||
            !hip_ok(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                             &per_cu, reinterpret_cast<const void *>(k_mis_async), kST, 0),
                    "hipOccupancyMaxActiveBlocksPerMultiprocessor"))
            return fail();
        mis_grid = (int)std::min<long long>((long long)cus * per_cu, (m + kST - 1) / kST);
    }
    if (!lower_to_full(ctx, ps, m, nred, rrp, rcol, rval, arp, acol, aval) ||
        !hip_ok(ctx, hipMemcpyAsync(d, sdiag, mv * sizeof(double), hipMemcpyDeviceToDevice, st),
                "copy") ||
        !hip_ok(ctx, hipMemsetAsync(alive, 0, (mv + 1) * sizeof(int), st), "memset"))
        return fail();
    // every node starts alive: alive = 1 on [0, m)
    if (m > 0) LAUNCH(k_fill_i32, m, alive, m, 1);
    // record and receiver arrays: at most m records, D m record entries, D m contributions
    // (each entry of a record is one contribution), D m receivers (D the degree cap)
    const int max_degree = elim_degree();
    const size_t rcap = mv + 1, ecap = (size_t)max_degree * mv + 1;
    if (!dev_alloc(ctx, &c->e_node, rcap) || !dev_alloc(ctx, &c->e_off, rcap + 1) ||
        !dev_alloc(ctx, &c->e_d, rcap) || !dev_alloc(ctx, &c->e_nbr, ecap) ||
        !dev_alloc(ctx, &c->e_a, ecap) || !dev_alloc(ctx, &c->rv_node, ecap) ||
        !dev_alloc(ctx, &c->rv_off, ecap + 1) || !dev_alloc(ctx, &c->rv_rec, ecap) ||
        !dev_alloc(ctx, &c->rv_a, ecap) || !dev_alloc(ctx, &c->fS, mv) ||
        !dev_alloc(ctx, &c->xS, mv))
        return fail();
    c->round_off.assign(1, 0);
    c->rv_round_off.assign(1, 0);
    int R0 = 0, E0 = 0, RV0 = 0, C0 = 0;
    std::vector<int> rs_host(3 + kMaxRounds + 1, 0);  // rs as of the last selection
    long long left = m;  // nodes not eliminated yet
    while (m > 0) {
        LAUNCH(k_cand, m, m, arp, d, sdiag, alive, state, max_degree);
        // passes in batches of kMisBatch launches, each counting the candidates left waiting
        // into its own word; a pass after one that left none returns at once
        int passes = 0;  // batch passes run (none when the cooperative set settles)
        auto mis_passes = [&]() -> bool {
            for (;;) {
                unsigned int und[kMisBatch];
                if (!hip_ok(ctx, hipMemsetAsync(undecided, 0, kMisBatch * sizeof(unsigned), st),
                            "memset"))
                    return false;
                for (int q = 0; q < kMisBatch; q++)
                    LAUNCH(k_mis_pass, m, m, arp, acol, state,
                           q > 0 ? undecided + q - 1 : (const unsigned *)nullptr, undecided + q);
                if (!hip_ok(ctx, hipGetLastError(), "k_mis_pass") ||
                    !d2h_sync(ctx, und, undecided, sizeof(und)))
                    return false;
                bool settled = false;
                for (int q = 0; q < kMisBatch && !settled; q++) {
                    passes++;
                    settled = und[q] == 0;
                }
                if (settled) return true;
                if (passes > kMaxMisPasses) {
                    set_error(ctx, "series rounds: independent set not settled after %d passes",
                              kMaxMisPasses);
                    return false;
                }
            }
        };
        // tot: 0 selected nodes, 1 their record entries, 2 a bound on the round's fill,
        // 3 the adjacency's entries, 4 whether the cooperative set gave up, 5 whether the
        // last round overflowed its arrays
        int tot[8];
        const int q = (int)c->round_off.size() - 1;  // this round's index
        auto select = [&]() -> bool {
            LAUNCH(k_sel_flags, mv + 1, m, state, arp, sel, c4);
            if (!cub.scan4(c4, c4s, m + 1)) return false;
            LAUNCH(k_unpack4, mv + 1, m + 1, c4s, selpos, degoff, (int *)nullptr,
                   (int *)nullptr);
            return pick_ints(ctx, totbuf,
                             {{selpos, m},
                              {degoff, m},
                              {reinterpret_cast<const int *>(c4s), 4 * m + 2},
                              {arp, m},
                              {gave_up, 0},
                              {rs, 0}},
                             tot, rs, q + 4, rs_host.data());
        };
        if (mis_grid > 0) {
            // the whole set in one cooperative launch; should it give up, the passes finish it
            unsigned sweeps = kMisMaxSweeps;
            void *args[] = {(void *)&m, (void *)&arp, (void *)&acol, (void *)&state,
                            (void *)&gave_up, (void *)&sweeps};
            if (!hip_ok(ctx, hipLaunchCooperativeKernel(reinterpret_cast<const void *>(k_mis_async),
                                                        dim3(mis_grid), dim3(kST), args, 0, st),
                        "hipLaunchCooperativeKernel (independent set)") ||
                !select())
                return fail();
            if (tot[4] != 0 &&
                (!hip_ok(ctx, hipMemsetAsync(gave_up, 0, sizeof(int), st), "memset") ||
                 !mis_passes() || !select()))
                return fail();
        } else if (!mis_passes() || !select()) {
            return fail();
        }
        if (tot[5] != 0) {
            set_error(ctx, "series rounds: receiver or merge arrays overflow");
            return fail();
        }
        const int nsel = tot[0], nent = tot[1];
        // a round that would take fewer than 1 % of the nodes left is not worth its pass over
        // the whole system: the rounds stop (the host form stops by the same rule)
        if (nsel == 0 || (long long)nsel * min_round_fraction() < left) break;
        left -= nsel;
        if ((size_t)R0 + nsel > rcap - 1 || (size_t)E0 + nent > ecap - 1 || q >= kMaxRounds) {
            set_error(ctx, "series rounds: record arrays overflow");
            return fail();
        }
        // the fill of a receiver is at most deg - 1 per eliminated neighbour of degree deg,
        // and a new row holds at most its old entries and its fill
        const long long fill_cap = tot[2], ub_cap = (long long)tot[3] + tot[2];
        LAUNCH(k_records, m, m, sel, selpos, degoff, arp, acol, aval, d, R0, E0, c->e_node,
               c->e_off, c->e_d, c->e_nbr, c->e_a, rec_of);
        LAUNCH(k_targets, mv + 1, m, alive, sel, arp, acol, tflag, c4);
        if (!cub.scan4(c4, c4s, m + 1)) return fail();
        LAUNCH(k_unpack4, mv + 1, m + 1, c4s, tpos, cpos, fpos, upos);
        int *fc, *oc;
        double *fvv, *ov;
        if (!ps.get(fc, (size_t)fill_cap) || !ps.get(fvv, (size_t)fill_cap) ||
            !ps.get(oc, (size_t)ub_cap) || !ps.get(ov, (size_t)ub_cap))
            return fail();
        hipLaunchKernelGGL(k_round_guard, dim3(1), dim3(64), 0, st, m, tpos, cpos, fpos, upos, rs,
                           (long long)ecap - 1, fill_cap, ub_cap);
        LAUNCH(k_receivers, m, m, tflag, tpos, cpos, arp, acol, aval, sel, rec_of, c->e_d, rs,
               c->rv_node, c->rv_off, c->rv_rec, c->rv_a, d);
        // the end of the round's last record (the merge reads e_off[rec + 1])
        hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, st, c->e_off + R0 + nsel, E0 + nent);
        LAUNCH(k_merge, mv + 1, m, alive, sel, tflag, arp, acol, aval, upos, fpos, rec_of, c->e_off,
               c->e_nbr, c->e_a, c->e_d, fc, fvv, oc, ov, len, rs);
        // the new rows hold at most ub_cap entries, so no read-back of their true total is
        // needed
        if (!cub.scan(len, brp, m + 1) || !ps.get(bcol, (size_t)std::max(ub_cap, 1ll)) ||
            !ps.get(bval, (size_t)std::max(ub_cap, 1ll)))
            return fail();
        LAUNCH(k_compact, m, m, upos, brp, oc, ov, bcol, bval, rs);
        hipLaunchKernelGGL(k_round_end, dim3(1), dim3(64), 0, st, m, tpos, cpos, rs, q);
        LAUNCH(k_kill, m, m, sel, alive);
        // return the old adjacency and the merge buffers to the pool, keep the new one
        for (void *p : {(void *)acol, (void *)aval, (void *)fc, (void *)fvv, (void *)oc, (void *)ov}) {
            ctx->pool.put(p);
            ps.got.erase(std::remove(ps.got.begin(), ps.got.end(), p), ps.got.end());
        }
        std::swap(arp, brp);
        acol = bcol;
        aval = bval;
        R0 += nsel;
        E0 += nent;
        c->round_off.push_back(R0);
        if (!hip_ok(ctx, hipGetLastError(), "series round")) return fail();
        if (clk.on) {
            int ex[8];
            if (!last_ints(ctx, totbuf, {tpos, fpos}, m, ex)) return fail();
            fprintf(stderr, "[cicy setup]   round %zu: %d eliminated, %d receivers, %d fill, "
                    "%d independent-set passes\n", c->round_off.size() - 1, nsel, ex[0], ex[1],
                    passes);
            clk.mark("device condense: round");
        }
    }
    // the receivers' round offsets and totals, read with the last round's selection (or, with
    // no round at all, zeros)
    const int rounds = (int)c->round_off.size() - 1;
    RV0 = rs_host[1];
    C0 = rs_host[2];
    c->rv_round_off.assign(rs_host.begin() + 3, rs_host.begin() + 3 + rounds + 1);
    // closing offsets of the record and receiver lists
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, st, c->e_off + R0, E0);
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, st, c->rv_off + RV0, C0);
    c->nrec = R0;
    clk.mark("device condense: series rounds");
    // the core
    int *cflag, *core_of, *clen, *cpos2;
    if (!ps.get(cflag, mv + 1) || !ps.get(core_of, mv + 1) || !ps.get(clen, mv + 1) ||
        !ps.get(cpos2, mv + 1))
        return fail();
    int mc = 0, nnz2 = 0;
    if (m > 0) {
        LAUNCH(k_core_len, mv + 1, m, alive, arp, cflag, clen);
        int ct[8];
        if (!cub.scan(cflag, core_of, m + 1) || !cub.scan(clen, cpos2, m + 1) ||
            !last_ints(ctx, totbuf, {core_of, cpos2}, m, ct))
            return fail();
        mc = ct[0];
        nnz2 = ct[1];
    }
    c->mc = mc;
    c->nnz_core = nnz2 / 2;
    if (mc > 0) {
        if (!dev_alloc(ctx, &c->core_nodes, (size_t)mc) || !dev_alloc(ctx, &c->d_frp, (size_t)mc + 1) ||
            !dev_alloc(ctx, &c->d_fcol, (size_t)nnz2) || !dev_alloc(ctx, &c->d_fval, (size_t)nnz2) ||
            !dev_alloc(ctx, &c->d_cdiag, (size_t)mc) || !dev_alloc(ctx, &c->d_orig, (size_t)mc) ||
            !dev_alloc(ctx, &c->core_b, (size_t)mc) || !dev_alloc(ctx, &c->core_x, (size_t)mc))
            return fail();
        LAUNCH(k_core_fill, m, m, alive, core_of, cpos2, arp, acol, aval, d, sdiag, c->anchors,
               c->core_nodes, c->d_frp, c->d_fcol, c->d_fval, c->d_cdiag, c->d_orig);
        hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, st, c->d_frp + mc, nnz2);
        // a core row with a vanishing diagonal that still has neighbours: no condensed
        // system, as on the host
        int bad = 0;
        if (!hip_ok(ctx, hipMemsetAsync(totbuf, 0, sizeof(int), st), "memset")) return fail();
        LAUNCH(k_nonpositive, mc, mc, c->d_cdiag, totbuf);
        if (!hip_ok(ctx, hipGetLastError(), "core") || !d2h_sync(ctx, &bad, totbuf, sizeof(int)))
            return fail();
        if (bad) {
            condensed_free(c);
            *why = "the core system has a non-positive diagonal (floating component)";
            return nullptr;
        }
    }
    if (!hip_ok(ctx, hipStreamSynchronize(st), "sync")) return fail();
    clk.mark("device condense: core");
    return c;
}

}  // namespace cicy
