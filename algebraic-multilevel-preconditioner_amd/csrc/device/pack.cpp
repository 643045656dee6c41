// pack.cpp -- supernode detection, task split, level schedule and value layout of the batched
// local solve.
//
// Input: the factors of all subdomains in compressed-column form (L, and U^T for LU), the form
// the reference's LocalSolver backends expose. Output: one table of supernodes per sweep; dense
// blocks (diagonal block column-major, rows below row-major); for the forward sweep the lists of
// products each row pulls from its descendants; and the tasks of the work queue: the supernodal
// elimination tree is cut into subtree tasks of balanced dense work and one task above them (the
// skeleton with the root). The queue lists the forward entries of all subtree tasks first, then
// the top tasks (both sweeps in one entry), then the subtrees' backward entries, all subdomains
// interleaved (a level-ordered work queue).
//
// Dense tops: the skeleton of a tree (the separators above its subtree tasks) is a chain of
// dependent wide supernodes in both sweeps. When its widths sum to at most dense_top_max, the
// skeleton is replaced by the dense inverse of its Schur complement, S^-1 = U_SS^-1 L_SS^-1: the
// top task pulls the subtrees' products into its S rows and multiplies by S^-1 (both sweeps of the
// skeleton in one product), its rows split over several workgroups. Its supernodes push nothing.
#include "pack.hpp"

#include <algorithm>
#include <atomic>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <thread>

namespace s2l {

std::vector<int32_t> detectSupernodes(int32_t n, const std::vector<int64_t>& cp, const std::vector<int32_t>& row)
{
    for (int32_t j = 0; j < n; ++j) {
        if (cp[j + 1] <= cp[j] || row[cp[j]] != j)
            throw std::invalid_argument("factor column " + std::to_string(j) + " does not start with its diagonal");
        for (int64_t p = cp[j] + 1; p < cp[j + 1]; ++p)
            if (row[p] <= row[p - 1] || row[p] >= n) throw std::invalid_argument("factor rows must be sorted and in range");
    }
    std::vector<int32_t> starts{0};
    for (int32_t j = 1; j < n; ++j) {
        const int64_t a0 = cp[j - 1], a1 = cp[j], b0 = cp[j], b1 = cp[j + 1];
        bool same = (a1 - a0) == (b1 - b0) + 1 && (a1 - a0) >= 2 && row[a0 + 1] == j;
        for (int64_t q = 0; same && q < (b1 - b0) - 1; ++q) same = row[a0 + 2 + q] == row[b0 + 1 + q];
        if (!same) starts.push_back(j);
    }
    starts.push_back(n);
    return starts;
}

namespace {

// Appends a level table, 7 ints per level: start and end of the level in lev_sn, start of its wide
// supernodes (the small ones come first), sum of the small widths, sum of the small heights,
// offset of its chunk list in lev_chunk and number of chunks. A chunk is a run of small
// supernodes processed together: at most kChunkSn of them and at most kChunkRows rows in total
// (the size of the kernel's staging area and result buffer in LDS). lev_wpre / lev_hpre hold, for
// every small supernode, the sum of the widths / heights of the small supernodes before it in its
// level.
//
// The same levels are also emitted as steps (StepDesc): every chunk, then every wide supernode of
// a level, in the order the kernel runs them. @p cptr / @p cend: the subdomain's slot ranges for
// the forward sweep (the pulled range [cptr[r], cend[r]) of every row), null for a backward sweep.
void appendLevels(Packed& P, const std::vector<std::vector<int32_t>>& byLevel, int32_t& lev_off, int32_t& nlev,
                  int32_t& step_off, int32_t& nsteps, const int32_t* cptr, const int32_t* cend)
{
    lev_off = static_cast<int32_t>(P.lev.size());
    nlev = static_cast<int32_t>(byLevel.size());
    step_off = static_cast<int32_t>(P.steps.size());
    for (const auto& list : byLevel) {
        const int32_t start = static_cast<int32_t>(P.lev_sn.size());
        int32_t wsum = 0, hsum = 0;
        for (int wide = 0; wide < 2; ++wide)
            for (int32_t s : list) {
                const SnDesc& d = P.sn[static_cast<std::size_t>(s)];
                if ((d.w > kSmallWidth) != (wide == 1)) continue;
                P.lev_sn.push_back(s);
                P.lev_wpre.push_back(wide ? 0 : wsum);
                P.lev_hpre.push_back(wide ? 0 : hsum);
                if (!wide) {
                    wsum += d.w;
                    hsum += d.h;
                }
            }
        const int32_t end = static_cast<int32_t>(P.lev_sn.size());
        int32_t split = start;
        while (split < end && P.sn[static_cast<std::size_t>(P.lev_sn[static_cast<std::size_t>(split)])].w <= kSmallWidth) ++split;
        const int32_t coff = static_cast<int32_t>(P.lev_chunk.size());
        int32_t q = start;
        while (q < split) {
            P.lev_chunk.push_back(q);
            int32_t rows = 0, cnt = 0;
            while (q < split && cnt < kChunkSn &&
                   rows + P.sn[static_cast<std::size_t>(P.lev_sn[static_cast<std::size_t>(q)])].w <= kChunkRows) {
                rows += P.sn[static_cast<std::size_t>(P.lev_sn[static_cast<std::size_t>(q)])].w;
                ++cnt;
                ++q;
            }
        }
        const int32_t nchunk = static_cast<int32_t>(P.lev_chunk.size()) - coff;
        P.lev_chunk.push_back(split);
        P.lev.insert(P.lev.end(), {start, end, split, wsum, hsum, coff, nchunk});
        // level-ordered descriptors, the rows of the small supernodes, and the steps
        const int32_t roff = static_cast<int32_t>(P.lev_row.size());
        for (int32_t q = start; q < end; ++q) P.lev_desc.push_back(P.sn[static_cast<std::size_t>(P.lev_sn[static_cast<std::size_t>(q)])]);
        if (cptr)
            for (int32_t q = start; q < split; ++q) {
                const SnDesc& d = P.sn[static_cast<std::size_t>(P.lev_sn[static_cast<std::size_t>(q)])];
                for (int32_t j = 0; j < d.w; ++j) {
                    P.lev_row.push_back(d.c0 + j);
                    P.lev_cb.push_back(cptr[d.c0 + j]);
                    P.lev_ce.push_back(cend[d.c0 + j]);
                }
            }
        for (int32_t ch = 0; ch < nchunk; ++ch) {
            const int32_t q0 = P.lev_chunk[static_cast<std::size_t>(coff + ch)], q1 = P.lev_chunk[static_cast<std::size_t>(coff + ch + 1)];
            StepDesc st{};
            st.kind = 0;
            st.q0 = q0;
            st.cnt = q1 - q0;
            st.wbase = P.lev_wpre[static_cast<std::size_t>(q0)];
            st.hbase = P.lev_hpre[static_cast<std::size_t>(q0)];
            st.W = (q1 < split ? P.lev_wpre[static_cast<std::size_t>(q1)] : wsum) - st.wbase;
            st.H = (q1 < split ? P.lev_hpre[static_cast<std::size_t>(q1)] : hsum) - st.hbase;
            st.roff = cptr ? roff + st.wbase : -1;
            P.steps.push_back(st);
        }
        for (int32_t q = split; q < end; ++q) {
            StepDesc st{};
            st.kind = 1;
            st.q0 = q;
            st.wide = P.sn[static_cast<std::size_t>(P.lev_sn[static_cast<std::size_t>(q)])];
            P.steps.push_back(st);
        }
    }
    nsteps = static_cast<int32_t>(P.steps.size()) - step_off;
}

} // namespace

// In-place inverse of the lower triangular m x m block stored column-major at B with leading
// dimension ld (B[c * ld + r] = L(r, c), r >= c), in FP64: column j of the inverse solves L x = e_j.
void invertLower(double* B, int ld, int m)
{
    std::vector<double> X(static_cast<std::size_t>(m) * m, 0.0);
    for (int j = 0; j < m; ++j) {
        double* x = X.data() + static_cast<std::size_t>(j) * m; // column j of the inverse
        x[j] = 1.0;
        for (int c = j; c < m; ++c) {
            x[c] /= B[static_cast<std::size_t>(c) * ld + c];
            const double xc = x[c];
            for (int r = c + 1; r < m; ++r) x[r] -= B[static_cast<std::size_t>(c) * ld + r] * xc;
        }
    }
    for (int c = 0; c < m; ++c)
        for (int r = 0; r < m; ++r) B[static_cast<std::size_t>(c) * ld + r] = (r >= c) ? X[static_cast<std::size_t>(c) * m + r] : 0.0;
}

// Same for an upper triangular block (B[c * ld + r] = U(r, c), r <= c): column j solves U x = e_j.
void invertUpper(double* B, int ld, int m)
{
    std::vector<double> X(static_cast<std::size_t>(m) * m, 0.0);
    for (int j = 0; j < m; ++j) {
        double* x = X.data() + static_cast<std::size_t>(j) * m;
        x[j] = 1.0;
        for (int c = j; c >= 0; --c) {
            x[c] /= B[static_cast<std::size_t>(c) * ld + c];
            const double xc = x[c];
            for (int r = 0; r < c; ++r) x[r] -= B[static_cast<std::size_t>(c) * ld + r] * xc;
        }
    }
    for (int c = 0; c < m; ++c)
        for (int r = 0; r < m; ++r) B[static_cast<std::size_t>(c) * ld + r] = (r <= c) ? X[static_cast<std::size_t>(c) * m + r] : 0.0;
}

namespace {

// Diagonal blocks as applied by the kernel: the inverse of the whole block, computed in FP64.
// A small supernode's stays column-major (a thread per row walks it: coalesced over the rows); a
// wide one's is stored row-major with the padded leading dimension wideLd(w), starting at a
// multiple of four entries (a wavefront reads a row in 16-byte pieces). The inversions (w^3 / 3
// flops each, most of the packing time for wide separators) run after the packing loop, spread
// over host threads: appendDiagonal reserves the block's place and keeps the block.
struct DiagJob {
    int64_t dst;
    int32_t w;
    bool lower;
};

inline std::size_t diagSize(int w) { return w <= kSmallWidth ? static_cast<std::size_t>(w) * w : static_cast<std::size_t>(w) * wideLd(w); }

// the block's place in out holds the block itself (column-major) until runDiagonalJobs
void appendDiagonal(std::vector<double>& out, std::vector<DiagJob>& jobs, const std::vector<double>& B, int w, bool lower)
{
    jobs.push_back(DiagJob{static_cast<int64_t>(out.size()), w, lower});
    out.resize(out.size() + diagSize(w), 0.0);
    std::copy(B.begin(), B.end(), out.begin() + jobs.back().dst);
}

void runDiagonalJobs(std::vector<double>& out, std::vector<DiagJob>& jobs)
{
    // widest first, so that the long inversions do not end up last on one thread
    std::stable_sort(jobs.begin(), jobs.end(), [](const DiagJob& a, const DiagJob& b) { return a.w > b.w; });
    std::atomic<std::size_t> next{0};
    auto work = [&] {
        std::vector<double> B;
        for (std::size_t k = next++; k < jobs.size(); k = next++) {
            const DiagJob& j = jobs[k];
            const int w = j.w;
            B.assign(out.begin() + j.dst, out.begin() + j.dst + static_cast<int64_t>(w) * w);
            if (j.lower)
                invertLower(B.data(), w, w);
            else
                invertUpper(B.data(), w, w);
            double* d = out.data() + j.dst;
            if (w <= kSmallWidth) {
                std::copy(B.begin(), B.end(), d);
                continue;
            }
            const int ld = wideLd(w);
            for (int r = 0; r < w; ++r)
                for (int c = 0; c < w; ++c) d[static_cast<std::size_t>(r) * ld + c] = B[static_cast<std::size_t>(c) * w + r];
        }
    };
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const std::size_t nt = std::min<std::size_t>(hw, std::max<std::size_t>(1, jobs.size() / 64));
    std::vector<std::thread> pool;
    for (std::size_t t = 1; t < nt; ++t) pool.emplace_back(work);
    work();
    for (std::thread& t : pool) t.join();
}

// Pads a value array to a multiple of four entries (the start of a wide supernode's blocks).
void alignFour(std::vector<double>& v)
{
    while (v.size() % 4 != 0) v.push_back(0.0);
}

// A dense top to compute: S^-1 of the skeleton rows @p rows (factor positions, increasing) of factor
// @p f, row-major with leading dimension wideLd(S) at diag[dst].
struct DenseJob {
    int64_t dst;
    int32_t f;
    std::vector<int32_t> rows;
};

// The S x S block of a factor in compressed-column form on the rows @p rows (column-major, lower
// triangular in their order); every entry of these columns must lie in these rows.
std::vector<double> skeletonBlock(int32_t n, const std::vector<int64_t>& cp, const std::vector<int32_t>& row,
                                  const std::vector<double>& val, const std::vector<int32_t>& rows)
{
    const int S = static_cast<int>(rows.size());
    std::vector<int32_t> qOf(static_cast<std::size_t>(n), -1);
    for (int q = 0; q < S; ++q) qOf[static_cast<std::size_t>(rows[static_cast<std::size_t>(q)])] = q;
    std::vector<double> B(static_cast<std::size_t>(S) * S, 0.0);
    for (int q = 0; q < S; ++q) {
        const int32_t j = rows[static_cast<std::size_t>(q)];
        for (int64_t p = cp[static_cast<std::size_t>(j)]; p < cp[static_cast<std::size_t>(j) + 1]; ++p) {
            const int32_t qr = qOf[static_cast<std::size_t>(row[static_cast<std::size_t>(p)])];
            if (qr < q) throw std::invalid_argument("dense top: a skeleton column has a row outside the skeleton");
            B[static_cast<std::size_t>(q) * S + qr] = val[static_cast<std::size_t>(p)];
        }
    }
    return B;
}

// S^-1 = U_SS^-1 L_SS^-1 in FP64 for every dense top, spread over host threads: X = L_SS^-1 and
// Y = (U_SS^T)^-1 (Y = X for Cholesky), both lower triangular, and S^-1(i, j) = sum over k >=
// max(i, j) of Y(k, i) X(k, j).
void runDenseJobs(std::vector<double>& out, const std::vector<DenseJob>& jobs, const std::vector<HostFactor>& fs)
{
    std::atomic<std::size_t> next{0};
    std::atomic<bool> failed{false};
    std::string error;
    auto work = [&] {
        try {
            for (std::size_t k = next++; k < jobs.size() && !failed; k = next++) {
                const DenseJob& jb = jobs[k];
                const HostFactor& F = fs[static_cast<std::size_t>(jb.f)];
                const int S = static_cast<int>(jb.rows.size()), ld = wideLd(S);
                std::vector<double> X = skeletonBlock(F.n, F.lcp, F.lr, F.lv, jb.rows);
                invertLower(X.data(), S, S);
                std::vector<double> Y;
                if (!F.cholesky) {
                    Y = skeletonBlock(F.n, F.ucp, F.ur, F.uv, jb.rows);
                    invertLower(Y.data(), S, S);
                }
                const std::vector<double>& Yc = F.cholesky ? X : Y;  // column-major: Y(k, i) at i S + k
                std::vector<double> Xr(static_cast<std::size_t>(S) * S);   // row-major: X(k, j) at k S + j
                for (int c = 0; c < S; ++c)
                    for (int r = 0; r < S; ++r) Xr[static_cast<std::size_t>(r) * S + c] = X[static_cast<std::size_t>(c) * S + r];
                std::vector<double> acc(static_cast<std::size_t>(S));
                for (int i = 0; i < S; ++i) {
                    std::fill(acc.begin(), acc.end(), 0.0);
                    for (int kk = i; kk < S; ++kk) {
                        const double y = Yc[static_cast<std::size_t>(i) * S + kk];
                        if (y == 0.0) continue;
                        const double* xr = Xr.data() + static_cast<std::size_t>(kk) * S;
                        for (int j = 0; j <= kk; ++j) acc[static_cast<std::size_t>(j)] += y * xr[j];
                    }
                    std::copy(acc.begin(), acc.end(), out.begin() + jb.dst + static_cast<int64_t>(i) * ld);
                }
            }
        } catch (const std::exception& e) {
            if (!failed.exchange(true)) error = e.what();
        }
    };
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const std::size_t nt = std::min<std::size_t>(hw, jobs.size());
    std::vector<std::thread> pool;
    for (std::size_t t = 1; t < nt; ++t) pool.emplace_back(work);
    if (!jobs.empty()) work();
    for (std::thread& t : pool) t.join();
    if (failed) throw std::invalid_argument(error);
}

} // namespace

Packed packFactors(const std::vector<HostFactor>& fs, const std::vector<int32_t>& map_off, int tasks_per_sub,
                   int dense_top_max)
{
    Packed P;
    std::vector<DiagJob> jobs;  // diagonal blocks to invert
    std::vector<DenseJob> djobs; // dense tops to compute
    // supernodes of every factor first, so that the value arrays are allocated once
    std::vector<std::vector<int32_t>> fstAll(fs.size()), bstAll(fs.size());
    {
        std::size_t ndiag = 0, nbelow = 0, nridx = 0, nsn = 0;
        for (std::size_t i = 0; i < fs.size(); ++i) {
            const HostFactor& F = fs[i];
            fstAll[i] = detectSupernodes(F.n, F.lcp, F.lr);
            if (!F.cholesky) bstAll[i] = detectSupernodes(F.n, F.ucp, F.ur);
            auto add = [&](const std::vector<int32_t>& st, const std::vector<int64_t>& cp, bool below) {
                for (std::size_t s = 0; s + 1 < st.size(); ++s) {
                    const int32_t w = st[s + 1] - st[s];
                    const int64_t h = cp[static_cast<std::size_t>(st[s]) + 1] - cp[static_cast<std::size_t>(st[s])] - w;
                    ndiag += diagSize(w) + 3;
                    if (below) {
                        nbelow += static_cast<std::size_t>(h) * static_cast<std::size_t>(w > kSmallWidth ? wideLd(w) : w) + 3;
                        nridx += static_cast<std::size_t>(h);
                    }
                }
                nsn += st.size() - 1;
            };
            add(fstAll[i], F.lcp, true);
            if (F.cholesky)
                nsn += fstAll[i].size() - 1;
            else
                add(bstAll[i], F.ucp, true);
        }
        P.diag.reserve(ndiag);
        P.below.reserve(nbelow);
        P.ridx.reserve(nridx);
        P.sn.reserve(nsn);
    }
    // queue groups, keyed by (phase, depth): forward groups in decreasing depth, then backward
    std::map<std::pair<int, int>, std::vector<int32_t>> groups; // key (0, -depth) / (1, depth)
    std::vector<std::vector<int32_t>> perSub(fs.size());
    for (std::size_t i = 0; i < fs.size(); ++i) {
        const HostFactor& F = fs[i];
        const int32_t n = F.n;
        SubDesc sd{};
        sd.n = n;
        sd.map_off = map_off[i];
        sd.t_off = P.t_total;
        P.max_n = std::max(P.max_n, n);
        const bool chol = F.cholesky;
        const std::vector<int64_t>& ucp = chol ? F.lcp : F.ucp;
        const std::vector<int32_t>& ur = chol ? F.lr : F.ur;
        const std::vector<double>& uv = chol ? F.lv : F.uv;

        // ---------------- forward sweep (L)
        const std::vector<int32_t>& fst = fstAll[i];
        const int32_t nsf = static_cast<int32_t>(fst.size()) - 1;
        std::vector<int32_t> snodeOf(static_cast<std::size_t>(n));
        for (int32_t s = 0; s < nsf; ++s)
            for (int32_t j = fst[s]; j < fst[s + 1]; ++j) snodeOf[static_cast<std::size_t>(j)] = s;
        const int32_t f_base = static_cast<int32_t>(P.sn.size());
        std::vector<double> blk;  // one diagonal block, column-major, before its inversion
        std::vector<int32_t> parent(static_cast<std::size_t>(nsf), -1);
        int32_t toff = 0;
        std::vector<int64_t> belowOff(static_cast<std::size_t>(nsf)), ridxOff(static_cast<std::size_t>(nsf));
        for (int32_t s = 0; s < nsf; ++s) {
            const int32_t c0 = fst[s], w = fst[s + 1] - c0;
            const int64_t p0 = F.lcp[c0];
            const int32_t h = static_cast<int32_t>(F.lcp[c0 + 1] - p0) - w;
            for (int32_t r = 0; r < w; ++r)
                if (F.lr[p0 + r] != c0 + r) throw std::invalid_argument("malformed supernode");
            if (w > kChunkRows) throw std::invalid_argument("supernode wider than " + std::to_string(kChunkRows) + " columns");
            SnDesc d{};
            d.c0 = c0;
            d.w = w;
            d.h = h;
            d.toff = toff;
            toff += h;
            const bool wide = w > kSmallWidth;
            if (wide) alignFour(P.diag);
            d.diag = static_cast<int64_t>(P.diag.size());
            blk.assign(static_cast<std::size_t>(w) * w, 0.0);
            for (int32_t c = 0; c < w; ++c)
                for (int32_t r = c; r < w; ++r) blk[static_cast<std::size_t>(c) * w + r] = F.lv[F.lcp[c0 + c] + (r - c)];
            appendDiagonal(P.diag, jobs, blk, w, true);
            if (wide) alignFour(P.below);
            d.below = static_cast<int64_t>(P.below.size());
            const int32_t ldb = wide ? wideLd(w) : w;
            for (int32_t k = 0; k < h; ++k) {
                for (int32_t c = 0; c < w; ++c) P.below.push_back(F.lv[F.lcp[c0 + c] + (w - c) + k]);
                P.below.insert(P.below.end(), static_cast<std::size_t>(ldb - w), 0.0);
            }
            d.ridx = static_cast<int64_t>(P.ridx.size());
            for (int32_t k = 0; k < h; ++k) P.ridx.push_back(F.lr[p0 + w + k]);
            belowOff[static_cast<std::size_t>(s)] = d.below;
            ridxOff[static_cast<std::size_t>(s)] = d.ridx;
            if (h > 0) parent[static_cast<std::size_t>(s)] = snodeOf[static_cast<std::size_t>(F.lr[p0 + w])];
            P.sn.push_back(d);
            P.entries_fwd += static_cast<int64_t>(w) * (w + 1) / 2 + static_cast<int64_t>(h) * w;
        }
        P.t_total += toff;
        // ---------------- task split (before the contributions: the supernodes of a dense top push nothing)
        const std::vector<int32_t>& bst = chol ? fst : bstAll[i];
        const bool samePartition = (bst == fst);
        std::vector<int32_t> depth(static_cast<std::size_t>(nsf), 0);
        for (int32_t s = nsf - 1; s >= 0; --s) {
            const int32_t p = parent[static_cast<std::size_t>(s)];
            depth[static_cast<std::size_t>(s)] = (p >= 0) ? depth[static_cast<std::size_t>(p)] + 1 : 0;
        }
        // Task split: the elimination tree is cut into a skeleton near the top, one task with the
        // root, and subtree tasks below it. Starting from the whole tree, the subtree with the
        // largest estimate is cut repeatedly (its root joins the skeleton, its children become
        // subtrees) while it is estimated above 1 / tasks_per_sub of the whole, for at most
        // kMaxSkeleton cuts. The subtrees hanging from one skeleton supernode are then packed
        // into as few tasks as the same bound allows (first fit by decreasing estimate): a task
        // may hold several sibling subtrees. The estimate of a subtree is its dense entries
        // (w (w + 1) / 2 + h w per supernode).
        std::vector<char> skel(static_cast<std::size_t>(nsf), 0);
        std::vector<std::vector<int32_t>> kids(static_cast<std::size_t>(nsf));
        std::vector<int64_t> entries(static_cast<std::size_t>(nsf), 0);
        for (int32_t s = 0; s < nsf; ++s) {
            const SnDesc& d = P.sn[static_cast<std::size_t>(f_base + s)];
            entries[static_cast<std::size_t>(s)] += static_cast<int64_t>(d.w) * (d.w + 1) / 2 + static_cast<int64_t>(d.h) * d.w;
            const int32_t p = parent[static_cast<std::size_t>(s)];
            if (p >= 0) {
                kids[static_cast<std::size_t>(p)].push_back(s);
                entries[static_cast<std::size_t>(p)] += entries[static_cast<std::size_t>(s)];
            }
        }
        auto estimate = [&](int32_t s) { return entries[static_cast<std::size_t>(s)]; };
        int64_t limit = std::numeric_limits<int64_t>::max();
        if (samePartition && tasks_per_sub > 1) {
            std::vector<std::pair<int64_t, int32_t>> heap;
            int64_t total = 0;
            for (int32_t s = 0; s < nsf; ++s)
                if (parent[static_cast<std::size_t>(s)] < 0) {
                    heap.emplace_back(estimate(s), s);
                    total += estimate(s);
                }
            std::make_heap(heap.begin(), heap.end());
            limit = total / tasks_per_sub;
            for (int32_t cuts = 0; !heap.empty() && cuts < kMaxSkeleton; ++cuts) {
                std::pop_heap(heap.begin(), heap.end());
                const auto top = heap.back();
                const auto& ch = kids[static_cast<std::size_t>(top.second)];
                if (top.first <= limit || ch.empty()) break;
                heap.pop_back();
                skel[static_cast<std::size_t>(top.second)] = 1;
                for (int32_t c : ch) {
                    heap.emplace_back(estimate(c), c);
                    std::push_heap(heap.begin(), heap.end());
                }
            }
        }
        // dense tops: the skeleton supernodes below every root that was cut, when their widths sum
        // to at most dense_top_max (a skeleton is closed upwards: every supernode above a skeleton
        // supernode belongs to it)
        std::vector<char> dense(static_cast<std::size_t>(nsf), 0);
        if (dense_top_max > 0) {
            std::vector<int32_t> rootOf(static_cast<std::size_t>(nsf));
            std::vector<int64_t> width(static_cast<std::size_t>(nsf), 0);
            for (int32_t s = nsf - 1; s >= 0; --s) {
                const int32_t p = parent[static_cast<std::size_t>(s)];
                rootOf[static_cast<std::size_t>(s)] = p < 0 ? s : rootOf[static_cast<std::size_t>(p)];
                if (skel[static_cast<std::size_t>(s)]) width[static_cast<std::size_t>(rootOf[static_cast<std::size_t>(s)])] += P.sn[static_cast<std::size_t>(f_base + s)].w;
            }
            for (int32_t s = 0; s < nsf; ++s)
                dense[static_cast<std::size_t>(s)] = skel[static_cast<std::size_t>(s)] &&
                                                     width[static_cast<std::size_t>(rootOf[static_cast<std::size_t>(s)])] <= dense_top_max;
        }
        const int32_t t_base = static_cast<int32_t>(P.task.size());
        std::vector<int32_t> taskOf(static_cast<std::size_t>(nsf), -1);
        std::vector<int32_t> taskRoot;
        for (int32_t s = nsf - 1; s >= 0; --s) {
            const int32_t p = parent[static_cast<std::size_t>(s)];
            if (!samePartition) {
                // the backward sweep uses another supernode partition (LU with another
                // structure): every subdomain is a single task
                if (taskRoot.empty()) taskRoot.push_back(nsf - 1);
                taskOf[static_cast<std::size_t>(s)] = 0;
                continue;
            }
            if (p < 0) {
                taskOf[static_cast<std::size_t>(s)] = static_cast<int32_t>(taskRoot.size());
                taskRoot.push_back(s);
            } else if (skel[static_cast<std::size_t>(s)] || !skel[static_cast<std::size_t>(p)]) {
                // the skeleton is one task with the root; a subtree below a skeleton supernode
                // was assigned when its parent was visited
                if (taskOf[static_cast<std::size_t>(s)] < 0) taskOf[static_cast<std::size_t>(s)] = taskOf[static_cast<std::size_t>(p)];
            }
            if (!skel[static_cast<std::size_t>(s)]) continue;
            // pack the subtrees below skeleton supernode s into tasks; a skeleton child gets its
            // own task when it is visited
            std::vector<int32_t> sub;
            for (int32_t c : kids[static_cast<std::size_t>(s)])
                if (!skel[static_cast<std::size_t>(c)]) sub.push_back(c);
            std::stable_sort(sub.begin(), sub.end(), [&](int32_t x, int32_t y) { return estimate(x) > estimate(y); });
            const int64_t cap = sub.empty() ? 0 : std::max(limit, estimate(sub.front()));
            std::vector<int64_t> load;
            std::vector<int32_t> binTask;
            for (int32_t c : sub) {
                std::size_t b = 0;
                while (b < load.size() && load[b] + estimate(c) > cap) ++b;
                if (b == load.size()) {
                    load.push_back(0);
                    binTask.push_back(static_cast<int32_t>(taskRoot.size()));
                    taskRoot.push_back(c);
                }
                load[b] += estimate(c);
                taskOf[static_cast<std::size_t>(c)] = binTask[b];
            }
        }
        const int32_t nt = static_cast<int32_t>(taskRoot.size());
        std::vector<std::vector<int32_t>> tsn(static_cast<std::size_t>(nt));
        for (int32_t s = 0; s < nsf; ++s) tsn[static_cast<std::size_t>(taskOf[static_cast<std::size_t>(s)])].push_back(s);
        // contributions: the product of row k below supernode d goes to a slot of its target row.
        // Slots are numbered task by task (a row's slots ordered by d), so a row pulls one
        // contiguous range [cptr, cend) and the slots of a task's rows are one contiguous range: a
        // task without child tasks, the only one pushing to its rows, can keep them in LDS. A task
        // numbers its rows from its top supernode down: when only a prefix of its slots fits in
        // LDS, that prefix holds the rows pulled most (the top supernodes receive the products of
        // the whole subtree). The supernodes of a dense top push nothing (their products are
        // inside S^-1).
        std::vector<int32_t> tslotBase(static_cast<std::size_t>(nt), 0), tslotCount(static_cast<std::size_t>(nt), 0);
        {
            std::vector<int32_t> cnt(static_cast<std::size_t>(n), 0);
            int64_t pushed = 0;
            for (int32_t s = 0; s < nsf; ++s) {
                if (dense[static_cast<std::size_t>(s)]) continue;
                const SnDesc& d = P.sn[static_cast<std::size_t>(f_base + s)];
                for (int32_t k = 0; k < d.h; ++k) ++cnt[static_cast<std::size_t>(P.ridx[static_cast<std::size_t>(d.ridx + k)])];
                pushed += d.h;
            }
            sd.cptr_off = static_cast<int32_t>(P.cptr.size());
            sd.dst_off = static_cast<int32_t>(P.cidx.size());
            std::vector<int32_t> beg(static_cast<std::size_t>(n), 0), end(static_cast<std::size_t>(n), 0);
            int32_t pos = 0;
            for (int32_t t = 0; t < nt; ++t) {
                tslotBase[static_cast<std::size_t>(t)] = pos;
                const std::vector<int32_t>& ts = tsn[static_cast<std::size_t>(t)];
                for (auto it = ts.rbegin(); it != ts.rend(); ++it)
                    for (int32_t r = fst[*it]; r < fst[*it + 1]; ++r) {
                        beg[static_cast<std::size_t>(r)] = pos;
                        pos += cnt[static_cast<std::size_t>(r)];
                        end[static_cast<std::size_t>(r)] = pos;
                    }
                tslotCount[static_cast<std::size_t>(t)] = pos - tslotBase[static_cast<std::size_t>(t)];
            }
            P.cptr.insert(P.cptr.end(), beg.begin(), beg.end());
            P.cend.insert(P.cend.end(), end.begin(), end.end());
            P.cidx.resize(static_cast<std::size_t>(sd.dst_off) + static_cast<std::size_t>(toff), 0);
            std::vector<int32_t> next(beg);
            for (int32_t s = 0; s < nsf; ++s) {
                if (dense[static_cast<std::size_t>(s)]) continue;
                const SnDesc& d = P.sn[static_cast<std::size_t>(f_base + s)];
                for (int32_t k = 0; k < d.h; ++k) {
                    const int32_t r = P.ridx[static_cast<std::size_t>(d.ridx + k)];
                    P.cidx[static_cast<std::size_t>(sd.dst_off + d.toff + k)] = next[static_cast<std::size_t>(r)]++;
                }
            }
            // slots read and written once, the slot table of every product, and per forward row
            // its factor position and slot range (lev_row, lev_cb, lev_ce; slot pointers for wide ones)
            P.bytes_index += pushed * 4 + static_cast<int64_t>(n) * 12;
        }

        // ---------------- backward sweep (U, given by U^T)
        const int32_t nsb = static_cast<int32_t>(bst.size()) - 1;
        std::vector<int32_t> snodeB(static_cast<std::size_t>(n));
        for (int32_t s = 0; s < nsb; ++s)
            for (int32_t j = bst[s]; j < bst[s + 1]; ++j) snodeB[static_cast<std::size_t>(j)] = s;
        const int32_t b_base = static_cast<int32_t>(P.sn.size());
        P.fwd_base.push_back(f_base);
        P.fwd_count.push_back(nsf);
        P.bwd_base.push_back(b_base);
        std::vector<int32_t> parentB(static_cast<std::size_t>(nsb), -1);
        for (int32_t s = 0; s < nsb; ++s) {
            const int32_t c0 = bst[s], w = bst[s + 1] - c0;
            const int64_t p0 = ucp[c0];
            const int32_t h = static_cast<int32_t>(ucp[c0 + 1] - p0) - w;
            if (w > kChunkRows) throw std::invalid_argument("supernode wider than " + std::to_string(kChunkRows) + " columns");
            SnDesc d{};
            d.c0 = c0;
            d.w = w;
            d.h = h;
            const bool wide = w > kSmallWidth;
            if (wide) alignFour(P.diag);
            d.diag = static_cast<int64_t>(P.diag.size());
            // G[c * w + r] = U(c0 + r, c0 + c) = U^T(c0 + c, c0 + r), r <= c
            blk.assign(static_cast<std::size_t>(w) * w, 0.0);
            for (int32_t c = 0; c < w; ++c)
                for (int32_t r = 0; r <= c; ++r) blk[static_cast<std::size_t>(c) * w + r] = uv[ucp[c0 + r] + (c - r)];
            appendDiagonal(P.diag, jobs, blk, w, false);
            if (chol) {
                d.below = belowOff[static_cast<std::size_t>(s)];
                d.ridx = ridxOff[static_cast<std::size_t>(s)];
            } else {
                if (wide) alignFour(P.below);
                d.below = static_cast<int64_t>(P.below.size());
                const int32_t ldb = wide ? wideLd(w) : w;
                for (int32_t k = 0; k < h; ++k) {
                    for (int32_t c = 0; c < w; ++c) P.below.push_back(uv[ucp[c0 + c] + (w - c) + k]);
                    P.below.insert(P.below.end(), static_cast<std::size_t>(ldb - w), 0.0);
                }
                d.ridx = static_cast<int64_t>(P.ridx.size());
                for (int32_t k = 0; k < h; ++k) P.ridx.push_back(ur[p0 + w + k]);
            }
            if (h > 0) parentB[static_cast<std::size_t>(s)] = snodeB[static_cast<std::size_t>(ur[p0 + w])];
            P.sn.push_back(d);
            P.entries_bwd += static_cast<int64_t>(w) * (w + 1) / 2 + static_cast<int64_t>(h) * w;
            P.bytes_index += static_cast<int64_t>(h) * 4;
        }
        P.bytes_index += static_cast<int64_t>(nsf + nsb) * static_cast<int64_t>(sizeof(SnDesc));

        std::vector<int32_t> heightT(static_cast<std::size_t>(nsf), 0);
        for (int32_t s = 0; s < nsf; ++s) {
            const int32_t p = parent[static_cast<std::size_t>(s)];
            if (p >= 0 && taskOf[static_cast<std::size_t>(p)] == taskOf[static_cast<std::size_t>(s)])
                heightT[static_cast<std::size_t>(p)] = std::max(heightT[static_cast<std::size_t>(p)], heightT[static_cast<std::size_t>(s)] + 1);
        }
        // backward level inside a task: depth in the backward tree relative to the task root
        std::vector<int32_t> depthB(static_cast<std::size_t>(nsb), 0);
        for (int32_t s = nsb - 1; s >= 0; --s) {
            const int32_t p = parentB[static_cast<std::size_t>(s)];
            depthB[static_cast<std::size_t>(s)] = (p >= 0) ? depthB[static_cast<std::size_t>(p)] + 1 : 0;
        }
        std::vector<std::vector<int32_t>> children(static_cast<std::size_t>(nt));
        int32_t qOff = -1;
        for (int32_t t = 0; t < nt; ++t) {
            const int32_t r = taskRoot[static_cast<std::size_t>(t)];
            TaskDesc td{};
            td.sub = static_cast<int32_t>(i);
            td.slot_base = tslotBase[static_cast<std::size_t>(t)];
            td.slot_count = tslotCount[static_cast<std::size_t>(t)];
            td.depth = samePartition ? depth[static_cast<std::size_t>(r)] : 0;
            const int32_t pr = samePartition ? parent[static_cast<std::size_t>(r)] : -1;
            td.parent = (pr >= 0) ? t_base + taskOf[static_cast<std::size_t>(pr)] : -1;
            if (pr >= 0) children[static_cast<std::size_t>(taskOf[static_cast<std::size_t>(pr)])].push_back(t_base + t);
            const auto& list = tsn[static_cast<std::size_t>(t)];
            td.sn_first = static_cast<int32_t>(P.task_sn.size());
            td.sn_count = static_cast<int32_t>(list.size());
            td.rows_first = static_cast<int32_t>(P.task_rows.size());
            for (int32_t s : list) {
                P.task_sn.push_back(f_base + s);
                for (int32_t j = fst[s]; j < fst[s + 1]; ++j) P.task_rows.push_back(j);
            }
            td.rows_count = static_cast<int32_t>(P.task_rows.size()) - td.rows_first;
            if (dense[static_cast<std::size_t>(r)]) {
                // a dense top: one step x_S = S^-1 (b_S - pulled), no levels and no backward steps;
                // its rows are split over nparts workgroups of at least 64 rows each
                const int32_t S = td.rows_count;
                td.nparts = std::max(1, std::min({tasks_per_sub, kMaxParts, (S + 63) / 64}));
                td.f_lev = td.b_lev = static_cast<int32_t>(P.lev.size());
                td.f_nlev = td.b_nlev = 0;
                alignFour(P.diag);
                StepDesc st{};
                st.kind = 2;
                st.W = S;
                st.wide.w = S;
                st.wide.diag = static_cast<int64_t>(P.diag.size());
                td.f_step = static_cast<int32_t>(P.steps.size());
                td.f_nsteps = 1;
                P.steps.push_back(st);
                td.b_step = static_cast<int32_t>(P.steps.size());
                td.b_nsteps = 0;
                td.ext_first = static_cast<int32_t>(P.ext_rows.size());
                td.ext_count = 0;
                P.diag.resize(P.diag.size() + static_cast<std::size_t>(S) * static_cast<std::size_t>(wideLd(S)), 0.0);
                DenseTop dt{};
                dt.sub = static_cast<int32_t>(i);
                dt.task = t_base + t;
                dt.S = S;
                dt.sn_first = static_cast<int32_t>(P.dtop_sn.size());
                dt.sn_count = static_cast<int32_t>(list.size());
                for (int32_t sk : list) {
                    P.dtop_sn.push_back(f_base + sk);
                    const SnDesc& d = P.sn[static_cast<std::size_t>(f_base + sk)];
                    const int64_t e = static_cast<int64_t>(d.w) * (d.w + 1) / 2 + static_cast<int64_t>(d.h) * d.w;
                    P.entries_skel_fwd += e;
                    P.entries_skel_bwd += e;  // the same partition: the same entries backward
                }
                if (qOff < 0) {  // the subdomain's position -> row map, shared by its dense tops
                    qOff = static_cast<int32_t>(P.dtop_q.size());
                    P.dtop_q.resize(P.dtop_q.size() + static_cast<std::size_t>(n), -1);
                }
                dt.q_off = qOff;
                for (int32_t q = 0; q < S; ++q) P.dtop_q[static_cast<std::size_t>(qOff + P.task_rows[static_cast<std::size_t>(td.rows_first + q)])] = q;
                dt.sinv = st.wide.diag;
                P.entries_dense += static_cast<int64_t>(S) * S;
                P.dtop.push_back(dt);
                djobs.push_back(DenseJob{dt.sinv, static_cast<int32_t>(i),
                                         std::vector<int32_t>(P.task_rows.begin() + td.rows_first, P.task_rows.end())});
                P.task.push_back(td);
                continue;
            }
            // forward levels: height inside the task
            int32_t hmax = 0;
            for (int32_t s : list) hmax = std::max(hmax, heightT[static_cast<std::size_t>(s)]);
            std::vector<std::vector<int32_t>> byLev(static_cast<std::size_t>(hmax) + 1);
            for (int32_t s : list) byLev[static_cast<std::size_t>(heightT[static_cast<std::size_t>(s)])].push_back(f_base + s);
            appendLevels(P, byLev, td.f_lev, td.f_nlev, td.f_step, td.f_nsteps, P.cptr.data() + sd.cptr_off,
                         P.cend.data() + sd.cptr_off);
            // backward levels: backward supernodes of the task (same columns when the partitions agree)
            std::vector<int32_t> blist;
            if (samePartition)
                blist = list;
            else
                for (int32_t s = 0; s < nsb; ++s) blist.push_back(s);
            int32_t dmin = 1 << 30, dmax = 0;
            for (int32_t s : blist) {
                dmin = std::min(dmin, depthB[static_cast<std::size_t>(s)]);
                dmax = std::max(dmax, depthB[static_cast<std::size_t>(s)]);
            }
            std::vector<std::vector<int32_t>> byLevB(static_cast<std::size_t>(dmax - dmin) + 1);
            for (int32_t s : blist) byLevB[static_cast<std::size_t>(depthB[static_cast<std::size_t>(s)] - dmin)].push_back(b_base + s);
            appendLevels(P, byLevB, td.b_lev, td.b_nlev, td.b_step, td.b_nsteps, nullptr, nullptr);
            // rows of other tasks read by the backward sweep
            td.ext_first = static_cast<int32_t>(P.ext_rows.size());
            std::vector<int32_t> ext;
            for (int32_t s : blist) {
                const SnDesc& d = P.sn[static_cast<std::size_t>(b_base + s)];
                for (int32_t k = 0; k < d.h; ++k) {
                    const int32_t row = P.ridx[static_cast<std::size_t>(d.ridx + k)];
                    if (taskOf[static_cast<std::size_t>(snodeOf[static_cast<std::size_t>(row)])] != t) ext.push_back(row);
                }
            }
            std::sort(ext.begin(), ext.end());
            ext.erase(std::unique(ext.begin(), ext.end()), ext.end());
            td.ext_count = static_cast<int32_t>(ext.size());
            P.ext_rows.insert(P.ext_rows.end(), ext.begin(), ext.end());
            P.task.push_back(td);
        }
        for (int32_t t = 0; t < nt; ++t) {
            TaskDesc& td = P.task[static_cast<std::size_t>(t_base + t)];
            td.child_first = static_cast<int32_t>(P.task_child.size());
            td.nchild = static_cast<int32_t>(children[static_cast<std::size_t>(t)].size());
            // other tasks push to the rows of a task with children: its slots stay in global memory
            if (td.nchild > 0) td.slot_count = 0;
            P.max_task_slots = std::max(P.max_task_slots, td.slot_count);
            P.task_child.insert(P.task_child.end(), children[static_cast<std::size_t>(t)].begin(), children[static_cast<std::size_t>(t)].end());
        }
        // tasks without child tasks (the subtrees) depend on no other task in the forward sweep:
        // they all go to the first forward group, and to the last backward group; the skeleton
        // tasks follow by decreasing depth (forward) and precede by increasing depth (backward)
        for (int32_t t = 0; t < nt; ++t) {
            const TaskDesc& td = P.task[static_cast<std::size_t>(t_base + t)];
            const int32_t id = t_base + t;
            const bool leaf = td.nchild == 0;
            if (id >= kMaxTasks) throw std::invalid_argument("more than " + std::to_string(kMaxTasks) + " local-solve tasks");
            auto& fg = groups[{0, leaf ? std::numeric_limits<int32_t>::min() : -td.depth}];
            for (int32_t part = 0; part < std::max(1, td.nparts); ++part) fg.push_back(queueCode(id, 0, part));
            if (td.parent >= 0) groups[{1, leaf ? std::numeric_limits<int32_t>::max() : td.depth}].push_back(queueCode(id, 1, 0));
        }
        P.max_levels_fwd = std::max(P.max_levels_fwd, 0);
        P.sub.push_back(sd);
    }
    // level-ordered queue: forward groups by decreasing depth, then backward groups
    for (const auto& kv : groups) {
        P.qgroup.push_back(static_cast<int32_t>(P.queue.size()));
        P.queue.insert(P.queue.end(), kv.second.begin(), kv.second.end());
    }
    P.qgroup.push_back(static_cast<int32_t>(P.queue.size()));
    // per-subdomain slices in the same order
    std::vector<std::vector<int32_t>> slice(fs.size());
    for (int32_t e : P.queue) slice[static_cast<std::size_t>(P.task[static_cast<std::size_t>(codeTask(e))].sub)].push_back(e);
    for (const auto& sl : slice) {
        P.qsub.push_back(static_cast<int32_t>(P.queue_sub.size()));
        P.queue_sub.insert(P.queue_sub.end(), sl.begin(), sl.end());
    }
    P.qsub.push_back(static_cast<int32_t>(P.queue_sub.size()));
    P.bytes_index += static_cast<int64_t>(P.steps.size()) * static_cast<int64_t>(sizeof(StepDesc));
    for (const TaskDesc& td : P.task) {
        P.max_levels_fwd = std::max(P.max_levels_fwd, td.f_nlev);
        P.max_levels_bwd = std::max(P.max_levels_bwd, td.b_nlev);
    }
    for (const StepDesc& sp : P.steps) {
        P.max_step_rows = std::max(P.max_step_rows, sp.kind == 0 ? sp.W : sp.wide.w);
        if (sp.kind == 0) P.max_step_sn = std::max(P.max_step_sn, sp.cnt);
    }
    runDiagonalJobs(P.diag, jobs);
    runDenseJobs(P.diag, djobs, fs);
    return P;
}

} // namespace s2l
