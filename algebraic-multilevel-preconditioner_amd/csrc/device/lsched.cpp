// lsched.cpp -- level-synchronous schedule and value layout of the batched local solves.
#include "lsched.hpp"

#include <algorithm>
#include <climits>
#include <stdexcept>
#include <string>

namespace s2l {

namespace {

int32_t ceilDiv(int32_t a, int32_t b) { return (a + b - 1) / b; }

/// Packed column-major triangle of an inverse: lower keeps rows c..w-1 of column c (offset
/// c * w - c * (c - 1) / 2), upper keeps rows 0..c (offset c * (c + 1) / 2).
void appendTriangle(std::vector<double>& out, const std::vector<double>& B, int32_t w, bool lower)
{
    for (int32_t c = 0; c < w; ++c) {
        if (lower)
            for (int32_t r = c; r < w; ++r) out.push_back(B[static_cast<std::size_t>(c) * w + r]);
        else
            for (int32_t r = 0; r <= c; ++r) out.push_back(B[static_cast<std::size_t>(c) * w + r]);
    }
}

struct Sweep {
    std::vector<int32_t> st;      // supernode starts
    std::vector<int32_t> parent;  // parent supernode or -1
};

Sweep structure(int32_t n, const std::vector<int64_t>& cp, const std::vector<int32_t>& row)
{
    Sweep S;
    S.st = detectSupernodes(n, cp, row);
    const int32_t ns = static_cast<int32_t>(S.st.size()) - 1;
    std::vector<int32_t> snodeOf(static_cast<std::size_t>(n));
    for (int32_t s = 0; s < ns; ++s)
        for (int32_t j = S.st[static_cast<std::size_t>(s)]; j < S.st[static_cast<std::size_t>(s) + 1]; ++j) snodeOf[static_cast<std::size_t>(j)] = s;
    S.parent.assign(static_cast<std::size_t>(ns), -1);
    for (int32_t s = 0; s < ns; ++s) {
        const int32_t c0 = S.st[static_cast<std::size_t>(s)], w = S.st[static_cast<std::size_t>(s) + 1] - c0;
        if (w > kLsMaxWidth) throw std::invalid_argument("supernode wider than " + std::to_string(kLsMaxWidth) + " columns");
        const int64_t p0 = cp[static_cast<std::size_t>(c0)];
        const int32_t h = static_cast<int32_t>(cp[static_cast<std::size_t>(c0) + 1] - p0) - w;
        if (h > 0) S.parent[static_cast<std::size_t>(s)] = snodeOf[static_cast<std::size_t>(row[static_cast<std::size_t>(p0 + w)])];
    }
    return S;
}

} // namespace

namespace {
bool isLarge(int32_t w, int32_t h) { return w > kLsRows || static_cast<int64_t>(w) * h > kLsLargeWork; }
} // namespace

LsPlan buildLsPlan(const std::vector<HostFactor>& fs, const std::vector<int32_t>& map_off)
{
    LsPlan P;
    const std::size_t nsub = fs.size();
    // supernodes of every level, all subdomains, kept apart until all are packed
    std::vector<std::vector<int32_t>> fwd, bwd;
    auto at = [](std::vector<std::vector<int32_t>>& v, int32_t lev) -> std::vector<int32_t>& {
        if (static_cast<int32_t>(v.size()) <= lev) v.resize(static_cast<std::size_t>(lev) + 1);
        return v[static_cast<std::size_t>(lev)];
    };
    // pushes (for the slot tables): every (row, push index) in push order
    std::vector<std::pair<int32_t, int64_t>> pushes;
    int64_t toff = 0;
    std::vector<double> blk;
    for (std::size_t i = 0; i < nsub; ++i) {
        const HostFactor& F = fs[i];
        const int32_t n = F.n;
        const int32_t row0 = map_off[i];
        const bool chol = F.cholesky;
        const std::vector<int64_t>& ucp = chol ? F.lcp : F.ucp;
        const std::vector<int32_t>& ur = chol ? F.lr : F.ur;
        const std::vector<double>& uv = chol ? F.lv : F.uv;

        // ---------------- forward sweep: L, levels by height (a supernode after all its children)
        const Sweep S = structure(n, F.lcp, F.lr);
        const int32_t nsf = static_cast<int32_t>(S.st.size()) - 1;
        std::vector<int32_t> height(static_cast<std::size_t>(nsf), 0);
        for (int32_t s = 0; s < nsf; ++s) {
            const int32_t p = S.parent[static_cast<std::size_t>(s)];
            if (p >= 0) {
                if (p <= s) throw std::invalid_argument("supernodal tree is not postordered");
                height[static_cast<std::size_t>(p)] = std::max(height[static_cast<std::size_t>(p)], height[static_cast<std::size_t>(s)] + 1);
            }
        }
        for (int32_t s = 0; s < nsf; ++s) {
            const int32_t c0 = S.st[static_cast<std::size_t>(s)], w = S.st[static_cast<std::size_t>(s) + 1] - c0;
            const int64_t p0 = F.lcp[static_cast<std::size_t>(c0)];
            const int32_t h = static_cast<int32_t>(F.lcp[static_cast<std::size_t>(c0) + 1] - p0) - w;
            for (int32_t r = 0; r < w; ++r)
                if (F.lr[static_cast<std::size_t>(p0 + r)] != c0 + r) throw std::invalid_argument("malformed supernode");
            LsSn d{};
            d.sub = static_cast<int32_t>(i);
            d.row0 = row0;
            d.c0 = c0;
            d.w = w;
            d.h = h;
            d.nrb = ceilDiv(w, kLsRows);
            d.nkc = 0;
            d.toff = static_cast<int32_t>(toff);
            d.poff = 0;
            blk.assign(static_cast<std::size_t>(w) * w, 0.0);
            for (int32_t c = 0; c < w; ++c)
                for (int32_t r = c; r < w; ++r)
                    blk[static_cast<std::size_t>(c) * w + r] = F.lv[static_cast<std::size_t>(F.lcp[static_cast<std::size_t>(c0 + c)] + (r - c))];
            invertLower(blk.data(), w, w);
            d.diag = static_cast<int64_t>(P.diag.size());
            appendTriangle(P.diag, blk, w, true);
            d.below = static_cast<int64_t>(P.belowF.size());
            for (int32_t c = 0; c < w; ++c)
                for (int32_t k = 0; k < h; ++k)
                    P.belowF.push_back(F.lv[static_cast<std::size_t>(F.lcp[static_cast<std::size_t>(c0 + c)] + (w - c) + k)]);
            d.ridx = static_cast<int64_t>(P.ridx.size());
            for (int32_t k = 0; k < h; ++k) P.ridx.push_back(F.lr[static_cast<std::size_t>(p0 + w + k)]);
            for (int32_t j = 0; j < d.nrb; ++j)
                for (int32_t k = 0; k < h; ++k) pushes.emplace_back(row0 + P.ridx[static_cast<std::size_t>(d.ridx + k)], toff + static_cast<int64_t>(j) * h + k);
            toff += static_cast<int64_t>(h) * d.nrb;
            at(fwd, height[static_cast<std::size_t>(s)]).push_back(static_cast<int32_t>(P.sn.size()));
            P.sn.push_back(d);
            P.entries += static_cast<int64_t>(w) * (w + 1) / 2 + static_cast<int64_t>(h) * w;
            // descriptor reads, row indices and slot indices of the pushes, the pulled rows' slot pointers
            P.bytes_index += static_cast<int64_t>(sizeof(LsSn)) * d.nrb + static_cast<int64_t>(h) * 4 * d.nrb +
                             static_cast<int64_t>(w) * 8;
        }

        // ---------------- backward sweep: U (given by U^T), levels by depth (a supernode after its
        // ancestors); a large one gets partial dot products with its ancestors' values first
        const Sweep B = chol ? S : structure(n, ucp, ur);
        const int32_t nsb = static_cast<int32_t>(B.st.size()) - 1;
        std::vector<int32_t> depth(static_cast<std::size_t>(nsb), 0);
        for (int32_t s = nsb - 1; s >= 0; --s) {
            const int32_t p = B.parent[static_cast<std::size_t>(s)];
            depth[static_cast<std::size_t>(s)] = p >= 0 ? depth[static_cast<std::size_t>(p)] + 1 : 0;
        }
        for (int32_t s = 0; s < nsb; ++s) {
            const int32_t c0 = B.st[static_cast<std::size_t>(s)], w = B.st[static_cast<std::size_t>(s) + 1] - c0;
            const int64_t p0 = ucp[static_cast<std::size_t>(c0)];
            const int32_t h = static_cast<int32_t>(ucp[static_cast<std::size_t>(c0) + 1] - p0) - w;
            LsSn d{};
            d.sub = static_cast<int32_t>(i);
            d.row0 = row0;
            d.c0 = c0;
            d.w = w;
            d.h = h;
            d.nrb = ceilDiv(w, kLsRows);
            d.nkc = (isLarge(w, h) && h > 0) ? ceilDiv(h, kLsKChunk) : 0;
            d.toff = 0;
            d.poff = P.partials;
            P.partials += static_cast<int64_t>(d.nkc) * w;
            // blk[c * w + r] = U(c0 + r, c0 + c) = U^T(c0 + c, c0 + r), r <= c
            blk.assign(static_cast<std::size_t>(w) * w, 0.0);
            for (int32_t c = 0; c < w; ++c)
                for (int32_t r = 0; r <= c; ++r)
                    blk[static_cast<std::size_t>(c) * w + r] = uv[static_cast<std::size_t>(ucp[static_cast<std::size_t>(c0 + r)] + (c - r))];
            invertUpper(blk.data(), w, w);
            d.diag = static_cast<int64_t>(P.diag.size());
            appendTriangle(P.diag, blk, w, false);
            // below[k * w + c] = U(c0 + c, ridx[k]) = U^T(ridx[k], c0 + c)
            d.below = static_cast<int64_t>(P.belowB.size());
            for (int32_t k = 0; k < h; ++k)
                for (int32_t c = 0; c < w; ++c)
                    P.belowB.push_back(uv[static_cast<std::size_t>(ucp[static_cast<std::size_t>(c0 + c)] + (w - c) + k)]);
            d.ridx = static_cast<int64_t>(P.ridx.size());
            for (int32_t k = 0; k < h; ++k) P.ridx.push_back(ur[static_cast<std::size_t>(p0 + w + k)]);
            at(bwd, depth[static_cast<std::size_t>(s)]).push_back(static_cast<int32_t>(P.sn.size()));
            P.sn.push_back(d);
            P.entries += static_cast<int64_t>(w) * (w + 1) / 2 + static_cast<int64_t>(h) * w;
            P.bytes_index += static_cast<int64_t>(h) * 4 + static_cast<int64_t>(sizeof(LsSn)) * d.nrb;
            P.bwd_below_rows += h;
        }
    }
    P.slots = toff;
    if (toff > INT32_MAX) throw std::invalid_argument("too many pushed products for 32-bit slot indices");

    // slot tables: the slots of one row are contiguous, in push order
    const int32_t nrows = map_off.empty() ? 0 : map_off.back();
    std::vector<int32_t> cnt(static_cast<std::size_t>(nrows) + 1, 0);
    for (const auto& pr : pushes) ++cnt[static_cast<std::size_t>(pr.first) + 1];
    P.cptr.assign(static_cast<std::size_t>(nrows) + 1, 0);
    for (int32_t r = 0; r < nrows; ++r) P.cptr[static_cast<std::size_t>(r) + 1] = P.cptr[static_cast<std::size_t>(r)] + cnt[static_cast<std::size_t>(r) + 1];
    P.cidx.assign(static_cast<std::size_t>(toff), 0);
    std::vector<int32_t> next(P.cptr.begin(), P.cptr.end() - 1);
    for (const auto& pr : pushes) P.cidx[static_cast<std::size_t>(pr.second)] = next[static_cast<std::size_t>(pr.first)]++;

    // launches
    auto launch = [&](int32_t kind, std::size_t first, std::size_t count, int32_t maxw) {
        if (count == 0) return;
        P.launch_kind.push_back(kind);
        P.launch_first.push_back(static_cast<int32_t>(first));
        P.launch_count.push_back(static_cast<int32_t>(count));
        P.launch_w.push_back(maxw);
    };
    // units of a level: a launch of the row blocks of its large supernodes, then a launch of its
    // small ones
    auto units = [&](const std::vector<int32_t>& list, bool forward) {
        std::size_t first = P.units.size();
        int32_t maxw = 0;
        std::vector<int32_t> small;
        for (int32_t id : list) {
            const LsSn& d = P.sn[static_cast<std::size_t>(id)];
            const bool large = forward ? isLarge(d.w, d.h) : (d.nkc > 0 || d.w > kLsRows);
            if (!large) {
                small.push_back(id);
                continue;
            }
            maxw = std::max(maxw, d.w);
            for (int32_t j = 0; j < d.nrb; ++j) {
                LsUnit u{};
                u.large = 1;
                u.part = j;
                u.sn[0] = id;
                for (int q = 1; q < kLsWaves; ++q) u.sn[q] = -1;
                P.units.push_back(u);
            }
        }
        launch(forward ? kLsForward : kLsBackward, first, P.units.size() - first, maxw);
        first = P.units.size();
        for (int32_t id : small) {
            LsUnit u{};
            u.large = 0;
            u.part = 0;
            u.sn[0] = id;
            for (int q = 1; q < kLsWaves; ++q) u.sn[q] = -1;
            P.units.push_back(u);
        }
        launch(forward ? kLsForwardSmall : kLsBackwardSmall, first, P.units.size() - first, kLsRows);
    };
    for (const auto& list : fwd) {
        // the pull launch covers the rows of the large supernodes (the small ones pull their own)
        const std::size_t first = P.pull_rows.size();
        for (int32_t id : list) {
            const LsSn& d = P.sn[static_cast<std::size_t>(id)];
            if (!isLarge(d.w, d.h)) continue;
            for (int32_t r = 0; r < d.w; ++r) P.pull_rows.push_back(d.row0 + d.c0 + r);
        }
        launch(kLsPull, first, P.pull_rows.size() - first, 0);
        units(list, true);
        P.launch_level.resize(P.launch_kind.size(), static_cast<int32_t>(&list - fwd.data()));
    }
    for (const auto& list : bwd) {
        const std::size_t first = P.items.size();
        int32_t maxw = 0;
        for (int32_t id : list) {
            const LsSn& d = P.sn[static_cast<std::size_t>(id)];
            for (int32_t j = 0; j < d.nrb && d.nkc > 0; ++j)
                for (int32_t kc = 0; kc < d.nkc; ++kc) P.items.push_back({id, j * d.nkc + kc});
            if (d.nkc > 0) maxw = std::max(maxw, d.w);
        }
        launch(kLsBackPartial, first, P.items.size() - first, maxw);
        units(list, false);
        P.launch_level.resize(P.launch_kind.size(), static_cast<int32_t>(fwd.size() + (&list - bwd.data())));
    }
    P.bytes_index += static_cast<int64_t>(P.pull_rows.size()) * 4 + static_cast<int64_t>(P.units.size()) * sizeof(LsUnit) +
                     static_cast<int64_t>(P.items.size()) * sizeof(LsItem);
    P.levels_fwd = static_cast<int32_t>(fwd.size());
    P.levels_bwd = static_cast<int32_t>(bwd.size());
    return P;
}

} // namespace s2l
