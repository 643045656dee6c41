// pack.cpp -- supernode detection, level schedule and value layout of the batched local solve.
//
// Input: the factors of all subdomains in compressed-column form (L, and U^T for LU), the form
// the reference's LocalSolver backends expose. Output: one table of supernodes per sweep, level
// lists (forward: height above the leaves of the supernodal elimination tree; backward: depth
// below the root), dense blocks (diagonal block column-major, rows below row-major) and, for the
// forward sweep, the lists of products each row pulls from its descendants.
#include "pack.hpp"

#include <algorithm>
#include <stdexcept>
#include <string>

namespace s2l {

namespace {

constexpr int32_t kSmallWidth = 64; // supernodes up to one wavefront wide are solved by one wave

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

struct SweepBuild {
    std::vector<int32_t> starts;
    std::vector<int32_t> parent;
};

} // namespace

Packed packFactors(const std::vector<HostFactor>& fs, const std::vector<int32_t>& map_off)
{
    Packed P;
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
        const std::vector<int32_t> fst = detectSupernodes(n, F.lcp, F.lr);
        const int32_t nsf = static_cast<int32_t>(fst.size()) - 1;
        std::vector<int32_t> snodeOf(static_cast<std::size_t>(n));
        for (int32_t s = 0; s < nsf; ++s)
            for (int32_t j = fst[s]; j < fst[s + 1]; ++j) snodeOf[static_cast<std::size_t>(j)] = s;
        const int32_t f_base = static_cast<int32_t>(P.sn.size());
        std::vector<int32_t> parentF(static_cast<std::size_t>(nsf), -1);
        int32_t toff = 0;
        // below blocks of the forward sweep, remembered for the Cholesky backward sweep
        std::vector<int64_t> belowOff(static_cast<std::size_t>(nsf)), ridxOff(static_cast<std::size_t>(nsf));
        for (int32_t s = 0; s < nsf; ++s) {
            const int32_t c0 = fst[s], w = fst[s + 1] - c0;
            const int64_t p0 = F.lcp[c0];
            const int32_t h = static_cast<int32_t>(F.lcp[c0 + 1] - p0) - w;
            for (int32_t r = 0; r < w; ++r)
                if (F.lr[p0 + r] != c0 + r) throw std::invalid_argument("malformed supernode");
            SnDesc d{};
            d.c0 = c0;
            d.w = w;
            d.h = h;
            d.toff = toff;
            toff += h;
            d.diag = static_cast<int64_t>(P.diag.size());
            for (int32_t c = 0; c < w; ++c)
                for (int32_t r = 0; r < w; ++r) P.diag.push_back(r >= c ? F.lv[F.lcp[c0 + c] + (r - c)] : 0.0);
            d.below = static_cast<int64_t>(P.below.size());
            for (int32_t k = 0; k < h; ++k)
                for (int32_t c = 0; c < w; ++c) P.below.push_back(F.lv[F.lcp[c0 + c] + (w - c) + k]);
            d.ridx = static_cast<int64_t>(P.ridx.size());
            for (int32_t k = 0; k < h; ++k) P.ridx.push_back(F.lr[p0 + w + k]);
            belowOff[static_cast<std::size_t>(s)] = d.below;
            ridxOff[static_cast<std::size_t>(s)] = d.ridx;
            if (h > 0) parentF[static_cast<std::size_t>(s)] = snodeOf[static_cast<std::size_t>(F.lr[p0 + w])];
            P.sn.push_back(d);
            P.entries_fwd += static_cast<int64_t>(w) * (w + 1) / 2 + static_cast<int64_t>(h) * w;
        }
        P.t_total += toff;
        // contributions: row R_d[k] pulls T[toff_d + k]
        {
            std::vector<int32_t> cnt(static_cast<std::size_t>(n) + 1, 0);
            for (int32_t s = 0; s < nsf; ++s) {
                const SnDesc& d = P.sn[static_cast<std::size_t>(f_base + s)];
                for (int32_t k = 0; k < d.h; ++k) ++cnt[static_cast<std::size_t>(P.ridx[static_cast<std::size_t>(d.ridx + k)]) + 1];
            }
            sd.cptr_off = static_cast<int32_t>(P.cptr.size());
            const int32_t cbase = static_cast<int32_t>(P.cidx.size());
            std::vector<int32_t> ptr(static_cast<std::size_t>(n) + 1, 0);
            for (int32_t r = 0; r < n; ++r) ptr[static_cast<std::size_t>(r) + 1] = ptr[static_cast<std::size_t>(r)] + cnt[static_cast<std::size_t>(r) + 1];
            for (int32_t r = 0; r <= n; ++r) P.cptr.push_back(cbase + ptr[static_cast<std::size_t>(r)]);
            P.cidx.resize(static_cast<std::size_t>(cbase) + static_cast<std::size_t>(ptr[static_cast<std::size_t>(n)]));
            std::vector<int32_t> next(ptr.begin(), ptr.end() - 1);
            for (int32_t s = 0; s < nsf; ++s) {
                const SnDesc& d = P.sn[static_cast<std::size_t>(f_base + s)];
                for (int32_t k = 0; k < d.h; ++k) {
                    const int32_t r = P.ridx[static_cast<std::size_t>(d.ridx + k)];
                    P.cidx[static_cast<std::size_t>(cbase + next[static_cast<std::size_t>(r)]++)] = d.toff + k;
                }
            }
            P.bytes_index += static_cast<int64_t>(ptr[static_cast<std::size_t>(n)]) * 4 + (static_cast<int64_t>(n) + 1) * 4;
        }
        // forward levels: height above the leaves
        {
            std::vector<int32_t> height(static_cast<std::size_t>(nsf), 0);
            int32_t nlev = 0;
            for (int32_t s = 0; s < nsf; ++s) {
                const int32_t p = parentF[static_cast<std::size_t>(s)];
                if (p >= 0) height[static_cast<std::size_t>(p)] = std::max(height[static_cast<std::size_t>(p)], height[static_cast<std::size_t>(s)] + 1);
                nlev = std::max(nlev, height[static_cast<std::size_t>(s)] + 1);
            }
            sd.f_lev = static_cast<int32_t>(P.lev.size());
            sd.f_nlev = nlev;
            for (int32_t L = 0; L < nlev; ++L) {
                P.lev.push_back(static_cast<int32_t>(P.lev_sn.size()));
                for (int pass = 0; pass < 2; ++pass) {
                    if (pass == 1) P.lev.push_back(static_cast<int32_t>(P.lev_sn.size()));
                    for (int32_t s = 0; s < nsf; ++s)
                        if (height[static_cast<std::size_t>(s)] == L && ((P.sn[static_cast<std::size_t>(f_base + s)].w > kSmallWidth) == (pass == 1)))
                            P.lev_sn.push_back(f_base + s);
                }
            }
            P.lev.push_back(static_cast<int32_t>(P.lev_sn.size()));
            P.max_levels_fwd = std::max(P.max_levels_fwd, nlev);
        }

        // ---------------- backward sweep (U, given by U^T)
        const std::vector<int32_t> bst = chol ? fst : detectSupernodes(n, ucp, ur);
        const int32_t nsb = static_cast<int32_t>(bst.size()) - 1;
        std::vector<int32_t> snodeB(static_cast<std::size_t>(n));
        for (int32_t s = 0; s < nsb; ++s)
            for (int32_t j = bst[s]; j < bst[s + 1]; ++j) snodeB[static_cast<std::size_t>(j)] = s;
        const int32_t b_base = static_cast<int32_t>(P.sn.size());
        std::vector<int32_t> parentB(static_cast<std::size_t>(nsb), -1);
        for (int32_t s = 0; s < nsb; ++s) {
            const int32_t c0 = bst[s], w = bst[s + 1] - c0;
            const int64_t p0 = ucp[c0];
            const int32_t h = static_cast<int32_t>(ucp[c0 + 1] - p0) - w;
            SnDesc d{};
            d.c0 = c0;
            d.w = w;
            d.h = h;
            d.toff = 0;
            d.diag = static_cast<int64_t>(P.diag.size());
            // G[c * w + r] = U(c0 + r, c0 + c) = U^T(c0 + c, c0 + r), r <= c
            for (int32_t c = 0; c < w; ++c)
                for (int32_t r = 0; r < w; ++r) P.diag.push_back(r <= c ? uv[ucp[c0 + r] + (c - r)] : 0.0);
            if (chol) {
                d.below = belowOff[static_cast<std::size_t>(s)];
                d.ridx = ridxOff[static_cast<std::size_t>(s)];
            } else {
                d.below = static_cast<int64_t>(P.below.size());
                for (int32_t k = 0; k < h; ++k)
                    for (int32_t c = 0; c < w; ++c) P.below.push_back(uv[ucp[c0 + c] + (w - c) + k]);
                d.ridx = static_cast<int64_t>(P.ridx.size());
                for (int32_t k = 0; k < h; ++k) P.ridx.push_back(ur[p0 + w + k]);
            }
            if (h > 0) parentB[static_cast<std::size_t>(s)] = snodeB[static_cast<std::size_t>(ur[p0 + w])];
            P.sn.push_back(d);
            P.entries_bwd += static_cast<int64_t>(w) * (w + 1) / 2 + static_cast<int64_t>(h) * w;
            P.bytes_index += static_cast<int64_t>(h) * 4;
        }
        // backward levels: depth below the root
        {
            std::vector<int32_t> depth(static_cast<std::size_t>(nsb), 0);
            int32_t nlev = 0;
            for (int32_t s = nsb - 1; s >= 0; --s) {
                const int32_t p = parentB[static_cast<std::size_t>(s)];
                depth[static_cast<std::size_t>(s)] = (p >= 0) ? depth[static_cast<std::size_t>(p)] + 1 : 0;
                nlev = std::max(nlev, depth[static_cast<std::size_t>(s)] + 1);
            }
            sd.b_lev = static_cast<int32_t>(P.lev.size());
            sd.b_nlev = nlev;
            for (int32_t L = 0; L < nlev; ++L) {
                P.lev.push_back(static_cast<int32_t>(P.lev_sn.size()));
                for (int pass = 0; pass < 2; ++pass) {
                    if (pass == 1) P.lev.push_back(static_cast<int32_t>(P.lev_sn.size()));
                    for (int32_t s = 0; s < nsb; ++s)
                        if (depth[static_cast<std::size_t>(s)] == L && ((P.sn[static_cast<std::size_t>(b_base + s)].w > kSmallWidth) == (pass == 1)))
                            P.lev_sn.push_back(b_base + s);
                }
            }
            P.lev.push_back(static_cast<int32_t>(P.lev_sn.size()));
            P.max_levels_bwd = std::max(P.max_levels_bwd, nlev);
        }
        P.bytes_index += static_cast<int64_t>(nsf + nsb) * static_cast<int64_t>(sizeof(SnDesc));
        P.sub.push_back(sd);
    }
    return P;
}

} // namespace s2l
