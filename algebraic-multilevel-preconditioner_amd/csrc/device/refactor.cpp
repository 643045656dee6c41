// refactor.cpp -- host-side plan of the device refactorization (see refactor.hpp).
#include "refactor.hpp"

#include <algorithm>
#include <stdexcept>
#include <string>

namespace s2l {

RfPlan buildRefactorPlan(const std::vector<HostFactor>& fs, const std::vector<int32_t>& local_ptr,
                         const std::vector<int32_t>& gid, const std::vector<int32_t>& rp, const std::vector<int32_t>& col,
                         const Packed& P)
{
    RfPlan R;
    const int nsub = static_cast<int>(fs.size());
    std::vector<int32_t> planBase(static_cast<std::size_t>(nsub), 0);  // plan index of every subdomain's first supernode
    const int32_t n_global = static_cast<int32_t>(rp.size()) - 1;
    std::vector<int32_t> g2l(static_cast<std::size_t>(n_global), -1);
    std::vector<int32_t> height;
    std::vector<std::vector<RfUpd>> updOf;
    std::vector<std::vector<int32_t>> relOf;
    for (int i = 0; i < nsub; ++i) {
        const HostFactor& F = fs[static_cast<std::size_t>(i)];
        if (!F.cholesky) throw std::invalid_argument("the device refactorization needs Cholesky factors");
        const int32_t n = F.n;
        const int32_t fb = P.fwd_base[static_cast<std::size_t>(i)], nsf = P.fwd_count[static_cast<std::size_t>(i)];
        const int32_t bb = P.bwd_base[static_cast<std::size_t>(i)];
        const int32_t base = static_cast<int32_t>(R.sn.size());
        planBase[static_cast<std::size_t>(i)] = base;
        std::vector<int32_t> snodeOf(static_cast<std::size_t>(n));
        for (int32_t s = 0; s < nsf; ++s) {
            const SnDesc& d = P.sn[static_cast<std::size_t>(fb + s)];
            for (int32_t j = d.c0; j < d.c0 + d.w; ++j) snodeOf[static_cast<std::size_t>(j)] = s;
            RfSn r{};
            r.c0 = d.c0;
            r.w = d.w;
            r.h = d.h;
            r.sub = i;
            r.wd = R.work;
            r.wb = R.work + static_cast<int64_t>(d.w) * d.w;
            R.work = r.wb + static_cast<int64_t>(d.h) * d.w;
            r.ridx = d.ridx;
            r.fdiag = d.diag;
            r.bdiag = P.sn[static_cast<std::size_t>(bb + s)].diag;
            r.below = d.below;
            R.inv_off.push_back(R.inv);
            R.inv += static_cast<int64_t>(d.w) * d.w;
            R.sn.push_back(r);
        }
        updOf.resize(R.sn.size());
        relOf.resize(R.sn.size());
        height.resize(R.sn.size(), 0);
        // descendants' updates: the rows below d, grouped by the ancestor holding them
        for (int32_t sd = 0; sd < nsf; ++sd) {
            const SnDesc& d = P.sn[static_cast<std::size_t>(fb + sd)];
            const int32_t* Rd = P.ridx.data() + d.ridx;
            if (d.h > 0) {
                const int32_t par = snodeOf[static_cast<std::size_t>(Rd[0])];
                height[static_cast<std::size_t>(base + par)] =
                    std::max(height[static_cast<std::size_t>(base + par)], height[static_cast<std::size_t>(base + sd)] + 1);
            }
            int32_t k = 0;
            while (k < d.h) {
                const int32_t s = snodeOf[static_cast<std::size_t>(Rd[k])];
                const SnDesc& t = P.sn[static_cast<std::size_t>(fb + s)];
                const int32_t c1 = t.c0 + t.w;
                int32_t k1 = k;
                while (k1 < d.h && Rd[k1] < c1) ++k1;
                RfUpd u{};
                u.d = base + sd;
                u.i0 = k;
                u.i1 = k1;
                std::vector<int32_t>& rel = relOf[static_cast<std::size_t>(base + s)];
                u.rel = static_cast<int64_t>(rel.size());  // rebased when flattened
                const int32_t* Rs = P.ridx.data() + t.ridx;
                for (int32_t q = k; q < d.h; ++q) {
                    const int32_t r = Rd[q];
                    if (r < c1) {
                        rel.push_back(r - t.c0);
                    } else {
                        const int32_t* it = std::lower_bound(Rs, Rs + t.h, r);
                        if (it == Rs + t.h || *it != r)
                            throw std::invalid_argument("refactorization plan: a row below a descendant is not below its ancestor");
                        rel.push_back(t.w + static_cast<int32_t>(it - Rs));
                    }
                }
                updOf[static_cast<std::size_t>(base + s)].push_back(u);
                k = k1;
            }
        }
        // the entries of A_ii(perm, perm) on and below the diagonal, by supernode
        const int32_t l0 = local_ptr[static_cast<std::size_t>(i)], nl = local_ptr[static_cast<std::size_t>(i) + 1] - l0;
        if (nl != n) throw std::invalid_argument("refactorization plan: factor size differs from the subdomain");
        std::vector<int32_t> iperm(static_cast<std::size_t>(n));
        for (int32_t p = 0; p < n; ++p) iperm[static_cast<std::size_t>(F.perm[static_cast<std::size_t>(p)])] = p;
        for (int32_t l = 0; l < n; ++l) g2l[static_cast<std::size_t>(gid[static_cast<std::size_t>(l0 + l)])] = l;
        std::vector<std::vector<std::pair<int64_t, int32_t>>> am(static_cast<std::size_t>(nsf));
        for (int32_t l = 0; l < n; ++l) {
            const int32_t g = gid[static_cast<std::size_t>(l0 + l)];
            for (int32_t q = rp[static_cast<std::size_t>(g)]; q < rp[static_cast<std::size_t>(g) + 1]; ++q) {
                const int32_t lc = g2l[static_cast<std::size_t>(col[static_cast<std::size_t>(q)])];
                if (lc < 0) continue;
                const int32_t fr = iperm[static_cast<std::size_t>(l)], fc = iperm[static_cast<std::size_t>(lc)];
                if (fr < fc) continue;
                const int32_t s = snodeOf[static_cast<std::size_t>(fc)];
                const RfSn& r = R.sn[static_cast<std::size_t>(base + s)];
                int64_t pos;
                if (fr < r.c0 + r.w) {
                    pos = r.wd + static_cast<int64_t>(fc - r.c0) * r.w + (fr - r.c0);
                } else {
                    const int32_t* Rs = P.ridx.data() + r.ridx;
                    const int32_t* it = std::lower_bound(Rs, Rs + r.h, fr);
                    if (it == Rs + r.h || *it != fr)
                        throw std::invalid_argument("refactorization plan: an entry of A_ii lies outside the factor's structure");
                    pos = r.wb + static_cast<int64_t>(fc - r.c0) * r.h + static_cast<int64_t>(it - Rs);
                }
                am[static_cast<std::size_t>(s)].emplace_back(pos, q);
            }
        }
        for (int32_t l = 0; l < n; ++l) g2l[static_cast<std::size_t>(gid[static_cast<std::size_t>(l0 + l)])] = -1;
        for (int32_t s = 0; s < nsf; ++s) {
            RfSn& r = R.sn[static_cast<std::size_t>(base + s)];
            r.amap_first = static_cast<int64_t>(R.amap_pos.size());
            r.amap_count = static_cast<int32_t>(am[static_cast<std::size_t>(s)].size());
            for (const auto& e : am[static_cast<std::size_t>(s)]) {
                R.amap_pos.push_back(e.first);
                R.amap_src.push_back(e.second);
            }
        }
    }
    // flatten the updates and their relative rows, target by target
    for (std::size_t s = 0; s < R.sn.size(); ++s) {
        RfSn& r = R.sn[s];
        r.upd_first = static_cast<int32_t>(R.upd.size());
        r.upd_count = static_cast<int32_t>(updOf[s].size());
        const int64_t rb = static_cast<int64_t>(R.rel.size());
        for (RfUpd u : updOf[s]) {
            u.rel += rb;
            R.upd.push_back(u);
        }
        R.rel.insert(R.rel.end(), relOf[s].begin(), relOf[s].end());
        std::vector<int32_t>().swap(relOf[s]);
    }
    // supernodes by height: a height's panels need only lower heights
    int32_t hmax = 0;
    for (int32_t h : height) hmax = std::max(hmax, h);
    R.level_ptr.assign(static_cast<std::size_t>(hmax) + 2, 0);
    for (int32_t h : height) ++R.level_ptr[static_cast<std::size_t>(h) + 1];
    for (int32_t h = 0; h <= hmax; ++h) R.level_ptr[static_cast<std::size_t>(h) + 1] += R.level_ptr[static_cast<std::size_t>(h)];
    R.order.assign(R.sn.size(), 0);
    std::vector<int32_t> next(R.level_ptr.begin(), R.level_ptr.end() - 1);
    for (std::size_t s = 0; s < R.sn.size(); ++s) R.order[static_cast<std::size_t>(next[static_cast<std::size_t>(height[s])]++)] = static_cast<int32_t>(s);
    // dense tops: the plan index of each of their supernodes, and S x S scratch blocks
    for (const DenseTop& dt : P.dtop) {
        for (int32_t k = dt.sn_first; k < dt.sn_first + dt.sn_count; ++k)
            R.dtop_rf.push_back(planBase[static_cast<std::size_t>(dt.sub)] + P.dtop_sn[static_cast<std::size_t>(k)] -
                                P.fwd_base[static_cast<std::size_t>(dt.sub)]);
        R.dtop_off.push_back(R.dtop_total);
        R.dtop_total += static_cast<int64_t>(dt.S) * dt.S;
    }
    return R;
}

} // namespace s2l
