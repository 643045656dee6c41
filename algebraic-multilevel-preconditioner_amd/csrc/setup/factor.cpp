// factor.cpp -- sparse direct factorization of the local matrices A_ii (host setup stand-in for
// the reference's Eigen SimplicialLLT / SparseLU / UMFPACK backends, src/linalg/local_solver.cpp).
//
// Ordering: nested dissection of G(A_ii + A_ii^T) with breadth-first level separators from a
// pseudo-peripheral vertex; every leaf of the dissection and every separator becomes one
// supernode (a dense block: diagonal triangle + dense rows below). Symbolic: block
// elimination (R_s = rows of A below the block united with the rows of the children).
// Numeric: left-looking supernodal Cholesky (L L^T) or LU without pivoting (L unit lower,
// U upper) in FP64. The factor is exported in compressed-column form, which is the form the
// reference's factorizations expose (Eigen matrixL() with a permutation).
#include "setup_internal.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>

namespace s2l {

namespace {

struct Orderer {
    const Graph& g;
    int leaf;
    std::vector<int32_t> mark;  // tag of the set a vertex belongs to; -tag while visited
    std::vector<int32_t> where; // scratch: 1 part1, 2 part2, 3 separator
    int tag = 0;
    std::vector<int32_t> order;    // new -> old
    std::vector<int32_t> sn_begin; // supernode starts in the new order

    Orderer(const Graph& gg, int lf) : g(gg), leaf(lf), mark(static_cast<std::size_t>(gg.n), 0), where(static_cast<std::size_t>(gg.n), 0) {}

    // BFS of the vertices marked `t` from s; returns order and levels; marks visited with -t.
    void bfs(int32_t s, int t, std::vector<int32_t>& ord, std::vector<int32_t>& lev)
    {
        ord.clear();
        lev.clear();
        ord.push_back(s);
        lev.push_back(0);
        mark[static_cast<std::size_t>(s)] = -t;
        for (std::size_t h = 0; h < ord.size(); ++h) {
            const int32_t v = ord[h];
            for (int64_t e = g.xadj[static_cast<std::size_t>(v)]; e < g.xadj[static_cast<std::size_t>(v) + 1]; ++e) {
                const int32_t u = g.adj[static_cast<std::size_t>(e)];
                if (mark[static_cast<std::size_t>(u)] == t) {
                    mark[static_cast<std::size_t>(u)] = -t;
                    ord.push_back(u);
                    lev.push_back(lev[h] + 1);
                }
            }
        }
    }

    void restore(const std::vector<int32_t>& ord, int t)
    {
        for (int32_t v : ord) mark[static_cast<std::size_t>(v)] = t;
    }

    void emit(const std::vector<int32_t>& verts)
    {
        if (verts.empty()) return;
        sn_begin.push_back(static_cast<int32_t>(order.size()));
        order.insert(order.end(), verts.begin(), verts.end());
    }

    // Orders the connected set `set` (all marked with a fresh tag by the caller).
    void dissect(std::vector<int32_t> set)
    {
        const int t = ++tag;
        for (int32_t v : set) mark[static_cast<std::size_t>(v)] = t;
        // Connected components: each is dissected on its own.
        std::vector<int32_t> ord, lev;
        std::vector<std::vector<int32_t>> comps;
        for (int32_t v : set) {
            if (mark[static_cast<std::size_t>(v)] != t) continue;
            bfs(v, t, ord, lev);
            comps.push_back(ord);
        }
        for (auto& c : comps) restore(c, t);
        if (comps.size() > 1) {
            for (auto& c : comps) {
                std::sort(c.begin(), c.end());
                dissect(std::move(c));
            }
            return;
        }
        // Pseudo-peripheral start vertex.
        int32_t s = *std::min_element(set.begin(), set.end());
        int depth = -1;
        for (int it = 0; it < 6; ++it) {
            bfs(s, t, ord, lev);
            restore(ord, t);
            if (lev.back() <= depth) break;
            depth = lev.back();
            s = ord.back();
        }
        bfs(s, t, ord, lev);
        restore(ord, t);
        const std::size_t sz = ord.size();
        const int L = lev.back();
        if (static_cast<int>(sz) <= leaf || L < 2) {
            emit(ord);
            return;
        }
        std::vector<std::size_t> cnt(static_cast<std::size_t>(L) + 1, 0);
        for (int32_t l : lev) ++cnt[static_cast<std::size_t>(l)];
        // Separator level: smallest level set with a part balance in [0.3, 0.7].
        std::size_t cum = 0;
        int best = -1;
        std::size_t bestSize = sz + 1;
        int half = -1;
        for (int l = 0; l <= L; ++l) {
            const std::size_t before = cum;
            cum += cnt[static_cast<std::size_t>(l)];
            if (l == 0 || l == L) continue;
            if (half < 0 && cum * 2 >= sz) half = l;
            const double f = static_cast<double>(before) / static_cast<double>(sz);
            const double f2 = static_cast<double>(sz - cum) / static_cast<double>(sz);
            if (f >= 0.3 && f2 >= 0.3 && cnt[static_cast<std::size_t>(l)] < bestSize) {
                best = l;
                bestSize = cnt[static_cast<std::size_t>(l)];
            }
        }
        if (best < 0) best = (half > 0 && half < L) ? half : L / 2;
        std::vector<int32_t> p1, p2, sep;
        for (std::size_t q = 0; q < sz; ++q) {
            const int32_t v = ord[q];
            const int l = lev[q];
            where[static_cast<std::size_t>(v)] = (l < best) ? 1 : (l > best ? 2 : 3);
        }
        // Thin the separator: a separator vertex with no neighbour in part 2 joins part 1.
        for (std::size_t q = 0; q < sz; ++q) {
            const int32_t v = ord[q];
            if (where[static_cast<std::size_t>(v)] != 3) continue;
            bool touches2 = false;
            for (int64_t e = g.xadj[static_cast<std::size_t>(v)]; e < g.xadj[static_cast<std::size_t>(v) + 1] && !touches2; ++e) {
                const int32_t u = g.adj[static_cast<std::size_t>(e)];
                touches2 = (mark[static_cast<std::size_t>(u)] == t && where[static_cast<std::size_t>(u)] == 2);
            }
            if (!touches2) where[static_cast<std::size_t>(v)] = 1;
        }
        for (std::size_t q = 0; q < sz; ++q) {
            const int32_t v = ord[q];
            const int w = where[static_cast<std::size_t>(v)];
            (w == 1 ? p1 : (w == 2 ? p2 : sep)).push_back(v);
        }
        if (p1.empty() || p2.empty()) {
            emit(ord);
            return;
        }
        std::sort(p1.begin(), p1.end());
        std::sort(p2.begin(), p2.end());
        dissect(std::move(p1));
        dissect(std::move(p2));
        emit(sep);
    }
};

inline double dot(const double* a, const double* b, int n)
{
    double s = 0.0;
    for (int k = 0; k < n; ++k) s += a[k] * b[k];
    return s;
}

} // namespace

Factor factorize(int32_t n, const int32_t* rowptr, const int32_t* col, const double* val, int kind, int leaf)
{
    if (n <= 0) throw std::invalid_argument("factorize: empty matrix");
    if (leaf <= 0) leaf = 16;
    // ---- Graph of A + A^T (no self loops).
    Graph g;
    g.n = n;
    {
        std::vector<std::vector<int32_t>> nb(static_cast<std::size_t>(n));
        for (int32_t i = 0; i < n; ++i)
            for (int32_t k = rowptr[i]; k < rowptr[i + 1]; ++k) {
                const int32_t j = col[k];
                if (j < 0 || j >= n) throw std::out_of_range("factorize: column index out of range");
                if (j == i) continue;
                nb[static_cast<std::size_t>(i)].push_back(j);
                nb[static_cast<std::size_t>(j)].push_back(i);
            }
        g.xadj.assign(static_cast<std::size_t>(n) + 1, 0);
        for (int32_t i = 0; i < n; ++i) {
            auto& v = nb[static_cast<std::size_t>(i)];
            std::sort(v.begin(), v.end());
            v.erase(std::unique(v.begin(), v.end()), v.end());
            g.xadj[static_cast<std::size_t>(i) + 1] = g.xadj[static_cast<std::size_t>(i)] + static_cast<int64_t>(v.size());
        }
        g.adj.reserve(static_cast<std::size_t>(g.xadj.back()));
        for (auto& v : nb) g.adj.insert(g.adj.end(), v.begin(), v.end());
    }
    // ---- Nested dissection.
    Orderer od(g, leaf);
    {
        std::vector<int32_t> all(static_cast<std::size_t>(n));
        std::iota(all.begin(), all.end(), 0);
        od.dissect(std::move(all));
    }
    Factor F;
    F.n = n;
    F.perm = od.order;
    F.sn_begin = od.sn_begin;
    F.sn_begin.push_back(n);
    const int32_t ns = static_cast<int32_t>(F.sn_begin.size()) - 1;
    std::vector<int32_t> iperm(static_cast<std::size_t>(n));
    for (int32_t k = 0; k < n; ++k) iperm[static_cast<std::size_t>(F.perm[static_cast<std::size_t>(k)])] = k;
    std::vector<int32_t> snode(static_cast<std::size_t>(n));
    for (int32_t s = 0; s < ns; ++s)
        for (int32_t j = F.sn_begin[static_cast<std::size_t>(s)]; j < F.sn_begin[static_cast<std::size_t>(s) + 1]; ++j) snode[static_cast<std::size_t>(j)] = s;

    // ---- Permuted matrix B = A(perm, perm) in CSC (lower part with the rows >= column) and, for
    //      LU, the transposed upper part (row j, columns >= j).
    std::vector<std::vector<std::pair<int32_t, double>>> lowerCol(static_cast<std::size_t>(n)), upperRow(static_cast<std::size_t>(n));
    for (int32_t i = 0; i < n; ++i)
        for (int32_t k = rowptr[i]; k < rowptr[i + 1]; ++k) {
            const int32_t pi = iperm[static_cast<std::size_t>(i)], pj = iperm[static_cast<std::size_t>(col[k])];
            if (pi >= pj) lowerCol[static_cast<std::size_t>(pj)].emplace_back(pi, val[k]); // B(pi, pj), pi >= pj
            if (pi <= pj) upperRow[static_cast<std::size_t>(pi)].emplace_back(pj, val[k]); // B(pi, pj), pj >= pi
        }

    // ---- Block symbolic factorization.
    std::vector<std::vector<int32_t>> R(static_cast<std::size_t>(ns));
    std::vector<int32_t> parent(static_cast<std::size_t>(ns), -1);
    std::vector<std::vector<int32_t>> children(static_cast<std::size_t>(ns));
    {
        std::vector<int32_t> seen(static_cast<std::size_t>(n), -1);
        for (int32_t s = 0; s < ns; ++s) {
            const int32_t c0 = F.sn_begin[static_cast<std::size_t>(s)], c1 = F.sn_begin[static_cast<std::size_t>(s) + 1];
            auto& rs = R[static_cast<std::size_t>(s)];
            auto addRow = [&](int32_t r) {
                if (r >= c1 && seen[static_cast<std::size_t>(r)] != s) {
                    seen[static_cast<std::size_t>(r)] = s;
                    rs.push_back(r);
                }
            };
            for (int32_t j = c0; j < c1; ++j) {
                for (const auto& e : lowerCol[static_cast<std::size_t>(j)]) addRow(e.first);
                for (const auto& e : upperRow[static_cast<std::size_t>(j)]) addRow(e.first);
            }
            for (int32_t c : children[static_cast<std::size_t>(s)])
                for (int32_t r : R[static_cast<std::size_t>(c)]) addRow(r);
            std::sort(rs.begin(), rs.end());
            if (!rs.empty()) {
                parent[static_cast<std::size_t>(s)] = snode[static_cast<std::size_t>(rs.front())];
                children[static_cast<std::size_t>(parent[static_cast<std::size_t>(s)])].push_back(s);
            }
        }
    }
    // Update lists: supernode d updates s with its rows R_d[p, q) (all inside s) and all rows after.
    struct Upd {
        int32_t d, p, q;
    };
    std::vector<std::vector<Upd>> upd(static_cast<std::size_t>(ns));
    for (int32_t d = 0; d < ns; ++d) {
        const auto& rd = R[static_cast<std::size_t>(d)];
        std::size_t p = 0;
        while (p < rd.size()) {
            const int32_t s = snode[static_cast<std::size_t>(rd[p])];
            std::size_t q = p;
            while (q < rd.size() && snode[static_cast<std::size_t>(rd[q])] == s) ++q;
            upd[static_cast<std::size_t>(s)].push_back({d, static_cast<int32_t>(p), static_cast<int32_t>(q)});
            p = q;
        }
    }

    // ---- Numeric factorization. Panels are (w + h) x w row-major: rows C_s then R_s.
    const bool lu0 = (kind == 1);
    auto run = [&](bool lu, std::vector<std::vector<double>>& Lp, std::vector<std::vector<double>>& Up, double& flops) -> bool {
        Lp.assign(static_cast<std::size_t>(ns), {});
        Up.assign(static_cast<std::size_t>(ns), {});
        flops = 0.0;
        std::vector<int32_t> rel(static_cast<std::size_t>(n), -1);
        for (int32_t s = 0; s < ns; ++s) {
            const int32_t c0 = F.sn_begin[static_cast<std::size_t>(s)], w = F.sn_begin[static_cast<std::size_t>(s) + 1] - c0;
            const auto& rs = R[static_cast<std::size_t>(s)];
            const int32_t h = static_cast<int32_t>(rs.size()), m = w + h;
            for (int32_t c = 0; c < w; ++c) rel[static_cast<std::size_t>(c0 + c)] = c;
            for (int32_t k = 0; k < h; ++k) rel[static_cast<std::size_t>(rs[static_cast<std::size_t>(k)])] = w + k;
            auto& L = Lp[static_cast<std::size_t>(s)];
            L.assign(static_cast<std::size_t>(m) * w, 0.0);
            std::vector<double>* Up_s = lu ? &Up[static_cast<std::size_t>(s)] : nullptr;
            if (lu) Up_s->assign(static_cast<std::size_t>(m) * w, 0.0);
            // Scatter A.
            for (int32_t c = 0; c < w; ++c) {
                const int32_t j = c0 + c;
                for (const auto& e : lowerCol[static_cast<std::size_t>(j)]) L[static_cast<std::size_t>(rel[static_cast<std::size_t>(e.first)]) * w + c] += e.second;
                if (lu)
                    for (const auto& e : upperRow[static_cast<std::size_t>(j)]) (*Up_s)[static_cast<std::size_t>(rel[static_cast<std::size_t>(e.first)]) * w + c] += e.second;
            }
            // Updates from descendants.
            for (const Upd& u : upd[static_cast<std::size_t>(s)]) {
                const int32_t d = u.d;
                const int32_t wd = F.sn_begin[static_cast<std::size_t>(d) + 1] - F.sn_begin[static_cast<std::size_t>(d)];
                const auto& rd = R[static_cast<std::size_t>(d)];
                const int32_t hd = static_cast<int32_t>(rd.size());
                const double* Ld = Lp[static_cast<std::size_t>(d)].data();
                const double* Ud = lu ? Up[static_cast<std::size_t>(d)].data() : Ld;
                for (int32_t a = u.p; a < hd; ++a) {
                    const int32_t ra = rel[static_cast<std::size_t>(rd[static_cast<std::size_t>(a)])];
                    const double* la = Ld + static_cast<std::size_t>(wd + a) * wd;
                    const double* ua = Ud + static_cast<std::size_t>(wd + a) * wd;
                    const int32_t bmax = std::min(a + 1, u.q);
                    for (int32_t b = u.p; b < bmax; ++b) {
                        const int32_t cb = rd[static_cast<std::size_t>(b)] - c0;
                        const double* ub = Ud + static_cast<std::size_t>(wd + b) * wd;
                        L[static_cast<std::size_t>(ra) * w + cb] -= dot(la, ub, wd);
                        if (lu) {
                            const double* lb = Ld + static_cast<std::size_t>(wd + b) * wd;
                            (*Up_s)[static_cast<std::size_t>(ra) * w + cb] -= dot(ua, lb, wd);
                        }
                    }
                    flops += 2.0 * wd * (bmax - u.p) * (lu ? 2 : 1);
                }
            }
            // Dense factorization of the panel.
            if (!lu) {
                for (int32_t j = 0; j < w; ++j) {
                    const double* lj = L.data() + static_cast<std::size_t>(j) * w;
                    const double dj = L[static_cast<std::size_t>(j) * w + j] - dot(lj, lj, j);
                    if (!(dj > 0.0) || !std::isfinite(dj)) return false;
                    const double ljj = std::sqrt(dj);
                    L[static_cast<std::size_t>(j) * w + j] = ljj;
                    for (int32_t i = j + 1; i < m; ++i) {
                        double* li = L.data() + static_cast<std::size_t>(i) * w;
                        li[j] = (li[j] - dot(li, lj, j)) / ljj;
                    }
                    flops += 2.0 * j * (m - j);
                }
            } else {
                auto& U = *Up_s;
                for (int32_t j = 0; j < w; ++j) {
                    const double* lj = L.data() + static_cast<std::size_t>(j) * w;
                    for (int32_t i = j; i < m; ++i) {
                        double* ui = U.data() + static_cast<std::size_t>(i) * w;
                        ui[j] -= dot(lj, ui, j);
                    }
                    const double ujj = U[static_cast<std::size_t>(j) * w + j];
                    if (ujj == 0.0 || !std::isfinite(ujj)) return false;
                    const double* uj = U.data() + static_cast<std::size_t>(j) * w;
                    for (int32_t i = j + 1; i < m; ++i) {
                        double* li = L.data() + static_cast<std::size_t>(i) * w;
                        li[j] = (li[j] - dot(li, uj, j)) / ujj;
                    }
                    L[static_cast<std::size_t>(j) * w + j] = 1.0;
                    for (int32_t c = j + 1; c < w; ++c) L[static_cast<std::size_t>(j) * w + c] = 0.0;
                    flops += 4.0 * j * (m - j);
                }
            }
            // The strictly upper part of the diagonal block is not part of the factor.
            for (int32_t r = 0; r < w; ++r)
                for (int32_t c = r + 1; c < w; ++c) {
                    L[static_cast<std::size_t>(r) * w + c] = 0.0;
                    if (lu) (*Up_s)[static_cast<std::size_t>(r) * w + c] = 0.0;
                }
        }
        return true;
    };
    std::vector<std::vector<double>> Lp, Up;
    double flops = 0.0;
    bool lu = lu0;
    if (!run(lu, Lp, Up, flops)) {
        if (lu) throw std::runtime_error("factorize: zero pivot in LU without pivoting");
        lu = true; // Cholesky failed: LU, as the reference's chain falls back
        if (!run(lu, Lp, Up, flops)) throw std::runtime_error("factorize: zero pivot in LU without pivoting");
    }
    F.kind = lu ? 1 : 0;
    F.flops = flops;

    // ---- Export CSC.
    auto exportCsc = [&](const std::vector<std::vector<double>>& P, std::vector<int64_t>& cp, std::vector<int32_t>& ri, std::vector<double>& vv) {
        cp.assign(static_cast<std::size_t>(n) + 1, 0);
        int64_t nnz = 0;
        for (int32_t s = 0; s < ns; ++s) {
            const int32_t c0 = F.sn_begin[static_cast<std::size_t>(s)], w = F.sn_begin[static_cast<std::size_t>(s) + 1] - c0;
            const int32_t h = static_cast<int32_t>(R[static_cast<std::size_t>(s)].size());
            for (int32_t c = 0; c < w; ++c) {
                nnz += (w - c) + h;
                cp[static_cast<std::size_t>(c0 + c) + 1] = nnz;
            }
        }
        ri.resize(static_cast<std::size_t>(nnz));
        vv.resize(static_cast<std::size_t>(nnz));
        for (int32_t s = 0; s < ns; ++s) {
            const int32_t c0 = F.sn_begin[static_cast<std::size_t>(s)], w = F.sn_begin[static_cast<std::size_t>(s) + 1] - c0;
            const auto& rs = R[static_cast<std::size_t>(s)];
            const int32_t h = static_cast<int32_t>(rs.size());
            const auto& M = P[static_cast<std::size_t>(s)];
            for (int32_t c = 0; c < w; ++c) {
                int64_t q = cp[static_cast<std::size_t>(c0 + c)];
                for (int32_t r = c; r < w; ++r, ++q) {
                    ri[static_cast<std::size_t>(q)] = c0 + r;
                    vv[static_cast<std::size_t>(q)] = M[static_cast<std::size_t>(r) * w + c];
                }
                for (int32_t k = 0; k < h; ++k, ++q) {
                    ri[static_cast<std::size_t>(q)] = rs[static_cast<std::size_t>(k)];
                    vv[static_cast<std::size_t>(q)] = M[static_cast<std::size_t>(w + k) * w + c];
                }
            }
        }
    };
    exportCsc(Lp, F.l_colptr, F.l_row, F.l_val);
    if (lu) exportCsc(Up, F.ut_colptr, F.ut_row, F.ut_val);
    return F;
}

} // namespace s2l
