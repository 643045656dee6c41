// partition.cpp -- partition of G(A + A^T) into subdomains (stand-in for the reference's METIS
// call, src/partition/graph_partitioner.cpp:17-94).
//
// Two deterministic methods: recursive coordinate bisection when cell centres are known (the
// generated problems), and recursive graph bisection in breadth-first order from a
// pseudo-peripheral vertex otherwise. Block unknowns of one cell stay together, as with the
// reference's cell-compressed graph (block_size > 1).
#include "setup_internal.hpp"

#include <algorithm>
#include <numeric>
#include <queue>
#include <stdexcept>

namespace s2l {

namespace {

// Splits `items` into parts [first, first + k) by recursive coordinate bisection along the
// longest extent; the first half gets round(|items| * floor(k/2) / k) items.
void rcb(std::vector<int64_t>& items, std::size_t lo, std::size_t hi, int k, int first, const double* xyz,
         std::vector<int32_t>& cellPart)
{
    if (k <= 1) {
        for (std::size_t q = lo; q < hi; ++q) cellPart[static_cast<std::size_t>(items[q])] = first;
        return;
    }
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (std::size_t q = lo; q < hi; ++q)
        for (int d = 0; d < 3; ++d) {
            const double v = xyz[3 * items[q] + d];
            mn[d] = std::min(mn[d], v);
            mx[d] = std::max(mx[d], v);
        }
    int ax = 0;
    for (int d = 1; d < 3; ++d)
        if (mx[d] - mn[d] > mx[ax] - mn[ax]) ax = d;
    const int k1 = k / 2;
    const std::size_t cnt = hi - lo;
    const std::size_t c1 = static_cast<std::size_t>((static_cast<double>(cnt) * k1) / k + 0.5);
    auto less = [&](int64_t a, int64_t b) {
        const double va = xyz[3 * a + ax], vb = xyz[3 * b + ax];
        return va < vb || (va == vb && a < b);
    };
    std::nth_element(items.begin() + static_cast<std::ptrdiff_t>(lo), items.begin() + static_cast<std::ptrdiff_t>(lo + c1),
                     items.begin() + static_cast<std::ptrdiff_t>(hi), less);
    rcb(items, lo, lo + c1, k1, first, xyz, cellPart);
    rcb(items, lo + c1, hi, k - k1, first + k1, xyz, cellPart);
}

// Breadth-first order of the vertices of `set` (marked with `tag` in `mark`) starting from
// `start`; unreached vertices start new searches in increasing order.
void bfsOrder(const Graph& g, const std::vector<int32_t>& set, int32_t start, std::vector<int32_t>& mark, int tag,
              std::vector<int32_t>& order, std::vector<int32_t>* level)
{
    order.clear();
    if (level) level->clear();
    std::vector<int32_t> seen;
    seen.reserve(set.size());
    auto visit = [&](int32_t s) {
        std::size_t head = order.size();
        order.push_back(s);
        if (level) level->push_back(0);
        mark[static_cast<std::size_t>(s)] = -tag;
        while (head < order.size()) {
            const int32_t v = order[head];
            const int32_t lv = level ? (*level)[head] : 0;
            ++head;
            for (int64_t e = g.xadj[static_cast<std::size_t>(v)]; e < g.xadj[static_cast<std::size_t>(v) + 1]; ++e) {
                const int32_t u = g.adj[static_cast<std::size_t>(e)];
                if (mark[static_cast<std::size_t>(u)] == tag) {
                    mark[static_cast<std::size_t>(u)] = -tag;
                    order.push_back(u);
                    if (level) level->push_back(lv + 1);
                }
            }
        }
    };
    visit(start);
    for (int32_t v : set)
        if (mark[static_cast<std::size_t>(v)] == tag) visit(v);
    for (int32_t v : set) mark[static_cast<std::size_t>(v)] = tag; // restore
}

int32_t pseudoPeripheral(const Graph& g, const std::vector<int32_t>& set, std::vector<int32_t>& mark, int tag)
{
    int32_t v = *std::min_element(set.begin(), set.end());
    std::vector<int32_t> order, level;
    int best = -1;
    for (int it = 0; it < 4; ++it) {
        bfsOrder(g, set, v, mark, tag, order, &level);
        // farthest vertex of the first component
        int32_t far = v;
        int lmax = -1;
        for (std::size_t q = 0; q < order.size(); ++q) {
            if (q > 0 && level[q] == 0) break;
            if (level[q] > lmax) {
                lmax = level[q];
                far = order[q];
            }
        }
        if (lmax <= best) break;
        best = lmax;
        v = far;
    }
    return v;
}

void graphBisect(const Graph& g, std::vector<int32_t>& set, int k, int first, std::vector<int32_t>& mark, int& tagCounter,
                 std::vector<int32_t>& part)
{
    if (k <= 1) {
        for (int32_t v : set) part[static_cast<std::size_t>(v)] = first;
        return;
    }
    const int tag = ++tagCounter;
    for (int32_t v : set) mark[static_cast<std::size_t>(v)] = tag;
    const int32_t s = pseudoPeripheral(g, set, mark, tag);
    std::vector<int32_t> order;
    bfsOrder(g, set, s, mark, tag, order, nullptr);
    const int k1 = k / 2;
    const std::size_t c1 = static_cast<std::size_t>((static_cast<double>(order.size()) * k1) / k + 0.5);
    std::vector<int32_t> a(order.begin(), order.begin() + static_cast<std::ptrdiff_t>(c1));
    std::vector<int32_t> b(order.begin() + static_cast<std::ptrdiff_t>(c1), order.end());
    std::sort(a.begin(), a.end());
    std::sort(b.begin(), b.end());
    set.clear();
    set.shrink_to_fit();
    graphBisect(g, a, k1, first, mark, tagCounter, part);
    graphBisect(g, b, k - k1, first + k1, mark, tagCounter, part);
}

} // namespace

std::vector<int32_t> partitionGraph(const Graph& g, int nparts, const double* xyz)
{
    std::vector<int32_t> part(static_cast<std::size_t>(g.n), 0);
    if (nparts <= 1 || g.n == 0) return part;
    if (nparts > g.n) throw std::invalid_argument("partition: more parts than vertices");
    if (xyz) {
        std::vector<int64_t> items(static_cast<std::size_t>(g.n));
        std::iota(items.begin(), items.end(), 0);
        rcb(items, 0, items.size(), nparts, 0, xyz, part);
        return part;
    }
    std::vector<int32_t> set(static_cast<std::size_t>(g.n)), mark(static_cast<std::size_t>(g.n), 0);
    std::iota(set.begin(), set.end(), 0);
    int tag = 0;
    graphBisect(g, set, nparts, 0, mark, tag, part);
    return part;
}

} // namespace s2l
