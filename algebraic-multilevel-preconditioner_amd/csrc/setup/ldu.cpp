// ldu.cpp -- OpenFOAM's lower-diagonal-upper (LDU) face storage to the global CSR matrix, in memory.
//
// The reference's OpenFOAM solver writes its pressure matrix to a Matrix Market file that the
// library then reads (openfoam/bubblyPressureFoam/bubblyPressureFoam.C:120-170): cell c carries the
// diagonal D[c] (with the boundary contributions), face f between the cells l[f] < u[f] carries
// upper[f] at (l[f], u[f]) and lower[f] at (u[f], l[f]), every cell is renumbered to gid[c], and the
// right-hand side is b[gid[c]] = source[c]. This builds the same matrix without the file: entries in
// the reader's order (the diagonal, then each face's pair), exact zeros dropped and every row sorted
// by column, so the result equals the write-then-read route bitwise. A symmetric lduMatrix stores
// no lower coefficients (lower = upper).
#include "setup_internal.hpp"

#include <stdexcept>
#include <string>
#include <vector>

namespace s2l {

Problem fromLdu(int64_t n, int64_t nfaces, const int64_t* l, const int64_t* u, const double* diag, const double* upper,
                const double* lower, const int64_t* gid, const double* source)
{
    if (n <= 0 || nfaces < 0) throw std::invalid_argument("from_ldu: invalid sizes");
    if (!diag || (nfaces > 0 && (!l || !u || !upper))) throw std::invalid_argument("from_ldu: missing arrays");
    std::vector<int64_t> g(static_cast<std::size_t>(n));
    std::vector<char> seen(static_cast<std::size_t>(n), 0);
    for (int64_t c = 0; c < n; ++c) {
        const int64_t x = gid ? gid[c] : c;
        if (x < 0 || x >= n || seen[static_cast<std::size_t>(x)])
            throw std::invalid_argument("from_ldu: the global cell numbers are not a permutation of 0..n-1");
        seen[static_cast<std::size_t>(x)] = 1;
        g[static_cast<std::size_t>(c)] = x;
    }
    std::vector<int64_t> rows, cols;
    std::vector<double> vals;
    rows.reserve(static_cast<std::size_t>(n + 2 * nfaces));
    cols.reserve(rows.capacity());
    vals.reserve(rows.capacity());
    auto add = [&](int64_t i, int64_t j, double v) {
        if (v == 0.0) return;
        rows.push_back(i);
        cols.push_back(j);
        vals.push_back(v);
    };
    for (int64_t c = 0; c < n; ++c) add(g[static_cast<std::size_t>(c)], g[static_cast<std::size_t>(c)], diag[c]);
    for (int64_t f = 0; f < nfaces; ++f) {
        const int64_t a = l[f], b = u[f];
        if (a < 0 || a >= n || b < 0 || b >= n || a == b)
            throw std::invalid_argument("from_ldu: face " + std::to_string(f) + " has invalid cells");
        const int64_t ga = g[static_cast<std::size_t>(a)], gb = g[static_cast<std::size_t>(b)];
        add(ga, gb, upper[f]);
        add(gb, ga, lower ? lower[f] : upper[f]);
    }
    Problem P;
    P.A = fromCoo(n, rows, cols, vals);
    P.symmetric = lower == nullptr;
    P.dim = 0;
    if (source) {
        P.rhs.assign(static_cast<std::size_t>(n), 0.0);
        for (int64_t c = 0; c < n; ++c) P.rhs[static_cast<std::size_t>(g[static_cast<std::size_t>(c)])] = source[c];
    }
    return P;
}

} // namespace s2l
