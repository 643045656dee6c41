// setup_internal.hpp -- shared types of the host setup stand-in.
#ifndef S2L_SETUP_INTERNAL_HPP
#define S2L_SETUP_INTERNAL_HPP

#include "s2l_setup.h"

#include <cstdint>
#include <string>
#include <vector>

namespace s2l {

/// Global matrix in CSR with int64 indices (the reference's CsrMatrix layout).
struct GlobalCsr {
    int64_t n = 0;
    std::vector<int64_t> rp, col;
    std::vector<double> val;
};

/// A generated problem.
struct Problem {
    GlobalCsr A;
    std::vector<double> rhs;
    std::vector<double> xyz; ///< cell centre of every unknown (3 per row)
    bool symmetric = false;
    int block_size = 1;
    int dim = 2;
};

GlobalCsr fromCoo(int64_t n, const std::vector<int64_t>& rows, const std::vector<int64_t>& cols,
                  const std::vector<double>& vals);
Problem generate(const std::string& name, const s2l_gen_opts& o);

/// Matrix Market input/output (mmio.cpp).
Problem readMatrixMarket(const std::string& path);
std::vector<double> readMatrixMarketVector(const std::string& path);
void writeMatrixMarket(const std::string& path, int64_t n, const int64_t* rp, const int64_t* col, const double* val);
void writeMatrixMarketVector(const std::string& path, int64_t n, const double* v);
std::string stepFileName(const std::string& pattern, int step);
uint64_t hashPattern(int64_t n, const int64_t* rp, const int64_t* col);
/// OpenFOAM LDU face storage to the global CSR matrix (ldu.cpp).
Problem fromLdu(int64_t n, int64_t nfaces, const int64_t* l, const int64_t* u, const double* diag, const double* upper,
                const double* lower, const int64_t* gid, const double* source);

/// Symmetric adjacency (no self loops) of an n-vertex graph, CSR with int32 indices.
struct Graph {
    int32_t n = 0;
    std::vector<int64_t> xadj;
    std::vector<int32_t> adj;
};

std::vector<int32_t> partitionGraph(const Graph& g, int nparts, const double* xyz);

/// Local factor: supernodal FP64, exported as CSC.
struct Factor {
    int32_t n = 0;
    int kind = 0; ///< 0 Cholesky, 1 LU
    std::vector<int32_t> perm;         ///< factor position -> local index
    std::vector<int32_t> sn_begin;     ///< supernode partition (size nsuper + 1)
    std::vector<int64_t> l_colptr, ut_colptr;
    std::vector<int32_t> l_row, ut_row;
    std::vector<double> l_val, ut_val;
    double flops = 0.0;
};

Factor factorize(int32_t n, const int32_t* rowptr, const int32_t* col, const double* val, int kind, int leaf);

} // namespace s2l

struct s2l_problem {
    s2l::Problem p;
};
struct s2l_factor {
    s2l::Factor f;
};

#endif
