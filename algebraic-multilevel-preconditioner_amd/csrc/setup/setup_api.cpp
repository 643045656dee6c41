// setup_api.cpp -- C-ABI of the host setup stand-in (declared in include/s2l_setup.h).
#include "setup_internal.hpp"

#include <algorithm>
#include <climits>
#include <cstring>
#include <exception>
#include <random>
#include <stdexcept>
#include <string>

namespace {
thread_local std::string g_error;

template <typename F> int guarded(F&& f)
{
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
    } catch (...) {
        g_error = "unknown error";
    }
    return -1;
}
} // namespace

extern "C" {

const char* s2l_setup_last_error(void) { return g_error.c_str(); }

void s2l_gen_defaults(s2l_gen_opts* o)
{
    std::memset(o, 0, sizeof(*o));
    o->nx = 64;
    o->nu = 1.0;
    o->ncomp = 4;
    o->pulse_period = 20;
    o->rho_ratio = 1000.0;
    o->nbubbles = 50;
    o->radius = 0.06;
    o->cfl = 0.5;
    o->seed = 2024;
    o->contrast = 1e6;
    o->channel_fraction = 0.1;
    o->aniso = 1e-3;
}

s2l_problem* s2l_generate(const char* name, const s2l_gen_opts* o)
{
    s2l_problem* out = nullptr;
    if (guarded([&] {
            auto* p = new s2l_problem;
            try {
                p->p = s2l::generate(name ? name : "", *o);
            } catch (...) {
                delete p;
                throw;
            }
            out = p;
        }) != 0)
        return nullptr;
    return out;
}

void s2l_problem_free(s2l_problem* p) { delete p; }
int64_t s2l_problem_n(const s2l_problem* p) { return p->p.A.n; }
int64_t s2l_problem_nnz(const s2l_problem* p) { return static_cast<int64_t>(p->p.A.col.size()); }
int s2l_problem_symmetric(const s2l_problem* p) { return p->p.symmetric ? 1 : 0; }
int s2l_problem_block_size(const s2l_problem* p) { return p->p.block_size; }
int s2l_problem_dim(const s2l_problem* p) { return p->p.dim; }

void s2l_problem_copy(const s2l_problem* p, int64_t* rowptr, int64_t* col, double* val, double* rhs, double* xyz)
{
    const auto& P = p->p;
    if (rowptr) std::memcpy(rowptr, P.A.rp.data(), P.A.rp.size() * sizeof(int64_t));
    if (col) std::memcpy(col, P.A.col.data(), P.A.col.size() * sizeof(int64_t));
    if (val) std::memcpy(val, P.A.val.data(), P.A.val.size() * sizeof(double));
    if (rhs) std::memcpy(rhs, P.rhs.data(), P.rhs.size() * sizeof(double));
    if (xyz) std::memcpy(xyz, P.xyz.data(), P.xyz.size() * sizeof(double));
}

void s2l_random_rhs(int64_t n, uint64_t seed, double* out)
{
    std::mt19937_64 gen(seed);
    std::uniform_real_distribution<double> dist(-1.0, 1.0);
    for (int64_t i = 0; i < n; ++i) out[i] = dist(gen);
}

void s2l_random_rhs_seq(int64_t n, uint64_t seed, int step, int corrector, double* out)
{
    std::seed_seq sq{static_cast<unsigned>(seed), static_cast<unsigned>(step), static_cast<unsigned>(corrector), 0x5eedu};
    std::mt19937_64 gen(sq);
    std::uniform_real_distribution<double> dist(-1.0, 1.0);
    for (int64_t i = 0; i < n; ++i) out[i] = dist(gen);
}

s2l_problem* s2l_read_matrix(const char* path)
{
    s2l_problem* out = nullptr;
    if (guarded([&] {
            if (!path) throw std::invalid_argument("read_matrix: null path");
            auto* p = new s2l_problem;
            try {
                p->p = s2l::readMatrixMarket(path);
            } catch (...) {
                delete p;
                throw;
            }
            out = p;
        }) != 0)
        return nullptr;
    return out;
}

int s2l_read_vector(const char* path, double* out, int64_t cap, int64_t* n)
{
    return guarded([&] {
        if (!path || !n) throw std::invalid_argument("read_vector: null argument");
        const std::vector<double> v = s2l::readMatrixMarketVector(path);
        *n = static_cast<int64_t>(v.size());
        if (out && cap > 0) std::memcpy(out, v.data(), sizeof(double) * static_cast<std::size_t>(std::min<int64_t>(cap, *n)));
    });
}

int s2l_write_matrix(const char* path, int64_t n, const int64_t* rowptr, const int64_t* col, const double* val)
{
    return guarded([&] {
        if (!path || n < 0 || !rowptr || (rowptr[n] > 0 && (!col || !val))) throw std::invalid_argument("write_matrix: bad arguments");
        s2l::writeMatrixMarket(path, n, rowptr, col, val);
    });
}

int s2l_write_vector(const char* path, int64_t n, const double* v)
{
    return guarded([&] {
        if (!path || n < 0 || (n > 0 && !v)) throw std::invalid_argument("write_vector: bad arguments");
        s2l::writeMatrixMarketVector(path, n, v);
    });
}

int64_t s2l_step_file_name(const char* pattern, int step, char* out, int64_t cap)
{
    int64_t len = -1;
    if (guarded([&] {
            if (!pattern) throw std::invalid_argument("step_file_name: null pattern");
            const std::string name = s2l::stepFileName(pattern, step);
            if (!out || cap < static_cast<int64_t>(name.size()) + 1)
                throw std::invalid_argument("step_file_name: buffer too short");
            std::memcpy(out, name.c_str(), name.size() + 1);
            len = static_cast<int64_t>(name.size());
        }) != 0)
        return -1;
    return len;
}

uint64_t s2l_hash_pattern(int64_t n, const int64_t* rowptr, const int64_t* col)
{
    return s2l::hashPattern(n, rowptr, col);
}

s2l_problem* s2l_from_ldu(int64_t n, int64_t nfaces, const int64_t* lower_addr, const int64_t* upper_addr,
                          const double* diag, const double* upper, const double* lower, const int64_t* gid,
                          const double* source)
{
    s2l_problem* out = nullptr;
    if (guarded([&] {
            auto* p = new s2l_problem;
            try {
                p->p = s2l::fromLdu(n, nfaces, lower_addr, upper_addr, diag, upper, lower, gid, source);
            } catch (...) {
                delete p;
                throw;
            }
            out = p;
        }) != 0)
        return nullptr;
    return out;
}

int64_t s2l_problem_rhs_size(const s2l_problem* p) { return static_cast<int64_t>(p->p.rhs.size()); }

int s2l_partition(int64_t n, const int64_t* rowptr, const int64_t* col, int nparts, const double* xyz, int block_size,
                  int32_t* part)
{
    return guarded([&] {
        const int bs = block_size > 0 ? block_size : 1;
        if (n % bs != 0) throw std::invalid_argument("partition: n is not a multiple of the block size");
        const int64_t nb = n / bs;
        if (nb > INT32_MAX) throw std::invalid_argument("partition: graph too large");
        s2l::Graph g;
        g.n = static_cast<int32_t>(nb);
        std::vector<std::vector<int32_t>> nbv(static_cast<std::size_t>(nb));
        for (int64_t i = 0; i < n; ++i)
            for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) {
                const int64_t I = i / bs, J = col[k] / bs;
                if (I == J) continue;
                nbv[static_cast<std::size_t>(I)].push_back(static_cast<int32_t>(J));
                nbv[static_cast<std::size_t>(J)].push_back(static_cast<int32_t>(I));
            }
        g.xadj.assign(static_cast<std::size_t>(nb) + 1, 0);
        for (int64_t I = 0; I < nb; ++I) {
            auto& v = nbv[static_cast<std::size_t>(I)];
            std::sort(v.begin(), v.end());
            v.erase(std::unique(v.begin(), v.end()), v.end());
            g.xadj[static_cast<std::size_t>(I) + 1] = g.xadj[static_cast<std::size_t>(I)] + static_cast<int64_t>(v.size());
        }
        for (auto& v : nbv) {
            g.adj.insert(g.adj.end(), v.begin(), v.end());
            std::vector<int32_t>().swap(v);
        }
        std::vector<double> cxyz;
        const double* cx = nullptr;
        if (xyz) {
            cxyz.resize(static_cast<std::size_t>(nb) * 3);
            for (int64_t I = 0; I < nb; ++I)
                for (int d = 0; d < 3; ++d) cxyz[static_cast<std::size_t>(I) * 3 + d] = xyz[(I * bs) * 3 + d];
            cx = cxyz.data();
        }
        const std::vector<int32_t> cp = s2l::partitionGraph(g, nparts, cx);
        for (int64_t i = 0; i < n; ++i) part[i] = cp[static_cast<std::size_t>(i / bs)];
    });
}

s2l_factor* s2l_factorize(int32_t n, const int32_t* rowptr, const int32_t* col, const double* val, int kind, int leaf)
{
    s2l_factor* out = nullptr;
    if (guarded([&] {
            auto* f = new s2l_factor;
            try {
                f->f = s2l::factorize(n, rowptr, col, val, kind, leaf);
            } catch (...) {
                delete f;
                throw;
            }
            out = f;
        }) != 0)
        return nullptr;
    return out;
}

void s2l_factor_free(s2l_factor* f) { delete f; }

void s2l_factor_info(const s2l_factor* f, int32_t* n, int64_t* nnz_l, int64_t* nnz_ut, int32_t* nsuper, int* kind,
                     double* flops)
{
    const auto& F = f->f;
    if (n) *n = F.n;
    if (nnz_l) *nnz_l = static_cast<int64_t>(F.l_row.size());
    if (nnz_ut) *nnz_ut = static_cast<int64_t>(F.kind == 1 ? F.ut_row.size() : F.l_row.size());
    if (nsuper) *nsuper = static_cast<int32_t>(F.sn_begin.size()) - 1;
    if (kind) *kind = F.kind;
    if (flops) *flops = F.flops;
}

void s2l_factor_copy(const s2l_factor* f, int32_t* perm, int64_t* l_colptr, int32_t* l_row, double* l_val, int64_t* ut_colptr,
                     int32_t* ut_row, double* ut_val)
{
    const auto& F = f->f;
    if (perm) std::memcpy(perm, F.perm.data(), F.perm.size() * sizeof(int32_t));
    if (l_colptr) std::memcpy(l_colptr, F.l_colptr.data(), F.l_colptr.size() * sizeof(int64_t));
    if (l_row) std::memcpy(l_row, F.l_row.data(), F.l_row.size() * sizeof(int32_t));
    if (l_val) std::memcpy(l_val, F.l_val.data(), F.l_val.size() * sizeof(double));
    const bool lu = (F.kind == 1);
    if (ut_colptr) std::memcpy(ut_colptr, (lu ? F.ut_colptr : F.l_colptr).data(), F.l_colptr.size() * sizeof(int64_t));
    if (ut_row) std::memcpy(ut_row, (lu ? F.ut_row : F.l_row).data(), (lu ? F.ut_row : F.l_row).size() * sizeof(int32_t));
    if (ut_val) std::memcpy(ut_val, (lu ? F.ut_val : F.l_val).data(), (lu ? F.ut_val : F.l_val).size() * sizeof(double));
}

void s2l_factor_supernodes(const s2l_factor* f, int32_t* sn_begin)
{
    std::memcpy(sn_begin, f->f.sn_begin.data(), f->f.sn_begin.size() * sizeof(int32_t));
}

} // extern "C"
