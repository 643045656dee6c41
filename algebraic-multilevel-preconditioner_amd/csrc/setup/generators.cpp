// generators.cpp -- matrix families of the apply-path workloads (host setup stand-in).
//
// The reference families (laplace/convdiff/coupled 2d/3d, bubbles 2d/3d) follow
// src/utils/matrix_generators.cpp:57-306 and include/utils/bubble_field.hpp:43-158
// step for step, because their matrices must come out bit-identical (the committed
// bubbles3d 24^3 fixture of the reference pins it). The arithmetic order of every
// coefficient, the COO -> CSR duplicate summation (src/partition/csr_matrix.cpp:13-64)
// and the libstdc++ random engines are kept for that reason. The extra families
// (poisson, aniso, highcontrast, per-bubble radii) have no reference counterpart.
#include "setup_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>

namespace s2l {

GlobalCsr fromCoo(int64_t n, const std::vector<int64_t>& rows, const std::vector<int64_t>& cols,
                  const std::vector<double>& vals)
{
    // Bucket by row in insertion order, then order each row by column and sum duplicates in the
    // order the sort leaves them (src/partition/csr_matrix.cpp:22-62).
    GlobalCsr A;
    A.n = n;
    const std::size_t nz = vals.size();
    std::vector<int64_t> ptr(static_cast<std::size_t>(n) + 1, 0);
    for (std::size_t k = 0; k < nz; ++k) {
        if (rows[k] < 0 || rows[k] >= n || cols[k] < 0 || cols[k] >= n) throw std::out_of_range("fromCoo: index out of range");
        ++ptr[static_cast<std::size_t>(rows[k]) + 1];
    }
    std::partial_sum(ptr.begin(), ptr.end(), ptr.begin());
    std::vector<int64_t> c(nz);
    std::vector<double> v(nz);
    {
        std::vector<int64_t> next(ptr.begin(), ptr.end() - 1);
        for (std::size_t k = 0; k < nz; ++k) {
            const int64_t p = next[static_cast<std::size_t>(rows[k])]++;
            c[static_cast<std::size_t>(p)] = cols[k];
            v[static_cast<std::size_t>(p)] = vals[k];
        }
    }
    A.rp.assign(static_cast<std::size_t>(n) + 1, 0);
    A.col.reserve(nz);
    A.val.reserve(nz);
    std::vector<std::size_t> perm;
    for (int64_t i = 0; i < n; ++i) {
        const auto b = static_cast<std::size_t>(ptr[static_cast<std::size_t>(i)]);
        const auto e = static_cast<std::size_t>(ptr[static_cast<std::size_t>(i) + 1]);
        perm.resize(e - b);
        std::iota(perm.begin(), perm.end(), b);
        std::sort(perm.begin(), perm.end(), [&](std::size_t x, std::size_t y) { return c[x] < c[y]; });
        int64_t last = -1;
        for (std::size_t q : perm) {
            if (c[q] == last) {
                A.val.back() += v[q];
            } else {
                A.col.push_back(c[q]);
                A.val.push_back(v[q]);
                last = c[q];
            }
        }
        A.rp[static_cast<std::size_t>(i) + 1] = static_cast<int64_t>(A.col.size());
    }
    return A;
}

namespace {

constexpr double kPi = 3.14159265358979323846;

double segmentDistance(double px, double py, double ax, double ay, double bx, double by)
{
    const double vx = bx - ax, vy = by - ay;
    const double t = std::clamp(((px - ax) * vx + (py - ay) * vy) / (vx * vx + vy * vy), 0.0, 1.0);
    const double dx = px - (ax + t * vx), dy = py - (ay + t * vy);
    return std::sqrt(dx * dx + dy * dy);
}

double P1(double s) { return s * s - s; }

// Heterogeneous diffusivity of the convdiff/coupled families (MatrixGenerator::kappa,
// src/utils/matrix_generators.cpp:33-44): one kappa = 2 box, two channels, a lattice of inclusions.
double kappaField(double x, double y)
{
    if (x >= 0.10 && x <= 0.45 && y >= 0.12 && y <= 0.30) return 2.0;
    if (segmentDistance(x, y, 0.15, 0.62, 0.75, 0.70) < 0.04) return 6e-2;
    if (segmentDistance(x, y, 0.55, 0.62, 0.90, 0.35) < 0.03) return 6e-2;
    const double px = std::fmod(x, 0.2) - 0.1, py = std::fmod(y, 0.2) - 0.1;
    if (std::abs(px) < 0.035 && std::abs(py) < 0.035) return 6e-2;
    return 1.0;
}

void fillCellCoords(Problem& P, int nx, int ny, int nz, int nc)
{
    const int64_t ncell = static_cast<int64_t>(nx) * ny * nz;
    P.xyz.resize(static_cast<std::size_t>(ncell * nc) * 3);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int64_t cell = i + static_cast<int64_t>(nx) * (j + static_cast<int64_t>(ny) * k);
                for (int c = 0; c < nc; ++c) {
                    const std::size_t q = static_cast<std::size_t>(cell * nc + c) * 3;
                    P.xyz[q] = (i + 0.5) / nx;
                    P.xyz[q + 1] = (j + 0.5) / ny;
                    P.xyz[q + 2] = (k + 0.5) / nz;
                }
            }
}

// Finite-volume laplace / convdiff / coupled (src/utils/matrix_generators.cpp:57-193).
Problem fvFamily(const std::string& kind, int dim, const s2l_gen_opts& opt)
{
    const int nx = opt.nx, ny = opt.ny;
    const int nz = (dim == 2) ? 1 : opt.nz;
    const bool convection = (kind != "laplace");
    const bool heterogeneous = (kind != "laplace");
    const int nc = (kind == "coupled") ? std::max(1, opt.ncomp) : 1;
    const double nuBase = (kind == "laplace") ? 1.0 : opt.nu;
    const double convFactor = 1.0 + opt.pulse * std::sin(2.0 * kPi * opt.time / std::max(1, opt.pulse_period));

    const double hx = 1.0 / nx, hy = 1.0 / ny, hz = (dim == 3) ? 1.0 / nz : 1.0;
    const int64_t ncell = static_cast<int64_t>(nx) * ny * nz;
    const int64_t n = ncell * nc;
    auto cid = [&](int i, int j, int k) { return static_cast<int64_t>(i) + nx * (static_cast<int64_t>(j) + static_cast<int64_t>(ny) * k); };
    auto dof = [&](int64_t cell, int c) { return cell * nc + c; };
    auto nuOf = [&](int c) { return nuBase * (1.0 + c); };
    auto kappaCell = [&](int i, int j) { return heterogeneous ? kappaField((i + 0.5) * hx, (j + 0.5) * hy) : 1.0; };

    std::vector<int64_t> rows, cols;
    std::vector<double> vals;
    const std::size_t est = static_cast<std::size_t>(n) * (2 * dim + 1 + (nc > 1 ? 1 : 0)) * 2;
    rows.reserve(est);
    cols.reserve(est);
    vals.reserve(est);
    Problem P;
    P.rhs.assign(static_cast<std::size_t>(n), 0.0);
    auto add = [&](int64_t r, int64_t c, double v) {
        rows.push_back(r);
        cols.push_back(c);
        vals.push_back(v);
    };
    auto addFace = [&](int64_t Pc, int64_t Nc, double geom, double F) {
        F *= convFactor;
        for (int c = 0; c < nc; ++c) {
            const double D = nuOf(c) * geom;
            const int64_t p = dof(Pc, c), q = dof(Nc, c);
            add(p, p, D + std::max(F, 0.0));
            add(p, q, -D - std::max(-F, 0.0));
            add(q, q, D + std::max(-F, 0.0));
            add(q, p, -D - std::max(F, 0.0));
        }
    };
    auto addDirichlet = [&](int64_t Pc, double geom, double ub) {
        for (int c = 0; c < nc; ++c) {
            const double D = nuOf(c) * geom;
            add(dof(Pc, c), dof(Pc, c), D);
            P.rhs[static_cast<std::size_t>(dof(Pc, c))] += D * ub;
        }
    };
    const double gammaCoupling = nuBase * ((dim == 2) ? hy / hx : hy * hz / hx);

    for (int k = 0; k < nz; ++k) {
        const double z0 = k * hz, z1 = (k + 1) * hz;
        for (int j = 0; j < ny; ++j) {
            const double y0 = j * hy, y1 = (j + 1) * hy;
            for (int i = 0; i < nx; ++i) {
                const double x0 = i * hx, x1 = (i + 1) * hx;
                const int64_t Pc = cid(i, j, k);
                const double kP = kappaCell(i, j);
                if (i + 1 < nx) {
                    const double kN = kappaCell(i + 1, j), kf = 2.0 * kP * kN / (kP + kN);
                    const double area = hy * hz, xf = x1;
                    double F = 0.0;
                    if (convection)
                        F = (dim == 2) ? xf * (1.0 - xf) * (P1(y1) - P1(y0))
                                       : 2.0 * xf * (1.0 - xf) * (P1(y1) - P1(y0)) * 0.5 * (z1 * z1 - z0 * z0);
                    addFace(Pc, cid(i + 1, j, k), kf * area / hx, F);
                }
                if (j + 1 < ny) {
                    const double kN = kappaCell(i, j + 1), kf = 2.0 * kP * kN / (kP + kN);
                    const double area = hx * hz, yf = y1;
                    double F = 0.0;
                    if (convection) F = -yf * (1.0 - yf) * (P1(x1) - P1(x0)) * ((dim == 2) ? 1.0 : (z1 - z0));
                    addFace(Pc, cid(i, j + 1, k), kf * area / hy, F);
                }
                if (dim == 3 && k + 1 < nz) {
                    const double area = hx * hy, zf = z1;
                    double F = 0.0;
                    if (convection) F = -zf * (1.0 - zf) * (P1(x1) - P1(x0)) * (P1(y1) - P1(y0));
                    addFace(Pc, cid(i, j, k + 1), kP * area / hz, F);
                }
                if (dim == 2) {
                    if (j == 0) addDirichlet(Pc, kP * hx / (0.5 * hy), 0.0);
                    if (j == ny - 1) addDirichlet(Pc, kP * hx / (0.5 * hy), 1.0);
                } else {
                    if (k == 0) addDirichlet(Pc, kP * hx * hy / (0.5 * hz), 1.0);
                    if (k == nz - 1) addDirichlet(Pc, kP * hx * hy / (0.5 * hz), 0.0);
                }
                if (nc > 1)
                    for (int c = 0; c < nc; ++c) {
                        add(dof(Pc, c), dof(Pc, c), gammaCoupling);
                        add(dof(Pc, c), dof(Pc, (c + 1) % nc), -gammaCoupling);
                    }
            }
        }
    }
    P.A = fromCoo(n, rows, cols, vals);
    P.symmetric = (kind == "laplace");
    P.block_size = nc;
    P.dim = dim;
    fillCellCoords(P, nx, ny, nz, nc);
    return P;
}

// Gas fraction of rising smoothed bubbles (include/utils/bubble_field.hpp:43-158), with an
// optional per-bubble radius drawn uniformly in [rmin, rmax] cells (no reference counterpart).
class Bubbles {
public:
    Bubbles(int dim, int nx, int ny, int nz, const s2l_gen_opts& o) : dim_(dim)
    {
        n_[0] = nx;
        n_[1] = ny;
        n_[2] = (dim == 3) ? nz : 1;
        h_[0] = 1.0 / nx;
        h_[1] = 1.0 / ny;
        h_[2] = (dim == 3) ? 1.0 / nz : 1.0;
        const int nv = (dim == 3) ? nz : ny;
        const double hv = 1.0 / nv;
        w_ = std::max(h_[0], std::max(h_[1], (dim == 3) ? h_[2] : h_[1]));
        const bool perBubble = o.rmax_cells > 0.0;
        std::mt19937_64 gen(o.seed);
        std::uniform_real_distribution<double> u01(0.0, 1.0);
        const double rise = o.cfl * hv * o.time;
        const int vdir = (dim == 3) ? 2 : 1;
        for (int b = 0; b < o.nbubbles; ++b) {
            double R = o.radius;
            if (perBubble) R = (o.rmin_cells + (o.rmax_cells - o.rmin_cells) * u01(gen)) * w_;
            const double reach = R + 4.0 * w_;
            const double x0 = R + (1.0 - 2.0 * R) * u01(gen), y0 = R + (1.0 - 2.0 * R) * u01(gen), z0 = u01(gen);
            const double phase = 2.0 * kPi * u01(gen);
            const double wob = 0.5 * R * std::sin(2.0 * kPi * o.time / 40.0 + phase);
            const double xw = x0 + wob;
            const double xc = (xw < R) ? R : ((1.0 - R < xw) ? 1.0 - R : xw);
            double c[3];
            if (dim == 3) {
                c[0] = xc;
                c[1] = y0;
                c[2] = z0 + rise;
            } else {
                c[0] = xc;
                c[1] = z0 + rise;
                c[2] = 0.5;
            }
            c[vdir] -= std::floor(c[vdir]);
            for (int image = -1; image <= 1; ++image) {
                Image im;
                im.R = R;
                im.c[0] = c[0];
                im.c[1] = c[1];
                im.c[2] = c[2];
                im.c[vdir] += image;
                bool empty = false;
                for (int d = 0; d < dim; ++d) {
                    im.lo[d] = std::max(0, static_cast<int>(std::floor((im.c[d] - reach) / h_[d])));
                    im.hi[d] = std::min(n_[d] - 1, static_cast<int>(std::floor((im.c[d] + reach) / h_[d])));
                    empty = empty || im.lo[d] > im.hi[d];
                }
                if (dim == 2) im.lo[2] = im.hi[2] = 0;
                if (!empty) images_.push_back(im);
            }
        }
    }

    void fill(std::vector<double>& alpha) const
    {
        alpha.assign(static_cast<std::size_t>(n_[0]) * static_cast<std::size_t>(n_[1]) * static_cast<std::size_t>(n_[2]), 0.0);
        for (const Image& im : images_)
            for (int k = im.lo[2]; k <= im.hi[2]; ++k)
                for (int j = im.lo[1]; j <= im.hi[1]; ++j)
                    for (int i = im.lo[0]; i <= im.hi[0]; ++i) {
                        double& al = alpha[static_cast<std::size_t>(i) +
                                           static_cast<std::size_t>(n_[0]) *
                                               (static_cast<std::size_t>(j) + static_cast<std::size_t>(n_[1]) * static_cast<std::size_t>(k))];
                        al = std::max(al, profile(im, i, j, k));
                    }
    }

private:
    struct Image {
        double c[3];
        double R;
        int lo[3], hi[3];
    };
    int dim_;
    int n_[3];
    double h_[3];
    double w_ = 0.0;
    std::vector<Image> images_;

    double profile(const Image& im, int i, int j, int k) const
    {
        const double dx = (i + 0.5) * h_[0] - im.c[0], dy = (j + 0.5) * h_[1] - im.c[1];
        const double dz = (dim_ == 3) ? (k + 0.5) * h_[2] - im.c[2] : 0.0;
        const double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
        return 0.5 * (1.0 - std::tanh((dist - im.R) / w_));
    }
};

// Bubbly pressure matrix (src/utils/matrix_generators.cpp:195-306): beta = 1/rho per cell,
// face coefficient = mean of beta times area/distance, p = 0 on the top face.
Problem bubblesFamily(int dim, const s2l_gen_opts& opt)
{
    const int nx = opt.nx, ny = opt.ny, nz = (dim == 3) ? opt.nz : 1;
    if (opt.rho_ratio <= 0.0 || opt.radius <= 0.0 || opt.nbubbles < 0)
        throw std::invalid_argument("bubbles: rho_ratio and radius must be > 0, nbubbles >= 0");
    const double hx = 1.0 / nx, hy = 1.0 / ny, hz = (dim == 3) ? 1.0 / nz : 1.0;
    const int64_t ncell = static_cast<int64_t>(nx) * ny * nz;
    auto cid = [&](int i, int j, int k) { return static_cast<int64_t>(i) + nx * (static_cast<int64_t>(j) + static_cast<int64_t>(ny) * k); };
    std::vector<double> alpha;
    Bubbles(dim, nx, ny, nz, opt).fill(alpha);
    const double rho_g = 1.0 / opt.rho_ratio;
    std::vector<double> beta(static_cast<std::size_t>(ncell));
    for (std::size_t q = 0; q < beta.size(); ++q) beta[q] = 1.0 / (alpha[q] * rho_g + (1.0 - alpha[q]));

    Problem P;
    GlobalCsr& A = P.A;
    A.n = ncell;
    A.rp.assign(static_cast<std::size_t>(ncell) + 1, 0);
    A.col.reserve(static_cast<std::size_t>(ncell) * (2 * dim + 1));
    A.val.reserve(A.col.capacity());
    P.rhs.assign(static_cast<std::size_t>(ncell), 0.0);
    const double vol = hx * hy * hz;
    const double gx = hy * hz / hx, gy = hx * hz / hy, gz = hx * hy / hz;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int64_t Pc = cid(i, j, k);
                const double bP = beta[static_cast<std::size_t>(Pc)];
                double diag = 0.0;
                std::size_t diag_pos = 0;
                auto face = [&](int64_t Nc, double geom) {
                    const double coef = 0.5 * (bP + beta[static_cast<std::size_t>(Nc)]) * geom;
                    A.col.push_back(Nc);
                    A.val.push_back(-coef);
                    diag += coef;
                };
                if (dim == 3 && k > 0) face(cid(i, j, k - 1), gz);
                if (j > 0) face(cid(i, j - 1, k), gy);
                if (i > 0) face(cid(i - 1, j, k), gx);
                diag_pos = A.col.size();
                A.col.push_back(Pc);
                A.val.push_back(0.0);
                if (i + 1 < nx) face(cid(i + 1, j, k), gx);
                if (j + 1 < ny) face(cid(i, j + 1, k), gy);
                if (dim == 3 && k + 1 < nz) face(cid(i, j, k + 1), gz);
                const bool top = (dim == 3) ? (k == nz - 1) : (j == ny - 1);
                if (top) diag += bP * ((dim == 3) ? hx * hy / (0.5 * hz) : hx * hz / (0.5 * hy));
                A.val[diag_pos] = diag;
                A.rp[static_cast<std::size_t>(Pc) + 1] = static_cast<int64_t>(A.col.size());
                P.rhs[static_cast<std::size_t>(Pc)] = vol * alpha[static_cast<std::size_t>(Pc)];
            }
    if (opt.buoyancy) {
        std::vector<double> rho(static_cast<std::size_t>(ncell));
        for (std::size_t q = 0; q < rho.size(); ++q) rho[q] = alpha[q] * rho_g + (1.0 - alpha[q]);
        std::fill(P.rhs.begin(), P.rhs.end(), 0.0);
        const double hv = (dim == 3) ? hz : hy;
        auto flux = [&](int64_t Pc, int64_t Nc, double geom, double zf) {
            const auto p = static_cast<std::size_t>(Pc), q = static_cast<std::size_t>(Nc);
            const double F = 0.5 * (beta[p] + beta[q]) * (1.0 - zf) * (rho[q] - rho[p]) * geom;
            P.rhs[p] += F;
            P.rhs[q] -= F;
        };
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const int64_t Pc = cid(i, j, k);
                    const double zc = (((dim == 3) ? k : j) + 0.5) * hv;
                    if (i + 1 < nx) flux(Pc, cid(i + 1, j, k), gx, zc);
                    if (j + 1 < ny) flux(Pc, cid(i, j + 1, k), gy, (dim == 3) ? zc : (j + 1) * hy);
                    if (dim == 3 && k + 1 < nz) flux(Pc, cid(i, j, k + 1), gz, (k + 1) * hz);
                }
    }
    P.symmetric = true;
    P.block_size = 1;
    P.dim = dim;
    fillCellCoords(P, nx, ny, nz, 1);
    return P;
}

// Finite differences with Dirichlet on every face (eliminated): 5-point 2D with coefficients
// (ax, ay) = (aniso, 1) or 7-point 3D. Rows are written with sorted columns directly.
Problem fdFamily(int dim, int nx, int ny, int nz, double ax, double ay, double az)
{
    if (dim == 2) nz = 1;
    const int64_t n = static_cast<int64_t>(nx) * ny * nz;
    Problem P;
    GlobalCsr& A = P.A;
    A.n = n;
    A.rp.assign(static_cast<std::size_t>(n) + 1, 0);
    A.col.reserve(static_cast<std::size_t>(n) * (2 * dim + 1));
    A.val.reserve(A.col.capacity());
    auto cid = [&](int i, int j, int k) { return static_cast<int64_t>(i) + nx * (static_cast<int64_t>(j) + static_cast<int64_t>(ny) * k); };
    const double diag = 2.0 * ax + 2.0 * ay + (dim == 3 ? 2.0 * az : 0.0);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                auto push = [&](int64_t c, double v) {
                    A.col.push_back(c);
                    A.val.push_back(v);
                };
                if (dim == 3 && k > 0) push(cid(i, j, k - 1), -az);
                if (j > 0) push(cid(i, j - 1, k), -ay);
                if (i > 0) push(cid(i - 1, j, k), -ax);
                push(cid(i, j, k), diag);
                if (i + 1 < nx) push(cid(i + 1, j, k), -ax);
                if (j + 1 < ny) push(cid(i, j + 1, k), -ay);
                if (dim == 3 && k + 1 < nz) push(cid(i, j, k + 1), -az);
                A.rp[static_cast<std::size_t>(cid(i, j, k)) + 1] = static_cast<int64_t>(A.col.size());
            }
    P.rhs.assign(static_cast<std::size_t>(n), 1.0);
    P.symmetric = true;
    P.block_size = 1;
    P.dim = dim;
    fillCellCoords(P, nx, ny, nz, 1);
    return P;
}

// 7-point finite volumes on the unit cube, kappa = 1 with axis-aligned square channels of
// kappa = contrast drawn from the seed until their volume fraction reaches the target;
// harmonic face means; u = 0 on every face (half-cell distance).
Problem highContrast3d(const s2l_gen_opts& o)
{
    const int nx = o.nx, ny = o.ny, nz = o.nz;
    const int64_t n = static_cast<int64_t>(nx) * ny * nz;
    std::vector<double> kap(static_cast<std::size_t>(n), 1.0);
    auto cid = [&](int i, int j, int k) { return static_cast<int64_t>(i) + nx * (static_cast<int64_t>(j) + static_cast<int64_t>(ny) * k); };
    std::mt19937_64 gen(o.seed);
    const int wmax = std::max(1, std::min({nx, ny, nz}) / 24);
    std::uniform_int_distribution<int> axis(0, 2), width(1, std::max(1, wmax));
    std::uniform_real_distribution<double> u01(0.0, 1.0);
    const double target = (o.channel_fraction > 0.0) ? o.channel_fraction : 0.1;
    int64_t inside = 0;
    int guard = 0;
    while (static_cast<double>(inside) < target * static_cast<double>(n) && guard++ < 100000) {
        const int a = axis(gen), w = width(gen);
        const int dims[3] = {nx, ny, nz};
        const int b1 = (a + 1) % 3, b2 = (a + 2) % 3;
        const int p1 = static_cast<int>(u01(gen) * (dims[b1] - w + 1)), p2 = static_cast<int>(u01(gen) * (dims[b2] - w + 1));
        for (int t = 0; t < dims[a]; ++t)
            for (int s1 = p1; s1 < p1 + w; ++s1)
                for (int s2 = p2; s2 < p2 + w; ++s2) {
                    int ijk[3];
                    ijk[a] = t;
                    ijk[b1] = s1;
                    ijk[b2] = s2;
                    double& kv = kap[static_cast<std::size_t>(cid(ijk[0], ijk[1], ijk[2]))];
                    if (kv != o.contrast) {
                        kv = o.contrast;
                        ++inside;
                    }
                }
    }
    const double hx = 1.0 / nx, hy = 1.0 / ny, hz = 1.0 / nz;
    const double gx = hy * hz / hx, gy = hx * hz / hy, gz = hx * hy / hz;
    Problem P;
    GlobalCsr& A = P.A;
    A.n = n;
    A.rp.assign(static_cast<std::size_t>(n) + 1, 0);
    A.col.reserve(static_cast<std::size_t>(n) * 7);
    A.val.reserve(A.col.capacity());
    auto hm = [](double a, double b) { return 2.0 * a * b / (a + b); };
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int64_t Pc = cid(i, j, k);
                const double kP = kap[static_cast<std::size_t>(Pc)];
                double diag = 0.0;
                std::size_t dpos = 0;
                auto face = [&](int64_t Nc, double geom) {
                    const double c = hm(kP, kap[static_cast<std::size_t>(Nc)]) * geom;
                    A.col.push_back(Nc);
                    A.val.push_back(-c);
                    diag += c;
                };
                if (k > 0) face(cid(i, j, k - 1), gz);
                if (j > 0) face(cid(i, j - 1, k), gy);
                if (i > 0) face(cid(i - 1, j, k), gx);
                dpos = A.col.size();
                A.col.push_back(Pc);
                A.val.push_back(0.0);
                if (i + 1 < nx) face(cid(i + 1, j, k), gx);
                if (j + 1 < ny) face(cid(i, j + 1, k), gy);
                if (k + 1 < nz) face(cid(i, j, k + 1), gz);
                const int nb = (i == 0) + (i == nx - 1);
                const int mb = (j == 0) + (j == ny - 1);
                const int lb = (k == 0) + (k == nz - 1);
                diag += kP * (nb * 2.0 * gx + mb * 2.0 * gy + lb * 2.0 * gz);
                A.val[dpos] = diag;
                A.rp[static_cast<std::size_t>(Pc) + 1] = static_cast<int64_t>(A.col.size());
            }
    P.rhs.assign(static_cast<std::size_t>(n), 1.0);
    P.symmetric = true;
    P.block_size = 1;
    P.dim = 3;
    fillCellCoords(P, nx, ny, nz, 1);
    return P;
}

} // namespace

Problem generate(const std::string& name, const s2l_gen_opts& in)
{
    s2l_gen_opts o = in;
    if (o.ny <= 0) o.ny = o.nx;
    if (o.nz <= 0) o.nz = o.nx;
    if (o.nx < 1) throw std::invalid_argument("generator: grid sizes must be positive");
    if (name.size() < 3) throw std::invalid_argument("unknown generator '" + name + "'");
    const std::string suffix = name.substr(name.size() - 2), kind = name.substr(0, name.size() - 2);
    if (suffix != "2d" && suffix != "3d") throw std::invalid_argument("unknown generator '" + name + "'");
    const int dim = (suffix == "3d") ? 3 : 2;
    if (kind == "laplace" || kind == "convdiff" || kind == "coupled") return fvFamily(kind, dim, o);
    if (kind == "bubbles") return bubblesFamily(dim, o);
    if (kind == "poisson") return fdFamily(dim, o.nx, o.ny, o.nz, 1.0, 1.0, 1.0);
    if (name == "aniso2d") return fdFamily(2, o.nx, o.ny, 1, (o.aniso > 0.0 ? o.aniso : 1e-3), 1.0, 0.0);
    if (name == "highcontrast3d") return highContrast3d(o);
    throw std::invalid_argument("unknown generator '" + name + "'");
}

} // namespace s2l
