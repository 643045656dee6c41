// mmio.cpp -- Matrix Market input/output and the step-file sequence of the host side
// (src/partition/matrix_market_io.cpp, src/partition/domain_decomposer.cpp:80-130).
//
// Matrices: `matrix coordinate real|integer|pattern general|symmetric|skew-symmetric`, 1-based
// indices, symmetric storage expanded (the mirror entry negated for skew-symmetric), explicit zeros
// dropped, duplicates summed (fromCoo). Vectors: `matrix array real general` with one column, or
// coordinate with one column (entries added). Complex and Hermitian files are refused. The file
// is read whole and parsed with strtoll / strtod: multi-GB CFD matrices parse at disk speed.
#include "setup_internal.hpp"

#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace s2l {

namespace {

std::string readWhole(const std::string& path)
{
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open file '" + path + "'");
    std::string buf;
    char chunk[1 << 16];
    std::size_t got;
    while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf.append(chunk, got);
    const bool bad = std::ferror(f) != 0;
    std::fclose(f);
    if (bad) throw std::runtime_error("read error on '" + path + "'");
    return buf;
}

struct Header {
    std::string format, field, symmetry;
};

std::string lowerWord(const char*& p, const char* end)
{
    while (p < end && (*p == ' ' || *p == '\t')) ++p;
    std::string w;
    while (p < end && !std::isspace(static_cast<unsigned char>(*p))) w.push_back(static_cast<char>(std::tolower(static_cast<unsigned char>(*p++))));
    return w;
}

/// Parses the banner line and skips the comment / empty lines after it; returns the offset of
/// the size line.
std::size_t parseHeader(const std::string& buf, Header& h)
{
    const std::size_t eol = buf.find('\n');
    const char* p = buf.data();
    const char* e = buf.data() + (eol == std::string::npos ? buf.size() : eol);
    const std::string banner = lowerWord(p, e), object = lowerWord(p, e);
    h.format = lowerWord(p, e);
    h.field = lowerWord(p, e);
    h.symmetry = lowerWord(p, e);
    if (banner != "%%matrixmarket" || object != "matrix") throw std::runtime_error("not a Matrix Market matrix file");
    if (h.field == "complex" || h.symmetry == "hermitian")
        throw std::runtime_error("complex Matrix Market files are not supported (real arithmetic only)");
    std::size_t pos = eol == std::string::npos ? buf.size() : eol + 1;
    while (pos < buf.size() && (buf[pos] == '%' || buf[pos] == '\n' || buf[pos] == '\r')) {
        const std::size_t q = buf.find('\n', pos);
        pos = q == std::string::npos ? buf.size() : q + 1;
    }
    return pos;
}

/// Sequential number reader over a NUL-terminated buffer.
struct Scanner {
    const char* p;
    const char* what;
    long long integer()
    {
        char* end = nullptr;
        errno = 0;
        const long long v = std::strtoll(p, &end, 10);
        if (end == p) throw std::runtime_error(std::string(what) + ": unexpected end of file");
        if (errno == ERANGE) throw std::runtime_error(std::string(what) + ": index out of range");
        p = end;
        return v;
    }
    double real()
    {
        char* end = nullptr;
        const double v = std::strtod(p, &end);
        if (end == p) throw std::runtime_error(std::string(what) + ": unexpected end of file");
        p = end;
        return v;
    }
};

} // namespace

Problem readMatrixMarket(const std::string& path)
{
    const std::string buf = readWhole(path);
    Header h;
    Scanner sc{buf.c_str() + parseHeader(buf, h), "readMatrix"};
    if (h.format != "coordinate") throw std::runtime_error("readMatrix: only coordinate format is supported");
    const long long m = sc.integer(), ncols = sc.integer(), nz = sc.integer();
    if (m <= 0 || ncols <= 0 || nz < 0) throw std::runtime_error("readMatrix: invalid size line");
    if (m != ncols) throw std::runtime_error("the matrix must be square");
    const bool pattern = h.field == "pattern";
    const bool sym = h.symmetry == "symmetric", skew = h.symmetry == "skew-symmetric";
    std::vector<int64_t> rows, cols;
    std::vector<double> vals;
    const std::size_t cap = static_cast<std::size_t>(nz) * (sym || skew ? 2 : 1);
    rows.reserve(cap);
    cols.reserve(cap);
    vals.reserve(cap);
    for (long long k = 0; k < nz; ++k) {
        const long long i = sc.integer() - 1, j = sc.integer() - 1;
        const double v = pattern ? 1.0 : sc.real();
        if (i < 0 || i >= m || j < 0 || j >= m) throw std::out_of_range("readMatrix: index out of range");
        if (v == 0.0) continue;  // explicit zeros carry no information
        rows.push_back(i);
        cols.push_back(j);
        vals.push_back(v);
        if ((sym || skew) && i != j) {
            rows.push_back(j);
            cols.push_back(i);
            vals.push_back(skew ? -v : v);
        }
    }
    Problem P;
    P.A = fromCoo(m, rows, cols, vals);
    P.symmetric = sym;
    P.dim = 0;
    return P;
}

std::vector<double> readMatrixMarketVector(const std::string& path)
{
    const std::string buf = readWhole(path);
    Header h;
    Scanner sc{buf.c_str() + parseHeader(buf, h), "readVector"};
    const long long m = sc.integer(), ncols = sc.integer();
    if (m < 0 || ncols < 0) throw std::runtime_error("readVector: invalid size line");
    if (h.format == "array") {
        if (ncols != 1) throw std::runtime_error("readVector: array must have one column");
        std::vector<double> v(static_cast<std::size_t>(m));
        for (auto& x : v) x = sc.real();
        return v;
    }
    if (h.format != "coordinate") throw std::runtime_error("readVector: unknown format '" + h.format + "'");
    const long long nz = sc.integer();
    std::vector<double> v(static_cast<std::size_t>(m), 0.0);
    for (long long k = 0; k < nz; ++k) {
        const long long i = sc.integer() - 1;
        (void)sc.integer();
        const double x = h.field == "pattern" ? 1.0 : sc.real();
        if (i < 0 || i >= m) throw std::out_of_range("readVector: index out of range");
        v[static_cast<std::size_t>(i)] += x;
    }
    return v;
}

void writeMatrixMarket(const std::string& path, int64_t n, const int64_t* rp, const int64_t* col, const double* val)
{
    std::FILE* f = std::fopen(path.c_str(), "w");
    if (!f) throw std::runtime_error("cannot write '" + path + "'");
    std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n%lld %lld %lld\n", static_cast<long long>(n),
                 static_cast<long long>(n), static_cast<long long>(rp[n]));
    for (int64_t i = 0; i < n; ++i)
        for (int64_t k = rp[i]; k < rp[i + 1]; ++k)
            std::fprintf(f, "%lld %lld %.17g\n", static_cast<long long>(i + 1), static_cast<long long>(col[k] + 1), val[k]);
    const bool bad = std::fclose(f) != 0;
    if (bad) throw std::runtime_error("write error on '" + path + "'");
}

void writeMatrixMarketVector(const std::string& path, int64_t n, const double* v)
{
    std::FILE* f = std::fopen(path.c_str(), "w");
    if (!f) throw std::runtime_error("cannot write '" + path + "'");
    std::fprintf(f, "%%%%MatrixMarket matrix array real general\n%lld 1\n", static_cast<long long>(n));
    for (int64_t i = 0; i < n; ++i) std::fprintf(f, "%.17g\n", v[i]);
    const bool bad = std::fclose(f) != 0;
    if (bad) throw std::runtime_error("write error on '" + path + "'");
}

std::string stepFileName(const std::string& pattern, int step)
{
    // One %d, %Nd or %0Nd conversion takes the step (domain_decomposer.cpp:80-97); no '%' = one file.
    const std::size_t pct = pattern.find('%');
    if (pct == std::string::npos) return pattern;
    std::size_t q = pct + 1;
    const bool zero = q < pattern.size() && pattern[q] == '0';
    if (zero) ++q;
    std::size_t width = 0;
    for (; q < pattern.size() && std::isdigit(static_cast<unsigned char>(pattern[q])); ++q)
        width = 10 * width + static_cast<std::size_t>(pattern[q] - '0');
    if (q >= pattern.size() || pattern[q] != 'd' || pattern.find('%', q + 1) != std::string::npos)
        throw std::invalid_argument("--matrix: a step pattern must contain exactly one %d / %0Nd conversion");
    std::string num = std::to_string(step);
    if (num.size() < width) num.insert(0, width - num.size(), zero ? '0' : ' ');
    return pattern.substr(0, pct) + num + pattern.substr(q + 1);
}

uint64_t hashPattern(int64_t n, const int64_t* rp, const int64_t* col)
{
    // FNV-1a steps over the little-endian bytes of n, the row pointers and the column indices
    // (domain_decomposer.cpp:116-130): a new step must keep the pattern of the first one. The
    // start value is the reference's 1469598103934665603 (not FNV's published offset basis
    // 14695981039346656037), so that both sides compute the same hash of a pattern.
    uint64_t h = 1469598103934665603ULL;
    auto mix = [&h](uint64_t v) {
        for (int b = 0; b < 64; b += 8) {
            h ^= (v >> b) & 0xffu;
            h *= 1099511628211ULL;
        }
    };
    mix(static_cast<uint64_t>(n));
    for (int64_t i = 0; i <= n; ++i) mix(static_cast<uint64_t>(rp[i]));
    for (int64_t k = 0; k < rp[n]; ++k) mix(static_cast<uint64_t>(col[k]));
    return h;
}

} // namespace s2l
