// eigen.cpp -- host-side tables of the device local eigensolver (see eigen.hpp).
#include "eigen.hpp"

#include <algorithm>
#include <stdexcept>

namespace s2l {

EigPlan buildEigPlan(const std::vector<int32_t>& local_ptr, const std::vector<int32_t>& gid,
                     const std::vector<int32_t>& ownership, const std::vector<int32_t>& rp, const std::vector<int32_t>& col,
                     const std::vector<double>& pou, bool boolean)
{
    EigPlan E;
    const int nsub = static_cast<int>(local_ptr.size()) - 1;
    const int32_t n = static_cast<int32_t>(rp.size()) - 1;
    const int32_t nl = local_ptr.back();
    std::vector<int32_t> g2l(static_cast<std::size_t>(n), -1);
    E.chunk_ptr.assign(static_cast<std::size_t>(nsub) + 1, 0);
    E.lrp.assign(static_cast<std::size_t>(nl) + 1, 0);
    E.orp.assign(static_cast<std::size_t>(nl) + 1, 0);
    E.ghost.assign(static_cast<std::size_t>(nl), 0);
    E.D.assign(static_cast<std::size_t>(nl), 0.0);
    for (int i = 0; i < nsub; ++i) {
        const int32_t l0 = local_ptr[static_cast<std::size_t>(i)], l1 = local_ptr[static_cast<std::size_t>(i) + 1];
        const int32_t nI = ownership[static_cast<std::size_t>(i) + 1] - ownership[static_cast<std::size_t>(i)];
        E.chunk_ptr[static_cast<std::size_t>(i)] = static_cast<int32_t>(E.chunk.size() / 3);
        for (int32_t r = l0; r < l1; r += kEigChunk) {
            E.chunk.push_back(i);
            E.chunk.push_back(r);
            E.chunk.push_back(std::min(l1, r + kEigChunk));
        }
        for (int32_t l = l0; l < l1; ++l) g2l[static_cast<std::size_t>(gid[static_cast<std::size_t>(l)])] = l;
        for (int32_t l = l0; l < l1; ++l) {
            const bool gh = l - l0 >= nI;
            E.ghost[static_cast<std::size_t>(l)] = gh ? 1 : 0;
            E.D[static_cast<std::size_t>(l)] = boolean ? (gh ? 0.0 : 1.0) : pou[static_cast<std::size_t>(l)];
            const int32_t g = gid[static_cast<std::size_t>(l)];
            for (int32_t q = rp[static_cast<std::size_t>(g)]; q < rp[static_cast<std::size_t>(g) + 1]; ++q) {
                const int32_t lc = g2l[static_cast<std::size_t>(col[static_cast<std::size_t>(q)])];
                if (lc >= 0) {
                    E.lcol.push_back(lc);
                    E.lsrc.push_back(q);
                } else if (gh) {
                    E.osrc.push_back(q);
                }
            }
            E.lrp[static_cast<std::size_t>(l) + 1] = static_cast<int32_t>(E.lcol.size());
            E.orp[static_cast<std::size_t>(l) + 1] = static_cast<int32_t>(E.osrc.size());
        }
        for (int32_t l = l0; l < l1; ++l) g2l[static_cast<std::size_t>(gid[static_cast<std::size_t>(l)])] = -1;
    }
    E.chunk_ptr[static_cast<std::size_t>(nsub)] = static_cast<int32_t>(E.chunk.size() / 3);
    if (E.lcol.size() > static_cast<std::size_t>(INT32_MAX)) throw std::invalid_argument("eigensolver plan: too many entries");
    return E;
}

} // namespace s2l
