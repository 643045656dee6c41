// pack.hpp -- host-side packing of the local factors into the layout of the batched
// supernodal triangular-solve kernel.
#ifndef S2L_PACK_HPP
#define S2L_PACK_HPP

#include <cstdint>
#include <vector>

namespace s2l {

/// One supernode of a sweep: columns [c0, c0 + w) of the factor, h rows below the diagonal block.
struct SnDesc {
    int32_t c0, w, h, toff; ///< toff: forward only, offset of its h products in the T buffer
    int64_t diag;           ///< offset of the w x w diagonal block (column-major)
    int64_t below;          ///< offset of the h x w block below (row-major)
    int64_t ridx;           ///< offset of the h row indices (factor positions)
};

/// Per-subdomain descriptor read by the kernel (one workgroup per subdomain).
struct SubDesc {
    int32_t n;        ///< local size n_i
    int32_t map_off;  ///< offset of the n_i entries of gmap / smap
    int32_t f_lev;    ///< offset of the forward level table (2 * nlev + 1 ints)
    int32_t f_nlev;
    int32_t b_lev;    ///< offset of the backward level table
    int32_t b_nlev;
    int32_t cptr_off; ///< offset of the n_i + 1 contribution row pointers
    int32_t n_owned;  ///< n_I (set by the context)
    int64_t t_off;    ///< offset of the subdomain's T buffer region
    int32_t begin;    ///< first owned global row (set by the context)
    int32_t pad;
};

/// Host copy of one uploaded factor (FP64 values).
struct HostFactor {
    int32_t n = 0;
    bool cholesky = true;
    std::vector<int32_t> perm;
    std::vector<int64_t> lcp, ucp;
    std::vector<int32_t> lr, ur;
    std::vector<double> lv, uv;
};

/// Packed factors of all subdomains (values still FP64).
struct Packed {
    std::vector<SnDesc> sn;           ///< forward and backward supernodes of all subdomains
    std::vector<int32_t> lev;         ///< level tables: [start0, split0, start1, split1, ..., start_nlev]
    std::vector<int32_t> lev_sn;      ///< supernode indices listed by level
    std::vector<double> diag, below;  ///< values
    std::vector<int32_t> ridx;        ///< below-row indices
    std::vector<int32_t> cptr, cidx;  ///< forward contributions: rows -> T indices
    std::vector<SubDesc> sub;
    int64_t t_total = 0;
    int64_t entries_fwd = 0, entries_bwd = 0;
    int32_t max_levels_fwd = 0, max_levels_bwd = 0, max_n = 0;
    int64_t bytes_index = 0;          ///< bytes of index data read per solve (ridx + contributions)
};

/// Packs factors; map_off[i] gives the offset of subdomain i in the gmap/smap arrays.
Packed packFactors(const std::vector<HostFactor>& f, const std::vector<int32_t>& map_off);

} // namespace s2l

#endif
