// lsched.hpp -- level-synchronous schedule of the batched local triangular solves.
//
// Every sweep of every subdomain is cut into levels of its supernodal elimination tree: the forward
// sweep by height (leaves first), the backward sweep by depth (roots first). Level k of all
// subdomains together is one launch whose work items are wavefront-sized:
//   * a small supernode (w <= kLsWide) is one item: one wavefront pulls its rows, applies the
//     inverse of its diagonal block and pushes its products below;
//   * a wide supernode is split into row blocks of 64 rows, one item each; every item pulls all
//     rows of the supernode (the pulls are cheap next to the dense block), computes its 64 rows of
//     the product with the inverse, and pushes the partial products of its columns to slots of
//     their own (forward), or combines the partial dot products of the backward pull phase.
//   * the backward pull of a wide supernode (rows times the rows of its ancestors) is split over
//     (row block, chunk of 64 ancestor rows) items of a launch of its own before the level.
// Products pushed to a row go to slots of that row (one slot per pushed product, in a fixed order)
// and are summed by the row when it is pulled: no atomics, bitwise repeatable results.
#ifndef S2L_LSCHED_HPP
#define S2L_LSCHED_HPP

#include <cstdint>
#include <vector>

#include "pack.hpp"

namespace s2l {

constexpr int32_t kLsWide = 64;   ///< supernodes wider than a wavefront are split into row blocks
constexpr int32_t kLsRows = 64;   ///< rows of one row block
constexpr int32_t kLsKChunk = 64; ///< ancestor rows of one backward partial dot product
constexpr int32_t kLsMaxWidth = 4096;

/// One supernode of one sweep of one subdomain.
struct LsSn {
    int32_t sub;     ///< subdomain
    int32_t row0;    ///< first row of the subdomain in the concatenated local buffers (map offset)
    int32_t c0, w;   ///< columns [c0, c0 + w) of the sweep's factor (factor positions)
    int32_t h;       ///< rows below the diagonal block
    int32_t nrb;     ///< row blocks of kLsRows rows (1 for a small supernode)
    int32_t nkc;     ///< backward wide supernodes: chunks of kLsKChunk ancestor rows (0 otherwise)
    int32_t toff;    ///< forward: first push slot index (h * nrb pushes: index toff + j * h + k)
    int64_t poff;    ///< backward wide: first partial (nkc * w values)
    int64_t diag;    ///< inverse of the diagonal block: packed column-major triangle (w (w + 1) / 2 values)
    int64_t below;   ///< forward: column-major h x w (below[c * h + k]); backward: row-major (below[k * w + c])
    int64_t ridx;    ///< h row indices below the block (factor positions)
};

/// A work item: supernode and part (small: 0; wide forward / backward combine: row block;
/// backward partial dots: row block * nkc + chunk).
struct LsItem {
    int32_t sn;
    int32_t part;
};

enum LsKind : int32_t { kLsForward = 0, kLsBackPartial = 1, kLsBackward = 2 };

struct LsPlan {
    std::vector<LsSn> sn;
    std::vector<LsItem> items;
    std::vector<int32_t> launch;      ///< item range of every launch: [launch[q], launch[q + 1])
    std::vector<int32_t> launch_kind; ///< LsKind of every launch
    std::vector<int32_t> launch_w;    ///< widest supernode of every launch (LDS per wavefront)
    std::vector<double> diag, belowF, belowB;
    std::vector<int32_t> ridx;
    std::vector<int32_t> cptr;        ///< slot range of every concatenated local row (+ end)
    std::vector<int32_t> cidx;        ///< slot of every push
    int64_t slots = 0;                ///< slots per right-hand side
    int64_t partials = 0;             ///< backward partials per right-hand side
    int64_t entries = 0;              ///< stored factor entries read per solve (both sweeps)
    int64_t bytes_index = 0;          ///< index bytes read per solve
    int64_t bwd_below_rows = 0;       ///< rows below the backward blocks (ancestor values read)
    int32_t levels_fwd = 0, levels_bwd = 0;
};

/// Builds the schedule of all subdomains' factors (rows of subdomain i start at map_off[i] in the
/// concatenated local buffers; map_off has nsub + 1 entries).
LsPlan buildLsPlan(const std::vector<HostFactor>& fs, const std::vector<int32_t>& map_off);

} // namespace s2l

#endif
