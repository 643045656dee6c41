// lsched.hpp -- level-synchronous schedule of the batched local triangular solves.
//
// Every sweep of every subdomain is cut into levels of its supernodal elimination tree: the forward
// sweep by height (leaves first), the backward sweep by depth (roots first). A level of all
// subdomains together is processed by launches over all of them:
//   * forward: a pull launch for the rows of its large supernodes (every row subtracts the
//     products its descendants pushed to its slots, four lanes per row) and their unit launch;
//     the unit launch of its small supernodes, which pull their own rows;
//   * backward: a launch of partial dot products for the large supernodes (rows times 64
//     ancestor values each) and their unit launch; the unit launch of its small supernodes.
// Within a level the large and the small supernodes are independent (two streams).
// A unit is the work of one workgroup: one row block (64 rows) of a large supernode, four
// wavefronts splitting its columns and the rows it pushes to, or one small supernode and one
// wavefront (launched apart from the large ones). Products pushed to a row go to slots of that row (one
// slot per pushed product, in a fixed order) and are summed when the row is pulled: no atomics,
// bitwise repeatable results.
#ifndef S2L_LSCHED_HPP
#define S2L_LSCHED_HPP

#include <cstdint>
#include <vector>

#include "pack.hpp"

namespace s2l {

constexpr int32_t kLsRows = 64;        ///< rows of one row block (a wavefront)
constexpr int32_t kLsWaves = 4;        ///< wavefronts of a unit's workgroup
constexpr int32_t kLsKChunk = 64;      ///< ancestor rows of one backward partial dot product
constexpr int32_t kLsPullLanes = 4;    ///< lanes per row of the forward pull
constexpr int32_t kLsMaxWidth = 4096;  ///< widest supernode handled
constexpr int64_t kLsLargeWork = 2048; ///< w * h above which a supernode of width <= 64 is large

/// One supernode of one sweep of one subdomain.
struct LsSn {
    int32_t sub;     ///< subdomain
    int32_t row0;    ///< first row of the subdomain in the concatenated local buffers (map offset)
    int32_t c0, w;   ///< columns [c0, c0 + w) of the sweep's factor (factor positions)
    int32_t h;       ///< rows below the diagonal block
    int32_t nrb;     ///< row blocks of kLsRows rows
    int32_t nkc;     ///< backward large supernodes: chunks of kLsKChunk ancestor rows (0 otherwise)
    int32_t toff;    ///< forward: first push slot index (h * nrb pushes: index toff + j * h + k)
    int64_t poff;    ///< backward large: first partial (nkc * w values)
    int64_t diag;    ///< inverse of the diagonal block: packed column-major triangle (w (w + 1) / 2 values)
    int64_t below;   ///< forward: column-major h x w (below[c * h + k]); backward: row-major (below[k * w + c])
    int64_t ridx;    ///< h row indices below the block (factor positions)
};

/// The work of one workgroup: large != 0: row block `part` of supernode sn[0] (four wavefronts);
/// else the small supernode sn[0] (one wavefront). sn[1..3] are -1.
struct LsUnit {
    int32_t large;
    int32_t part;
    int32_t sn[kLsWaves];
};

/// A backward partial item: supernode and part = row block * nkc + chunk.
struct LsItem {
    int32_t sn;
    int32_t part;
};

/// Launch kinds: the forward pull; units of large supernodes (a workgroup of four wavefronts per
/// row block) and of small ones (a workgroup of one wavefront per supernode) of both sweeps; the
/// backward partial products.
enum LsKind : int32_t {
    kLsPull = 0,
    kLsForward = 1,
    kLsForwardSmall = 2,
    kLsBackPartial = 3,
    kLsBackward = 4,
    kLsBackwardSmall = 5
};

struct LsPlan {
    std::vector<LsSn> sn;
    std::vector<LsUnit> units;
    std::vector<LsItem> items;         ///< backward partial items
    std::vector<int32_t> pull_rows;    ///< rows (concatenated local) of every forward level
    std::vector<int32_t> launch_kind;  ///< LsKind of every launch
    std::vector<int32_t> launch_first; ///< first row / unit / item of every launch
    std::vector<int32_t> launch_count; ///< rows / units / items of every launch
    std::vector<int32_t> launch_w;     ///< widest supernode of every launch (LDS of a unit)
    std::vector<int32_t> launch_level; ///< level of every launch (forward levels, then backward)
    std::vector<double> diag, belowF, belowB;
    std::vector<int32_t> ridx;
    std::vector<int32_t> cptr;         ///< slot range of every concatenated local row (+ end)
    std::vector<int32_t> cidx;         ///< slot of every push
    int64_t slots = 0;                 ///< slots per right-hand side
    int64_t partials = 0;              ///< backward partials per right-hand side
    int64_t entries = 0;               ///< stored factor entries read per solve (both sweeps)
    int64_t bytes_index = 0;           ///< index bytes read per solve
    int64_t bwd_below_rows = 0;        ///< rows below the backward blocks (ancestor values read)
    int32_t levels_fwd = 0, levels_bwd = 0;
};

/// Builds the schedule of all subdomains' factors (rows of subdomain i start at map_off[i] in the
/// concatenated local buffers; map_off has nsub + 1 entries).
LsPlan buildLsPlan(const std::vector<HostFactor>& fs, const std::vector<int32_t>& map_off);

} // namespace s2l

#endif
