// api.hip -- device context, uploads, preconditioner application and device Krylov loops behind
// the C-ABI of include/s2l_hip.h.
#include "s2l_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <functional>
#include <future>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "device.hpp"
#include "refactor.hpp"
#include "eigen.hpp"

#include <rocsolver/rocsolver.h>

namespace s2l {
template <typename T, typename TF> const void* localSolveKernel(int threads);
// two right-hand sides per workgroup: a float pair (FP32 arithmetic) or a double pair (FP64
// arithmetic of FP32 entries) per row
template <typename T> constexpr bool kIsPair = std::is_same_v<T, float2> || std::is_same_v<T, double2>;
}

namespace {

using namespace s2l;

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct ApiError : std::runtime_error {
    int code;
    ApiError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHECK(x)                                                                                         \
    do {                                                                                                    \
        hipError_t e_ = (x);                                                                                \
        if (e_ != hipSuccess) throw HipError(std::string(#x) + ": " + hipGetErrorString(e_));               \
    } while (0)

/// Byte counter of the context whose API call runs on this thread (set by guarded()).
thread_local int64_t* t_device_bytes = nullptr;

/// Device buffer (hipMalloc), grown on demand. Its bytes count toward the device footprint of
/// the context that allocated it.
template <typename T> struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    int64_t* acct = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) {
            (void)hipFree(p);
            if (acct) *acct -= static_cast<int64_t>(n * sizeof(T));
        }
        p = nullptr;
        n = 0;
        acct = nullptr;
    }
    void reserve(size_t cnt)
    {
        if (cnt <= n && p) return;
        release();
        const size_t bytes = std::max<size_t>(cnt, 1) * sizeof(T);
        HIPCHECK(hipMalloc(reinterpret_cast<void**>(&p), bytes));
        n = std::max<size_t>(cnt, 1);
        acct = t_device_bytes;
        if (acct) *acct += static_cast<int64_t>(bytes);
    }
    void upload(const T* h, size_t cnt, hipStream_t s)
    {
        reserve(cnt);
        if (cnt) HIPCHECK(hipMemcpyAsync(p, h, cnt * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
};

/// Pinned host staging buffer.
template <typename T> struct PinBuf {
    T* p = nullptr;
    size_t n = 0;
    ~PinBuf()
    {
        if (p) (void)hipHostFree(p);
    }
    void reserve(size_t cnt)
    {
        if (cnt <= n && p) return;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        HIPCHECK(hipHostMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(cnt, 1) * sizeof(T), hipHostMallocDefault));
        n = std::max<size_t>(cnt, 1);
    }
};

// ---- dense FP64 factorizations of A_00 on the host (row-major)
bool denseCholesky(int n, std::vector<double>& A)
{
    for (int j = 0; j < n; ++j) {
        double* lj = A.data() + static_cast<size_t>(j) * n;
        double d = lj[j];
        for (int k = 0; k < j; ++k) d -= lj[k] * lj[k];
        if (!(d > 0.0) || !std::isfinite(d)) return false;
        const double ljj = std::sqrt(d);
        lj[j] = ljj;
        for (int i = j + 1; i < n; ++i) {
            double* li = A.data() + static_cast<size_t>(i) * n;
            double s = li[j];
            for (int k = 0; k < j; ++k) s -= li[k] * lj[k];
            li[j] = s / ljj;
        }
    }
    return true;
}

// Inverse of the lower triangular part of L (unit diagonal if @p unit), row by row.
std::vector<double> lowerInverse(int n, const std::vector<double>& L, bool unit)
{
    std::vector<double> X(static_cast<size_t>(n) * n, 0.0);
    for (int i = 0; i < n; ++i) {
        double* xi = X.data() + static_cast<size_t>(i) * n;
        const double* li = L.data() + static_cast<size_t>(i) * n;
        for (int k = 0; k < i; ++k) {
            const double f = li[k];
            if (f == 0.0) continue;
            const double* xk = X.data() + static_cast<size_t>(k) * n;
            for (int j = 0; j <= k; ++j) xi[j] -= f * xk[j];
        }
        xi[i] += 1.0;
        const double d = unit ? 1.0 : li[i];
        for (int j = 0; j <= i; ++j) xi[j] /= d;
    }
    return X;
}

bool denseLu(int n, std::vector<double>& A, std::vector<int32_t>& p)
{
    p.resize(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) p[static_cast<size_t>(i)] = i;
    for (int j = 0; j < n; ++j) {
        int piv = j;
        double amax = std::fabs(A[static_cast<size_t>(j) * n + j]);
        for (int i = j + 1; i < n; ++i)
            if (std::fabs(A[static_cast<size_t>(i) * n + j]) > amax) {
                amax = std::fabs(A[static_cast<size_t>(i) * n + j]);
                piv = i;
            }
        if (amax == 0.0 || !std::isfinite(amax)) return false;
        if (piv != j) {
            std::swap_ranges(A.begin() + static_cast<ptrdiff_t>(j) * n, A.begin() + static_cast<ptrdiff_t>(j + 1) * n,
                             A.begin() + static_cast<ptrdiff_t>(piv) * n);
            std::swap(p[static_cast<size_t>(j)], p[static_cast<size_t>(piv)]);
        }
        const double d = A[static_cast<size_t>(j) * n + j];
        const double* aj = A.data() + static_cast<size_t>(j) * n;
        for (int i = j + 1; i < n; ++i) {
            double* ai = A.data() + static_cast<size_t>(i) * n;
            const double l = ai[j] / d;
            ai[j] = l;
            if (l != 0.0)
                for (int c = j + 1; c < n; ++c) ai[c] -= l * aj[c];
        }
    }
    return true;
}

} // namespace

namespace {
struct GmresRun;
}

// Scratch of s2l_apply_noalloc: one array per GPU, held until the process ends. Host code uses the
// device address that hipGetSymbolAddress returns for the current device (s2l_ctx::scratch),
// never the array's name.
constexpr size_t kScratchBytes = size_t(128) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr size_t kMaxGraphs = 160;  // cached apply graphs per context (GMRES(m): m inputs, m outputs if flexible)

struct s2l_ctx {
    int device = 0;
    int64_t device_bytes = 0;   ///< bytes of the context's device buffers (DevBuf)
    hipStream_t stream = nullptr;
    // second stream of the level-synchronous local solve: the large supernodes of a level run on
    // it beside the small ones (fork and join by events on the caller's stream)
    hipStream_t aux = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool ls_fork = false;  // measured: the cross-stream fork and join cost as much as the overlap gains
    std::string err;
    size_t max_lds = 0;

    // operator
    int32_t n = 0, nsub = 0;
    int64_t nnz = 0;
    std::vector<int32_t> ownership;
    std::vector<int32_t> rp_h, col_h;  ///< uploaded pattern (columns in the owners' local order)
    std::vector<int32_t> inpos_h;      ///< stored position -> position in the uploaded arrays
    DevBuf<int32_t> rp, col;
    DevBuf<double> val;
    bool system_ready = false;

    // subdomains
    std::vector<int32_t> local_ptr, gid_h;
    std::vector<double> pou_h;
    bool boolean = true;
    int one_level = 0;
    DevBuf<int32_t> rg_ptr, rg_idx;
    DevBuf<double> pou_d;
    bool subs_ready = false;

    // local factors
    std::vector<HostFactor> hf;
    int precision = 0;
    bool local_ready = false;
    int launch_mode = 0;
    // level-synchronous schedule (launch mode 2)
    LsPlan ls;
    DevBuf<LsSn> ls_sn;
    DevBuf<LsUnit> ls_units;
    DevBuf<LsItem> ls_items;
    DevBuf<unsigned char> ls_diag, ls_bf, ls_bb, ls_slots, ls_part, ls_b, ls_yf, ls_xb;
    DevBuf<int32_t> ls_ridx, ls_cptr, ls_cidx, ls_owned, ls_begin, ls_pull;
    // task-queue schedule (launch modes 0 and 1); needs a subdomain vector in the LDS of a workgroup
    bool queue_ok = false;
    Packed pk;
    DevBuf<SubDesc> sub_d;
    DevBuf<int32_t> levw_d, levh_d, ridx_d, cptr_d, cend_d, cidx_d, gmap_d, lperm_d;
    DevBuf<SnDesc> levdesc_d;
    DevBuf<int32_t> levrow_d, levcb_d, levce_d;
    DevBuf<StepDesc> steps_d;
    size_t y_off = 0, st_off = 0, step_off = 0;
    int32_t st_cap = 0;  // supernode descriptors of the staging area (the most of a chunk)
    size_t slot_off = 0, slot_off2 = 0;  // LDS slots of a task without child tasks (one rhs / pair)
    int32_t slot_cap = 0, slot_cap2 = 0;
    // two right-hand sides per workgroup (a float2 or double2 vector in LDS): its LDS layout
    size_t y_off2 = 0, st_off2 = 0, step_off2 = 0, lds_bytes2 = 0;
    bool pair_ok = false;
    DevBuf<float> diag_f, below_f;
    DevBuf<double> diag_dd, below_dd;
    DevBuf<unsigned char> tbuf, xgbuf;
    DevBuf<TaskDesc> task_d;
    DevBuf<int32_t> tchild_d, trows_d, ext_d, queue_d, qsubq_d;
    std::vector<int32_t> qsub_h;
    // dense tops in a launch of their own: per skeleton row (factor position, global row, local
    // index, first slot), its end of slots, each task's offset in the table (-1: not a dense top);
    // per depth group of the queue whether all its entries are dense tops; LDS of the widest b_S
    DevBuf<int4> dtrow_d;
    DevBuf<int32_t> dtce_d, dtb_d;
    std::vector<char> qgroup_dense;
    int32_t dt_maxS = 0;
    int tasks_per_sub = 0;  // 0: one subtree task per compute unit (compute units / subdomains)
    int dense_top_max = kDenseTopMax;  // widest skeleton replaced by a dense top (0: none)
    int local_threads_req = kLocalThreads; // threads of a local-solve workgroup asked for: 1024, or 512 (two per compute unit)
    int local_threads = kLocalThreads;     // in use (512 falls back to 1024 when half the LDS does not hold the layout)
    int cus = 0;
    int64_t local_total = 0;
    size_t lds_bytes = 0;

    // device refactorization (s2l_refactorize): its plan, device tables and work arrays
    bool rf_ready = false;
    bool ls_built = false;  ///< the level-synchronous layout exists (built for launch mode 2 only)
    bool ls_stale = false;  ///< the level-synchronous layout holds older factor values
    RfPlan rf;
    bool rf_planned = false;            ///< rf holds the plan of the current factors and pattern
    std::future<RfPlan> rf_pending;     ///< the plan, built on a host thread after finalize_local
    std::string rf_plan_err;            ///< why that plan could not be built
    /// Waits for the background plan and keeps it (or its error). Every call that changes what the
    /// plan reads (pattern, subdomains, factors, packing) calls this first.
    void rfJoin()
    {
        if (!rf_pending.valid()) return;
        try {
            rf = rf_pending.get();
            rf_planned = true;
        } catch (const std::exception& e) {
            rf_plan_err = e.what();
            rf_planned = false;
        }
    }
    void rfForget()
    {
        rfJoin();
        rf_ready = rf_planned = false;
        rf_plan_err.clear();
        eg_ready = false;
    }

    // device local eigensolver (s2l_local_eigensolve): its tables, shifted factors and Krylov bases
    bool eg_ready = false;
    EigPlan eg;
    DevBuf<int32_t> eg_chunk, eg_cptr, eg_lrp, eg_lcol, eg_lsrc, eg_orp, eg_osrc, eg_ghost, eg_lptr, eg_done, eg_conv,
        eg_live, eg_info;
    DevBuf<double> eg_D, eg_adj, eg_rowsum, eg_sigma, eg_diag, eg_below, eg_V, eg_AV, eg_H, eg_Hw, eg_theta, eg_E,
        eg_part, eg_hb, eg_hb2, eg_beta, eg_u, eg_w, eg_Aw, eg_lam;
    DevBuf<RfSn> rf_sn;
    DevBuf<RfUpd> rf_upd;
    DevBuf<int32_t> rf_rel, rf_src, rf_order, rf_status, rf_inpos;
    DevBuf<int64_t> rf_pos, rf_invoff;
    DevBuf<double> rf_work, rf_inv, rf_raw, rf_E, rf_Y, rf_X, rf_AX, rf_C, rf_part, rf_loc;
    rocblas_handle rf_blas = nullptr;
    // dense tops of the refactorized factors: their tables, L_SS / S^-1 scratch
    DevBuf<DenseTop> dt_d;
    DevBuf<int32_t> dt_rf_d, dt_q_d;
    DevBuf<int64_t> dt_off_d;
    DevBuf<double> dt_L, dt_C;
    DevBuf<rocblas_int> dt_info;
    double rf_ms[4] = {};   ///< last refactorization: values, local factors, A_00, total [ms]

    // coarse
    int32_t n0 = 0;
    int correction = 0;
    bool coarse_ready = false;
    CoarseArgs ca{};
    std::vector<int64_t> off_h;
    DevBuf<int64_t> off_d, voff_d;
    DevBuf<int32_t> vrows_d, vrowptr_d, vrowg_d, vbase_d, blksub_d, blkrow_d, cperm_d;
    DevBuf<double> V_d, T1_d, T2_d;  // inverted factors of A00, row-major with leading dimension tri_ld
    DevBuf<double> Einv_d;           // A00^-1 = T2 T1 (the rows of the permuted columns for LU), same layout
    int32_t tri_ld = 0;
    bool coarse_perm = false;

    // work: the apply's scratch buffers (grown by ensureWork) and the pointers the kernels use,
    // which s2l_apply_noalloc points at the static scratch array for one call
    DevBuf<double> x0, w, loc, cvec, tvec, yvec, cpart;
    int work_nrhs = 0;
    struct Work {
        double *x0, *w, *loc, *cvec, *tvec, *yvec, *cpart;
        unsigned char *tbuf, *xgbuf, *ls_slots, *ls_part, *ls_b, *ls_yf, *ls_xb;
    };
    Work work{};
    void* scratch = nullptr;  ///< this device's copy of the static scratch array (s2l_apply_noalloc)

    // HIP graphs of one apply + SpMV in the Krylov loops, one per (input, output) pair, captured on
    // first use; every upload, schedule change or work-buffer growth invalidates them (gen)
    struct ApplyGraph {
        hipGraphExec_t exec = nullptr;
        uint64_t gen = 0;
        const double* in = nullptr;
        double* z = nullptr;
        double* w = nullptr;
        const int32_t* skip = nullptr;
        int nrhs = 0;
        hipStream_t s = nullptr;
    };
    std::vector<ApplyGraph> graphs;
    uint64_t gen = 1;
    bool use_graphs = false;  // measured: replays end within noise of eager launches (see s2l_set_graphs)
    void invalidateGraphs()
    {
        ++gen;
        for (auto& g : graphs)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
        graphs.clear();
    }

    // per-phase timing with HIP events (s2l_phase_timing_enable): pairs recorded on the work's
    // stream, summed when read
    struct PhaseMark {
        int id;
        hipEvent_t e0, e1;
    };
    bool phase_on = false;
    std::vector<PhaseMark> phase_pending;
    double phase_ms[S2L_NPHASES] = {};
    int64_t phase_n[S2L_NPHASES] = {};

    // timing
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pending;
    int64_t launches = 0;
    double ms_total = 0.0;

    // Krylov: the GMRES run of s2l_krylov_solve (kept for the next solve of the same shape) and
    // the reduction buffers of BiCGStab
    std::shared_ptr<GmresRun> gm_cache;
    DevBuf<double> kpart, kred;
    PinBuf<double> pred;
    // BiCGStab
    DevBuf<double> bR, bRh, bP, bV, bY, bS, bZ, bT, bcoef, bones, bhist;
    DevBuf<int32_t> bmask, blive;
    DevBuf<BiState> bst;
    PinBuf<int32_t> plive;
    PinBuf<BiState> pst;
    PinBuf<double> phist;

    void ensureWork(int nrhs)
    {
        if (nrhs <= work_nrhs) return;
        invalidateGraphs();
        x0.reserve(static_cast<size_t>(n) * nrhs);
        w.reserve(static_cast<size_t>(n) * nrhs);
        loc.reserve(static_cast<size_t>(local_total) * nrhs);
        cvec.reserve(static_cast<size_t>(std::max(n0, 1)) * nrhs);
        tvec.reserve(static_cast<size_t>(std::max(n0, 1)) * nrhs);
        yvec.reserve(static_cast<size_t>(std::max(n0, 1)) * nrhs);
        cpart.reserve(static_cast<size_t>(std::max(nsub, 1)) * std::max(ca.rchunks, 1) * std::max(ca.mmax, 1) * nrhs);
        const size_t esz = precision == 1 ? sizeof(float) : sizeof(double);  // arithmetic type
        const int nr = (nrhs + 1) / 2 * 2;  // FP32 pairs of right-hand sides: an even count
        tbuf.reserve(static_cast<size_t>(std::max<int64_t>(pk.t_total, 1)) * nr * esz);
        xgbuf.reserve(static_cast<size_t>(std::max<int64_t>(local_total, 1)) * nr * esz);
        ls_slots.reserve(static_cast<size_t>(std::max<int64_t>(ls.slots, 1)) * nrhs * esz);
        ls_part.reserve(static_cast<size_t>(std::max<int64_t>(ls.partials, 1)) * nrhs * esz);
        ls_b.reserve(static_cast<size_t>(std::max<int64_t>(local_total, 1)) * nrhs * esz);
        ls_yf.reserve(static_cast<size_t>(std::max<int64_t>(local_total, 1)) * nrhs * esz);
        ls_xb.reserve(static_cast<size_t>(std::max<int64_t>(local_total, 1)) * nrhs * esz);
        work_nrhs = nrhs;
        work = Work{x0.p, w.p, loc.p, cvec.p, tvec.p, yvec.p, cpart.p, tbuf.p, xgbuf.p, ls_slots.p, ls_part.p, ls_b.p, ls_yf.p, ls_xb.p};
    }
};

namespace {

thread_local std::string g_err;

template <typename F> int guarded(s2l_ctx* ctx, F&& f)
{
    struct Scope {
        int64_t* prev;
        ~Scope() { t_device_bytes = prev; }
    } scope{t_device_bytes};
    t_device_bytes = ctx ? &ctx->device_bytes : nullptr;
    try {
        f();
        return S2L_OK;
    } catch (const ApiError& e) {
        (ctx ? ctx->err : g_err) = e.what();
        return e.code;
    } catch (const HipError& e) {
        (ctx ? ctx->err : g_err) = e.what();
        return S2L_EHIP;
    } catch (const std::exception& e) {
        (ctx ? ctx->err : g_err) = e.what();
        return S2L_EINVAL;
    }
}

void require(bool c, int code, const std::string& m)
{
    if (!c) throw ApiError(code, m);
}

hipStream_t pick(s2l_ctx* ctx, void* s) { return s ? static_cast<hipStream_t>(s) : ctx->stream; }

// ---- local solve launch: the task-queue kernel, one launch per depth group over every
//      subdomain's tasks (mode 0) or one single-workgroup launch per subdomain over its slice of
//      the queue (mode 1)
template <typename T, typename TF>
LocalArgs<T, TF> localArgs(s2l_ctx* c, const double* rhs, const double* x0, double* out, int mode, int nrhs, const int32_t* skip)
{
    LocalArgs<T, TF> a{};
    a.sub = c->sub_d.p;
    a.lev_wpre = c->levw_d.p;
    a.lev_hpre = c->levh_d.p;
    const bool pair = kIsPair<T>;
    a.y_off = static_cast<int32_t>(pair ? c->y_off2 : c->y_off);
    a.st_off = static_cast<int32_t>(pair ? c->st_off2 : c->st_off);
    a.step_off = static_cast<int32_t>(pair ? c->step_off2 : c->step_off);
    a.st_cap = c->st_cap;
    a.lev_desc = c->levdesc_d.p;
    a.lev_row = c->levrow_d.p;
    a.lev_cb = c->levcb_d.p;
    a.lev_ce = c->levce_d.p;
    a.steps = c->steps_d.p;
    if constexpr (std::is_same_v<TF, float>) {
        a.diag = c->diag_f.p;
        a.below = c->below_f.p;
    } else {
        a.diag = c->diag_dd.p;
        a.below = c->below_dd.p;
    }
    a.ridx = c->ridx_d.p;
    a.cptr = c->cptr_d.p;
    a.cend = c->cend_d.p;
    a.slot_off = static_cast<int32_t>(pair ? c->slot_off2 : c->slot_off);
    a.slot_cap = pair ? c->slot_cap2 : c->slot_cap;
    a.cidx = c->cidx_d.p;
    a.tbuf = reinterpret_cast<T*>(c->work.tbuf);
    a.gmap = c->gmap_d.p;
    a.lperm = c->lperm_d.p;
    a.rhs = rhs;
    a.x0 = x0;
    a.out = out;
    a.pou = (mode == 1 && c->one_level == 0 && !c->boolean) ? c->pou_d.p : nullptr;
    a.skip = skip;
    a.n_global = c->n;
    a.n_local_total = c->local_total;
    a.t_total = c->pk.t_total;
    a.mode = mode;
    a.task = c->task_d.p;
    a.task_child = c->tchild_d.p;
    a.task_rows = c->trows_d.p;
    a.ext_rows = c->ext_d.p;
    a.nrhs = nrhs;
    a.xg = reinterpret_cast<T*>(c->work.xgbuf);
    return a;
}

// One launch over @p count queue entries (x nrhs): a grid of one workgroup per entry, or a
// single workgroup running them in order.
template <typename T, typename TF>
void launchQueue(s2l_ctx* c, LocalArgs<T, TF>& a, const int32_t* q, int32_t count, bool serial, hipStream_t s)
{
    if (count <= 0) return;
    a.queue = q;
    a.qcount = count;
    a.serial = serial ? 1 : 0;
    const bool pair = kIsPair<T>;  // a workgroup per entry and pair of right-hand sides
    const int grid = serial ? 1 : count * (pair ? (a.nrhs + 1) / 2 : a.nrhs);
    launchLocalSolve<T, TF>(a, grid, pair ? c->lds_bytes2 : c->lds_bytes, c->local_threads, s);
}

template <typename T> LsArgs<T> lsArgs(s2l_ctx* c, const double* rhs, const double* x0, double* out, int mode,
                                       const int32_t* skip)
{
    LsArgs<T> a{};
    a.sn = c->ls_sn.p;
    a.units = c->ls_units.p;
    a.items = c->ls_items.p;
    a.pull_rows = c->ls_pull.p;
    a.diag = reinterpret_cast<const T*>(c->ls_diag.p);
    a.belowF = reinterpret_cast<const T*>(c->ls_bf.p);
    a.belowB = reinterpret_cast<const T*>(c->ls_bb.p);
    a.ridx = c->ls_ridx.p;
    a.cptr = c->ls_cptr.p;
    a.cidx = c->ls_cidx.p;
    a.gmap = c->gmap_d.p;
    a.lperm = c->lperm_d.p;
    a.owned = c->ls_owned.p;
    a.begin = c->ls_begin.p;
    a.rhs = rhs;
    a.x0 = x0;
    a.out = out;
    a.pou = (mode == 1 && c->one_level == 0 && !c->boolean) ? c->pou_d.p : nullptr;
    a.skip = skip;
    a.slots = reinterpret_cast<T*>(c->work.ls_slots);
    a.part = reinterpret_cast<T*>(c->work.ls_part);
    a.bbuf = reinterpret_cast<T*>(c->work.ls_b);
    a.yf = reinterpret_cast<T*>(c->work.ls_yf);
    a.xb = reinterpret_cast<T*>(c->work.ls_xb);
    a.n_global = c->n;
    a.n_local_total = c->local_total;
    a.nslots = c->ls.slots;
    a.npart = c->ls.partials;
    a.mode = mode;
    return a;
}

template <typename T> void lsLaunches(s2l_ctx* c, const LsArgs<T>& a, int nrhs, hipStream_t s)
{
    const LsPlan& P = c->ls;
    const size_t nl = P.launch_kind.size();
    auto onAux = [](int kind) { return kind == kLsPull || kind == kLsForward || kind == kLsBackPartial || kind == kLsBackward; };
    for (size_t q = 0; q < nl;) {
        size_t q1 = q;
        bool aux = false, main = false;
        while (q1 < nl && P.launch_level[q1] == P.launch_level[q]) {
            (onAux(P.launch_kind[q1]) ? aux : main) = true;
            ++q1;
        }
        const bool fork = aux && main && c->ls_fork;
        if (fork) {
            HIPCHECK(hipEventRecord(c->ev_fork, s));
            HIPCHECK(hipStreamWaitEvent(c->aux, c->ev_fork, 0));
        }
        for (size_t k = q; k < q1; ++k)
            launchLs<T>(a, P.launch_kind[k], P.launch_first[k], P.launch_count[k], P.launch_w[k], nrhs,
                        (fork && onAux(P.launch_kind[k])) ? c->aux : s);
        if (fork) {
            HIPCHECK(hipEventRecord(c->ev_join, c->aux));
            HIPCHECK(hipStreamWaitEvent(s, c->ev_join, 0));
        }
        q = q1;
    }
}

template <typename T, typename TF>
void localSolve(s2l_ctx* c, const double* rhs, const double* x0, double* out, int mode, int nrhs, const int32_t* skip,
                hipStream_t s)
{
    LocalArgs<T, TF> a = localArgs<T, TF>(c, rhs, x0, out, mode, nrhs, skip);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->timing) {
        if (c->ev_pool.size() < 2) {
            for (int q = 0; q < 64; ++q) {
                hipEvent_t e;
                HIPCHECK(hipEventCreate(&e));
                c->ev_pool.push_back(e);
            }
        }
        e0 = c->ev_pool.back();
        c->ev_pool.pop_back();
        e1 = c->ev_pool.back();
        c->ev_pool.pop_back();
        HIPCHECK(hipEventRecord(e0, s));
    }
    if (c->launch_mode == 2) {
        // level-synchronous: one launch per level of all subdomains' trees (its factors are stored
        // in the arithmetic type)
        if constexpr (std::is_same_v<T, TF>)
            lsLaunches<T>(c, lsArgs<T>(c, rhs, x0, out, mode, skip), nrhs, s);
        else
            require(false, S2L_ESTATE, "the level-synchronous schedule has no FP32-storage / FP64-arithmetic mode");
    } else if (c->launch_mode == 0) {
        // level-ordered queue: one launch per depth group, every subdomain batched
        for (size_t g = 0; g + 1 < c->pk.qgroup.size(); ++g) {
            const int32_t* q = c->queue_d.p + c->pk.qgroup[g];
            const int32_t count = c->pk.qgroup[g + 1] - c->pk.qgroup[g];
            if (c->qgroup_dense[g]) {
                // a group of dense tops: the dense-top kernel (fewer registers, more wavefronts)
                a.queue = q;
                a.qcount = count;
                a.serial = 0;
                const bool pair = kIsPair<T>;
                launchDenseTop<T, TF>(a, count * (pair ? (a.nrhs + 1) / 2 : a.nrhs), c->dtrow_d.p, c->dtce_d.p, c->dtb_d.p,
                                      static_cast<size_t>(c->dt_maxS) * sizeof(T), c->cus, s);
            } else {
                launchQueue<T, TF>(c, a, q, count, false, s);
            }
        }
    } else {
        for (int i = 0; i < c->nsub; ++i)
            launchQueue<T, TF>(c, a, c->qsubq_d.p + c->qsub_h[static_cast<size_t>(i)],
                               c->qsub_h[static_cast<size_t>(i) + 1] - c->qsub_h[static_cast<size_t>(i)], true, s);
    }
    HIPCHECK(hipGetLastError());
    if (c->timing) {
        HIPCHECK(hipEventRecord(e1, s));
        c->ev_pending.emplace_back(e0, e1);
        c->launches += 1;
    }
}


hipEvent_t takeEvent(s2l_ctx* c)
{
    if (c->ev_pool.empty()) {
        for (int q = 0; q < 64; ++q) {
            hipEvent_t e;
            HIPCHECK(hipEventCreate(&e));
            c->ev_pool.push_back(e);
        }
    }
    hipEvent_t e = c->ev_pool.back();
    c->ev_pool.pop_back();
    return e;
}

// Phase @p id starts on stream s (a token for phaseEnd; -1 when phase timing is off).
int phaseBegin(s2l_ctx* c, int id, hipStream_t s)
{
    if (!c->phase_on) return -1;
    s2l_ctx::PhaseMark m{id, takeEvent(c), takeEvent(c)};
    HIPCHECK(hipEventRecord(m.e0, s));
    c->phase_pending.push_back(m);
    return static_cast<int>(c->phase_pending.size()) - 1;
}

void phaseEnd(s2l_ctx* c, int tok, hipStream_t s)
{
    if (tok < 0) return;
    HIPCHECK(hipEventRecord(c->phase_pending[static_cast<size_t>(tok)].e1, s));
}

void collectPhases(s2l_ctx* c)
{
    for (auto& m : c->phase_pending) {
        HIPCHECK(hipEventSynchronize(m.e1));
        float ms = 0.f;
        HIPCHECK(hipEventElapsedTime(&ms, m.e0, m.e1));
        c->phase_ms[m.id] += ms;
        c->phase_n[m.id] += 1;
        c->ev_pool.push_back(m.e0);
        c->ev_pool.push_back(m.e1);
    }
    c->phase_pending.clear();
}

void collectTiming(s2l_ctx* c)
{
    for (auto& pr : c->ev_pending) {
        HIPCHECK(hipEventSynchronize(pr.second));
        float ms = 0.f;
        HIPCHECK(hipEventElapsedTime(&ms, pr.first, pr.second));
        c->ms_total += ms;
        c->ev_pool.push_back(pr.first);
        c->ev_pool.push_back(pr.second);
    }
    c->ev_pending.clear();
}

// one-level part: z = x0 + M1^{-1} rhs on the owned rows
void oneLevel(s2l_ctx* c, const double* rhs, const double* x0, double* z, int nrhs, const int32_t* skip, hipStream_t s)
{
    const bool direct = (c->one_level == 0 && c->boolean);
    const int mode = direct ? 0 : 1;
    double* out = direct ? z : c->work.loc;
    if (c->precision == 1 && nrhs >= 2 && c->pair_ok && c->launch_mode == 0)
        localSolve<float2, float>(c, rhs, x0, out, mode, nrhs, skip, s);
    else if (c->precision == 2 && nrhs >= 2 && c->pair_ok && c->launch_mode == 0)
        localSolve<double2, float>(c, rhs, x0, out, mode, nrhs, skip, s);
    else if (c->precision == 1)
        localSolve<float, float>(c, rhs, x0, out, mode, nrhs, skip, s);
    else if (c->precision == 2)
        localSolve<double, float>(c, rhs, x0, out, mode, nrhs, skip, s);
    else
        localSolve<double, double>(c, rhs, x0, out, mode, nrhs, skip, s);
    if (!direct) launchRevGather(c->n, c->rg_ptr.p, c->rg_idx.p, c->work.loc, c->local_total, x0, z, nrhs, skip, s);
}

// The parts of one application are timed under the reference's timer names when phase timing is
// on (two_level_schwarz.cpp:144-159, coarse_space.cpp:237-244).
void applyImpl(s2l_ctx* c, const double* r, double* z, int nrhs, const int32_t* skip, hipStream_t s, bool grow = true)
{
    require(c->system_ready && c->subs_ready && c->local_ready, S2L_ESTATE, "apply: upload the system, subdomains and factors first");
    require(c->coarse_ready, S2L_ESTATE, "apply: upload the coarse space first (n0 = 0 for one level)");
    if (grow) c->ensureWork(nrhs);
    if (c->n0 == 0) {
        const int tl = phaseBegin(c, S2L_PHASE_PREC_LOCAL, s);
        oneLevel(c, r, nullptr, z, nrhs, skip, s);
        phaseEnd(c, tl, s);
        return;
    }
    // x0 = R_0^T A_00^{-1} R_0 r
    const int tc = phaseBegin(c, S2L_PHASE_PREC_COARSE, s);
    int t = phaseBegin(c, S2L_PHASE_PREC_COARSE_RESTRICT, s);
    const s2l_ctx::Work& W = c->work;
    launchCoarseRestrict(c->ca, r, c->n, W.cvec, W.cpart, nrhs, skip, s);
    phaseEnd(c, t, s);
    t = phaseBegin(c, S2L_PHASE_PREC_COARSE_SOLVE, s);
    // y = A00^-1 c in one product with the inverse (T2 T1): one launch and no dependent second pass
    launchTriGemv(c->n0, c->tri_ld, c->Einv_d.p, 2, c->coarse_perm ? c->cperm_d.p : nullptr, W.cvec, W.yvec, nrhs, skip, s);
    phaseEnd(c, t, s);
    t = phaseBegin(c, S2L_PHASE_PREC_COARSE_PROLONG, s);
    if (c->boolean) {
        launchCoarseProlong(c->ca, W.yvec, W.x0, c->n, 1, nrhs, skip, s);
    } else {
        launchCoarseProlong(c->ca, W.yvec, W.loc, c->local_total, 0, nrhs, skip, s);
        launchRevGather(c->n, c->rg_ptr.p, c->rg_idx.p, W.loc, c->local_total, nullptr, W.x0, nrhs, skip, s);
    }
    phaseEnd(c, t, s);
    phaseEnd(c, tc, s);
    if (c->correction == 1) {
        t = phaseBegin(c, S2L_PHASE_PREC_LOCAL, s);
        oneLevel(c, r, W.x0, z, nrhs, skip, s); // z = x0 + M1^{-1} r
        phaseEnd(c, t, s);
    } else {
        t = phaseBegin(c, S2L_PHASE_PREC_SPMV, s);
        launchSpmv(c->n, c->rp.p, c->col.p, c->val.p, W.x0, W.w, r, nrhs, skip, s); // w = r - A x0
        phaseEnd(c, t, s);
        t = phaseBegin(c, S2L_PHASE_PREC_LOCAL, s);
        oneLevel(c, W.w, W.x0, z, nrhs, skip, s);                                       // z = x0 + M1^{-1} w
        phaseEnd(c, t, s);
    }
    HIPCHECK(hipGetLastError());
}

void spmvImpl(s2l_ctx* c, const double* x, double* y, const double* b, int nrhs, const int32_t* skip, hipStream_t s)
{
    require(c->system_ready, S2L_ESTATE, "spmv: upload the system first");
    launchSpmv(c->n, c->rp.p, c->col.p, c->val.p, x, y, b, nrhs, skip, s);
    HIPCHECK(hipGetLastError());
}

// z = M^{-1} in, w = A z (w null: the apply alone): one launch of a HIP graph of all their kernels
// (about a dozen launches otherwise), captured on the first call for these buffers and replayed
// afterwards. Without graphs (or while the local solve is timed with events) the kernels are
// launched one by one.
void applySpmv(s2l_ctx* c, const double* in, double* z, double* w, int nrhs, const int32_t* skip, hipStream_t s)
{
    if (!c->use_graphs || c->timing || c->phase_on) {
        int t = phaseBegin(c, S2L_PHASE_PRECONDITIONER, s);  // gmres.cpp:67
        applyImpl(c, in, z, nrhs, skip, s);
        phaseEnd(c, t, s);
        if (w) {
            t = phaseBegin(c, S2L_PHASE_SPMV, s);            // gmres.cpp:71
            spmvImpl(c, z, w, nullptr, nrhs, skip, s);
            phaseEnd(c, t, s);
        }
        return;
    }
    c->ensureWork(nrhs);  // no allocation while capturing
    for (const auto& g : c->graphs)
        if (g.gen == c->gen && g.in == in && g.z == z && g.w == w && g.skip == skip && g.nrhs == nrhs && g.s == s) {
            HIPCHECK(hipGraphLaunch(g.exec, s));
            return;
        }
    if (c->graphs.size() >= kMaxGraphs) c->invalidateGraphs();
    hipGraph_t graph = nullptr;
    HIPCHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
        applyImpl(c, in, z, nrhs, skip, s);
        if (w) spmvImpl(c, z, w, nullptr, nrhs, skip, s);
    } catch (...) {
        (void)hipStreamEndCapture(s, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
    }
    HIPCHECK(hipStreamEndCapture(s, &graph));
    s2l_ctx::ApplyGraph g;
    const hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHECK(e);
    g.gen = c->gen;
    g.in = in;
    g.z = z;
    g.w = w;
    g.skip = skip;
    g.nrhs = nrhs;
    g.s = s;
    c->graphs.push_back(g);
    HIPCHECK(hipGraphLaunch(g.exec, s));
}

// ---- device GMRES (gmres.cpp:18-161) over `slots` right-hand sides sharing every launch. Each
// slot is an independent solve: its sums, its masks and its decisions are its own, so its results
// do not depend on which other slots run beside it. Slots enter (begin) and leave between restart
// cycles, which is what a serving loop needs to admit requests into a running solve.
struct GmresRun {
    s2l_ctx* c = nullptr;
    s2l_krylov_opts o{};
    int cap = 0;          ///< slots = right-hand sides of every launch
    int64_t n = 0;
    bool of = false, flex = false;
    GmArgs a{};
    DevBuf<double> V, Z, w, r, z, u, H, cs, sn, g, h, h2, y, hist, part, red, ones, coef;
    DevBuf<GmState> st;
    DevBuf<int32_t> flags, mask;
    PinBuf<GmState> pst;
    PinBuf<int32_t> pflags, pmask;
    PinBuf<double> pred, pred0, pcoef;  ///< pred0: the sums of ||b||^2 at the start of a solve
    struct Slot {
        bool live = false;      ///< holds a solve
        bool conv = false, breakdown = false;
        bool nonfinite = false; ///< a residual was not finite: the solve stops (no counterpart)
        double rate = 0.0;      ///< residual estimate's factor per iteration so far (0 = unknown)
        int iters = 0;
        double bnorm = 0, beta = 0, rel = 0;
        double normF = 1, l1_0 = 0, l1 = 0, thr = 0;  ///< OpenFOAM test
        std::vector<double> hist;
        bool active(int maxit) const { return live && !conv && !nonfinite && iters < maxit; }
    };
    std::vector<Slot> sl;
    std::vector<char> fresh;  ///< admitted since the last cycle (begin() pending)
    double rate_hint = 0.0;   ///< residual factor per iteration of the last converged slot (0 = none yet)

    GmresRun(s2l_ctx* ctx, const s2l_krylov_opts& opts, int slots) : c(ctx) { configure(opts, slots); }
    /// options and slot count of the next solves; the device buffers only grow
    void configure(const s2l_krylov_opts& opts, int slots);
    double sum(int q, int k) const { return pred.p[3 * q + k]; }
    /// per-slot sums (kind of launchReduce) of the slots whose mask entry is 0, on the host
    void reduce(int kind, const double* x, const double* yv, hipStream_t s)
    {
        launchReduce(kind, n, x, yv, part.p, red.p, cap, mask.p, s);
        HIPCHECK(hipMemcpyAsync(pred.p, red.p, sizeof(double) * 3 * cap, hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
    }
    void setMask(const std::vector<char>& on, hipStream_t s)
    {
        HIPCHECK(hipStreamSynchronize(s));  // earlier copies have read the staging buffer
        for (int q = 0; q < cap; ++q) pmask.p[q] = on[static_cast<size_t>(q)] ? 0 : 1;
        HIPCHECK(hipMemcpyAsync(mask.p, pmask.p, sizeof(int32_t) * cap, hipMemcpyHostToDevice, s));
    }
    void begin(const std::vector<char>& on, const double* d_b, double* d_x, hipStream_t s);
    bool cycle(const double* d_b, double* d_x, hipStream_t s);
    void result(int q, s2l_krylov_result* R, double* history) const;
};

void GmresRun::configure(const s2l_krylov_opts& opts, int slots)
{
    o = opts;
    cap = slots;
    n = c->n;
    const int m = o.restart;
    require(m >= 1 && m + 1 <= kGmMaxK, S2L_ELIMIT, "GMRES restart must be in [1, 64] on the device");
    require(o.stop_norm == 0 || o.stop_norm == 1, S2L_EINVAL, "stop_norm must be 0 (l2) or 1 (openfoam)");
    require(o.maxit >= 0, S2L_EINVAL, "maxit must be >= 0");
    require(o.atol == 0.0 || o.stop_norm == 1, S2L_EINVAL, "an absolute tolerance needs the OpenFOAM stopping norm");
    require(slots >= 1, S2L_EINVAL, "at least one right-hand side");
    of = o.stop_norm == 1;
    flex = o.flexible != 0;
    const int nb = gmBlocks(n);
    const size_t N = static_cast<size_t>(n) * cap;
    V.reserve(static_cast<size_t>(m + 1) * N);
    if (flex) Z.reserve(static_cast<size_t>(m) * N);
    w.reserve(N);
    r.reserve(N);
    z.reserve(N);
    u.reserve(N);
    H.reserve(static_cast<size_t>(m + 1) * m * cap);
    cs.reserve(static_cast<size_t>(m) * cap);
    sn.reserve(static_cast<size_t>(m) * cap);
    g.reserve(static_cast<size_t>(m + 1) * cap);
    h.reserve(static_cast<size_t>(m + 1) * cap);
    h2.reserve(static_cast<size_t>(m + 1) * cap);
    y.reserve(static_cast<size_t>(m) * cap);
    hist.reserve(static_cast<size_t>(o.maxit + 1) * cap);
    part.reserve(static_cast<size_t>(nb) * kGmMaxK * cap);
    red.reserve(static_cast<size_t>(3) * cap);
    coef.reserve(static_cast<size_t>(4) * cap);
    st.reserve(static_cast<size_t>(cap));
    flags.reserve(static_cast<size_t>(2) * cap);
    mask.reserve(static_cast<size_t>(cap));
    pst.reserve(static_cast<size_t>(cap));
    pflags.reserve(static_cast<size_t>(2) * cap);
    pmask.reserve(static_cast<size_t>(cap));
    pred.reserve(static_cast<size_t>(3) * cap);
    pred0.reserve(static_cast<size_t>(3) * cap);
    pcoef.reserve(static_cast<size_t>(4) * cap);
    sl.assign(static_cast<size_t>(cap), Slot{});
    fresh.assign(static_cast<size_t>(cap), 0);
    a.n = n;
    a.m = m;
    a.nrhs = cap;
    a.V = V.p;
    a.Z = flex ? Z.p : nullptr;
    a.w = w.p;
    a.H = H.p;
    a.cs = cs.p;
    a.sn = sn.p;
    a.g = g.p;
    a.h = h.p;
    a.h2 = h2.p;
    a.y = y.p;
    a.st = st.p;
    a.flags = flags.p;
    a.hist = hist.p;
    a.hist_ld = o.maxit + 1;
    a.rtol = o.rtol;
    a.partial = part.p;
    a.nblocks = nb;
    a.of = of ? 1 : 0;
}

// Start of the solves in the slots marked in @p on (gmres.cpp:26-55): ||b||, r = b - A x0, the
// OpenFOAM normalisation and threshold, and the test of x0. Other slots are untouched.
void GmresRun::begin(const std::vector<char>& on, const double* d_b, double* d_x, hipStream_t s)
{
    bool any = false;
    for (int q = 0; q < cap; ++q) {
        if (!on[static_cast<size_t>(q)]) continue;
        any = true;
        sl[static_cast<size_t>(q)] = Slot{};
        sl[static_cast<size_t>(q)].live = true;
    }
    if (!any) return;
    setMask(on, s);
    if (!of) {
        // ||b||^2, r = b - A x0 and ||r||^2 behind one synchronization (a slot with b = 0 computes
        // an r it does not use)
        launchReduce(0, n, d_b, nullptr, part.p, red.p, cap, mask.p, s);
        HIPCHECK(hipMemcpyAsync(pred0.p, red.p, sizeof(double) * 3 * cap, hipMemcpyDeviceToHost, s));
        spmvImpl(c, d_x, r.p, d_b, cap, mask.p, s);
        reduce(0, r.p, nullptr, s);
    } else {
        reduce(0, d_b, nullptr, s);
        std::memcpy(pred0.p, pred.p, sizeof(double) * 3 * cap);
    }
    std::vector<char> nz(static_cast<size_t>(cap), 0);
    for (int q = 0; q < cap; ++q) {
        if (!on[static_cast<size_t>(q)]) continue;
        Slot& S = sl[static_cast<size_t>(q)];
        S.bnorm = std::sqrt(pred0.p[3 * q]);
        if (S.bnorm == 0.0) {  // x = 0, converged (gmres.cpp:28-33)
            HIPCHECK(hipMemsetAsync(d_x + static_cast<int64_t>(q) * n, 0, sizeof(double) * n, s));
            S.conv = true;
            S.hist.push_back(0.0);
        } else {
            nz[static_cast<size_t>(q)] = 1;
        }
    }
    bool anynz = false;
    for (char v : nz) anynz = anynz || v;
    if (!anynz) return;
    if (of) {
        // F = sum|A x0 - xm A 1| + sum|b - xm A 1| + 1e-20, xm = mean(x0) (krylov.cpp:47-57)
        spmvImpl(c, d_x, r.p, nullptr, cap, mask.p, s);  // r = A x0
        reduce(5, d_x, nullptr, s);
        std::vector<double> xm(static_cast<size_t>(cap), 0.0);
        std::vector<char> zero_m(static_cast<size_t>(cap), 0), with_m(static_cast<size_t>(cap), 0);
        for (int q = 0; q < cap; ++q) {
            if (!nz[static_cast<size_t>(q)]) continue;
            xm[static_cast<size_t>(q)] = n ? sum(q, 1) / static_cast<double>(n) : 0.0;
            (xm[static_cast<size_t>(q)] == 0.0 ? zero_m : with_m)[static_cast<size_t>(q)] = 1;
        }
        bool az = false, aw = false;
        for (int q = 0; q < cap; ++q) {
            az = az || zero_m[static_cast<size_t>(q)];
            aw = aw || with_m[static_cast<size_t>(q)];
        }
        if (az) {
            setMask(zero_m, s);
            reduce(1, r.p, nullptr, s);
            for (int q = 0; q < cap; ++q)
                if (zero_m[static_cast<size_t>(q)]) sl[static_cast<size_t>(q)].normF = sum(q, 1);
            reduce(1, d_b, nullptr, s);
            for (int q = 0; q < cap; ++q)
                if (zero_m[static_cast<size_t>(q)]) sl[static_cast<size_t>(q)].normF += sum(q, 1) + 1e-20;
        }
        if (aw) {
            setMask(with_m, s);
            ones.reserve(static_cast<size_t>(n) * 2);  // A 1 of the current values
            launchFill(n, 1.0, ones.p, s);
            spmvImpl(c, ones.p, ones.p + n, nullptr, 1, nullptr, s);
            for (int q = 0; q < cap; ++q) {
                pcoef.p[4 * q] = xm[static_cast<size_t>(q)];
                pcoef.p[4 * q + 1] = pcoef.p[4 * q + 2] = pcoef.p[4 * q + 3] = 0.0;
            }
            HIPCHECK(hipMemcpyAsync(coef.p, pcoef.p, sizeof(double) * 4 * cap, hipMemcpyHostToDevice, s));
            ofNorm(n, r.p, d_b, ones.p + n, coef.p, part.p, red.p, cap, mask.p, s);
            HIPCHECK(hipMemcpyAsync(pred.p, red.p, sizeof(double) * 3 * cap, hipMemcpyDeviceToHost, s));
            HIPCHECK(hipStreamSynchronize(s));
            for (int q = 0; q < cap; ++q)
                if (with_m[static_cast<size_t>(q)]) sl[static_cast<size_t>(q)].normF = sum(q, 0) + sum(q, 1) + 1e-20;
        }
        setMask(nz, s);
        launchAxpby(n, 1.0, d_b, -1.0, r.p, cap, mask.p, s);  // r = b - A x0
        reduce(1, r.p, nullptr, s);
    } else {
        setMask(nz, s);  // the slots that iterate
    }
    for (int q = 0; q < cap; ++q) {
        if (!nz[static_cast<size_t>(q)]) continue;
        Slot& S = sl[static_cast<size_t>(q)];
        S.beta = std::sqrt(sum(q, 0));
        S.rel = S.beta / S.bnorm;
        S.hist.push_back(S.rel);
        if (!std::isfinite(S.rel)) {  // a non-finite right-hand side or initial guess
            S.nonfinite = true;
            continue;
        }
        if (of) {
            S.l1_0 = S.l1 = sum(q, 1);
            S.thr = std::max(o.rtol * S.l1_0, o.atol * S.normF);
            if (S.l1_0 <= S.thr) S.conv = true;  // no iteration if x0 satisfies the test
        } else if (S.beta <= o.rtol * S.bnorm) {
            S.conv = true;
        }
    }
}

// One restart cycle (gmres.cpp:57-157) of every slot that holds an unfinished solve; slots that
// finish early are masked while the others continue. Returns false when no slot had work.
bool GmresRun::cycle(const double* d_b, double* d_x, hipStream_t s)
{
    const int m = o.restart, maxit = o.maxit;
    bool any = false;
    std::vector<char> act(static_cast<size_t>(cap), 0);
    for (int q = 0; q < cap; ++q) {
        act[static_cast<size_t>(q)] = sl[static_cast<size_t>(q)].active(maxit) ? 1 : 0;
        any = any || act[static_cast<size_t>(q)];
    }
    if (!any) return false;
    HIPCHECK(hipStreamSynchronize(s));  // earlier copies have read the staging buffers
    for (int q = 0; q < cap; ++q) {
        const Slot& S = sl[static_cast<size_t>(q)];
        GmState& G = pst.p[q];
        std::memset(&G, 0, sizeof(G));
        G.bnorm = S.bnorm;
        G.beta = S.beta;
        G.iters = S.iters;
        G.maxit = maxit;
        G.thr = S.thr;
        G.done = act[static_cast<size_t>(q)] ? 0 : 1;
        pflags.p[q] = act[static_cast<size_t>(q)] ? 0 : 1;
        pflags.p[cap + q] = act[static_cast<size_t>(q)] ? 0 : 1;
    }
    HIPCHECK(hipMemcpyAsync(st.p, pst.p, sizeof(GmState) * cap, hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemcpyAsync(flags.p, pflags.p, sizeof(int32_t) * 2 * cap, hipMemcpyHostToDevice, s));
    gmStartCycle(a, r.p, s);
    const int32_t* live = flags.p;
    const int32_t* idl = flags.p + cap;
    const int64_t N = static_cast<int64_t>(n) * cap;
    // Iterations between host polls of the flags: as many as the slowest active slot still needs to
    // reach the tolerance at the rate its residual estimate has fallen so far (a slot without
    // iterations yet takes the rate of the last converged slot of this run), 1 to kMaxBatch
    // (kMaxUnknown while no rate is known, and with the OpenFOAM test). A poll costs one
    // synchronization; every iteration launched past a slot's convergence costs the launches of an
    // iteration that find it finished (a restart cycle after the true-residual check often needs
    // one or two iterations). The results do not depend on the batches.
    constexpr int kMaxBatch = 8, kMaxUnknown = 4;
    auto itersNeeded = [&](double rel, double rate) {
        if (of || !(rel > 0.0) || !(rate > 0.0 && rate < 1.0)) return kMaxUnknown;
        if (rel <= o.rtol) return 1;
        const double k = std::ceil(std::log(o.rtol / rel) / std::log(rate));
        return static_cast<int>(std::min<double>(kMaxBatch, std::max(1.0, k)));
    };
    int batch = 1;
    for (int q = 0; q < cap; ++q)
        if (act[static_cast<size_t>(q)]) {
            const Slot& S = sl[static_cast<size_t>(q)];
            batch = std::max(batch, itersNeeded(S.bnorm > 0.0 ? S.beta / S.bnorm : 0.0, S.rate > 0.0 ? S.rate : rate_hint));
        }
    // one right-hand side whose rows fit one per thread of the GMRES grid: the SpMV of the Arnoldi
    // step and the first CGS2 pass in one launch (the same bits as the two launches)
    const bool fuse = cap == 1 && static_cast<int64_t>(a.nblocks) * kGmThreads >= n;
    for (int j0 = 0; j0 < m;) {
        const int j1 = std::min(m, j0 + batch);
        for (int j = j0; j < j1; ++j) {
            const double* vj = V.p + static_cast<int64_t>(j) * N;
            double* zj = flex ? Z.p + static_cast<int64_t>(j) * N : z.p;
            if (fuse) {
                applySpmv(c, vj, zj, nullptr, cap, live, s);  // z = M^{-1} v_j (kept: flexible)
                const int ts = phaseBegin(c, S2L_PHASE_SPMV, s);  // gmres.cpp:71, with CGS2 pass 1's products
                gmSpmvDots(a, j, c->rp.p, c->col.p, c->val.p, zj, s);  // w = A z, partials of V^T w
                phaseEnd(c, ts, s);
            } else {
                applySpmv(c, vj, zj, w.p, cap, live, s);  // z = M^{-1} v_j (kept: flexible); w = A z
            }
            const int t = phaseBegin(c, S2L_PHASE_ORTHOGONALIZATION, s);  // gmres.cpp:77
            if (fuse)
                gmFinalH(a, j, s);                    // CGS2 pass 1: h
            else
                gmDots(a, j, s);                      // CGS2 pass 1
            gmUpdateDots(a, j, s);                    // w -= V h; pass 2
            gmUpdateNorm(a, j, s);                    // w -= V h2; partials of ||w||^2
            gmNormalize(a, j, s);                     // ||w||, Givens, v_{j+1} = w / ||w||
            phaseEnd(c, t, s);
            if (of) {
                const int ts = phaseBegin(c, S2L_PHASE_STOPPING_TEST, s);  // gmres.cpp:108
                gmOpenFoamTest(a, j, s);              // sum|r_j| against the threshold
                phaseEnd(c, ts, s);
            }
        }
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(pflags.p, flags.p, sizeof(int32_t) * cap, hipMemcpyDeviceToHost, s));
        HIPCHECK(hipMemcpyAsync(pst.p, st.p, sizeof(GmState) * cap, hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        bool all = true;
        batch = 1;
        for (int q = 0; q < cap; ++q) {
            all = all && pflags.p[q];
            if (pflags.p[q]) continue;
            Slot& S = sl[static_cast<size_t>(q)];
            const GmState& G = pst.p[q];
            const double rel0 = S.bnorm > 0.0 ? S.beta / S.bnorm : 0.0;
            if (G.k > 0 && rel0 > 0.0 && G.rel > 0.0) S.rate = std::pow(G.rel / rel0, 1.0 / G.k);
            batch = std::max(batch, itersNeeded(G.rel, S.rate));
        }
        if (all) break;
        j0 = j1;
    }
    // x += M^{-1} V_k y_k; true residual (gmres.cpp:135-156)
    gmLsq(a, s);
    if (flex) {
        gmCombine(a, Z.p, z.p, s);      // x += Z y
    } else {
        gmCombine(a, V.p, u.p, s);      // x += M^{-1}(V y)
        const int t = phaseBegin(c, S2L_PHASE_PRECONDITIONER, s);  // gmres.cpp:139
        applyImpl(c, u.p, z.p, cap, idl, s);
        phaseEnd(c, t, s);
    }
    launchAxpby(n, 1.0, z.p, 1.0, d_x, cap, idl, s);
    spmvImpl(c, d_x, r.p, d_b, cap, idl, s);
    HIPCHECK(hipMemcpyAsync(pst.p, st.p, sizeof(GmState) * cap, hipMemcpyDeviceToHost, s));
    launchReduce(of ? 1 : 0, n, r.p, nullptr, part.p, red.p, cap, idl, s);
    HIPCHECK(hipMemcpyAsync(pred.p, red.p, sizeof(double) * 3 * cap, hipMemcpyDeviceToHost, s));
    std::vector<double> hc(static_cast<size_t>(maxit + 1) * cap);
    HIPCHECK(hipMemcpyAsync(hc.data(), hist.p, sizeof(double) * hc.size(), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    for (int q = 0; q < cap; ++q) {
        if (!act[static_cast<size_t>(q)]) continue;
        Slot& S = sl[static_cast<size_t>(q)];
        const GmState& G = pst.p[q];
        for (int it = S.iters + 1; it <= G.iters; ++it) S.hist.push_back(hc[static_cast<size_t>(q) * (maxit + 1) + it]);
        S.iters = G.iters;
        S.rel = G.rel;
        const double b2 = std::sqrt(sum(q, 0));
        S.beta = b2;
        if (of) S.l1 = sum(q, 1);
        bool cv = G.conv != 0;
        const bool within = of ? (S.l1 <= S.thr) : (b2 <= o.rtol * S.bnorm);
        const bool within_slack = of ? (S.l1 <= 1.1 * S.thr) : (b2 <= 1.1 * o.rtol * S.bnorm);
        if (G.breakdown) cv = within || cv;
        // the estimate can be optimistic for a preconditioner that is not exactly linear (FP32
        // factors): accept only a true residual within 10 % (gmres.cpp:151-155)
        if (cv && !within_slack) cv = false;
        if (b2 == 0.0) cv = true;
        S.conv = cv;
        if (cv && S.rate > 0.0 && S.rate < 1.0) rate_hint = S.rate;
        S.breakdown = S.breakdown || G.breakdown;
        // a non-finite residual (the Arnoldi estimate or the true one) ends the solve unconverged
        // instead of restarting until maxit, which is what the reference's loop would do
        if (!std::isfinite(b2) || !std::isfinite(G.rel)) {
            S.conv = false;
            S.nonfinite = true;
        }
    }
    return true;
}

void GmresRun::result(int q, s2l_krylov_result* R, double* history) const
{
    const Slot& S = sl[static_cast<size_t>(q)];
    std::memset(R, 0, sizeof(*R));
    R->converged = S.conv;
    R->breakdown = S.nonfinite ? 1 : 0;
    R->iterations = S.iters;
    R->rel_residual = S.rel;
    R->true_rel_residual = S.bnorm > 0.0 ? S.beta / S.bnorm : 0.0;
    if (of && S.bnorm > 0.0) {
        R->initial_residual_of = S.l1_0 / S.normF;
        R->final_residual_of = S.l1 / S.normF;
    }
    R->history_len = static_cast<int>(S.hist.size());
    if (history)
        for (size_t k = 0; k < S.hist.size() && k < static_cast<size_t>(o.maxit + 1); ++k) history[k] = S.hist[k];
}

// s2l_krylov_solve with GMRES: every right-hand side in its own slot, cycles until all finish. The
// run (and its workspace) is kept by the context for the next solve of the same shape.
int gmresImpl(s2l_ctx* c, const s2l_krylov_opts* o, const double* d_b, double* d_x, int nrhs, s2l_krylov_result* res,
              double* history, hipStream_t s)
{
    std::shared_ptr<GmresRun>& run = c->gm_cache;
    if (!run)
        run = std::make_shared<GmresRun>(c, *o, nrhs);
    else
        run->configure(*o, nrhs);
    std::vector<char> all(static_cast<size_t>(nrhs), 1);
    run->begin(all, d_b, d_x, s);
    while (run->cycle(d_b, d_x, s)) {
    }
    for (int q = 0; q < nrhs; ++q) run->result(q, res + q, history ? history + static_cast<size_t>(q) * (o->maxit + 1) : nullptr);
    return S2L_OK;
}


// ---- device BiCGStab (bicgstab.cpp:15-154), nrhs independent solves sharing every launch. The
// scalar recurrences and every decision run on the device (k_bicg_control, FP64, the reference's
// order of operations), between the fused vector kernels, whose launches each right-hand side joins
// or skips through per-stage masks the control kernel writes. The host reads one word per loop
// iteration: how many right-hand sides are still running.
int bicgImpl(s2l_ctx* c, const s2l_krylov_opts* o, const double* d_b, double* d_x, int nrhs, s2l_krylov_result* res,
             double* history, hipStream_t s)
{
    require(o->maxit >= 0, S2L_EINVAL, "maxit must be >= 0");
    require(o->stop_norm == 0 || o->stop_norm == 1, S2L_EINVAL, "stop_norm must be 0 (l2) or 1 (openfoam)");
    require(o->flexible == 0, S2L_EINVAL, "flexible applies to GMRES only");
    const bool of = o->stop_norm == 1;
    const int64_t n = c->n;
    const int maxit = o->maxit;
    const size_t N = static_cast<size_t>(n) * nrhs;
    const size_t H = static_cast<size_t>(maxit + 1) * nrhs;
    for (DevBuf<double>* b : {&c->bR, &c->bRh, &c->bP, &c->bV, &c->bY, &c->bS, &c->bZ, &c->bT}) b->reserve(N);
    c->bcoef.reserve(static_cast<size_t>(kBiMasks) * 4 * nrhs);
    c->bmask.reserve(static_cast<size_t>(kBiMasks) * nrhs);
    c->bst.reserve(static_cast<size_t>(nrhs));
    c->bhist.reserve(H);
    c->blive.reserve(1);
    c->plive.reserve(1);
    c->pst.reserve(static_cast<size_t>(nrhs));
    c->phist.reserve(H);
    c->kpart.reserve(static_cast<size_t>(reduceBlocks(n)) * 3 * nrhs);
    c->kred.reserve(static_cast<size_t>(3) * nrhs);
    double* r = c->bR.p;
    double* rh = c->bRh.p;
    double* p = c->bP.p;
    double* v = c->bV.p;
    double* y = c->bY.p;
    double* sv = c->bS.p;
    double* z = c->bZ.p;
    double* t = c->bT.p;
    BiArgs A{};
    A.st = c->bst.p;
    A.red = c->kred.p;
    A.mask = c->bmask.p;
    A.coef = c->bcoef.p;
    A.hist = c->bhist.p;
    A.live = c->blive.p;
    A.n = n;
    A.rtol = o->rtol;
    A.atol = o->atol;
    A.nrhs = nrhs;
    A.maxit = maxit;
    A.of = of ? 1 : 0;
    auto mask = [&](int m) { return static_cast<const int32_t*>(c->bmask.p + static_cast<size_t>(m) * nrhs); };
    auto coef = [&](int m) { return static_cast<const double*>(c->bcoef.p + static_cast<size_t>(m) * 4 * nrhs); };
    auto ctl = [&](int stage) { launchBicgControl(A, stage, s); };
    double* part = c->kpart.p;
    double* red = c->kred.p;

    // ||b||; x = 0 where b = 0
    launchReduce(0, n, d_b, nullptr, part, red, nrhs, nullptr, s);
    ctl(kBiBnorm);
    launchFillMasked(n, 0.0, d_x, nrhs, mask(kBiZero), s);
    // r = A x0; the OpenFOAM normalisation from A x0; r = b - r
    spmvImpl(c, d_x, r, nullptr, nrhs, mask(kBiLive), s);
    if (of) {
        launchReduce(5, n, d_x, nullptr, part, red, nrhs, mask(kBiLive), s);
        ctl(kBiXmean);
        launchReduce(1, n, r, nullptr, part, red, nrhs, mask(kBiNom), s);  // xm = 0: F = |A x0|_1 + |b|_1 + 1e-20
        ctl(kBiNomAx);
        launchReduce(1, n, d_b, nullptr, part, red, nrhs, mask(kBiNom), s);
        ctl(kBiNomB);
        c->bones.reserve(static_cast<size_t>(n) * 2);
        launchFill(n, 1.0, c->bones.p, s);
        spmvImpl(c, c->bones.p, c->bones.p + n, nullptr, 1, nullptr, s);
        ofNorm(n, r, d_b, c->bones.p + n, coef(kBiWithm), part, red, nrhs, mask(kBiWithm), s);
        ctl(kBiWithmNorm);
    }
    launchAxpby(n, 1.0, d_b, -1.0, r, nrhs, mask(kBiLive), s);
    launchReduce(1, n, r, nullptr, part, red, nrhs, mask(kBiLive), s);
    ctl(kBiR0);

    // every loop iteration ends a cycle or advances the iteration count of each running solve
    for (int it = 0; it <= maxit + 1; ++it) {
        ctl(kBiCycle);  // cycle starts: convergence / budget tests, r^ = r (bicgstab.cpp:73-82)
        launchAxpby(n, 1.0, r, 0.0, rh, nrhs, mask(kBiStart), s);
        // first half: p, y = M p, v = A y, alpha = rho / (r^, v)
        bicgP(n, r, p, v, coef(kBiP), nrhs, mask(kBiP), s);
        int tp = phaseBegin(c, S2L_PHASE_PRECONDITIONER, s);
        applyImpl(c, p, y, nrhs, mask(kBiP), s);
        phaseEnd(c, tp, s);
        tp = phaseBegin(c, S2L_PHASE_SPMV, s);
        spmvImpl(c, y, v, nullptr, nrhs, mask(kBiP), s);
        phaseEnd(c, tp, s);
        launchReduce(3, n, v, rh, part, red, nrhs, mask(kBiP), s);
        ctl(kBiAlpha);
        // s = r - alpha v and its norms: the half step may end the cycle (x += alpha y; r = s)
        bicgS(n, r, v, sv, coef(kBiHalf), part, red, nrhs, mask(kBiHalf), s);
        ctl(kBiHalfStep);
        bicgHalf(n, d_x, y, sv, r, coef(kBiHx), nrhs, mask(kBiHx), s);
        // second half: z = M s, t = A z, omega = (t, s) / (t, t), x and r
        tp = phaseBegin(c, S2L_PHASE_PRECONDITIONER, s);
        applyImpl(c, sv, z, nrhs, mask(kBiFull), s);
        phaseEnd(c, tp, s);
        tp = phaseBegin(c, S2L_PHASE_SPMV, s);
        spmvImpl(c, z, t, nullptr, nrhs, mask(kBiFull), s);
        phaseEnd(c, tp, s);
        launchReduce(4, n, t, sv, part, red, nrhs, mask(kBiFull), s);
        ctl(kBiOmega);
        bicgXR(n, d_x, y, z, sv, t, r, rh, coef(kBiFull), part, red, nrhs, mask(kBiFull), s);
        ctl(kBiXr);
        // end of a cycle: true residual r = b - A x (bicgstab.cpp:138-151)
        spmvImpl(c, d_x, r, d_b, nrhs, mask(kBiEnd), s);
        launchReduce(1, n, r, nullptr, part, red, nrhs, mask(kBiEnd), s);
        ctl(kBiEndCycle);
        HIPCHECK(hipMemcpyAsync(c->plive.p, c->blive.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        if (c->plive.p[0] == 0) break;
    }
    HIPCHECK(hipMemcpyAsync(c->pst.p, c->bst.p, sizeof(BiState) * nrhs, hipMemcpyDeviceToHost, s));
    HIPCHECK(hipMemcpyAsync(c->phist.p, c->bhist.p, sizeof(double) * H, hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    for (int q = 0; q < nrhs; ++q) {
        const BiState& st = c->pst.p[q];
        const double* hq = c->phist.p + static_cast<size_t>(q) * (maxit + 1);
        const int nh = std::min(st.nhist, maxit + 1);
        s2l_krylov_result& out = res[q];
        std::memset(&out, 0, sizeof(out));
        out.converged = st.conv != 0;
        out.breakdown = st.brk != 0;
        out.iterations = st.its;
        out.restarts = st.restarts;
        out.rel_residual = nh > 0 ? hq[nh - 1] : 0.0;
        out.true_rel_residual = st.bnorm > 0.0 ? st.rn / st.bnorm : 0.0;
        if (of && st.bnorm > 0.0) {
            out.initial_residual_of = st.l1_0 / st.normF;
            out.final_residual_of = st.l1 / st.normF;
        }
        out.history_len = st.nhist;
        if (history)
            for (int k = 0; k < nh; ++k) history[static_cast<size_t>(q) * (maxit + 1) + k] = hq[k];
    }
    return S2L_OK;
}
// The level-synchronous schedule (launch mode 2) of the uploaded factors with values in @p precision.
void buildLs(s2l_ctx* c, int precision)
{
    c->ls = buildLsPlan(c->hf, c->local_ptr);
    {
        LsPlan& L = c->ls;
        c->ls_sn.upload(L.sn, c->stream);
        c->ls_units.upload(L.units, c->stream);
        c->ls_items.upload(L.items, c->stream);
        c->ls_pull.upload(L.pull_rows, c->stream);
        c->ls_ridx.upload(L.ridx, c->stream);
        c->ls_cptr.upload(L.cptr, c->stream);
        c->ls_cidx.upload(L.cidx, c->stream);
        std::vector<int32_t> owned(static_cast<size_t>(c->nsub)), begin(static_cast<size_t>(c->nsub));
        for (int i = 0; i < c->nsub; ++i) {
            owned[static_cast<size_t>(i)] = c->ownership[i + 1] - c->ownership[i];
            begin[static_cast<size_t>(i)] = c->ownership[i];
        }
        c->ls_owned.upload(owned, c->stream);
        c->ls_begin.upload(begin, c->stream);
        auto values = [&](DevBuf<unsigned char>& dst, std::vector<double>& src) {
            if (precision != 0) {
                std::vector<float> f(src.begin(), src.end());
                dst.upload(reinterpret_cast<const unsigned char*>(f.data()), f.size() * sizeof(float), c->stream);
                HIPCHECK(hipStreamSynchronize(c->stream));
            } else {
                dst.upload(reinterpret_cast<const unsigned char*>(src.data()), src.size() * sizeof(double), c->stream);
                HIPCHECK(hipStreamSynchronize(c->stream));
            }
            std::vector<double>().swap(src);
        };
        values(c->ls_diag, L.diag);
        values(c->ls_bf, L.belowF);
        values(c->ls_bb, L.belowB);
    }
    c->ls_built = true;
    c->work_nrhs = 0;  // the scratch of the level-synchronous schedule is sized again
}

// The plan of the device refactorization and its device tables (built once per pattern, subdomains
// and factors; the host part usually comes from the thread started by finalize_local).
void ensureRfPlan(s2l_ctx* c, hipStream_t s)
{
    if (c->rf_ready) return;
    c->rfJoin();
    if (!c->rf_plan_err.empty()) throw std::invalid_argument(c->rf_plan_err);
    if (!c->rf_planned) c->rf = buildRefactorPlan(c->hf, c->local_ptr, c->gid_h, c->rp_h, c->col_h, c->pk);
    c->rf_planned = true;
    RfPlan& R = c->rf;
    c->rf_sn.upload(R.sn, s);
    c->rf_upd.upload(R.upd, s);
    c->rf_rel.upload(R.rel, s);
    c->rf_pos.upload(R.amap_pos, s);
    c->rf_src.upload(R.amap_src, s);
    c->rf_order.upload(R.order, s);
    c->rf_invoff.upload(R.inv_off, s);
    c->rf_inpos.upload(c->inpos_h, s);
    c->rf_work.reserve(static_cast<size_t>(std::max<int64_t>(R.work, 1)));
    c->rf_inv.reserve(static_cast<size_t>(std::max<int64_t>(R.inv, 1)));
    if (!c->pk.dtop.empty()) {
        c->dt_d.upload(c->pk.dtop, s);
        c->dt_rf_d.upload(R.dtop_rf, s);
        c->dt_q_d.upload(c->pk.dtop_q, s);
        c->dt_off_d.upload(R.dtop_off, s);
        c->dt_L.reserve(static_cast<size_t>(R.dtop_total));
        c->dt_C.reserve(static_cast<size_t>(R.dtop_total));
        c->dt_info.reserve(1);
    }
    c->rf_status.reserve(1);
    HIPCHECK(hipStreamSynchronize(s));
    c->rf_ready = true;
}

// The dense tops of refactorized factors (after launchRfPack): L_SS assembled from the panels,
// X = L_SS^-1 (rocSOLVER), S^-1 = X^T X (rocBLAS), then the solve layout in @p diag.
void rfDenseTops(s2l_ctx* c, const RfArgs& ra, void* diag, bool fp32, hipStream_t s)
{
    const std::vector<DenseTop>& D = c->pk.dtop;
    if (D.empty()) return;
    const RfPlan& R = c->rf;
    const int32_t nd = static_cast<int32_t>(D.size());
    launchDtAssemble(ra, c->dt_d.p, nd, c->dt_rf_d.p, c->dt_q_d.p, c->dt_off_d.p, c->dt_L.p, s);
    HIPCHECK(hipGetLastError());
    if (!c->rf_blas)
        require(rocblas_create_handle(&c->rf_blas) == rocblas_status_success, S2L_EHIP, "dense tops: rocblas_create_handle failed");
    require(rocblas_set_stream(c->rf_blas, s) == rocblas_status_success, S2L_EHIP, "dense tops: rocblas_set_stream failed");
    const double one = 1.0, zero = 0.0;
    for (int32_t t = 0; t < nd; ++t) {
        const int S = D[static_cast<size_t>(t)].S;
        double* X = c->dt_L.p + R.dtop_off[static_cast<size_t>(t)];
        require(rocsolver_dtrtri(c->rf_blas, rocblas_fill_lower, rocblas_diagonal_non_unit, S, X, S, c->dt_info.p) ==
                    rocblas_status_success,
                S2L_EHIP, "dense tops: rocsolver_dtrtri failed");
        require(rocblas_dgemm(c->rf_blas, rocblas_operation_transpose, rocblas_operation_none, S, S, S, &one, X, S, X, S, &zero,
                              c->dt_C.p + R.dtop_off[static_cast<size_t>(t)], S) == rocblas_status_success,
                S2L_EHIP, "dense tops: rocblas_dgemm failed");
    }
    launchDtLayout(c->dt_d.p, nd, c->dt_off_d.p, c->dt_C.p, diag, fp32, s);
    HIPCHECK(hipGetLastError());
}

RfArgs rfArgs(s2l_ctx* c)
{
    RfArgs a{};
    a.sn = c->rf_sn.p;
    a.upd = c->rf_upd.p;
    a.rel = c->rf_rel.p;
    a.ridx = c->ridx_d.p;
    a.amap_pos = c->rf_pos.p;
    a.amap_src = c->rf_src.p;
    a.val = c->val.p;
    a.work = c->rf_work.p;
    a.inv = c->rf_inv.p;
    a.inv_off = c->rf_invoff.p;
    a.status = c->rf_status.p;
    return a;
}

} // namespace

// A GMRES run over a fixed number of right-hand-side slots with its own right-hand sides and
// solutions on the device: requests enter free slots between restart cycles (continuous batching).
struct s2l_solver {
    s2l_ctx* c = nullptr;
    std::unique_ptr<GmresRun> run;
    DevBuf<double> B, X;
    std::vector<char> pending;  ///< admitted, started at the next step
    std::vector<char> used;     ///< holds a solve (pending, running, or finished and not taken)
};

// ============================================================================ C-ABI
// A_00 = W^T A W of the uploaded operator and coarse basis into rf_E (column-major n0 x n0), 16
// columns at a time: prolongation of unit vectors, SpMV, restriction (coarse_space.cpp:95-105).
void formCoarseMatrix(s2l_ctx* c, hipStream_t s)
{
    const int32_t n0 = c->n0, G = 16;
    c->rf_E.reserve(static_cast<size_t>(n0) * n0);
    c->rf_Y.reserve(static_cast<size_t>(n0) * G);
    c->rf_C.reserve(static_cast<size_t>(n0) * G);
    c->rf_X.reserve(static_cast<size_t>(c->n) * G);
    c->rf_AX.reserve(static_cast<size_t>(c->n) * G);
    c->rf_part.reserve(static_cast<size_t>(std::max(c->nsub, 1)) * std::max(c->ca.rchunks, 1) * std::max(c->ca.mmax, 1) * G);
    if (!c->boolean) c->rf_loc.reserve(static_cast<size_t>(c->local_total) * G);
    for (int32_t k0 = 0; k0 < n0; k0 += G) {
        const int32_t cnt = std::min(G, n0 - k0);
        launchUnitColumns(c->rf_Y.p, n0, k0, cnt, s);
        if (c->boolean) {
            launchCoarseProlong(c->ca, c->rf_Y.p, c->rf_X.p, c->n, 1, cnt, nullptr, s);
        } else {
            launchCoarseProlong(c->ca, c->rf_Y.p, c->rf_loc.p, c->local_total, 0, cnt, nullptr, s);
            launchRevGather(c->n, c->rg_ptr.p, c->rg_idx.p, c->rf_loc.p, c->local_total, nullptr, c->rf_X.p, cnt, nullptr, s);
        }
        launchSpmv(c->n, c->rp.p, c->col.p, c->val.p, c->rf_X.p, c->rf_AX.p, nullptr, cnt, nullptr, s);
        launchCoarseRestrict(c->ca, c->rf_AX.p, c->n, c->rf_C.p, c->rf_part.p, cnt, nullptr, s);
        launchStoreColumns(c->rf_C.p, n0, k0, cnt, c->rf_E.p, n0, s);
    }
    HIPCHECK(hipGetLastError());
}

// A00^-1 = T2 T1 on the device (rocBLAS dgemm over the row-major factors: in column-major terms
// C = T1^T T2^T = (T2 T1)^T, which is T2 T1 row-major), the padding columns zero.
void coarseInverse(s2l_ctx* c, hipStream_t s)
{
    const int32_t n0 = c->n0, ld = c->tri_ld;
    c->Einv_d.reserve(static_cast<size_t>(n0) * ld);
    HIPCHECK(hipMemsetAsync(c->Einv_d.p, 0, sizeof(double) * static_cast<size_t>(n0) * ld, s));
    if (!c->rf_blas)
        require(rocblas_create_handle(&c->rf_blas) == rocblas_status_success, S2L_EHIP, "coarse inverse: rocblas_create_handle failed");
    require(rocblas_set_stream(c->rf_blas, s) == rocblas_status_success, S2L_EHIP, "coarse inverse: rocblas_set_stream failed");
    const double one = 1.0, zero = 0.0;
    require(rocblas_dgemm(c->rf_blas, rocblas_operation_none, rocblas_operation_none, n0, n0, n0, &one, c->T1_d.p, ld, c->T2_d.p,
                          ld, &zero, c->Einv_d.p, ld) == rocblas_status_success,
            S2L_EHIP, "coarse inverse: rocblas_dgemm failed");
    HIPCHECK(hipGetLastError());
}

// Packs the uploaded local factors for the local solve in @p precision (the body of
// s2l_finalize_local; the autotuner repacks through it).
void finalizeLocal(s2l_ctx* c, int precision)
{
    if (c) c->invalidateGraphs();
    require(c && c->subs_ready, S2L_ESTATE, "finalize_local: upload the subdomains first");
    c->rfForget();
    require(precision >= 0 && precision <= 2, S2L_EINVAL,
            "precision must be 0 (FP64), 1 (FP32) or 2 (FP32 factors, FP64 arithmetic)");
    for (int i = 0; i < c->nsub; ++i) require(c->hf[static_cast<size_t>(i)].n > 0, S2L_ESTATE, "factor of subdomain " + std::to_string(i) + " missing");
    HIPCHECK(hipSetDevice(c->device));
    std::vector<int32_t> moff(c->local_ptr.begin(), c->local_ptr.end() - 1);
    const size_t esz = precision == 1 ? sizeof(float) : sizeof(double);  // arithmetic type (LDS)
    // the level-synchronous schedule is built when launch mode 2 needs it
    c->ls_built = false;
    c->ls = LsPlan{};
    for (DevBuf<unsigned char>* b : {&c->ls_diag, &c->ls_bf, &c->ls_bb}) b->release();
    // task-queue schedule (launch modes 0 and 1): only when a subdomain vector fits the LDS
    // the subtree tasks of all subdomains run in one launch, one workgroup per compute unit
    const int tasks = c->tasks_per_sub > 0 ? c->tasks_per_sub : std::max(1, c->cus / std::max(1, c->nsub));
    c->pk = packFactors(c->hf, moff, tasks, c->dense_top_max);
    Packed& P = c->pk;
    for (int i = 0; i < c->nsub; ++i) {
        P.sub[static_cast<size_t>(i)].n_owned = c->ownership[i + 1] - c->ownership[i];
        P.sub[static_cast<size_t>(i)].begin = c->ownership[i];
    }
    // LDS: the subdomain vector, the result buffer of one step, the staged descriptors, sized to the
    // packed steps (the rest holds slots). The result buffer keeps at least kMinStepRows rows: the
    // row groups of a backward step (G x w <= threads of a workgroup) use it too.
    const size_t yRows = static_cast<size_t>(std::max(P.max_step_rows, kMinStepRows));
    c->st_cap = std::max(P.max_step_sn, 1);
    const size_t stBytes = static_cast<size_t>(c->st_cap) * (sizeof(SnDesc) + 2 * sizeof(int32_t));
    c->y_off = (static_cast<size_t>(P.max_n) * esz + 15) / 16 * 16;
    c->st_off = c->y_off + (yRows * esz + 15) / 16 * 16;
    c->step_off = (c->st_off + stBytes + 15) / 16 * 16;
    c->lds_bytes = c->step_off + static_cast<size_t>(kStepBlock) * sizeof(StepDesc);
    // two 512-thread workgroups share a compute unit's LDS; when half of it does not hold the
    // layout, the workgroups take 1024 threads
    c->local_threads = c->local_threads_req;
    if (c->local_threads == 512 && c->lds_bytes > c->max_lds / 2) c->local_threads = 1024;
    const size_t lds_max = c->local_threads == 512 ? c->max_lds / 2 : c->max_lds;
    c->queue_ok = c->lds_bytes <= lds_max;
    // the slots of the largest task without child tasks that fit in the rest of the LDS
    auto slotRoom = [&](size_t base, size_t elem, size_t& off, int32_t& cap) {
        off = (base + 15) / 16 * 16;
        const size_t room = lds_max > off ? (lds_max - off) / elem : 0;
        cap = static_cast<int32_t>(std::min<size_t>(room, static_cast<size_t>(P.max_task_slots)));
        return off + static_cast<size_t>(cap) * elem;
    };
    if (c->queue_ok) c->lds_bytes = slotRoom(c->lds_bytes, esz, c->slot_off, c->slot_cap);
    // pairs of right-hand sides: a float pair per row (FP32), a double pair (FP64 arithmetic of FP32 entries)
    const size_t psz = precision == 2 ? sizeof(double2) : sizeof(float2);
    c->y_off2 = (static_cast<size_t>(P.max_n) * psz + 15) / 16 * 16;
    c->st_off2 = c->y_off2 + (yRows * psz + 15) / 16 * 16;
    c->step_off2 = (c->st_off2 + stBytes + 15) / 16 * 16;
    c->lds_bytes2 = c->step_off2 + static_cast<size_t>(kStepBlock) * sizeof(StepDesc);
    c->pair_ok = (precision == 1 || precision == 2) && c->queue_ok && c->lds_bytes2 <= lds_max;
    if (c->pair_ok) c->lds_bytes2 = slotRoom(c->lds_bytes2, psz, c->slot_off2, c->slot_cap2);
    if (c->pair_ok)
        HIPCHECK(hipFuncSetAttribute(precision == 2 ? localSolveKernel<double2, float>(c->local_threads)
                                                    : localSolveKernel<float2, float>(c->local_threads),
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     static_cast<int>(c->lds_bytes2)));
    if (c->queue_ok) {
        const void* kern = precision == 1   ? localSolveKernel<float, float>(c->local_threads)
                           : precision == 2 ? localSolveKernel<double, float>(c->local_threads)
                                            : localSolveKernel<double, double>(c->local_threads);
        HIPCHECK(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(c->lds_bytes)));
    } else if (c->launch_mode != 2) {
        require(precision != 2, S2L_ELIMIT,
                "FP32 factors with FP64 arithmetic need the task queue, whose subdomain vector does not fit the LDS");
        c->launch_mode = 2;  // the task queue cannot hold these subdomains in LDS
    }
    if (c->launch_mode == 2) buildLs(c, precision);
    // maps: gmap[p] = global row of factor position p, lperm[p] = its local index
    std::vector<int32_t> gmap(c->gid_h.size()), lperm(c->gid_h.size());
    for (int i = 0; i < c->nsub; ++i) {
        const int32_t off = c->local_ptr[i];
        const HostFactor& F = c->hf[static_cast<size_t>(i)];
        for (int32_t p = 0; p < F.n; ++p) {
            lperm[static_cast<size_t>(off + p)] = F.perm[static_cast<size_t>(p)];
            gmap[static_cast<size_t>(off + p)] = c->gid_h[static_cast<size_t>(off + F.perm[static_cast<size_t>(p)])];
        }
    }
    c->sub_d.upload(P.sub, c->stream);
    c->levw_d.upload(P.lev_wpre, c->stream);
    c->levh_d.upload(P.lev_hpre, c->stream);
    c->levdesc_d.upload(P.lev_desc, c->stream);
    c->levrow_d.upload(P.lev_row, c->stream);
    c->levcb_d.upload(P.lev_cb, c->stream);
    c->levce_d.upload(P.lev_ce, c->stream);
    c->steps_d.upload(P.steps, c->stream);
    c->ridx_d.upload(P.ridx, c->stream);
    c->cptr_d.upload(P.cptr, c->stream);
    c->cend_d.upload(P.cend, c->stream);
    c->cidx_d.upload(P.cidx, c->stream);
    c->gmap_d.upload(gmap, c->stream);
    c->lperm_d.upload(lperm, c->stream);
    c->task_d.upload(P.task, c->stream);
    c->tchild_d.upload(P.task_child, c->stream);
    c->trows_d.upload(P.task_rows, c->stream);
    c->ext_d.upload(P.ext_rows, c->stream);
    c->queue_d.upload(P.queue, c->stream);
    c->qsubq_d.upload(P.queue_sub, c->stream);
    c->qsub_h = P.qsub;
    {
        // the row table of the dense tops and the depth groups made of dense tops only
        std::vector<int4> dt;
        std::vector<int32_t> dtce, dtb(P.task.size(), -1);
        int32_t maxS = 0;
        for (const DenseTop& d : P.dtop) {
            const TaskDesc& td = P.task[static_cast<size_t>(d.task)];
            const SubDesc& sd = P.sub[static_cast<size_t>(td.sub)];
            dtb[static_cast<size_t>(d.task)] = static_cast<int32_t>(dt.size());
            for (int32_t q = 0; q < td.rows_count; ++q) {
                const int32_t row = P.task_rows[static_cast<size_t>(td.rows_first + q)];
                const size_t m = static_cast<size_t>(sd.map_off + row);
                dt.push_back(int4{row, gmap[m], lperm[m], P.cptr[static_cast<size_t>(sd.cptr_off + row)]});
                dtce.push_back(P.cend[static_cast<size_t>(sd.cptr_off + row)]);
            }
            maxS = std::max(maxS, td.rows_count);
        }
        c->qgroup_dense.assign(P.qgroup.size() > 0 ? P.qgroup.size() - 1 : 0, 0);
        for (size_t g = 0; g + 1 < P.qgroup.size(); ++g) {
            bool all = P.qgroup[g + 1] > P.qgroup[g];
            for (int32_t k = P.qgroup[g]; k < P.qgroup[g + 1] && all; ++k) {
                const int32_t tk = codeTask(P.queue[static_cast<size_t>(k)]);
                all = dtb[static_cast<size_t>(tk)] >= 0 && P.task[static_cast<size_t>(tk)].parent < 0;
            }
            c->qgroup_dense[g] = all ? 1 : 0;
        }
        c->dtrow_d.upload(dt, c->stream);
        c->dtce_d.upload(dtce, c->stream);
        c->dtb_d.upload(dtb, c->stream);
        c->dt_maxS = maxS;
    }
    if (!c->queue_ok) {
        c->diag_f.release();
        c->below_f.release();
        c->diag_dd.release();
        c->below_dd.release();
    } else if (precision != 0) {
        std::vector<float> d(P.diag.begin(), P.diag.end()), b(P.below.begin(), P.below.end());
        c->diag_f.upload(d, c->stream);
        c->below_f.upload(b, c->stream);
        HIPCHECK(hipStreamSynchronize(c->stream));
        c->diag_dd.release();
        c->below_dd.release();
    } else {
        c->diag_dd.upload(P.diag, c->stream);
        c->below_dd.upload(P.below, c->stream);
        HIPCHECK(hipStreamSynchronize(c->stream));
        c->diag_f.release();
        c->below_f.release();
    }
    // the host copies of the values are no longer needed
    std::vector<double>().swap(P.diag);
    std::vector<double>().swap(P.below);
    c->precision = precision;
    c->work_nrhs = 0;
    c->local_ready = true;
    c->ls_stale = false;
    // the symbolic plan of a later device refactorization, on a host thread meanwhile
    bool chol = true;
    for (const HostFactor& F : c->hf) chol = chol && F.cholesky;
    if (chol && c->queue_ok && !c->rp_h.empty())
        c->rf_pending = std::async(std::launch::async, [c] {
            return buildRefactorPlan(c->hf, c->local_ptr, c->gid_h, c->rp_h, c->col_h, c->pk);
        });
}

extern "C" {

const char* s2l_last_error(const s2l_ctx* ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

int s2l_ctx_create(int device, s2l_ctx** out)
{
    return guarded(nullptr, [&] {
        require(out != nullptr, S2L_EINVAL, "out is null");
        int count = 0;
        HIPCHECK(hipGetDeviceCount(&count));
        require(device >= 0 && device < count, S2L_EINVAL, "no HIP device " + std::to_string(device));
        HIPCHECK(hipSetDevice(device));
        auto* c = new s2l_ctx;
        c->device = device;
        try {
            {
                // the device address of this GPU's scratch array, looked up once per device
                static std::mutex mu;
                static std::vector<void*> per_device;
                std::lock_guard<std::mutex> lock(mu);
                if (per_device.size() <= static_cast<size_t>(device)) per_device.resize(static_cast<size_t>(device) + 1, nullptr);
                if (!per_device[static_cast<size_t>(device)]) {
                    void* ptr = nullptr;
                    HIPCHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
                    per_device[static_cast<size_t>(device)] = ptr;
                }
                c->scratch = per_device[static_cast<size_t>(device)];
            }
            HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
            HIPCHECK(hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking));
            HIPCHECK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
            HIPCHECK(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
            int lds = 0, cus = 0;
            HIPCHECK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
            HIPCHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
            c->max_lds = static_cast<size_t>(lds);
            c->cus = cus;
        } catch (...) {
            delete c;
            throw;
        }
        *out = c;
    });
}

int s2l_ctx_destroy(s2l_ctx* ctx)
{
    if (!ctx) return S2L_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto& pr : ctx->ev_pending) {
        (void)hipEventDestroy(pr.first);
        (void)hipEventDestroy(pr.second);
    }
    for (auto& m : ctx->phase_pending) {
        (void)hipEventDestroy(m.e0);
        (void)hipEventDestroy(m.e1);
    }
    for (auto e : ctx->ev_pool) (void)hipEventDestroy(e);
    ctx->rfJoin();
    ctx->invalidateGraphs();
    if (ctx->aux) (void)hipStreamSynchronize(ctx->aux);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    if (ctx->aux) (void)hipStreamDestroy(ctx->aux);
    if (ctx->rf_blas) (void)rocblas_destroy_handle(ctx->rf_blas);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return S2L_OK;
}

int s2l_upload_system(s2l_ctx* c, int32_t n, const int32_t* rowptr, const int32_t* col, const double* val, int32_t nsub,
                      const int32_t* ownership)
{
    return guarded(c, [&] {
        if (c) c->invalidateGraphs();
        require(c && n > 0 && rowptr && col && val && nsub >= 1 && ownership, S2L_EINVAL, "upload_system: bad arguments");
        require(ownership[0] == 0 && ownership[nsub] == n, S2L_EINVAL, "ownership must cover [0, n)");
        for (int p = 0; p < nsub; ++p) require(ownership[p + 1] > ownership[p], S2L_EINVAL, "empty subdomain in ownership");
        HIPCHECK(hipSetDevice(c->device));
        const int64_t nnz = rowptr[n];
        std::vector<int32_t> cl(static_cast<size_t>(nnz)), inpos(static_cast<size_t>(nnz));
        std::vector<double> vl(static_cast<size_t>(nnz));
        std::vector<std::pair<int64_t, int64_t>> key;
        int p = 0;
        for (int32_t i = 0; i < n; ++i) {
            while (i >= ownership[p + 1]) ++p;
            const int32_t b = ownership[p], e = ownership[p + 1];
            key.clear();
            for (int32_t k = rowptr[i]; k < rowptr[i + 1]; ++k) {
                const int32_t cc = col[k];
                require(cc >= 0 && cc < n, S2L_EINVAL, "column index out of range");
                const bool own = cc >= b && cc < e;
                key.emplace_back((own ? 0 : (int64_t)1 << 40) + cc, k);
            }
            std::sort(key.begin(), key.end());
            int64_t q = rowptr[i];
            for (const auto& kv : key) {
                cl[static_cast<size_t>(q)] = col[kv.second];
                vl[static_cast<size_t>(q)] = val[kv.second];
                inpos[static_cast<size_t>(q)] = static_cast<int32_t>(kv.second);
                ++q;
            }
        }
        // the same pattern (row pointers and every column) and ownership refresh the values only;
        // anything else is a new system whose subdomains, factors and coarse space follow
        if (c->system_ready && c->n == n && c->nnz == nnz && c->nsub == nsub &&
            std::equal(c->ownership.begin(), c->ownership.end(), ownership) &&
            std::equal(c->rp_h.begin(), c->rp_h.end(), rowptr) && c->col_h == cl) {
            c->val.upload(vl, c->stream);
        } else {
            c->n = n;
            c->nnz = nnz;
            c->nsub = nsub;
            c->ownership.assign(ownership, ownership + nsub + 1);
            c->rp_h.assign(rowptr, rowptr + n + 1);
            c->col_h = cl;
            c->inpos_h = std::move(inpos);
            c->rfForget();
            c->rp.upload(rowptr, static_cast<size_t>(n) + 1, c->stream);
            c->col.upload(cl, c->stream);
            c->val.upload(vl, c->stream);
            c->subs_ready = c->local_ready = c->coarse_ready = false;
            c->work_nrhs = 0;
            c->gm_cache.reset();
        }
        HIPCHECK(hipStreamSynchronize(c->stream));
        c->system_ready = true;
    });
}

int s2l_upload_subdomains(s2l_ctx* c, int32_t nsub, const int32_t* local_ptr, const int32_t* gid, const double* pou,
                          int one_level)
{
    return guarded(c, [&] {
        if (c) c->invalidateGraphs();
        require(c && c->system_ready, S2L_ESTATE, "upload_subdomains: upload the system first");
        c->rfForget();
        require(nsub == c->nsub && local_ptr && gid, S2L_EINVAL, "upload_subdomains: bad arguments");
        require(one_level == 0 || one_level == 1, S2L_EINVAL, "one_level must be 0 (RAS) or 1 (ASM)");
        HIPCHECK(hipSetDevice(c->device));
        const int64_t tot = local_ptr[nsub];
        c->local_ptr.assign(local_ptr, local_ptr + nsub + 1);
        c->gid_h.assign(gid, gid + tot);
        // owned rows first, in order; ghosts sorted, outside the owned range
        for (int i = 0; i < nsub; ++i) {
            const int32_t b = c->ownership[i], e = c->ownership[i + 1], l0 = local_ptr[i], l1 = local_ptr[i + 1];
            require(l1 - l0 >= e - b, S2L_EINVAL, "subdomain smaller than its owned rows");
            for (int32_t l = 0; l < e - b; ++l) require(gid[l0 + l] == b + l, S2L_EINVAL, "owned rows must come first, in order");
            for (int32_t l = l0 + (e - b); l < l1; ++l) {
                require(gid[l] >= 0 && gid[l] < c->n && (gid[l] < b || gid[l] >= e), S2L_EINVAL, "bad ghost id");
                if (l > l0 + (e - b)) require(gid[l] > gid[l - 1], S2L_EINVAL, "ghosts must be sorted by global id");
            }
        }
        c->boolean = (pou == nullptr);
        c->one_level = one_level;
        c->local_total = tot;
        if (pou) {
            c->pou_h.assign(pou, pou + tot);
            c->pou_d.upload(c->pou_h, c->stream);
        }
        // reverse-add plan: every global row lists the owner's copy, then the ghost copies of the
        // other subdomains in increasing subdomain order (halo_exchange.cpp:98-106)
        std::vector<int32_t> cnt(static_cast<size_t>(c->n) + 1, 1);
        cnt[0] = 0;
        for (int i = 0; i < nsub; ++i) {
            const int32_t nI = c->ownership[i + 1] - c->ownership[i];
            for (int32_t l = local_ptr[i] + nI; l < local_ptr[i + 1]; ++l) ++cnt[static_cast<size_t>(gid[l]) + 1];
        }
        std::vector<int32_t> ptr(static_cast<size_t>(c->n) + 1, 0);
        for (int32_t g = 0; g < c->n; ++g) ptr[static_cast<size_t>(g) + 1] = ptr[static_cast<size_t>(g)] + cnt[static_cast<size_t>(g) + 1];
        std::vector<int32_t> idx(static_cast<size_t>(ptr[static_cast<size_t>(c->n)]));
        std::vector<int32_t> next(ptr.begin(), ptr.end() - 1);
        for (int i = 0; i < nsub; ++i) {
            const int32_t b = c->ownership[i], nI = c->ownership[i + 1] - b;
            for (int32_t l = 0; l < nI; ++l) idx[static_cast<size_t>(next[static_cast<size_t>(b + l)]++)] = local_ptr[i] + l;
        }
        for (int i = 0; i < nsub; ++i) {
            const int32_t nI = c->ownership[i + 1] - c->ownership[i];
            for (int32_t l = local_ptr[i] + nI; l < local_ptr[i + 1]; ++l) idx[static_cast<size_t>(next[static_cast<size_t>(gid[l])]++)] = l;
        }
        c->rg_ptr.upload(ptr, c->stream);
        c->rg_idx.upload(idx, c->stream);
        HIPCHECK(hipStreamSynchronize(c->stream));
        c->hf.assign(static_cast<size_t>(nsub), HostFactor{});
        c->subs_ready = true;
        c->local_ready = c->coarse_ready = false;
        c->work_nrhs = 0;
    });
}

int s2l_upload_local_factor(s2l_ctx* c, int32_t i, int32_t n, const int32_t* perm, const int64_t* l_colptr, const int32_t* l_row,
                            const double* l_val, const int64_t* ut_colptr, const int32_t* ut_row, const double* ut_val)
{
    return guarded(c, [&] {
        if (c) c->invalidateGraphs();
        require(c && c->subs_ready, S2L_ESTATE, "upload_local_factor: upload the subdomains first");
        c->rfForget();
        require(i >= 0 && i < c->nsub, S2L_EINVAL, "subdomain index out of range");
        require(n == c->local_ptr[i + 1] - c->local_ptr[i], S2L_EINVAL, "factor size differs from n_i");
        require(perm && l_colptr && l_row && l_val, S2L_EINVAL, "upload_local_factor: null factor");
        HostFactor& F = c->hf[static_cast<size_t>(i)];
        F.n = n;
        F.perm.assign(perm, perm + n);
        std::vector<char> seen(static_cast<size_t>(n), 0);
        for (int32_t k = 0; k < n; ++k) {
            require(perm[k] >= 0 && perm[k] < n && !seen[static_cast<size_t>(perm[k])], S2L_EINVAL, "perm is not a permutation");
            seen[static_cast<size_t>(perm[k])] = 1;
        }
        F.lcp.assign(l_colptr, l_colptr + n + 1);
        F.lr.assign(l_row, l_row + l_colptr[n]);
        F.lv.assign(l_val, l_val + l_colptr[n]);
        for (double v : F.lv) require(std::isfinite(v), S2L_ENUMERIC, "upload_local_factor: non-finite factor entry");
        F.cholesky = (ut_colptr == nullptr);
        if (!F.cholesky) {
            require(ut_row && ut_val, S2L_EINVAL, "upload_local_factor: incomplete U^T");
            F.ucp.assign(ut_colptr, ut_colptr + n + 1);
            F.ur.assign(ut_row, ut_row + ut_colptr[n]);
            F.uv.assign(ut_val, ut_val + ut_colptr[n]);
            for (double v : F.uv) require(std::isfinite(v), S2L_ENUMERIC, "upload_local_factor: non-finite factor entry");
        } else {
            F.ucp.clear();
            F.ur.clear();
            F.uv.clear();
        }
        c->local_ready = false;
    });
}

int s2l_finalize_local(s2l_ctx* c, int precision)
{
    return guarded(c, [&] { finalizeLocal(c, precision); });
}

int s2l_upload_coarse(s2l_ctx* c, int32_t n0, const int64_t* offsets, const double* V, const double* E, int symmetric,
                      int correction)
{
    return guarded(c, [&] {
        if (c) c->invalidateGraphs();
        require(c && c->subs_ready, S2L_ESTATE, "upload_coarse: upload the subdomains first");
        require(n0 >= 0 && offsets && (correction == 0 || correction == 1), S2L_EINVAL, "upload_coarse: bad arguments");
        require(offsets[0] == 0 && offsets[c->nsub] == n0, S2L_EINVAL, "coarse offsets must cover [0, n0)");
        HIPCHECK(hipSetDevice(c->device));
        c->n0 = n0;
        c->correction = correction;
        c->off_h.assign(offsets, offsets + c->nsub + 1);
        if (n0 > 0) {
            require(V, S2L_EINVAL, "upload_coarse: null V");
            require(E || c->system_ready, S2L_ESTATE, "upload_coarse: E = NULL forms A_00 from the operator: upload the system first");
            std::vector<int64_t> voff(static_cast<size_t>(c->nsub));
            std::vector<int32_t> vrows(static_cast<size_t>(c->nsub)), vrowptr(static_cast<size_t>(c->nsub)), vrowg, bsub, brow;
            int64_t vtot = 0;
            int mmax = 0;
            for (int i = 0; i < c->nsub; ++i) {
                const int32_t nI = c->ownership[i + 1] - c->ownership[i];
                const int32_t rows = c->boolean ? nI : (c->local_ptr[i + 1] - c->local_ptr[i]);
                const int64_t m = offsets[i + 1] - offsets[i];
                require(m >= 0, S2L_EINVAL, "coarse offsets must be nondecreasing");
                mmax = std::max<int>(mmax, static_cast<int>(m));
                voff[static_cast<size_t>(i)] = vtot;
                vrows[static_cast<size_t>(i)] = rows;
                vrowptr[static_cast<size_t>(i)] = static_cast<int32_t>(vrowg.size());
                for (int32_t l = 0; l < rows; ++l) vrowg.push_back(c->gid_h[static_cast<size_t>(c->local_ptr[i] + l)]);
                vtot += static_cast<int64_t>(rows) * m;
                // every subdomain gets prolongation blocks: those without coarse vectors write zeros
                for (int32_t r0 = 0; r0 < rows; r0 += 256) {
                    bsub.push_back(i);
                    brow.push_back(r0);
                }
            }
            c->V_d.upload(V, static_cast<size_t>(vtot), c->stream);
            c->off_d.upload(c->off_h, c->stream);
            c->voff_d.upload(voff, c->stream);
            c->vrows_d.upload(vrows, c->stream);
            c->vrowptr_d.upload(vrowptr, c->stream);
            c->vrowg_d.upload(vrowg, c->stream);
            {
                // blocks whose rows are consecutive global rows (the owned rows of a Boolean
                // partition after the renumbering): the restriction reads r without the row map
                std::vector<int32_t> vbase(static_cast<size_t>(c->nsub), -1);
                for (int i = 0; i < c->nsub; ++i) {
                    const int32_t p0 = vrowptr[static_cast<size_t>(i)], rows = vrows[static_cast<size_t>(i)];
                    bool run = rows > 0;
                    for (int32_t l = 1; l < rows && run; ++l)
                        run = vrowg[static_cast<size_t>(p0 + l)] == vrowg[static_cast<size_t>(p0)] + l;
                    if (run) vbase[static_cast<size_t>(i)] = vrowg[static_cast<size_t>(p0)];
                }
                c->vbase_d.upload(vbase, c->stream);
            }
            c->blksub_d.upload(bsub, c->stream);
            c->blkrow_d.upload(brow, c->stream);
            CoarseArgs& a = c->ca;
            a.nsub = c->nsub;
            a.n0 = n0;
            a.mmax = mmax;
            int32_t maxrows = 0;
            for (int32_t v : vrows) maxrows = std::max(maxrows, v);
            a.rchunks = std::max(1, (maxrows + 255) / 256);
            a.off = c->off_d.p;
            a.voff = c->voff_d.p;
            a.vrows = c->vrows_d.p;
            a.vrow_ptr = c->vrowptr_d.p;
            a.vrow_g = c->vrowg_d.p;
            a.vbase = c->vbase_d.p;
            a.V = c->V_d.p;
            a.blk_sub = c->blksub_d.p;
            a.blk_row = c->blkrow_d.p;
            a.nblk = static_cast<int32_t>(bsub.size());
            // A_00 given (row-major), or formed here on the device from the operator and the basis
            std::vector<double> Eh;
            if (!E) {
                formCoarseMatrix(c, c->stream);
                std::vector<double> Ec(static_cast<size_t>(n0) * n0);
                HIPCHECK(hipMemcpyAsync(Ec.data(), c->rf_E.p, sizeof(double) * Ec.size(), hipMemcpyDeviceToHost, c->stream));
                HIPCHECK(hipStreamSynchronize(c->stream));
                Eh.resize(Ec.size());
                for (int32_t i = 0; i < n0; ++i)
                    for (int32_t j = 0; j < n0; ++j)
                        Eh[static_cast<size_t>(i) * n0 + j] = Ec[static_cast<size_t>(j) * n0 + i];  // row-major
                E = Eh.data();
            }
            // A_00: Cholesky (symmetric) or LU with partial pivoting, FP64; the device applies the
            // inverses of the triangular factors as two triangular matrix-vector products.
            std::vector<double> M(E, E + static_cast<size_t>(n0) * n0);
            for (double v : M) require(std::isfinite(v), S2L_ENUMERIC, "A_00 has non-finite entries");
            std::vector<double> T1, T2(static_cast<size_t>(n0) * n0, 0.0);
            bool chol = symmetric && denseCholesky(n0, M);
            c->coarse_perm = false;
            if (chol) {
                T1 = lowerInverse(n0, M, false);
                for (int i = 0; i < n0; ++i)
                    for (int j = i; j < n0; ++j) T2[static_cast<size_t>(i) * n0 + j] = T1[static_cast<size_t>(j) * n0 + i];
            } else {
                M.assign(E, E + static_cast<size_t>(n0) * n0);
                std::vector<int32_t> piv;
                require(denseLu(n0, M, piv), S2L_ENUMERIC, "A_00 is singular");
                T1 = lowerInverse(n0, M, true);
                // U^{-1} = (L'^{-1})^T with L' = U^T (lower)
                std::vector<double> Ut(static_cast<size_t>(n0) * n0, 0.0);
                for (int i = 0; i < n0; ++i)
                    for (int j = i; j < n0; ++j) Ut[static_cast<size_t>(j) * n0 + i] = M[static_cast<size_t>(i) * n0 + j];
                const std::vector<double> X = lowerInverse(n0, Ut, false);
                for (int i = 0; i < n0; ++i)
                    for (int j = i; j < n0; ++j) T2[static_cast<size_t>(i) * n0 + j] = X[static_cast<size_t>(j) * n0 + i];
                c->cperm_d.upload(piv, c->stream);
                c->coarse_perm = true;
            }
            for (double v : T1) require(std::isfinite(v), S2L_ENUMERIC, "A_00 factor is not finite");
            // rows padded with zeros to a multiple of eight entries (64-byte aligned rows)
            c->tri_ld = triLd(n0);
            std::vector<double> Tp(static_cast<size_t>(n0) * c->tri_ld, 0.0);
            for (int i = 0; i < n0; ++i)
                std::copy(T1.begin() + static_cast<size_t>(i) * n0, T1.begin() + static_cast<size_t>(i + 1) * n0,
                          Tp.begin() + static_cast<size_t>(i) * c->tri_ld);
            c->T1_d.upload(Tp, c->stream);
            HIPCHECK(hipStreamSynchronize(c->stream));
            for (int i = 0; i < n0; ++i)
                std::copy(T2.begin() + static_cast<size_t>(i) * n0, T2.begin() + static_cast<size_t>(i + 1) * n0,
                          Tp.begin() + static_cast<size_t>(i) * c->tri_ld);
            c->T2_d.upload(Tp, c->stream);
            coarseInverse(c, c->stream);
            HIPCHECK(hipStreamSynchronize(c->stream));
        }
        c->work_nrhs = 0;
        c->coarse_ready = true;
    });
}

int s2l_local_eigensolve(s2l_ctx* c, int32_t nev, int32_t ncv, double shift_rel, double tol, int32_t max_restarts,
                         const int32_t* skip, double* lambda, double* Z, int32_t* converged, int32_t* info)
{
    return guarded(c, [&] {
        require(c && lambda && Z, S2L_EINVAL, "local_eigensolve: bad arguments");
        require(c->system_ready && c->subs_ready && c->local_ready, S2L_ESTATE,
                "local_eigensolve: upload the system, subdomains and local factors first");
        for (const HostFactor& F : c->hf) require(F.cholesky, S2L_ESTATE, "local_eigensolve: needs Cholesky factors (symmetric A)");
        require(c->queue_ok && c->launch_mode != 2, S2L_ESTATE, "local_eigensolve: needs the task-queue schedule (launch modes 0 and 1)");
        require(c->precision == 0, S2L_ESTATE, "local_eigensolve: needs FP64 local factors (precision 0)");
        require(nev >= 1 && ncv >= nev + 2 && ncv <= kEigMaxM, S2L_EINVAL,
                "local_eigensolve: need 1 <= nev, nev + 2 <= ncv <= " + std::to_string(kEigMaxM));
        require(tol > 0.0 && shift_rel > 0.0 && max_restarts >= 1, S2L_EINVAL, "local_eigensolve: tol, shift_rel > 0, max_restarts >= 1");
        const int nsub = c->nsub;
        for (int i = 0; i < nsub; ++i) {
            const int32_t ni = c->local_ptr[static_cast<size_t>(i) + 1] - c->local_ptr[static_cast<size_t>(i)];
            require((skip && skip[i]) || ni >= ncv + 1, S2L_EINVAL,
                    "local_eigensolve: subdomain " + std::to_string(i) + " has " + std::to_string(ni) +
                        " rows; the Krylov basis needs ncv + 1 (skip it)");
        }
        HIPCHECK(hipSetDevice(c->device));
        hipStream_t s = c->stream;
        c->invalidateGraphs();
        ensureRfPlan(c, s);
        const int64_t nl = c->local_total;
        if (!c->eg_ready) {
            c->eg = buildEigPlan(c->local_ptr, c->gid_h, c->ownership, c->rp_h, c->col_h, c->pou_h, c->boolean);
            EigPlan& E = c->eg;
            c->eg_chunk.upload(E.chunk, s);
            c->eg_cptr.upload(E.chunk_ptr, s);
            c->eg_lrp.upload(E.lrp, s);
            c->eg_lcol.upload(E.lcol, s);
            c->eg_lsrc.upload(E.lsrc, s);
            c->eg_orp.upload(E.orp, s);
            c->eg_osrc.upload(E.osrc, s);
            c->eg_ghost.upload(E.ghost, s);
            c->eg_D.upload(E.D, s);
            c->eg_lptr.upload(c->local_ptr, s);
            HIPCHECK(hipStreamSynchronize(s));
            c->eg_ready = true;
        }
        const int m = ncv, kk = std::max(nev + 1, (nev + ncv) / 2);
        const int nchunk = static_cast<int>(c->eg.chunk.size() / 3);
        const size_t ld = static_cast<size_t>(kEigMaxM) + 1;
        c->eg_adj.reserve(static_cast<size_t>(nl));
        c->eg_rowsum.reserve(static_cast<size_t>(nl));
        c->eg_sigma.reserve(static_cast<size_t>(nsub));
        c->eg_diag.reserve(c->diag_dd.n);
        c->eg_below.reserve(c->below_dd.n);
        c->eg_V.reserve(static_cast<size_t>(m + 1) * nl);
        c->eg_AV.reserve(static_cast<size_t>(m + 1) * nl);
        c->eg_H.reserve(static_cast<size_t>(nsub) * m * m);
        c->eg_Hw.reserve(static_cast<size_t>(nsub) * m * m);
        c->eg_theta.reserve(static_cast<size_t>(nsub) * m);
        c->eg_E.reserve(static_cast<size_t>(nsub) * m);
        c->eg_part.reserve(static_cast<size_t>(nchunk) * ld);
        c->eg_hb.reserve(static_cast<size_t>(nsub) * ld);
        c->eg_hb2.reserve(static_cast<size_t>(nsub) * ld);
        c->eg_beta.reserve(static_cast<size_t>(nsub));
        c->eg_u.reserve(static_cast<size_t>(nl));
        c->eg_w.reserve(static_cast<size_t>(nl));
        c->eg_Aw.reserve(static_cast<size_t>(nl));
        c->eg_lam.reserve(static_cast<size_t>(nsub) * nev);
        c->eg_done.reserve(static_cast<size_t>(nsub));
        c->eg_conv.reserve(static_cast<size_t>(nsub));
        c->eg_live.reserve(1);
        c->eg_info.reserve(static_cast<size_t>(nsub));
        std::vector<int32_t> done0(static_cast<size_t>(nsub), 0);
        for (int i = 0; i < nsub; ++i) done0[static_cast<size_t>(i)] = (skip && skip[i]) ? 1 : 0;
        HIPCHECK(hipMemcpyAsync(c->eg_done.p, done0.data(), sizeof(int32_t) * nsub, hipMemcpyHostToDevice, s));
        HIPCHECK(hipMemsetAsync(c->eg_H.p, 0, sizeof(double) * nsub * m * m, s));
        HIPCHECK(hipMemsetAsync(c->eg_V.p, 0, sizeof(double) * static_cast<size_t>(m + 1) * nl, s));  // skipped subdomains stay 0
        HIPCHECK(hipMemsetAsync(c->eg_lam.p, 0, sizeof(double) * nsub * nev, s));
        EigArgs A{};
        A.nsub = nsub;
        A.nchunk = nchunk;
        A.m = m;
        A.nl = nl;
        A.lptr = c->eg_lptr.p;
        A.chunk = c->eg_chunk.p;
        A.chunk_ptr = c->eg_cptr.p;
        A.lrp = c->eg_lrp.p;
        A.lcol = c->eg_lcol.p;
        A.lsrc = c->eg_lsrc.p;
        A.orp = c->eg_orp.p;
        A.osrc = c->eg_osrc.p;
        A.ghost = c->eg_ghost.p;
        A.D = c->eg_D.p;
        A.val = c->val.p;
        A.adj = c->eg_adj.p;
        A.done = c->eg_done.p;

        // factors of A~_ii + sigma_i I on the structure of A_ii's (local_eigensolver.cpp:285-303:
        // the shift grows by 1e3 when a factorization fails, at most four attempts)
        RfArgs ra = rfArgs(c);
        ra.diag_adj = c->eg_adj.p;
        ra.lperm = c->lperm_d.p;
        ra.sub_off = c->eg_lptr.p;
        const RfPlan& R = c->rf;
        int attempt = 0;
        for (;; ++attempt) {
            launchEigLump(A, c->eg_rowsum.p, s);
            launchEigShift(A, c->eg_rowsum.p, shift_rel, c->eg_sigma.p, s);
            launchEigAddShift(A, c->eg_sigma.p, std::pow(1e3, attempt), s);
            const int32_t zero = 0;
            HIPCHECK(hipMemcpyAsync(c->rf_status.p, &zero, sizeof(int32_t), hipMemcpyHostToDevice, s));
            for (size_t h = 0; h + 1 < R.level_ptr.size(); ++h)
                launchRfLevel(ra, c->rf_order.p + R.level_ptr[h], R.level_ptr[h + 1] - R.level_ptr[h], s);
            launchRfPack(ra, static_cast<int32_t>(R.sn.size()), c->eg_diag.p, c->eg_below.p, false, s);
            HIPCHECK(hipGetLastError());
            int32_t status = 0;
            HIPCHECK(hipMemcpyAsync(&status, c->rf_status.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
            HIPCHECK(hipStreamSynchronize(s));
            if (status == 0) break;
            require(attempt < 3, S2L_ENUMERIC, "local_eigensolve: cannot factorize A~_ii + sigma I (supernode " + std::to_string(status - 1) + ")");
        }
        rfDenseTops(c, ra, c->eg_diag.p, false, s);

        // the operator's local solve: local input and output, the shifted factors
        c->ensureWork(1);
        LocalArgs<double, double> la = localArgs<double, double>(c, c->eg_u.p, nullptr, c->eg_w.p, 1, 1, nullptr);
        la.diag = c->eg_diag.p;
        la.below = c->eg_below.p;
        la.pou = nullptr;
        la.local_in = 1;
        auto solve = [&] {
            for (size_t g = 0; g + 1 < c->pk.qgroup.size(); ++g)
                launchQueue<double, double>(c, la, c->queue_d.p + c->pk.qgroup[g], c->pk.qgroup[g + 1] - c->pk.qgroup[g], false, s);
        };
        double* V = c->eg_V.p;
        double* AV = c->eg_AV.p;
        double* u = c->eg_u.p;
        double* w = c->eg_w.p;
        double* Aw = c->eg_Aw.p;
        // v_0: a seeded vector normalized in the (A~ + sigma I) norm
        launchEigStart(A, w, s);
        launchEigSpmv(A, 0, w, Aw, s);
        launchEigDots(A, Aw, 1, w, c->eg_part.p, s);
        launchEigDotsFinal(A, c->eg_part.p, 1, 2, nullptr, nullptr, nullptr, 0, c->eg_beta.p, s);
        launchEigScale(A, V, AV, 0, w, Aw, c->eg_beta.p, s);
        if (!c->rf_blas)
            require(rocblas_create_handle(&c->rf_blas) == rocblas_status_success, S2L_EHIP, "local_eigensolve: rocblas_create_handle failed");
        require(rocblas_set_stream(c->rf_blas, s) == rocblas_status_success, S2L_EHIP, "local_eigensolve: rocblas_set_stream failed");
        int k = 0, restarts = 0, steps = 0;
        int32_t live = nsub;
        for (;;) {
            // Lanczos steps j = k .. m-1 with full reorthogonalization (two classical Gram-Schmidt
            // passes in the A~ + sigma I inner product): w = (A~ + sigma I)^{-1} B v_j
            for (int j = k; j < m; ++j, ++steps) {
                launchEigSpmv(A, 1, V + static_cast<int64_t>(j) * nl, u, s);
                solve();
                launchEigDots(A, AV, j + 1, w, c->eg_part.p, s);
                launchEigDotsFinal(A, c->eg_part.p, j + 1, 0, c->eg_hb.p, c->eg_hb2.p, c->eg_H.p, j, nullptr, s);
                launchEigUpdate(A, V, j + 1, c->eg_hb.p, w, s);
                launchEigDots(A, AV, j + 1, w, c->eg_part.p, s);
                launchEigDotsFinal(A, c->eg_part.p, j + 1, 1, c->eg_hb.p, c->eg_hb2.p, c->eg_H.p, j, nullptr, s);
                launchEigUpdate(A, V, j + 1, c->eg_hb2.p, w, s);
                launchEigSpmv(A, 0, w, Aw, s);
                launchEigDots(A, Aw, 1, w, c->eg_part.p, s);
                launchEigDotsFinal(A, c->eg_part.p, 1, 2, nullptr, nullptr, nullptr, 0, c->eg_beta.p, s);
                launchEigScale(A, V, AV, j + 1, w, Aw, c->eg_beta.p, s);
            }
            HIPCHECK(hipGetLastError());
            // Ritz pairs of the projected matrix, convergence, thick restart with kk Ritz vectors
            HIPCHECK(hipMemcpyAsync(c->eg_Hw.p, c->eg_H.p, sizeof(double) * nsub * m * m, hipMemcpyDeviceToDevice, s));
            require(rocsolver_dsyevd_strided_batched(c->rf_blas, rocblas_evect_original, rocblas_fill_upper, m, c->eg_Hw.p, m,
                                                     static_cast<rocblas_stride>(m) * m, c->eg_theta.p, m, c->eg_E.p, m,
                                                     c->eg_info.p, nsub) == rocblas_status_success,
                    S2L_EHIP, "local_eigensolve: rocsolver_dsyevd_strided_batched failed");
            launchEigCheck(A, c->eg_Hw.p, c->eg_theta.p, c->eg_beta.p, nev, tol, c->eg_conv.p, c->eg_lam.p, c->eg_live.p, s);
            launchEigRotate(A, V, AV, c->eg_Hw.p, kk, s);
            launchEigReset(A, c->eg_H.p, c->eg_theta.p, kk, c->eg_conv.p, c->eg_done.p, s);
            HIPCHECK(hipGetLastError());
            ++restarts;
            HIPCHECK(hipMemcpyAsync(&live, c->eg_live.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
            HIPCHECK(hipStreamSynchronize(s));
            if (live == 0 || restarts >= max_restarts) break;
            k = kk;
        }
        // the nev leading Ritz vectors are the first columns of every basis
        HIPCHECK(hipMemcpyAsync(Z, V, sizeof(double) * static_cast<size_t>(nev) * nl, hipMemcpyDeviceToHost, s));
        HIPCHECK(hipMemcpyAsync(lambda, c->eg_lam.p, sizeof(double) * static_cast<size_t>(nsub) * nev, hipMemcpyDeviceToHost, s));
        std::vector<int32_t> done(static_cast<size_t>(nsub));
        HIPCHECK(hipMemcpyAsync(done.data(), c->eg_done.p, sizeof(int32_t) * nsub, hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        int nconv = 0;
        for (int i = 0; i < nsub; ++i) {
            const bool ok = done[static_cast<size_t>(i)] && !(skip && skip[i]);
            if (converged) converged[i] = ok ? 1 : 0;
            nconv += ok ? 1 : 0;
        }
        if (info) {
            info[0] = restarts;
            info[1] = steps;
            info[2] = attempt + 1;
            info[3] = nconv;
        }
    });
}

int s2l_refactorize(s2l_ctx* c, const double* val, int host)
{
    return guarded(c, [&] {
        require(c && val, S2L_EINVAL, "refactorize: bad arguments");
        require(c->system_ready && c->subs_ready && c->local_ready && c->coarse_ready, S2L_ESTATE,
                "refactorize: upload the system, subdomains, factors and coarse space first");
        for (const HostFactor& F : c->hf) require(F.cholesky, S2L_ESTATE, "refactorize: the device refactorization needs Cholesky factors");
        require(c->queue_ok && c->launch_mode != 2, S2L_ESTATE, "refactorize: needs the task-queue schedule (launch modes 0 and 1)");
        require(c->n0 == 0 || !c->coarse_perm, S2L_ESTATE, "refactorize: needs the Cholesky factor of A_00 (a symmetric coarse matrix)");
        HIPCHECK(hipSetDevice(c->device));
        hipStream_t s = c->stream;
        c->invalidateGraphs();
        using clk = std::chrono::steady_clock;
        const auto t_start = clk::now();
        auto ms_since = [](clk::time_point t) { return std::chrono::duration<double, std::milli>(clk::now() - t).count(); };
        ensureRfPlan(c, s);
        RfPlan& R = c->rf;
        // the new values of A in the stored order
        c->rf_raw.reserve(static_cast<size_t>(c->nnz));
        HIPCHECK(hipMemcpyAsync(c->rf_raw.p, val, sizeof(double) * c->nnz, host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, s));
        launchGather(c->rf_raw.p, c->rf_inpos.p, c->val.p, c->nnz, s);
        HIPCHECK(hipStreamSynchronize(s));
        c->rf_ms[0] = ms_since(t_start);
        const auto t_local = clk::now();
        // local factors: one launch per height of the elimination trees, then the solve layout
        const int32_t zero = 0;
        HIPCHECK(hipMemcpyAsync(c->rf_status.p, &zero, sizeof(int32_t), hipMemcpyHostToDevice, s));
        RfArgs a = rfArgs(c);
        for (size_t h = 0; h + 1 < R.level_ptr.size(); ++h)
            launchRfLevel(a, c->rf_order.p + R.level_ptr[h], R.level_ptr[h + 1] - R.level_ptr[h], s);
        const bool f32 = c->precision != 0;
        launchRfPack(a, static_cast<int32_t>(R.sn.size()), f32 ? static_cast<void*>(c->diag_f.p) : static_cast<void*>(c->diag_dd.p),
                     f32 ? static_cast<void*>(c->below_f.p) : static_cast<void*>(c->below_dd.p), f32, s);
        HIPCHECK(hipGetLastError());
        int32_t status = 0;
        HIPCHECK(hipMemcpyAsync(&status, c->rf_status.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        require(status == 0, S2L_ENUMERIC, "refactorize: a local factor lost positive definiteness at supernode " + std::to_string(status - 1));
        rfDenseTops(c, a, f32 ? static_cast<void*>(c->diag_f.p) : static_cast<void*>(c->diag_dd.p), f32, s);
        c->ls_stale = true;
        c->rf_ms[1] = ms_since(t_local);
        const auto t_coarse = clk::now();
        if (c->n0 > 0) {
            // A_00 = W^T A W on the device
            const int32_t n0 = c->n0;
            formCoarseMatrix(c, s);
            // Cholesky factor and inverse of A_00 (dense FP64 library routines), then the layout of
            // the coarse solve
            if (!c->rf_blas) {
                require(rocblas_create_handle(&c->rf_blas) == rocblas_status_success, S2L_EHIP, "refactorize: rocblas_create_handle failed");
            }
            require(rocblas_set_stream(c->rf_blas, s) == rocblas_status_success, S2L_EHIP, "refactorize: rocblas_set_stream failed");
            rocblas_int* info = reinterpret_cast<rocblas_int*>(c->rf_status.p);
            require(rocsolver_dpotrf(c->rf_blas, rocblas_fill_lower, n0, c->rf_E.p, n0, info) == rocblas_status_success, S2L_EHIP,
                    "refactorize: rocsolver_dpotrf failed");
            int32_t pinfo = 0;
            HIPCHECK(hipMemcpyAsync(&pinfo, info, sizeof(int32_t), hipMemcpyDeviceToHost, s));
            HIPCHECK(hipStreamSynchronize(s));
            require(pinfo == 0, S2L_ENUMERIC, "refactorize: A_00 is not positive definite (leading minor " + std::to_string(pinfo) + ")");
            require(rocsolver_dtrtri(c->rf_blas, rocblas_fill_lower, rocblas_diagonal_non_unit, n0, c->rf_E.p, n0, info) ==
                        rocblas_status_success,
                    S2L_EHIP, "refactorize: rocsolver_dtrtri failed");
            launchTriLayout(c->rf_E.p, n0, c->tri_ld, c->T1_d.p, c->T2_d.p, s);
            HIPCHECK(hipGetLastError());
            coarseInverse(c, s);
            HIPCHECK(hipMemcpyAsync(&pinfo, info, sizeof(int32_t), hipMemcpyDeviceToHost, s));
            HIPCHECK(hipStreamSynchronize(s));
            require(pinfo == 0, S2L_ENUMERIC, "refactorize: the factor of A_00 is singular");
        }
        c->rf_ms[2] = ms_since(t_coarse);
        c->rf_ms[3] = ms_since(t_start);
    });
}

int s2l_apply(s2l_ctx* c, const double* d_r, double* d_z, int32_t nrhs, void* stream)
{
    return guarded(c, [&] {
        require(c && d_r && d_z && nrhs >= 1, S2L_EINVAL, "apply: bad arguments");
        HIPCHECK(hipSetDevice(c->device));
        applyImpl(c, d_r, d_z, nrhs, nullptr, pick(c, stream));
    });
}

// The apply's scratch for nrhs right-hand sides carved from @p base (256-byte aligned pieces, only
// those the context's schedule and coarse space use; the others null); returns the bytes needed.
size_t scratchLayout(const s2l_ctx* c, int nrhs, unsigned char* base, s2l_ctx::Work* w)
{
    size_t off = 0;
    auto take = [&](size_t bytes, bool used) -> unsigned char* {
        if (!used) return nullptr;
        unsigned char* p = base + off;
        off += (std::max<size_t>(bytes, 1) + 255) / 256 * 256;
        return p;
    };
    const size_t R = static_cast<size_t>(nrhs);
    const size_t n = static_cast<size_t>(c->n), lt = static_cast<size_t>(std::max<int64_t>(c->local_total, 1));
    const size_t n0 = static_cast<size_t>(std::max(c->n0, 1));
    const bool coarse = c->n0 > 0;
    const bool direct = c->one_level == 0 && c->boolean;
    const bool lsm = c->launch_mode == 2;
    const size_t esz = c->precision == 1 ? sizeof(float) : sizeof(double);
    const size_t nr = (R + 1) / 2 * 2;
    *w = s2l_ctx::Work{};
    w->x0 = reinterpret_cast<double*>(take(n * R * 8, coarse));
    w->w = reinterpret_cast<double*>(take(n * R * 8, coarse && c->correction == 0));
    w->loc = reinterpret_cast<double*>(take(lt * R * 8, !direct || (coarse && !c->boolean)));
    w->cvec = reinterpret_cast<double*>(take(n0 * R * 8, coarse));
    w->tvec = reinterpret_cast<double*>(take(n0 * R * 8, coarse));
    w->yvec = reinterpret_cast<double*>(take(n0 * R * 8, coarse));
    w->cpart = reinterpret_cast<double*>(take(static_cast<size_t>(std::max(c->nsub, 1)) * std::max(c->ca.rchunks, 1) *
                                                  std::max(c->ca.mmax, 1) * R * 8, coarse));
    w->tbuf = take(static_cast<size_t>(std::max<int64_t>(c->pk.t_total, 1)) * nr * esz, !lsm);
    w->xgbuf = take(lt * nr * esz, !lsm);
    w->ls_slots = take(static_cast<size_t>(std::max<int64_t>(c->ls.slots, 1)) * R * esz, lsm);
    w->ls_part = take(static_cast<size_t>(std::max<int64_t>(c->ls.partials, 1)) * R * esz, lsm);
    w->ls_b = take(lt * R * esz, lsm);
    w->ls_yf = take(lt * R * esz, lsm);
    w->ls_xb = take(lt * R * esz, lsm);
    return off;
}

int s2l_apply_noalloc(s2l_ctx* c, const double* d_r, double* d_z, int32_t nrhs, void* stream)
{
    return guarded(c, [&] {
        require(c && d_r && d_z && nrhs >= 1, S2L_EINVAL, "apply_noalloc: bad arguments");
        require(c->system_ready && c->subs_ready && c->local_ready && c->coarse_ready, S2L_ESTATE,
                "apply_noalloc: upload the system, subdomains, factors and coarse space first");
        require(c->scratch != nullptr, S2L_ESTATE, "apply_noalloc: no scratch address for this device");
        HIPCHECK(hipSetDevice(c->device));
        s2l_ctx::Work w{};
        const size_t need = scratchLayout(c, nrhs, static_cast<unsigned char*>(c->scratch), &w);
        require(need <= kScratchBytes, S2L_ELIMIT,
                "apply_noalloc: " + std::to_string(nrhs) + " right-hand sides need " + std::to_string(need) +
                    " bytes of scratch; the maximum is " + std::to_string(kScratchBytes));
        // one call runs with the scratch pointers, without event timing (no event creation)
        const s2l_ctx::Work saved = c->work;
        const bool timing = c->timing, phases = c->phase_on;
        c->work = w;
        c->timing = c->phase_on = false;
        try {
            applyImpl(c, d_r, d_z, nrhs, nullptr, pick(c, stream), false);
        } catch (...) {
            c->work = saved;
            c->timing = timing;
            c->phase_on = phases;
            throw;
        }
        c->work = saved;
        c->timing = timing;
        c->phase_on = phases;
    });
}

int s2l_spmv(s2l_ctx* c, const double* d_x, double* d_y, int32_t nrhs, void* stream)
{
    return guarded(c, [&] {
        require(c && d_x && d_y && nrhs >= 1, S2L_EINVAL, "spmv: bad arguments");
        HIPCHECK(hipSetDevice(c->device));
        spmvImpl(c, d_x, d_y, nullptr, nrhs, nullptr, pick(c, stream));
    });
}

int s2l_krylov_solve(s2l_ctx* c, const s2l_krylov_opts* o, const double* d_b, double* d_x, int32_t nrhs,
                     s2l_krylov_result* results, double* history, void* stream)
{
    int rc = S2L_OK;
    const int g = guarded(c, [&] {
        require(c && o && d_b && d_x && results && nrhs >= 1, S2L_EINVAL, "krylov_solve: bad arguments");
        require(o->method == 0 || o->method == 1, S2L_EINVAL, "krylov_solve: method must be 0 (GMRES) or 1 (BiCGStab)");
        HIPCHECK(hipSetDevice(c->device));
        rc = o->method == 0 ? gmresImpl(c, o, d_b, d_x, nrhs, results, history, pick(c, stream))
                            : bicgImpl(c, o, d_b, d_x, nrhs, results, history, pick(c, stream));
    });
    return g != S2L_OK ? g : rc;
}

int s2l_solver_create(s2l_ctx* c, const s2l_krylov_opts* o, int32_t slots, s2l_solver** out)
{
    return guarded(c, [&] {
        require(c && o && out && slots >= 1, S2L_EINVAL, "solver_create: bad arguments");
        require(o->method == 0, S2L_EINVAL, "solver_create: the slot solver runs GMRES (method 0)");
        require(c->system_ready && c->subs_ready && c->local_ready && c->coarse_ready, S2L_ESTATE,
                "solver_create: upload the system, subdomains, factors and coarse space first");
        HIPCHECK(hipSetDevice(c->device));
        auto sv = std::make_unique<s2l_solver>();
        sv->c = c;
        sv->run = std::make_unique<GmresRun>(c, *o, slots);
        sv->B.reserve(static_cast<size_t>(c->n) * slots);
        sv->X.reserve(static_cast<size_t>(c->n) * slots);
        sv->pending.assign(static_cast<size_t>(slots), 0);
        sv->used.assign(static_cast<size_t>(slots), 0);
        c->ensureWork(slots);
        *out = sv.release();
    });
}

int s2l_solver_destroy(s2l_solver* sv)
{
    if (!sv) return S2L_OK;
    s2l_ctx* c = sv->c;
    return guarded(c, [&] {
        HIPCHECK(hipSetDevice(c->device));
        HIPCHECK(hipStreamSynchronize(c->stream));
        delete sv;
    });
}

int s2l_solver_admit(s2l_solver* sv, int32_t slot, const double* b, const double* x0, int host)
{
    s2l_ctx* c = sv ? sv->c : nullptr;
    return guarded(c, [&] {
        require(sv && b && slot >= 0 && slot < sv->run->cap, S2L_EINVAL, "solver_admit: bad arguments");
        require(!sv->used[static_cast<size_t>(slot)], S2L_ESTATE, "solver_admit: slot " + std::to_string(slot) + " is busy");
        HIPCHECK(hipSetDevice(c->device));
        const int64_t n = c->n;
        const hipMemcpyKind k = host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
        double* bs = sv->B.p + static_cast<int64_t>(slot) * n;
        double* xs = sv->X.p + static_cast<int64_t>(slot) * n;
        HIPCHECK(hipMemcpyAsync(bs, b, sizeof(double) * n, k, c->stream));
        if (x0)
            HIPCHECK(hipMemcpyAsync(xs, x0, sizeof(double) * n, k, c->stream));
        else
            launchFill(n, 0.0, xs, c->stream);
        if (host) HIPCHECK(hipStreamSynchronize(c->stream));  // the caller may reuse its buffers
        sv->used[static_cast<size_t>(slot)] = 1;
        sv->pending[static_cast<size_t>(slot)] = 1;
    });
}

int s2l_solver_step(s2l_solver* sv, int32_t* finished)
{
    s2l_ctx* c = sv ? sv->c : nullptr;
    return guarded(c, [&] {
        require(sv != nullptr, S2L_EINVAL, "solver_step: null solver");
        HIPCHECK(hipSetDevice(c->device));
        GmresRun& R = *sv->run;
        R.begin(sv->pending, sv->B.p, sv->X.p, c->stream);
        std::fill(sv->pending.begin(), sv->pending.end(), 0);
        R.cycle(sv->B.p, sv->X.p, c->stream);
        if (finished)
            for (int q = 0; q < R.cap; ++q)
                finished[q] = (sv->used[static_cast<size_t>(q)] && !R.sl[static_cast<size_t>(q)].active(R.o.maxit)) ? 1 : 0;
    });
}

int s2l_solver_take(s2l_solver* sv, int32_t slot, s2l_krylov_result* res, double* history, double* x, int host)
{
    s2l_ctx* c = sv ? sv->c : nullptr;
    return guarded(c, [&] {
        require(sv && res && slot >= 0 && slot < sv->run->cap, S2L_EINVAL, "solver_take: bad arguments");
        GmresRun& R = *sv->run;
        require(sv->used[static_cast<size_t>(slot)] && !sv->pending[static_cast<size_t>(slot)] &&
                    !R.sl[static_cast<size_t>(slot)].active(R.o.maxit),
                S2L_ESTATE, "solver_take: slot " + std::to_string(slot) + " holds no finished solve");
        HIPCHECK(hipSetDevice(c->device));
        R.result(slot, res, history);
        if (x) {
            const int64_t n = c->n;
            HIPCHECK(hipMemcpyAsync(x, sv->X.p + static_cast<int64_t>(slot) * n, sizeof(double) * n,
                                    host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, c->stream));
            HIPCHECK(hipStreamSynchronize(c->stream));
        }
        sv->used[static_cast<size_t>(slot)] = 0;
        R.sl[static_cast<size_t>(slot)].live = false;
    });
}

int s2l_apply_host(s2l_ctx* c, const double* r, double* z, int32_t nrhs)
{
    return guarded(c, [&] {
        require(c && r && z && nrhs >= 1, S2L_EINVAL, "apply_host: bad arguments");
        HIPCHECK(hipSetDevice(c->device));
        DevBuf<double> dr, dz;
        dr.upload(r, static_cast<size_t>(c->n) * nrhs, c->stream);
        dz.reserve(static_cast<size_t>(c->n) * nrhs);
        applyImpl(c, dr.p, dz.p, nrhs, nullptr, c->stream);
        HIPCHECK(hipMemcpyAsync(z, dz.p, sizeof(double) * c->n * nrhs, hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(hipStreamSynchronize(c->stream));
    });
}

int s2l_spmv_host(s2l_ctx* c, const double* x, double* y, int32_t nrhs)
{
    return guarded(c, [&] {
        require(c && x && y && nrhs >= 1, S2L_EINVAL, "spmv_host: bad arguments");
        HIPCHECK(hipSetDevice(c->device));
        DevBuf<double> dx, dy;
        dx.upload(x, static_cast<size_t>(c->n) * nrhs, c->stream);
        dy.reserve(static_cast<size_t>(c->n) * nrhs);
        spmvImpl(c, dx.p, dy.p, nullptr, nrhs, nullptr, c->stream);
        HIPCHECK(hipMemcpyAsync(y, dy.p, sizeof(double) * c->n * nrhs, hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(hipStreamSynchronize(c->stream));
    });
}

int s2l_krylov_solve_host(s2l_ctx* c, const s2l_krylov_opts* o, const double* b, double* x, int32_t nrhs,
                          s2l_krylov_result* results, double* history)
{
    int rc = S2L_OK;
    const int g = guarded(c, [&] {
        require(c && o && b && x && results && nrhs >= 1, S2L_EINVAL, "krylov_solve_host: bad arguments");
        require(o->method == 0 || o->method == 1, S2L_EINVAL, "krylov_solve: method must be 0 (GMRES) or 1 (BiCGStab)");
        HIPCHECK(hipSetDevice(c->device));
        DevBuf<double> db, dx;
        db.upload(b, static_cast<size_t>(c->n) * nrhs, c->stream);
        dx.upload(x, static_cast<size_t>(c->n) * nrhs, c->stream);
        rc = o->method == 0 ? gmresImpl(c, o, db.p, dx.p, nrhs, results, history, c->stream)
                            : bicgImpl(c, o, db.p, dx.p, nrhs, results, history, c->stream);
        HIPCHECK(hipMemcpyAsync(x, dx.p, sizeof(double) * c->n * nrhs, hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(hipStreamSynchronize(c->stream));
    });
    return g != S2L_OK ? g : rc;
}

int s2l_local_solve_host(s2l_ctx* c, int32_t i, const double* b, double* y)
{
    return guarded(c, [&] {
        require(c && c->local_ready && b && y && i >= 0 && i < c->nsub, S2L_EINVAL, "local_solve_host: bad arguments");
        HIPCHECK(hipSetDevice(c->device));
        c->ensureWork(1);
        // Scatter b into a global-length vector through the local -> global map, solve, read the
        // local buffer (mode 1 without scaling).
        const int32_t l0 = c->local_ptr[i], nl = c->local_ptr[i + 1] - l0;
        std::vector<double> g(static_cast<size_t>(c->n), 0.0);
        for (int32_t l = 0; l < nl; ++l) g[static_cast<size_t>(c->gid_h[static_cast<size_t>(l0 + l)])] = b[l];
        DevBuf<double> dg;
        dg.upload(g, c->stream);
        auto run = [&](auto tag) {
            using T = decltype(tag);
            if (c->launch_mode == 2) {
                // every subdomain is solved; the others see their share of the scattered b
                LsArgs<T> a = lsArgs<T>(c, dg.p, nullptr, c->work.loc, 1, nullptr);
                a.pou = nullptr;
                lsLaunches<T>(c, a, 1, c->stream);
                return;
            }
            auto queue = [&](auto ftag) {
                using F = decltype(ftag);
                LocalArgs<T, F> a = localArgs<T, F>(c, dg.p, nullptr, c->work.loc, 1, 1, nullptr);
                a.pou = nullptr; // y by local index, no partition of unity
                launchQueue<T, F>(c, a, c->qsubq_d.p + c->qsub_h[static_cast<size_t>(i)],
                                  c->qsub_h[static_cast<size_t>(i) + 1] - c->qsub_h[static_cast<size_t>(i)], true, c->stream);
            };
            if (c->precision == 2)
                queue(float{});
            else
                queue(T{});
        };
        if (c->precision == 1)
            run(float{});
        else
            run(double{});
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipStreamSynchronize(c->stream));
        HIPCHECK(hipMemcpyAsync(y, c->work.loc + l0, sizeof(double) * nl, hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(hipStreamSynchronize(c->stream));
    });
}

int s2l_device_alloc(s2l_ctx* c, int64_t bytes, void** out)
{
    return guarded(c, [&] {
        require(c && out && bytes >= 0, S2L_EINVAL, "device_alloc: bad arguments");
        HIPCHECK(hipSetDevice(c->device));
        HIPCHECK(hipMalloc(out, static_cast<size_t>(std::max<int64_t>(bytes, 1))));
    });
}

int s2l_device_free(s2l_ctx* c, void* p)
{
    return guarded(c, [&] {
        require(c != nullptr, S2L_EINVAL, "device_free: null context");
        if (p) HIPCHECK(hipFree(p));
    });
}

int s2l_memcpy_h2d(s2l_ctx* c, void* dst, const void* src, int64_t bytes)
{
    return guarded(c, [&] {
        require(c && dst && src && bytes >= 0, S2L_EINVAL, "memcpy_h2d: bad arguments");
        HIPCHECK(hipMemcpyAsync(dst, src, static_cast<size_t>(bytes), hipMemcpyHostToDevice, c->stream));
        HIPCHECK(hipStreamSynchronize(c->stream));
    });
}

int s2l_memcpy_d2h(s2l_ctx* c, void* dst, const void* src, int64_t bytes)
{
    return guarded(c, [&] {
        require(c && dst && src && bytes >= 0, S2L_EINVAL, "memcpy_d2h: bad arguments");
        HIPCHECK(hipMemcpyAsync(dst, src, static_cast<size_t>(bytes), hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(hipStreamSynchronize(c->stream));
    });
}

int s2l_copy_d2d(s2l_ctx* c, double* dst, const double* src, int64_t count, void* stream)
{
    return guarded(c, [&] {
        require(c && dst && src && count >= 0, S2L_EINVAL, "copy_d2d: bad arguments");
        if (count == 0) return;
        launchCopy(count, src, dst, 1, stream ? static_cast<hipStream_t>(stream) : c->stream);
        HIPCHECK(hipGetLastError());
    });
}

int s2l_synchronize(s2l_ctx* c)
{
    return guarded(c, [&] {
        require(c != nullptr, S2L_EINVAL, "synchronize: null context");
        HIPCHECK(hipStreamSynchronize(c->stream));
    });
}

int s2l_device_footprint(s2l_ctx* c, int64_t* bytes)
{
    return guarded(c, [&] {
        require(c && bytes, S2L_EINVAL, "device_footprint: bad arguments");
        *bytes = c->device_bytes;
    });
}

// Slots the task queue keeps in LDS: the first slot_cap of every task without child tasks.
static int64_t ldsSlots(const s2l_ctx* c)
{
    if (!c->queue_ok) return 0;
    int64_t k = 0;
    for (const TaskDesc& t : c->pk.task) k += std::min(t.slot_count, c->slot_cap);
    return k;
}

int s2l_get_stats(s2l_ctx* c, s2l_stats* o)
{
    return guarded(c, [&] {
        require(c && o, S2L_EINVAL, "get_stats: bad arguments");
        std::memset(o, 0, sizeof(*o));
        o->n = c->n;
        o->nsub = c->nsub;
        o->n0 = c->n0;
        o->precision = c->precision;
        o->nnz_a = c->nnz;
        const bool lsm = c->launch_mode == 2;
        o->factor_entries_fwd = lsm ? c->ls.entries : c->pk.entries_fwd;
        o->factor_entries_bwd = lsm ? 0 : c->pk.entries_bwd;
        o->local_rows = c->local_total;
        int64_t crows = 0;
        for (int i = 0; i < c->nsub && c->n0 > 0; ++i) {
            const int32_t nI = c->ownership[i + 1] - c->ownership[i];
            const int32_t rows = c->boolean ? nI : (c->local_ptr[i + 1] - c->local_ptr[i]);
            crows += static_cast<int64_t>(rows) * (c->off_h[static_cast<size_t>(i) + 1] - c->off_h[static_cast<size_t>(i)]);
        }
        o->coarse_rows = crows;
        o->levels_fwd_max = lsm ? c->ls.levels_fwd : c->pk.max_levels_fwd;
        o->levels_bwd_max = lsm ? c->ls.levels_bwd : c->pk.max_levels_bwd;
        o->supernodes = static_cast<int32_t>(lsm ? c->ls.sn.size() : c->pk.sn.size());
        const double sv = c->precision == 0 ? 8.0 : 4.0;    // factor entries
        const double st = c->precision == 1 ? 4.0 : 8.0;    // arithmetic type: slots, partials, LDS
        double local = 0.0;
        if (lsm) {
            // level-synchronous: factor values of both sweeps (triangles of the inverted diagonal
            // blocks and the blocks below), index data (descriptors, row indices, slot tables),
            // rhs gather (8 B value + 4 B map), slots and backward partials written and read once,
            // forward results written and read, backward results written and read by the
            // descendants (one read per row index below a block), permutation and owned-row output
            // (x0 read 8 B + z write 8 B)
            const double lt = static_cast<double>(c->local_total);
            local = static_cast<double>(c->ls.entries) * sv + static_cast<double>(c->ls.bytes_index) + lt * (8.0 + 4.0 + 4.0) +
                    static_cast<double>(c->n) * 16.0 + static_cast<double>(c->ls.slots + c->ls.partials) * st * 2.0 + lt * st * 3.0 +
                    static_cast<double>(c->ls.bwd_below_rows) * st;
        } else {
            // task queue: factor values of both sweeps, index data, rhs gather (8 B value + 4 B map),
            // permutation and owned-row output (read x0 8 B + write z 8 B), the slots outside LDS
            // written and read once
            local = (c->pk.entries_fwd + c->pk.entries_bwd) * sv + static_cast<double>(c->pk.bytes_index) +
                    static_cast<double>(c->local_total) * (8.0 + 4.0 + 4.0) + static_cast<double>(c->n) * 16.0 +
                    static_cast<double>(c->pk.t_total - ldsSlots(c)) * st * 2.0;
            o->lds_slot_cap = c->queue_ok ? c->slot_cap : 0;
            o->max_task_slots = c->pk.max_task_slots;
            o->lds_slots = ldsSlots(c);
        }
        o->bytes_local_solve = local;
        const double spmv = static_cast<double>(c->nnz) * 12.0 + (static_cast<double>(c->n) + 1) * 4.0 + static_cast<double>(c->n) * 16.0;
        o->bytes_spmv = spmv;
        double apply = local;
        if (c->n0 > 0) {
            apply += static_cast<double>(crows) * 8.0 * 2.0 + static_cast<double>(c->n0) * (c->n0 + 1) * 8.0 +
                     static_cast<double>(c->n) * 8.0 * 2.0; // V twice, both triangles, r read + x0 write
            if (c->correction == 0) apply += spmv + static_cast<double>(c->n) * 8.0;
        }
        o->bytes_apply = apply;
        for (int q = 0; q < 4; ++q) o->refactor_ms[q] = c->rf_ms[q];
    });
}

int s2l_timing_enable(s2l_ctx* c, int on)
{
    return guarded(c, [&] {
        if (c) c->invalidateGraphs();
        require(c != nullptr, S2L_EINVAL, "timing_enable: null context");
        c->timing = (on != 0);
    });
}

int s2l_timing_read(s2l_ctx* c, int64_t* launches, double* ms_total, int reset)
{
    return guarded(c, [&] {
        require(c != nullptr, S2L_EINVAL, "timing_read: null context");
        collectTiming(c);
        if (launches) *launches = c->launches;
        if (ms_total) *ms_total = c->ms_total;
        if (reset) {
            c->launches = 0;
            c->ms_total = 0.0;
        }
    });
}

const char* s2l_phase_name(int phase)
{
    static const char* names[S2L_NPHASES] = {
        "solve.preconditioner",         "solve.preconditioner.coarse",         "solve.preconditioner.coarse.restrict",
        "solve.preconditioner.coarse.solve", "solve.preconditioner.coarse.prolong", "solve.preconditioner.spmv",
        "solve.preconditioner.local",   "solve.spmv",                          "solve.orthogonalization",
        "solve.stopping_test"};
    return phase >= 0 && phase < S2L_NPHASES ? names[phase] : nullptr;
}

int s2l_phase_timing_enable(s2l_ctx* c, int on)
{
    return guarded(c, [&] {
        require(c != nullptr, S2L_EINVAL, "phase_timing_enable: null context");
        c->invalidateGraphs();
        c->phase_on = on != 0;
    });
}

int s2l_phase_timing_read(s2l_ctx* c, double* ms, int64_t* counts, int reset)
{
    return guarded(c, [&] {
        require(c != nullptr, S2L_EINVAL, "phase_timing_read: null context");
        collectPhases(c);
        for (int q = 0; q < S2L_NPHASES; ++q) {
            if (ms) ms[q] = c->phase_ms[q];
            if (counts) counts[q] = c->phase_n[q];
            if (reset) {
                c->phase_ms[q] = 0.0;
                c->phase_n[q] = 0;
            }
        }
    });
}

int s2l_set_tasks_per_subdomain(s2l_ctx* c, int tasks)
{
    return guarded(c, [&] {
        if (c) c->invalidateGraphs();
        require(c && tasks >= 0, S2L_EINVAL, "tasks per subdomain must be >= 0 (0 = automatic)");
        c->tasks_per_sub = tasks;
        c->local_ready = false;
    });
}

int s2l_set_local_threads(s2l_ctx* c, int32_t threads)
{
    return guarded(c, [&] {
        if (c) c->invalidateGraphs();
        require(c && (threads == 512 || threads == 1024), S2L_EINVAL, "local_threads: 512 or 1024");
        c->local_threads_req = threads;
        c->local_ready = false;
    });
}

int s2l_set_dense_top(s2l_ctx* c, int32_t max_width)
{
    return guarded(c, [&] {
        if (c) c->invalidateGraphs();
        require(c && max_width >= 0 && max_width <= kDenseTopMax, S2L_EINVAL,
                "dense_top: the widest dense top must be in [0, " + std::to_string(kDenseTopMax) + "]");
        c->dense_top_max = max_width;
        c->local_ready = false;
    });
}

int s2l_autotune_local(s2l_ctx* c, int32_t nrhs, int32_t applies, s2l_tune_entry* entries, int32_t max_entries,
                       int32_t* count, int32_t* best)
{
    return guarded(c, [&] {
        if (c) c->invalidateGraphs();
        require(c && c->local_ready && c->system_ready, S2L_ESTATE, "autotune_local: upload the system and the factors first");
        require(nrhs >= 1 && applies >= 1 && entries && max_entries >= 1 && count && best, S2L_EINVAL,
                "autotune_local: bad arguments");
        HIPCHECK(hipSetDevice(c->device));
        hipStream_t s = c->stream;
        const int precision = c->precision;
        // candidates: tasks per subdomain around one workgroup per compute unit, each with and
        // without dense tops, and the level-synchronous schedule
        const int b = std::max(1, c->cus / std::max(1, c->nsub));
        std::vector<int> tset{1, std::max(1, b / 2), std::max(1, b - 1), b, b + 1, 2 * b};
        std::sort(tset.begin(), tset.end());
        tset.erase(std::unique(tset.begin(), tset.end()), tset.end());
        std::vector<s2l_tune_entry> cand;
        for (int t : tset)
            for (int d : {static_cast<int>(kDenseTopMax), 0}) {
                if (t == 1 && d != 0) continue;  // one task per subdomain: no skeleton
                cand.push_back(s2l_tune_entry{t, d, 0, 1024, 0.0});
            }
        // two 512-thread workgroups per compute unit: b, 3b/2 and 2b tasks per subdomain
        std::vector<int> tset2{b, b + b / 2, 2 * b};
        tset2.erase(std::unique(tset2.begin(), tset2.end()), tset2.end());
        for (int t : tset2)
            for (int d : {static_cast<int>(kDenseTopMax), 0}) {
                if (t == 1 && d != 0) continue;
                cand.push_back(s2l_tune_entry{t, d, 0, 512, 0.0});
            }
        if (precision != 2) cand.push_back(s2l_tune_entry{b, static_cast<int>(kDenseTopMax), 2, 1024, 0.0});
        require(static_cast<int32_t>(cand.size()) <= max_entries, S2L_ELIMIT,
                "autotune_local: " + std::to_string(cand.size()) + " candidates, max_entries too small");
        // input: a fixed vector per right-hand side; output: the owned rows
        std::vector<double> h(static_cast<size_t>(c->n) * nrhs);
        for (size_t q = 0; q < h.size(); ++q) h[q] = std::cos(1e-3 * static_cast<double>(q));
        DevBuf<double> in, out;
        in.upload(h, s);
        out.reserve(h.size());
        hipEvent_t e0, e1;
        HIPCHECK(hipEventCreate(&e0));
        HIPCHECK(hipEventCreate(&e1));
        auto run = [&](int cnt) {
            for (int k = 0; k < cnt; ++k) {
                if (precision == 1 && nrhs >= 2 && c->pair_ok && c->launch_mode == 0)
                    localSolve<float2, float>(c, in.p, nullptr, out.p, 0, nrhs, nullptr, s);
                else if (precision == 2 && nrhs >= 2 && c->pair_ok && c->launch_mode == 0)
                    localSolve<double2, float>(c, in.p, nullptr, out.p, 0, nrhs, nullptr, s);
                else if (precision == 1)
                    localSolve<float, float>(c, in.p, nullptr, out.p, 0, nrhs, nullptr, s);
                else if (precision == 2)
                    localSolve<double, float>(c, in.p, nullptr, out.p, 0, nrhs, nullptr, s);
                else
                    localSolve<double, double>(c, in.p, nullptr, out.p, 0, nrhs, nullptr, s);
            }
        };
        const bool timing = c->timing;
        c->timing = false;
        int32_t ibest = -1;
        try {
            for (size_t k = 0; k < cand.size(); ++k) {
                s2l_tune_entry& e = cand[k];
                c->tasks_per_sub = e.tasks;
                c->dense_top_max = e.dense_top;
                c->local_threads_req = e.threads;
                c->launch_mode = 0;
                finalizeLocal(c, precision);
                e.threads = c->local_threads;  // 512 may fall back to 1024
                if (e.launch_mode == 2 || !c->queue_ok) {
                    if (!c->ls_built) buildLs(c, precision);
                    c->launch_mode = 2;
                    e.launch_mode = 2;
                }
                c->ensureWork(nrhs);
                run(3);
                HIPCHECK(hipEventRecord(e0, s));
                run(applies);
                HIPCHECK(hipEventRecord(e1, s));
                HIPCHECK(hipEventSynchronize(e1));
                float ms = 0.f;
                HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
                e.us = 1e3 * static_cast<double>(ms) / applies;
                if (ibest < 0 || e.us < cand[static_cast<size_t>(ibest)].us) ibest = static_cast<int32_t>(k);
            }
        } catch (...) {
            c->timing = timing;
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        c->timing = timing;
        HIPCHECK(hipEventDestroy(e0));
        HIPCHECK(hipEventDestroy(e1));
        // keep the fastest
        const s2l_tune_entry& w = cand[static_cast<size_t>(ibest)];
        c->tasks_per_sub = w.tasks;
        c->dense_top_max = w.dense_top;
        c->local_threads_req = w.threads;
        c->launch_mode = 0;
        finalizeLocal(c, precision);
        if (w.launch_mode == 2) {
            if (!c->ls_built) buildLs(c, precision);
            c->launch_mode = 2;
        }
        std::copy(cand.begin(), cand.end(), entries);
        *count = static_cast<int32_t>(cand.size());
        *best = ibest;
    });
}

int s2l_set_graphs(s2l_ctx* c, int on)
{
    return guarded(c, [&] {
        require(c != nullptr, S2L_EINVAL, "set_graphs: null context");
        c->invalidateGraphs();
        c->use_graphs = on != 0;
    });
}

int s2l_set_local_launch_mode(s2l_ctx* c, int mode)
{
    return guarded(c, [&] {
        if (c) c->invalidateGraphs();
        require(c && mode >= 0 && mode <= 2, S2L_EINVAL, "launch mode must be 0, 1 or 2");
        require(mode != 2 || !c->local_ready || c->precision != 2, S2L_EINVAL,
                "the level-synchronous schedule has no FP32-storage / FP64-arithmetic mode");
        require(mode == 2 || !c->local_ready || c->queue_ok, S2L_ELIMIT,
                "launch modes 0 and 1 need a subdomain vector in the LDS of one workgroup");
        require(mode != 2 || !c->ls_stale, S2L_ESTATE,
                "the level-synchronous layout predates the device refactorization; upload the factors again");
        if (mode == 2 && c->local_ready && !c->ls_built) buildLs(c, c->precision);
        c->launch_mode = mode;
    });
}

} // extern "C"
