// Host side of the library: index planning, the level loop of the
// factorization, the apply sweeps, PCG, and the C ABI of include/hifde_amd.h.
//
// The level loop follows HIFFactor::Factor (src/Factor.cpp:41-63): at each
// level every cell is eliminated (LevelFactorSchur, :184-213), the level
// matrix is updated (UpdateMatrixAndDOFs, :262-311), then every face is
// skeletonized (LevelFactorSkel, :232-260) and the matrix updated again.  All
// boxes of a half level read the same pre-level matrix, so each half level is
// one batch of boxes on the GPU.

#include <dlfcn.h>
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/hifde_amd.h"
#include "hifde_internal.hpp"

namespace hifde {
namespace {
thread_local Stager* t_stager = nullptr;
}
Stager* active_stager() { return t_stager; }
void set_active_stager(Stager* st) { t_stager = st; }
namespace {
thread_local KernelTimer* t_timer = nullptr;
thread_local hipStream_t t_owner = nullptr;
}
KernelTimer* active_timer() { return t_timer; }
void set_active_timer(KernelTimer* t) { t_timer = t; }
hipStream_t owner_stream() { return t_owner; }
void set_owner_stream(hipStream_t s) { t_owner = s; }
// Marks the calling thread as working for the object that owns stream s.
struct OwnerGuard {
    hipStream_t prev;
    explicit OwnerGuard(hipStream_t s) : prev(owner_stream()) { set_owner_stream(s); }
    ~OwnerGuard() { set_owner_stream(prev); }
};
// Stream of the calling thread for the batch free functions (no factor
// object): one non-blocking stream per thread and device, so that these
// calls never touch the legacy null stream or wait for other threads' work.
namespace {
struct AuxStreams {
    std::vector<hipStream_t> s;
    ~AuxStreams() {
        for (hipStream_t x : s)
            if (x) (void)hipStreamDestroy(x);
    }
};
}  // namespace
hipStream_t aux_stream() {
    thread_local AuxStreams t;
    int dev = 0;
    HIFDE_CHECK(hipGetDevice(&dev));
    if ((int)t.s.size() <= dev) t.s.resize(dev + 1, nullptr);
    if (!t.s[dev]) HIFDE_CHECK(hipStreamCreateWithFlags(&t.s[dev], hipStreamNonBlocking));
    return t.s[dev];
}

// Scratch of the allocation-free apply (hifde_apply_device_noalloc): the
// single-vector sweeps' diagonal-stage vector and two pass-flag side
// vectors, 3 NC^3 doubles, for grids up to NC^3 = 128^3 (48 MiB).  Every
// GPU the library runs on holds its own copy of this array until the
// process ends.  In host code the array's name is not a device address:
// host code uses only the per-device pointer from hipGetSymbolAddress.
constexpr long long kNoAllocMaxCells = 1LL << 21;  // NC^3 <= 128^3
static __device__ __align__(256) double scratch[3 * kNoAllocMaxCells];
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace
// Resolves the calling thread's current device's copy of the scratch array
// (once per device, with that device current); returns its address.
double* resolve_scratch(int dev) {
    if (dev < 0 || dev >= kMaxDevices) throw Error(HIFDE_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIFDE_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    return g_scratch_ptr[dev];
}

const void* Stager::stage(const void* src, size_t bytes) {
    const size_t need = (bytes + 255) & ~(size_t)255;
    if (used + need > cap) {
        // Continue in a larger buffer instead of waiting for the stream: the
        // full one is kept until the next synchronization.
        if (buf) retired.push_back(buf);
        buf = nullptr;
        used = 0;
        cap = std::max(need, std::max(2 * cap, (size_t)64 << 20));
        HIFDE_CHECK(hipHostMalloc(reinterpret_cast<void**>(&buf), cap));
    }
    if (bytes >= ((size_t)4 << 20)) {
        // large uploads: the copy into pinned memory is split over the host cores
        const size_t chunk = (size_t)1 << 20;
        const long long nch = (long long)((bytes + chunk - 1) / chunk);
        unsigned char* dst = buf + used;
        const unsigned char* sp = static_cast<const unsigned char*>(src);
#pragma omp parallel for schedule(static)
        for (long long c = 0; c < nch; ++c) {
            const size_t o = (size_t)c * chunk;
            std::memcpy(dst + o, sp + o, std::min(chunk, bytes - o));
        }
    } else {
        std::memcpy(buf + used, src, bytes);
    }
    const void* p = buf + used;
    used += need;
    return p;
}
}  // namespace hifde

using namespace hifde;

namespace {

thread_local std::string g_last_error;

enum Stage {
    ST_PLAN = 0,
    ST_UPLOAD,
    ST_GATHER,
    ST_INVERSE,
    ST_GEMM,
    ST_ID,
    ST_TRANSFORM,
    ST_SCATTER,
    ST_COMPACT,
    ST_DOWNLOAD,
    ST_COUNT
};
const char* kStageNames[ST_COUNT] = {"plan",   "upload",    "gather",  "inverse", "gemm",
                                     "interp_decomp", "face_transform", "scatter", "compact",
                                     "download"};

enum FaceId { TOP = 0, BOTTOM, RIGHT, LEFT, FRONT, BACK };  // include/data.hpp:7

using clk = std::chrono::steady_clock;

template <class V>
std::vector<long long> prefix_ll(const V& a, const V& b) {
    std::vector<long long> off(a.size() + 1, 0);
    for (size_t i = 0; i < a.size(); ++i) off[i + 1] = off[i] + (long long)a[i] * b[i];
    return off;
}
std::vector<int> prefix_i(const std::vector<int>& a) {
    std::vector<int> off(a.size() + 1, 0);
    for (size_t i = 0; i < a.size(); ++i) off[i + 1] = off[i] + a[i];
    return off;
}

}  // namespace

struct hifde_factor {
    int N, P, NC, L;
    double eps;
    int precision;
    bool quirk;
    hipStream_t stream = nullptr;
    std::vector<unsigned char> remaining;
    DevBuf<unsigned char> d_alive;
    LevelTable tabs[2];  // ping-pong level matrices (compaction writes the other)
    int cur = 0;
    DevBuf<unsigned long long> tab_counter;
    size_t tab_inserted = 0;  // upper bound of the entries in tabs[cur]
    bool op_symmetric = false;  // the assembled operator equals its transpose bitwise
    int sym_option = -1;        // -1 auto, 0 off
    bool sym_on = false;        // symmetric fast path of the current factorization
    LevelTable& table_ref() { return tabs[cur]; }
    bool have_matrix = false, factored = false;
    // operator kept for SpMV / PCG (CSR)
    // Assembled operator: CSR on the device (d_rp, d_ci, d_v, d_rr) and a host
    // copy (h_*), which device assembly leaves stale until a host-side use
    // (the reference-form SpMV, save) downloads it.
    std::vector<int> h_rp, h_ci;
    std::vector<double> h_v;
    long long nnz_op = 0;
    bool host_csr_valid = false;
    void ensure_host_csr() {
        if (host_csr_valid) return;
        const long long nd = (long long)NC * NC * NC;
        h_rp.resize(nd + 1);
        h_ci.resize(nnz_op);
        h_v.resize(nnz_op);
        HIFDE_CHECK(hipMemcpyAsync(h_rp.data(), d_rp.p, (nd + 1) * sizeof(int),
                                   hipMemcpyDeviceToHost, stream));
        if (nnz_op) {
            HIFDE_CHECK(hipMemcpyAsync(h_ci.data(), d_ci.p, nnz_op * sizeof(int),
                                       hipMemcpyDeviceToHost, stream));
            HIFDE_CHECK(hipMemcpyAsync(h_v.data(), d_v.p, nnz_op * sizeof(double),
                                       hipMemcpyDeviceToHost, stream));
        }
        HIFDE_CHECK(hipStreamSynchronize(stream));
        host_csr_valid = true;
    }
    DevBuf<int> d_rp, d_ci, d_rr;
    // slot-major copy of the CSR (ell_width entries per row; 0: none)
    int ell_width = 0;
    DevBuf<double> d_ell_v;
    DevBuf<int> d_ell_c;
    DevBuf<double> d_v;
    // per-kernel-family device time (HIP events), enabled by hifde_set_profiling
    struct KStat {
        long long launches = 0;
        double sec = 0, flops = 0, bytes = 0;
    };
    struct PendingRec {
        int fam;
        hipEvent_t e0, e1;
        double flops, bytes;
    };
    bool profiling = false;
    std::string profile_only;  // non-empty: time only this kernel's launches
    std::vector<std::string> fam_names;
    std::vector<KStat> fam_stats;
    std::vector<PendingRec> pending;
    std::vector<hipEvent_t> ev_pool;
    std::vector<BoxBatch> cells, faces;
    std::vector<int> face_red_count;
    // workspaces
    DevBuf<unsigned char> ws_a, ws_b, ws_c;
    DevBuf<double> ws_M, ws_C, ws_W, ws_T;
    DevBuf<int> ws_perm, ws_k, ws_info;
    GJWork gj_ws;
    GemmWork gemm_ws;
    GramWork gram_ws;
    DevBuf<unsigned char> ws_mode;
    DevBuf<int> ws_gm, ws_gk;
    DevBuf<long long> ws_goffM, ws_goffC;
    DevBuf<double> ws_G;
    DevBuf<SliceSum> d_gsum;
    static constexpr long long kGramSplitTiles = 1536;  // three waves of 512 workgroups
    bool sym_gj = true;           // symmetric level matrix: half-flop block inverses
    bool split_schur = false;     // FP32 levels: Schur products as bf16x3 split products
    int host_threads = 0;         // OpenMP threads of the host planning (0: the caller's)
    bool keep_workspaces = true;  // keep the factorization workspaces for the next one
    bool ahead_option = true;     // plan the level-1 cells during the level-0 face IDs
    // Keep every box's A22 beside its inverse (the forward Apply F reads it).
    // Off: A22 is gathered straight into the inverse's buffer and inverted
    // there (no copy), and A22 buffers of earlier factorizations are freed;
    // Apply F^{-1}, PCG and every solve are unchanged, Apply F is refused.
    bool keep_a22 = true;
    bool a22_kept = false;  // the current factor holds A22 (set by factor / load)
    int gram_min_n = 16;          // faces with more columns use the Gram path
    double gram_min_eps = 1e-7;   // below this tolerance every face uses Householder
    DevBuf<double> pcg_buf, pcg_part, pcg_scalar, apply_tmp, apply_side;
    double stage_t[ST_COUNT] = {0};
    std::vector<double> stage_lt;  // per level: [level * ST_COUNT + stage]
    int cur_level = 0;
    bool stage_sync = false;  // synchronize the stream at every stage boundary
    int lu_info = 0;
    DevBuf<int> d_status;     // first nonzero block-inverse status of the factorization
    Stager stager;
    // Element storage of the level matrix (level_elem.hip), used when every
    // coupling of the operator lies inside a level-0 cell closure.
    struct CellTab {
        std::vector<int> lo, ext, tab;  // per cell: closure box and frontal position by coordinate
        std::vector<long long> off;
    } ctab;
    bool local_ok = false;   // operator couplings stay inside level-0 cell closures
    int elem_option = -1;    // -1 auto, 0 off (hash-table level matrix)
    bool elem_on = false;
    // FP32 elements when every level is FP32 (-1 auto, 0 off: FP64 elements)
    int elem_f32_option = -1;
    bool elem32 = false;
    bool plan0_valid = false;  // level-0 cell and face index built and on the device
    DevBuf<double> elem[2];  // FP64 element blocks of the previous / current level
    int elem_cur = 0;
    DevBuf<InitDesc> d_init;
    DevBuf<ExtendDesc> d_ext;
    DevBuf<FaceGatherDesc> d_fg;
    DevBuf<FaceAddDesc> d_fa;
    DevBuf<MirrorDesc> d_mir;
    DevBuf<long long> d_eoff;
    DevBuf<int> d_emap, d_spos, d_etab, d_egid;

    // Device buffers handed in by the caller: the factor's stream waits for
    // the work queued on the caller's stream before the call, without
    // blocking the host.  The caller's stream is the one named by
    // hifde_set_stream (PyTorch's current stream through the Python binding);
    // until one is named it is the legacy default stream.  Naming a stream of
    // the caller's own keeps the entry points off the legacy stream, which a
    // multithreaded caller needs while other threads capture graphs.
    hipEvent_t entry_ev = nullptr;
    hipStream_t caller_stream = nullptr;
    int device = 0;                  // the HIP device current at creation
    double* scratch_dev = nullptr;   // that device's copy of the no-allocation scratch
    bool no_alloc_call = false;      // inside hifde_apply_device_noalloc: no event timing
    void after_caller_stream() {
        if (!entry_ev) HIFDE_CHECK(hipEventCreateWithFlags(&entry_ev, hipEventDisableTiming));
        HIFDE_CHECK(hipEventRecord(entry_ev, caller_stream));
        HIFDE_CHECK(hipStreamWaitEvent(stream, entry_ev, 0));
    }

    ~hifde_factor() {
        if (entry_ev) (void)hipEventDestroy(entry_ev);
        clear_graphs();
        for (PendingRec& r : pending) {
            (void)hipEventDestroy(r.e0);
            (void)hipEventDestroy(r.e1);
        }
        for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
    }

    int width(int level) const { return (1 << level) * P; }
    int lin(int i, int j, int k) const { return i + NC * j + NC * NC * k; }
    int fp32_levels = -1;  // mixed precision: levels factored in FP32 (-1: all)

    int family(const std::string& name) {
        for (size_t i = 0; i < fam_names.size(); ++i)
            if (fam_names[i] == name) return (int)i;
        fam_names.push_back(name);
        fam_stats.emplace_back();
        return (int)fam_names.size() - 1;
    }
    template <class T>
    static std::string fam(const char* base) {
        return std::string(base) + (sizeof(T) == sizeof(float) ? "_f32" : "_f64");
    }
    hipEvent_t take_event() {
        if (!ev_pool.empty()) {
            hipEvent_t e = ev_pool.back();
            ev_pool.pop_back();
            return e;
        }
        hipEvent_t e;
        HIFDE_CHECK(hipEventCreate(&e));
        return e;
    }
    // Records one launch (or a short launch sequence) of a kernel family on
    // the factor's stream, with its algorithmic flops and bytes.
    struct Prof {
        hifde_factor* f;
        int fam = -1;
        hipEvent_t e0 = nullptr;
        double flops, bytes;
        Prof(hifde_factor* f_, const std::string& name, double fl, double by)
            : f(f_), flops(fl), bytes(by) {
            if (!f->profiling || !f->profile_only.empty() || f->no_alloc_call) return;
            fam = f->family(name);
            e0 = f->take_event();
            HIFDE_CHECK(hipEventRecord(e0, f->stream));
        }
        ~Prof() noexcept(false) {
            if (fam < 0) return;
            hipEvent_t e1 = f->take_event();
            HIFDE_CHECK(hipEventRecord(e1, f->stream));
            f->pending.push_back({fam, e0, e1, flops, bytes});
        }
    };
    // Per-launch kernel timing (families named "kernel:<name>").
    // With profile_only set, only launches of that kernel are timed and the
    // family records are off, so the timing itself costs little.
    struct LaunchTimer : KernelTimer {
        hifde_factor* f;
        hipEvent_t e0 = nullptr;
        int fam = -1;
        explicit LaunchTimer(hifde_factor* f_) : f(f_) {}
        void begin(hipStream_t s, const char* kernel) override {
            fam = -1;
            if (!f->profile_only.empty() && f->profile_only != kernel) return;
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            HIFDE_CHECK(hipStreamIsCapturing(s, &cs));
            if (cs != hipStreamCaptureStatusNone) return;  // not timed inside a graph
            fam = f->family(std::string("kernel:") + kernel);
            e0 = f->take_event();
            HIFDE_CHECK(hipEventRecord(e0, s));
        }
        void end(hipStream_t s, double flops, double bytes) override {
            if (fam < 0) return;
            hipEvent_t e1 = f->take_event();
            HIFDE_CHECK(hipEventRecord(e1, s));
            f->pending.push_back({fam, e0, e1, flops, bytes});
            fam = -1;
        }
    };
    LaunchTimer launch_timer{this};
    // Installs the launch timer for the duration of a profiled call.
    // Per call of a factor method: its kernel timer (when profiling) and its
    // stream as the owner of the buffers it grows.
    struct TimerGuard {
        KernelTimer* prev;
        hipStream_t prev_owner;
        TimerGuard(hifde_factor* f) : prev(active_timer()), prev_owner(owner_stream()) {
            if (f->profiling) set_active_timer(&f->launch_timer);
            set_owner_stream(f->stream);
        }
        ~TimerGuard() {
            set_active_timer(prev);
            set_owner_stream(prev_owner);
        }
    };

    void resolve_profile() {
        if (pending.empty()) return;
        HIFDE_CHECK(hipStreamSynchronize(stream));
        for (PendingRec& r : pending) {
            float ms = 0.f;
            HIFDE_CHECK(hipEventElapsedTime(&ms, r.e0, r.e1));
            KStat& k = fam_stats[r.fam];
            k.launches += 1;
            k.sec += ms * 1e-3;
            k.flops += r.flops;
            k.bytes += r.bytes;
            ev_pool.push_back(r.e0);
            ev_pool.push_back(r.e1);
        }
        pending.clear();
    }

    // Host time per stage; with stage_sync the stream is drained first, so
    // the time includes the stage's device work.
    void timed(Stage st, clk::time_point& t0) {
        if (stage_sync) HIFDE_CHECK(hipStreamSynchronize(stream));
        auto t1 = clk::now();
        const double dt = std::chrono::duration<double>(t1 - t0).count();
        stage_t[st] += dt;
        if ((size_t)(cur_level * ST_COUNT + st) < stage_lt.size())
            stage_lt[cur_level * ST_COUNT + st] += dt;
        t0 = t1;
    }

    // ------------------------------------------------------------------
    // Index planning (integer, host).  Orders follow the reference exactly:
    // cells scanned i -> j -> k (InteriorCellIndexData, src/Factor.cpp:359-392);
    // face columns/rows as in SkelInteractionIndexData (:645-755) built from
    // InteriorFaceDOFs (:402-521) and InteriorEdgeDOFs (:523-643).
    // ------------------------------------------------------------------
    struct AxisClass {
        std::vector<unsigned char> interior;  // per coordinate 0..NC
    };
    AxisClass axis_class(int level) const {
        AxisClass a;
        a.interior.assign(NC + 1, 0);
        int w = width(level);
        for (int x = 1; x <= N; ++x) a.interior[x] = (x % w) != 0;
        return a;
    }

    void plan_cells(int level, BoxBatch& B) {
        const int w = width(level), c = NC / w;
        const AxisClass ac = axis_class(level);
        const unsigned char* I = ac.interior.data();
        B.reset_host();
        B.nbox = c * c * c;
        B.r.assign(B.nbox, 0);
        B.s.assign(B.nbox, 0);
        B.o.assign(B.nbox, 0);
        auto box_range = [&](int cc, int& lo, int& hi) {
            lo = std::max(w * cc, 1);
            hi = std::min(w * (cc + 1), N);
        };
        // The boxes are walked in memory order (i fastest); the lists are
        // then put in the reference's scan order (i, j, k lexicographic, k
        // fastest) by sorting the live points.
#pragma omp parallel for schedule(dynamic, 1)
        for (int b = 0; b < B.nbox; ++b) {
            int cc[3] = {b / (c * c), (b / c) % c, b % c};
            int lo[3], hi[3];
            for (int a = 0; a < 3; ++a) box_range(cc[a], lo[a], hi[a]);
            int r = 0, s = 0, o = 0;
            for (int k = lo[2]; k <= hi[2]; ++k)
                for (int j = lo[1]; j <= hi[1]; ++j) {
                    const unsigned char* row = remaining.data() + lin(0, j, k);
                    for (int i = lo[0]; i <= hi[0]; ++i) {
                        if (!row[i]) continue;
                        int ni = I[i] + I[j] + I[k];
                        r += ni == 3;
                        s += ni == 2;
                        o += ni < 2;
                    }
                }
            B.r[b] = r;
            B.s[b] = s;
            B.o[b] = o;
        }
        B.red_off = prefix_i(B.r);
        B.skel_off = prefix_i(B.s);
        B.other_off = prefix_i(B.o);
        B.red_gid.resize(B.red_off[B.nbox]);
        B.skel_gid.resize(B.skel_off[B.nbox]);
        B.other_gid.resize(B.other_off[B.nbox]);
        B.skel_pass.resize(B.skel_off[B.nbox]);
        if (elem_on) {
            // per cell: frontal position [red | skel | other] by local coordinate
            ctab.lo.assign(3 * B.nbox, 0);
            ctab.ext.assign(3 * B.nbox, 0);
            ctab.off.assign(B.nbox + 1, 0);
            for (int b = 0; b < B.nbox; ++b) {
                int cc[3] = {b / (c * c), (b / c) % c, b % c};
                long long vol = 1;
                for (int a = 0; a < 3; ++a) {
                    int lo, hi;
                    box_range(cc[a], lo, hi);
                    ctab.lo[3 * b + a] = lo;
                    ctab.ext[3 * b + a] = hi - lo + 1;
                    vol *= hi - lo + 1;
                }
                ctab.off[b + 1] = ctab.off[b] + vol;
            }
            // every position starts as -1 (no live DOF); filled box by box
            // in parallel, the buffer only ever grows
            if ((long long)ctab.tab.size() < ctab.off.back()) ctab.tab.resize(ctab.off.back());
#pragma omp parallel for schedule(dynamic, 1)
            for (int b = 0; b < B.nbox; ++b)
                std::fill(ctab.tab.begin() + ctab.off[b], ctab.tab.begin() + ctab.off[b + 1], -1);
        }
        long long min_face = 0;  // skeleton entries on min faces
#pragma omp parallel reduction(+ : min_face)
        {
            std::vector<std::pair<long long, int>> pts;  // (scan key, gid)
#pragma omp for schedule(dynamic, 1)
            for (int b = 0; b < B.nbox; ++b) {
                int cc[3] = {b / (c * c), (b / c) % c, b % c};
                int lo[3], hi[3];
                for (int a = 0; a < 3; ++a) box_range(cc[a], lo[a], hi[a]);
                const long long ey = hi[1] - lo[1] + 1, ez = hi[2] - lo[2] + 1;
                pts.clear();
                for (int k = lo[2]; k <= hi[2]; ++k)
                    for (int j = lo[1]; j <= hi[1]; ++j) {
                        const int g0 = lin(0, j, k);
                        const unsigned char* row = remaining.data() + g0;
                        for (int i = lo[0]; i <= hi[0]; ++i)
                            if (row[i])
                                pts.push_back({((i - lo[0]) * ey + (j - lo[1])) * ez + (k - lo[2]),
                                               g0 + i});
                    }
                std::sort(pts.begin(), pts.end());
                int* rg = B.red_gid.data() + B.red_off[b];
                int* sg = B.skel_gid.data() + B.skel_off[b];
                int* og = B.other_gid.data() + B.other_off[b];
                unsigned char* sp = B.skel_pass.data() + B.skel_off[b];
                int* tb = elem_on ? ctab.tab.data() + ctab.off[b] : nullptr;
                const int ex = hi[0] - lo[0] + 1;
                int pr = 0, ps = B.r[b], po = B.r[b] + B.s[b];
                for (const auto& pt : pts) {
                    const int g = pt.second;
                    const int i = g % NC, j = (g / NC) % NC, k = g / (NC * NC);
                    int ni = I[i] + I[j] + I[k];
                    int pos;
                    if (ni == 3) {
                        *rg++ = g;
                        pos = pr++;
                    } else if (ni == 2) {
                        *sg++ = g;
                        int x = !I[i] ? i : (!I[j] ? j : k);
                        int a = !I[i] ? 0 : (!I[j] ? 1 : 2);
                        *sp = (x == w * cc[a]) ? 1 : 0;  // on a min face
                        min_face += *sp++;
                        pos = ps++;
                    } else {
                        *og++ = g;
                        pos = po++;
                    }
                    if (tb) tb[(i - lo[0]) + ex * ((j - lo[1]) + ey * (k - lo[2]))] = pos;
                }
            }
        }
        // A cell's skeleton DOFs lie on its faces, and a face DOF is on the
        // min face of one cell and the max face of the other (the domain's
        // own faces hold no unknowns): the pass flags pair up.
        B.sides_paired = B.skel_off[B.nbox] > 0 && 2 * min_face == B.skel_off[B.nbox];
        B.off_rr = prefix_ll(B.r, B.r);
        B.off_rs = prefix_ll(B.r, B.s);
        B.off_ss = prefix_ll(B.s, B.s);
        B.tot_rr = B.off_rr.back();
        B.tot_rs = B.off_rs.back();
        B.tot_ss = B.off_ss.back();
        std::vector<int> e(B.nbox);
        for (int b = 0; b < B.nbox; ++b) e[b] = B.s[b] + B.o[b];
        B.off_ee = prefix_ll(e, e);
        B.tot_ee = B.off_ee.back();
    }

    // Frontal position of DOF g in cell b of the current level's table, or -1.
    // g / NC as a multiply and shift: with mul = floor(2^40 / NC) + 1 the
    // quotient is exact for g < 2^40 / NC, which holds for every grid index
    // when NC < 1024 (checked where the factorization is created).
    unsigned div_nc(unsigned g) const {
        const unsigned long long mul = (1ULL << 40) / (unsigned)NC + 1;
        return (unsigned)(((unsigned long long)g * mul) >> 40);
    }
    int cell_pos(int b, int g) const {
        const unsigned q1 = div_nc((unsigned)g), q2 = div_nc(q1);
        const int x = g - (int)q1 * NC, y = (int)q1 - (int)q2 * NC, z = (int)q2;
        const int lx = x - ctab.lo[3 * b], ly = y - ctab.lo[3 * b + 1], lz = z - ctab.lo[3 * b + 2];
        const int ex = ctab.ext[3 * b], ey = ctab.ext[3 * b + 1], ez = ctab.ext[3 * b + 2];
        if (lx < 0 || ly < 0 || lz < 0 || lx >= ex || ly >= ey || lz >= ez) return -1;
        return ctab.tab[ctab.off[b] + lx + (long long)ex * (ly + (long long)ey * lz)];
    }

    // DOFs of face `face` of cell (ci, cj, ck) with `want` interior coordinates
    // (2: face-interior, 1: edge-interior).  count-only when out == nullptr.
    int face_dofs(int ci, int cj, int ck, int face, int want, const unsigned char* I, int w,
                  int* out) const {
        int lo[3] = {std::max(w * ci, 1), std::max(w * cj, 1), std::max(w * ck, 1)};
        int hi[3] = {std::min(w * (ci + 1), N), std::min(w * (cj + 1), N),
                     std::min(w * (ck + 1), N)};
        int cnt = 0;
        auto take = [&](int i, int j, int k) {
            int g = lin(i, j, k);
            if (remaining[g] && I[i] + I[j] + I[k] == want) {
                if (out) out[cnt] = g;
                ++cnt;
            }
        };
        int fixed;
        switch (face) {
        case TOP:
        case BOTTOM:
            fixed = face == TOP ? lo[0] : hi[0];
            if (fixed == (face == TOP ? 1 : N)) return 0;
            for (int j = lo[1]; j <= hi[1]; ++j)
                for (int k = lo[2]; k <= hi[2]; ++k) take(fixed, j, k);
            return cnt;
        case LEFT:
        case RIGHT:
            fixed = face == LEFT ? lo[1] : hi[1];
            if (fixed == (face == LEFT ? 1 : N)) return 0;
            for (int i = lo[0]; i <= hi[0]; ++i)
                for (int k = lo[2]; k <= hi[2]; ++k) take(i, fixed, k);
            return cnt;
        default:  // FRONT / BACK
            fixed = face == FRONT ? lo[2] : hi[2];
            if (fixed == (face == FRONT ? (quirk ? 0 : 1) : N)) return 0;
            for (int i = lo[0]; i <= hi[0]; ++i)
                for (int j = lo[1]; j <= hi[1]; ++j) take(i, j, fixed);
            return cnt;
        }
    }

    struct FacePlan {
        int nface = 0;
        std::vector<int> m, n, rows, cols, rows_off, cols_off;
        std::vector<int> c1, c2;  // the face's cell (max side) and its neighbour
    };

    // Per cell of a level: the face-interior DOFs of its six faces and the
    // edge DOFs of its three min faces (face_dofs), each list scanned once.
    // The rows of a face's ID matrix are its edges, the cell's other five
    // faces, then the neighbour cell's five faces, in the order of
    // src/Factor.cpp:665-685, 694-714, 723-743; they are copied from these
    // lists.
    // Slots: 0-5 face-interior lists by FaceId, 6-8 edge lists of TOP, LEFT,
    // FRONT.
    std::vector<long long> flist_off;
    std::vector<int> flist;
    std::vector<int> ext_maps;  // extend-add maps of the current level (capacity kept)
    std::vector<std::vector<ExtendDesc>> ext_pass;  // its descriptors, eight passes
    int planned_ahead = -1;  // level whose cells plan_cells_ahead built (-1: none)
    int deferred_faces = -1;  // level whose face boxes wait for finish_deferred_faces
    std::vector<int> def_kk, def_perm;  // that level's ranks and pivots
    static int edge_slot(int face) { return face == TOP ? 6 : (face == LEFT ? 7 : 8); }

    void plan_faces(int level, FacePlan& F) {
        const int w = width(level), c = NC / w;
        const AxisClass ac = axis_class(level);
        const unsigned char* I = ac.interior.data();
        const int nb = c * c * c;
        const int ncand = 3 * nb;
        static const int order[3] = {TOP, FRONT, LEFT};  // src/Factor.cpp:249-254
        static const int own_top[5] = {BOTTOM, LEFT, RIGHT, FRONT, BACK};
        static const int nbr_top[5] = {TOP, LEFT, RIGHT, FRONT, BACK};
        static const int own_left[5] = {TOP, BOTTOM, RIGHT, FRONT, BACK};
        static const int nbr_left[5] = {TOP, BOTTOM, LEFT, FRONT, BACK};
        static const int own_front[5] = {TOP, BOTTOM, LEFT, RIGHT, BACK};
        static const int nbr_front[5] = {TOP, BOTTOM, LEFT, RIGHT, FRONT};
        auto slot_face = [](int q) { return q < 6 ? q : (q == 6 ? TOP : (q == 7 ? LEFT : FRONT)); };
        flist_off.assign((size_t)nb * 9 + 1, 0);
#pragma omp parallel for schedule(dynamic, 16)
        for (int b = 0; b < nb; ++b) {
            const int ci = b / (c * c), cj = (b / c) % c, ck = b % c;
            for (int q = 0; q < 9; ++q)
                flist_off[(size_t)b * 9 + q + 1] =
                    face_dofs(ci, cj, ck, slot_face(q), q < 6 ? 2 : 1, I, w, nullptr);
        }
        for (size_t t = 0; t < (size_t)nb * 9; ++t) flist_off[t + 1] += flist_off[t];
        flist.resize((size_t)flist_off.back());
#pragma omp parallel for schedule(dynamic, 16)
        for (int b = 0; b < nb; ++b) {
            const int ci = b / (c * c), cj = (b / c) % c, ck = b % c;
            for (int q = 0; q < 9; ++q)
                face_dofs(ci, cj, ck, slot_face(q), q < 6 ? 2 : 1, I, w,
                          flist.data() + flist_off[(size_t)b * 9 + q]);
        }
        auto len = [&](int b, int q) {
            return (int)(flist_off[(size_t)b * 9 + q + 1] - flist_off[(size_t)b * 9 + q]);
        };
        // face t = 3 b + (TOP, FRONT, LEFT): its cell b, neighbour across the face
        auto nbr_cell = [&](int b, int f) { return b - (f == TOP ? c * c : (f == LEFT ? c : 1)); };
        auto lists = [&](int f, const int*& own, const int*& nbr) {
            own = f == TOP ? own_top : (f == LEFT ? own_left : own_front);
            nbr = f == TOP ? nbr_top : (f == LEFT ? nbr_left : nbr_front);
        };
        std::vector<int> cm(ncand, 0), cn(ncand, 0);
#pragma omp parallel for schedule(static)
        for (int t = 0; t < ncand; ++t) {
            const int b = t / 3, f = order[t % 3];
            const int ci = b / (c * c), cj = (b / c) % c, ck = b % c;
            const int cc = f == TOP ? ci : (f == LEFT ? cj : ck);
            if (cc == 0) continue;  // lo == 1: no face (src/Factor.cpp:661,690,719)
            cn[t] = len(b, f);
            if (cn[t] == 0) continue;
            const int* own;
            const int* nbr;
            lists(f, own, nbr);
            const int bn = nbr_cell(b, f);
            int m = len(b, edge_slot(f));
            for (int u = 0; u < 5; ++u) m += len(b, own[u]) + len(bn, nbr[u]);
            cm[t] = m;
        }
        F.nface = 0;
        for (std::vector<int>* v : {&F.m, &F.n, &F.c1, &F.c2}) v->clear();
        std::vector<int> slot(ncand, -1);
        for (int t = 0; t < ncand; ++t)
            if (cn[t] > 0) {
                slot[t] = F.nface++;
                F.m.push_back(cm[t]);
                F.n.push_back(cn[t]);
                const int b = t / 3, f = order[t % 3];
                F.c1.push_back(b);
                F.c2.push_back(nbr_cell(b, f));
            }
        F.rows_off = prefix_i(F.m);
        F.cols_off = prefix_i(F.n);
        F.rows.resize(F.rows_off.back());
        F.cols.resize(F.cols_off.back());
#pragma omp parallel for schedule(dynamic, 16)
        for (int t = 0; t < ncand; ++t) {
            if (slot[t] < 0) continue;
            const int b = t / 3, f = order[t % 3];
            const int sidx = slot[t];
            auto put = [&](int bb, int q, int* dst) {
                const int* src = flist.data() + flist_off[(size_t)bb * 9 + q];
                const int n = len(bb, q);
                std::copy(src, src + n, dst);
                return n;
            };
            put(b, f, F.cols.data() + F.cols_off[sidx]);
            const int* own;
            const int* nbr;
            lists(f, own, nbr);
            const int bn = nbr_cell(b, f);
            int* r = F.rows.data() + F.rows_off[sidx];
            r += put(b, edge_slot(f), r);
            for (int u = 0; u < 5; ++u) r += put(b, own[u], r);
            for (int u = 0; u < 5; ++u) r += put(bn, nbr[u], r);
        }
    }

    // ------------------------------------------------------------------
    void upload_batch_index(BoxBatch& B) {
        finalize_tiles(B);
        B.d_r.upload(B.r, stream);
        B.d_s.upload(B.s, stream);
        B.d_red_off.upload(B.red_off, stream);
        B.d_skel_off.upload(B.skel_off, stream);
        B.d_red_gid.upload(B.red_gid, stream);
        B.d_skel_gid.upload(B.skel_gid, stream);
        if (!B.skel_pass.empty()) B.d_skel_pass.upload(B.skel_pass, stream);
        B.d_off_rr.upload(B.off_rr, stream);
        B.d_off_rs.upload(B.off_rs, stream);
        B.d_off_ss.upload(B.off_ss, stream);
    }

    // A22 (r x r) is in B.A22.  Computes A22inv, X = A22^{-1} A21 and
    // S = -A12 X, written to S + s_off[b] with leading dimension s_ld[b]
    // (accumulate: added to what is there).
    template <class T, class TS = T>
    void schur_batch(BoxBatch& B, const T* A21, const T* A12, TS* S, const long long* s_off,
                     const int* s_ld, bool accumulate, clk::time_point& t0) {
        const int nb = B.nbox;
        int max_r = 0;
        for (int x : B.r) max_r = std::max(max_r, x);
        if (keep_a22) {
            B.A22inv.reserve(B.tot_rr * sizeof(T));
            if (B.tot_rr)
                HIFDE_CHECK(hipMemcpyAsync(B.A22inv.p, B.A22.p, B.tot_rr * sizeof(T),
                                           hipMemcpyDeviceToDevice, stream));
        }  // else A22 was gathered into B.A22inv (a22_dest) and is inverted there
        B.X.reserve(B.tot_rs * sizeof(T));
        T* Ainv = reinterpret_cast<T*>(B.A22inv.p);
        T* X = reinterpret_cast<T*>(B.X.p);
        ws_info.reserve(nb);
        HIFDE_CHECK(hipMemsetAsync(ws_info.p, 0, nb * sizeof(int), stream));
        const double w = sizeof(T);
        double fl_inv = 0, by_inv = 0, fl_x = 0, by_x = 0, fl_s = 0, by_s = 0;
        for (int b = 0; b < nb; ++b) {
            double r = B.r[b], sd = B.s[b];
            fl_inv += 2 * r * r * r;
            by_inv += 2 * r * r * w;
            fl_x += 2 * r * r * sd;
            by_x += (r * r + 2 * r * sd) * w;
            fl_s += 2 * r * sd * sd;
            by_s += (2 * r * sd + sd * sd) * w;
        }
        {
            Prof pr(this, fam<T>("gj_inverse"), fl_inv, by_inv);
            // with the symmetric level matrix every A22 is symmetric
            batched_gj_inverse<T>(nb, B.r.data(), B.off_rr.data(), B.d_r.p, B.d_off_rr.p, Ainv,
                                  ws_info.p, gj_ws, gemm_ws, stream, sym_on && sym_gj);
        }
        fold_info(nb, ws_info.p, d_status.p, stream);
        timed(ST_INVERSE, t0);
        // FP32 levels may take the Schur products as bf16x3 split products
        struct SplitGuard {
            GemmPlan& p;
            SplitGuard(GemmPlan& q, bool on) : p(q) { p.split = on; }
            ~SplitGuard() { p.split = false; }
        } split_guard(gemm_ws.plan, split_schur && sizeof(T) == sizeof(float));
        {
            Prof pr(this, fam<T>("gemm_x"), fl_x, by_x);
            batched_gemm<T>(false, nb, B.r.data(), B.s.data(), B.r.data(), B.off_rr.data(),
                            B.off_rs.data(), B.off_rs.data(), Ainv, A21, X, (T)1, gemm_ws,
                            stream);
        }
        {
            // S = -A12 X; symmetric: S = -A21^T X, only the tiles on or below
            // the diagonal.  A box with r == 0 still writes S = 0 (the
            // reference's k == 0 GEMM) unless S accumulates.
            std::vector<GemmDesc> d;
            d.reserve(nb);
            for (int b = 0; b < nb; ++b) {
                const int r = B.r[b], sb = B.s[b];
                if (sb == 0 || (accumulate && r == 0)) continue;
                GemmDesc g{B.off_rs[b], B.off_rs[b], s_off[b], sb, sb, r,
                           std::max(sym_on ? r : sb, 1), std::max(r, 1), s_ld[b]};
                g.lower = sym_on ? 1 : 0;
                d.push_back(g);
            }
            Prof pr(this, fam<T>("gemm_s"), sym_on ? fl_s / 2 : fl_s, by_s);
            batched_gemm_desc<T, TS>(sym_on, d, sym_on ? A21 : A12, X, S, (T)-1,
                                     accumulate ? (T)1 : (T)0, gemm_ws, stream);
        }
        timed(ST_GEMM, t0);
    }

    // Where a batch's A22 blocks are gathered: their own buffer, or with
    // keep_a22 off the inverse's buffer, which is then inverted in place.
    template <class T>
    T* a22_dest(BoxBatch& B) {
        DevBuf<unsigned char>& d = keep_a22 ? B.A22 : B.A22inv;
        d.reserve(B.tot_rr * sizeof(T));
        return reinterpret_cast<T*>(d.p);
    }

    void mark_eliminated(const BoxBatch& B) {
        for (int g : B.red_gid) remaining[g] = 0;  // UpdateRemainingDOFs (:91-111)
        if (!elem_on) d_alive.upload(remaining, stream);  // read by the table compaction
    }

    // Number of S entries written in the second scatter pass (both DOFs on
    // min faces of their cell).
    double pass2_entries(const BoxBatch& B) const {
        double tot = 0;
        for (int b = 0; b < B.nbox; ++b) {
            double m = 0;
            for (int t = B.skel_off[b]; t < B.skel_off[b + 1]; ++t) m += B.skel_pass[t];
            tot += m * m;
        }
        return tot;
    }

    template <class T>
    void scatter(BoxBatch& B, const T* S, bool two_pass, clk::time_point& t0) {
        // DeleteRow/DeleteCol of the eliminated DOFs, then Sparse::Add of S
        // (src/Factor.cpp:301-309).  Compaction drops the deleted pairs.
        // Entries of eliminated DOFs are never looked up again, so they may
        // stay in the table; it is compacted only when the coming inserts
        // would push the load factor above one half.
        const long long skel_tot = B.skel_off.empty() ? 0 : B.skel_off.back();
        const size_t adds = sym_on ? (size_t)((B.tot_ss + skel_tot) / 2) : (size_t)B.tot_ss;
        if ((tab_inserted + adds) * 2 > tabs[cur].cap) {
            Prof pr(this, "compact", 0.0, (double)tabs[cur].cap * 16 * 2);
            tab_inserted = table_compact(tabs[cur], tabs[1 - cur], d_alive.p, adds, tab_counter,
                                         stream);
            cur = 1 - cur;
        }
        tab_inserted += adds;
        timed(ST_COMPACT, t0);
        // One profiling record per launch: algorithmic bytes are the S values
        // read plus one 16-byte slot read and written per entry added.
        if (two_pass) {
            for (int pass = 0; pass < 2; ++pass) {
                double cnt = pass == 0 ? (double)B.tot_ss - pass2_entries(B) : pass2_entries(B);
                Prof pr(this, fam<T>("scatter"), 0.0, (double)B.tot_ss * (4 + 4) + cnt * (sizeof(T) + 32));
                table_scatter<T>(tabs[cur], B.nbox, B.d_s.p, B.d_skel_off.p, B.d_skel_gid.p,
                                 B.d_skel_pass.p, B.d_off_ss.p, B.tot_ss, S, pass, stream);
            }
        } else {
            Prof pr(this, fam<T>("scatter"), 0.0, (double)B.tot_ss * (sizeof(T) + 32));
            table_scatter<T>(tabs[cur], B.nbox, B.d_s.p, B.d_skel_off.p, B.d_skel_gid.p, nullptr,
                             B.d_off_ss.p, B.tot_ss, S, -1, stream);
        }
        timed(ST_SCATTER, t0);
    }

    template <class T>
    void level_cells(int level) {
        auto t0 = clk::now();
        BoxBatch& B = cells[level];
        // the level-0 cells depend on the grid only (kept with the element storage)
        const bool cached = level == 0 && elem_on && plan0_valid;
        // or planned during the previous level's face IDs (plan_cells_ahead)
        const bool ahead = planned_ahead == level;
        if (!cached && !ahead) plan_cells(level, B);
        B.esize = sizeof(T);
        timed(ST_PLAN, t0);
        if (!cached && !ahead) upload_batch_index(B);
        timed(ST_UPLOAD, t0);
        if (elem_on) {
            if constexpr (std::is_same<T, float>::value) {
                if (elem32) {
                    level_cells_elem<float, float>(level, t0);
                    return;
                }
            }
            level_cells_elem<T, double>(level, t0);
            return;
        }
        T* A22 = a22_dest<T>(B);
        ws_a.reserve(B.tot_rs * sizeof(T));
        ws_b.reserve(B.tot_rs * sizeof(T));
        ws_c.reserve(B.tot_ss * sizeof(T));
        T* A21 = reinterpret_cast<T*>(ws_a.p);
        T* A12 = reinterpret_cast<T*>(ws_b.p);
        T* S = reinterpret_cast<T*>(ws_c.p);
        // DenseSubmatrix of the cell's DOFs (include/Schur.hpp:44-61), only
        // the three blocks Schur reads.
        // One profiling record per launch: algorithmic bytes are one 16-byte
        // slot read and one value written per element.
        {
            Prof pr(this, fam<T>("gather"), 0.0, (double)B.tot_rr * (sizeof(T) + 16));
            table_gather<T>(tabs[cur], B.nbox, B.d_red_gid.p, B.d_red_off.p, B.d_red_gid.p,
                            B.d_red_off.p, B.d_off_rr.p, B.tot_rr, A22, stream);
        }
        {
            Prof pr(this, fam<T>("gather"), 0.0, (double)B.tot_rs * (sizeof(T) + 16));
            table_gather<T>(tabs[cur], B.nbox, B.d_red_gid.p, B.d_red_off.p, B.d_skel_gid.p,
                            B.d_skel_off.p, B.d_off_rs.p, B.tot_rs, A21, stream);
        }
        if (!sym_on) {  // A12 = A21^T exactly when the level matrix is stored symmetric
            Prof pr(this, fam<T>("gather"), 0.0, (double)B.tot_rs * (sizeof(T) + 16));
            table_gather<T>(tabs[cur], B.nbox, B.d_skel_gid.p, B.d_skel_off.p, B.d_red_gid.p,
                            B.d_red_off.p, B.d_off_rs.p, B.tot_rs, A12, stream);
        }
        timed(ST_GATHER, t0);
        schur_batch<T>(B, A21, A12, S, B.off_ss.data(), B.s.data(), false, t0);
        mark_eliminated(B);
        scatter<T>(B, S, true, t0);
    }

    // Level-0 frontal assembly index: per cell its DOF list [red | skel |
    // other], position table and block offsets, on the device.
    void init_level0_index(const BoxBatch& B) {
        const int nb = B.nbox, c = NC / P;
        std::vector<InitDesc> desc(nb);
        std::vector<int> gid_off(nb + 1, 0);
        for (int b = 0; b < nb; ++b) gid_off[b + 1] = gid_off[b] + B.r[b] + B.s[b] + B.o[b];
        std::vector<int> gids(gid_off[nb]);
#pragma omp parallel for schedule(static)
        for (int b = 0; b < nb; ++b) {
            InitDesc& d = desc[b];
            d.rr = B.off_rr[b];
            d.rs = B.off_rs[b];
            d.ee = B.off_ee[b];
            d.tab = ctab.off[b];
            d.gid = gid_off[b];
            d.r = B.r[b];
            d.s = B.s[b];
            d.o = B.o[b];
            const int cc[3] = {b / (c * c), (b / c) % c, b % c};
            for (int a = 0; a < 3; ++a) {
                d.lo[a] = ctab.lo[3 * b + a];
                d.ext[a] = ctab.ext[3 * b + a];
                d.cc[a] = cc[a];
            }
            int* g = gids.data() + gid_off[b];
            g = std::copy(B.red_gid.begin() + B.red_off[b], B.red_gid.begin() + B.red_off[b + 1], g);
            g = std::copy(B.skel_gid.begin() + B.skel_off[b], B.skel_gid.begin() + B.skel_off[b + 1],
                          g);
            std::copy(B.other_gid.begin() + B.other_off[b], B.other_gid.begin() + B.other_off[b + 1],
                      g);
        }
        d_init.upload(desc, stream);
        d_etab.upload(ctab.tab.data(), ctab.off.back(), stream);
        d_egid.upload(gids, stream);
    }

    // Extend-add plan of a level (> 0) with the element storage: for every
    // (parent, child) pair the parent position of each entry of the child's
    // element (ext_maps, uploaded to d_emap) and one descriptor per pair in
    // eight passes by child position (ext_pass).  Reads the level's cell plan
    // and position table (plan_cells) and the previous level's cell lists.
    void plan_extend(int level) {
        const BoxBatch& B = cells[level];
        const BoxBatch& Cb = cells[level - 1];
        const int nb = B.nbox;
        const int c = NC / width(level), cc = 2 * c;
        // child k of a parent: (2 pi + k/4, 2 pj + k/2 % 2, 2 pk + k % 2)
        auto child = [&](int b, int k) {
            const int pi = b / (c * c), pj = (b / c) % c, pk = b % c;
            return (2 * pi + (k >> 2)) * cc * cc + (2 * pj + ((k >> 1) & 1)) * cc +
                   (2 * pk + (k & 1));
        };
        std::vector<long long> moff((size_t)nb * 8 + 1, 0);
        for (int b = 0; b < nb; ++b)
            for (int k = 0; k < 8; ++k) {
                const int ch = child(b, k);
                moff[(size_t)b * 8 + k + 1] = Cb.s[ch] + Cb.o[ch];
            }
        for (size_t t = 0; t < (size_t)nb * 8; ++t) moff[t + 1] += moff[t];
        if (moff.back() > (long long)INT32_MAX) throw Error(HIFDE_ERR_SIZE, "extend-add maps");
        std::vector<int>& maps = ext_maps;  // every entry is written below
        if ((long long)maps.size() < moff.back()) maps.resize(moff.back());
        ext_pass.assign(8, std::vector<ExtendDesc>(nb));
#pragma omp parallel for schedule(dynamic, 1)
        for (int b = 0; b < nb; ++b)
            for (int k = 0; k < 8; ++k) {
                const int ch = child(b, k);
                ExtendDesc& d = ext_pass[k][b];
                d.src = Cb.off_ee[ch];
                d.rr = B.off_rr[b];
                d.rs = B.off_rs[b];
                d.ee = B.off_ee[b];
                d.e = Cb.s[ch] + Cb.o[ch];
                d.map = (int)moff[(size_t)b * 8 + k];
                d.r = B.r[b];
                d.s = B.s[b];
                d.ep = B.s[b] + B.o[b];
                int* m = maps.data() + d.map;
                for (int t = 0; t < Cb.s[ch]; ++t)
                    *m++ = cell_pos(b, Cb.skel_gid[Cb.skel_off[ch] + t]);
                for (int t = 0; t < Cb.o[ch]; ++t)
                    *m++ = cell_pos(b, Cb.other_gid[Cb.other_off[ch] + t]);
            }
        d_emap.upload(maps.data(), moff.back(), stream);
    }

    // The next level's cells planned while the device runs this level's
    // face IDs: the cell plan, its device index and (element storage) the
    // extend-add plan, built from the remaining-DOF mask as it stands before
    // the faces' redundant DOFs are removed.  level_faces keeps it only when
    // its faces eliminate nothing (every face of full numerical rank, as the
    // level-0 faces at P = 4 and eps 1e-4), in which case the mask the next
    // level sees is that same mask.  The previous level's position table is
    // not read after its faces are planned, so it is overwritten here.
    void plan_cells_ahead(int level) {
        BoxBatch& B = cells[level];
        plan_cells(level, B);
        upload_batch_index(B);
        plan_extend(level);
        planned_ahead = level;
    }

    // Cells of a level with the element storage: the frontal blocks A22,
    // A21 (A12) and the element over [skel | other] are assembled from the
    // operator (level 0) or by extend-add of the eight children's elements;
    // S = -A12 A22^{-1} A21 is then added to the element by the Schur GEMM.
    // The element buffers hold TE values (elements are FP32 only when every
    // level is FP32): reserved in doubles, rounded up.
    template <class TE>
    TE* elem_ptr(int k, long long count = -1) {
        if (count >= 0)
            elem[k].reserve(std::max<long long>((count * (long long)sizeof(TE) + 7) / 8, 1));
        return reinterpret_cast<TE*>(elem[k].p);
    }

    template <class T, class TE>
    void level_cells_elem(int level, clk::time_point& t0) {
        BoxBatch& B = cells[level];
        const int nb = B.nbox;
        T* A22 = a22_dest<T>(B);
        ws_a.reserve(B.tot_rs * sizeof(T));
        if (!sym_on) ws_b.reserve(B.tot_rs * sizeof(T));
        const int ne = 1 - elem_cur;
        T* A21 = reinterpret_cast<T*>(ws_a.p);
        T* A12 = sym_on ? nullptr : reinterpret_cast<T*>(ws_b.p);
        TE* E = elem_ptr<TE>(ne, B.tot_ee);
        const double w = sizeof(T), we = sizeof(TE);
        const double fr_bytes =
            ((double)B.tot_rr + (sym_on ? 1.0 : 2.0) * B.tot_rs) * w + (double)B.tot_ee * we;
        if (level == 0) {
            const int c = NC / P;
            if (!plan0_valid) init_level0_index(B);
            Prof pr(this, fam<T>("assemble"), 0.0, fr_bytes + (double)nnz_op * 12.0);
            frontal_init<T, TE>(nb, d_init.p, d_etab.p, d_egid.p, d_rp.p, d_ci.p, d_v.p, NC, P, c,
                            A22, A21, A12, E, stream);
        } else {
            const BoxBatch& Cb = cells[level - 1];
            if (planned_ahead != level) plan_extend(level);
            timed(ST_PLAN, t0);
            Prof pr(this, fam<T>("assemble"), 0.0, fr_bytes + (double)Cb.tot_ee * (we + 4.0));
            fill_zero<T>(A22, B.tot_rr, stream);
            fill_zero<T>(A21, B.tot_rs, stream);
            if (A12) fill_zero<T>(A12, B.tot_rs, stream);
            fill_zero<TE>(E, B.tot_ee, stream);
            extend_add<T, TE>(ext_pass, d_ext, d_eoff, d_emap.p, elem_ptr<TE>(elem_cur), A22, A21,
                              A12, E, stream);
        }
        elem_cur = ne;
        timed(ST_GATHER, t0);
        std::vector<int> eld(nb);
        for (int b = 0; b < nb; ++b) eld[b] = std::max(B.s[b] + B.o[b], 1);
        schur_batch<T, TE>(B, A21, A12, E, B.off_ee.data(), eld.data(), true, t0);
        if (sym_on) {
            std::vector<MirrorDesc> md;
            for (int b = 0; b < nb; ++b)
                if (B.s[b] > 1 && B.r[b] > 0) md.push_back({B.off_ee[b], B.s[b], eld[b]});
            mirror_lower(md, d_mir, d_eoff, E, stream);
        }
        timed(ST_SCATTER, t0);
        mark_eliminated(B);
    }

    // Index of one level's faces: the plan and its device copies.
    struct FaceDev {
        FacePlan F;
        std::vector<long long> offM, offC;
        std::vector<int> cpos1;  // element position of each face DOF in the face's own cell
        DevBuf<int> rows, rows_off, cols, cols_off, m, n;
        DevBuf<long long> offMd, offCd;
        DevBuf<unsigned> rpos, cpos;  // packed positions in the two adjacent elements
        DevBuf<FaceGatherDesc> gd;
        DevBuf<long long> gd_off;
        long long gd_cols = 0;
        // host side of the positions, computed with the plan
        std::vector<unsigned> h_rpos, h_cpos;
        std::vector<FaceGatherDesc> h_gd;
    };
    FaceDev fdev[2];  // [0]: level 0, kept while plan0_valid; [1]: the other levels

    // The face plan of a level: plan_faces and the element positions.
    void plan_faces_host(int level, FaceDev& fd) {
        plan_faces(level, fd.F);
        fd.offM = prefix_ll(fd.F.m, fd.F.n);
        fd.offC = prefix_ll(fd.F.n, fd.F.n);
        if (elem_on) face_positions_host(level, fd);
    }

    // Element positions of the face rows and columns in the two adjacent
    // cells, packed as low / high halves (kNoPos when absent).
    void face_positions_host(int level, FaceDev& fd) {
        const FacePlan& F = fd.F;
        const BoxBatch& Cb = cells[level];
        const int nf = F.nface;
        std::vector<unsigned>& rpos = fd.h_rpos;
        std::vector<unsigned>& cpos = fd.h_cpos;
        rpos.resize(F.rows_off.back());
        cpos.resize(F.cols_off.back());
        fd.cpos1.assign(F.cols_off.back(), 0);
        std::vector<FaceGatherDesc>& gd = fd.h_gd;
        gd.resize(nf);
        int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(max : bad)
        for (int f = 0; f < nf; ++f) {
            const int c1 = F.c1[f], c2 = F.c2[f], m = F.m[f], n = F.n[f];
            auto epos = [&](int cell, int g) -> unsigned {
                const int p = cell_pos(cell, g);
                return p >= Cb.r[cell] ? (unsigned)(p - Cb.r[cell]) : kNoPos;
            };
            auto pack = [&](unsigned p1, unsigned p2) {
                if ((p1 != kNoPos && p1 >= kNoPos) || (p2 != kNoPos && p2 >= kNoPos)) bad = 1;
                return (p1 & 0xFFFFu) | (p2 << 16);
            };
            for (int i = 0; i < m; ++i) {
                const int g = F.rows[F.rows_off[f] + i];
                rpos[F.rows_off[f] + i] = pack(epos(c1, g), epos(c2, g));
            }
            for (int j = 0; j < n; ++j) {
                const int g = F.cols[F.cols_off[f] + j];
                const unsigned p1 = epos(c1, g), p2 = epos(c2, g);
                if (p1 == kNoPos || p2 == kNoPos) bad = 2;
                cpos[F.cols_off[f] + j] = pack(p1, p2);
                fd.cpos1[F.cols_off[f] + j] = (int)p1;
            }
            FaceGatherDesc& d = gd[f];
            d.e1 = Cb.off_ee[c1];
            d.e2 = Cb.off_ee[c2];
            d.offM = fd.offM[f];
            d.offC = fd.offC[f];
            d.ld1 = Cb.s[c1] + Cb.o[c1];
            d.ld2 = Cb.s[c2] + Cb.o[c2];
            d.m = m;
            d.n = n;
            d.rows = F.rows_off[f];
            d.cols = F.cols_off[f];
        }
        if (bad == 1) throw Error(HIFDE_ERR_SIZE, "element larger than 65535 DOFs");
        if (bad == 2) throw Error(HIFDE_ERR_STATE, "face DOF outside its cells");
    }
    void face_positions_upload(FaceDev& fd) {
        fd.rpos.upload(fd.h_rpos, stream);
        fd.cpos.upload(fd.h_cpos, stream);
        fd.gd_cols = face_gather_prepare(fd.h_gd, fd.gd, fd.gd_off, stream);
    }

    // Face boxes of a level from the ID's ranks and pivots: red = the ID's
    // redundant columns, skel = its skeleton columns (Skeletonize,
    // src/Factor.cpp:171-178), and their device index.
    void face_boxes(int level, const std::vector<int>& kk, const std::vector<int>& perm) {
        const FacePlan& F = fdev[level == 0 ? 0 : 1].F;
        BoxBatch& B = faces[level];
        const int nf = F.nface;
        B.r.resize(nf);
        B.s.resize(nf);
        for (int f = 0; f < nf; ++f) {
            B.s[f] = kk[f];
            B.r[f] = F.n[f] - kk[f];
        }
        B.red_off = prefix_i(B.r);
        B.skel_off = prefix_i(B.s);
        B.red_gid.resize(B.red_off.back());
        B.skel_gid.resize(B.skel_off.back());
        for (int f = 0; f < nf; ++f) {
            const int* p = perm.data() + F.cols_off[f];
            const int* cg = F.cols.data() + F.cols_off[f];
            for (int t = 0; t < B.s[f]; ++t) B.skel_gid[B.skel_off[f] + t] = cg[p[t]];
            for (int t = 0; t < B.r[f]; ++t) B.red_gid[B.red_off[f] + t] = cg[p[B.s[f] + t]];
        }
        B.off_rr = prefix_ll(B.r, B.r);
        B.off_rs = prefix_ll(B.r, B.s);
        B.off_ss = prefix_ll(B.s, B.s);
        B.tot_rr = B.off_rr.back();
        B.tot_rs = B.off_rs.back();
        B.tot_ss = B.off_ss.back();
        upload_batch_index(B);
    }
    // The face boxes level_faces left for later (deferred_faces >= 0).
    void finish_deferred_faces() {
        if (deferred_faces < 0) return;
        const int level = deferred_faces;
        deferred_faces = -1;
        face_boxes(level, def_kk, def_perm);
    }

    template <class T>
    void level_faces(int level) {
        auto t0 = clk::now();
        // The level-0 faces depend on the grid only: with the element
        // storage their index is built once and kept.
        const bool cached = level == 0 && elem_on && plan0_valid;
        FaceDev& fd = fdev[level == 0 ? 0 : 1];
        const FacePlan& F = fd.F;
        if (!cached) plan_faces_host(level, fd);
        timed(ST_PLAN, t0);
        const int nf = F.nface;
        BoxBatch& B = faces[level];
        B.reset_host();
        B.nbox = nf;
        B.esize = sizeof(T);
        if (nf == 0) return;
        const std::vector<long long>& offM = fd.offM;
        const std::vector<long long>& offC = fd.offC;
        if (!cached) {
            if (!elem_on) {
                fd.rows.upload(F.rows, stream);
                fd.rows_off.upload(F.rows_off, stream);
                fd.cols.upload(F.cols, stream);
            }
            fd.cols_off.upload(F.cols_off, stream);
            fd.m.upload(F.m, stream);
            fd.n.upload(F.n, stream);
            fd.offMd.upload(offM, stream);
            fd.offCd.upload(offC, stream);
            if (elem_on) face_positions_upload(fd);
            if (level == 0 && elem_on) plan0_valid = true;
        }
        timed(ST_UPLOAD, t0);
        ws_M.reserve(offM.back());
        ws_C.reserve(offC.back());
        ws_W.reserve(offC.back());
        ws_T.reserve(offC.back());
        if (elem_on) {
            // A(rows, F) and A(F, F) summed from the two adjacent cells' elements
            Prof pr(this, "gather_f64", 0.0,
                    (double)(offM.back() + offC.back()) * (8 + 2.0 * (elem32 ? 4 : 8)));
            if (elem32)
                face_gather<float>(nf, fd.gd_cols, fd.gd.p, fd.gd_off.p, fd.rpos.p, fd.cpos.p,
                                   elem_ptr<float>(elem_cur), ws_M.p, ws_C.p, stream);
            else
                face_gather<double>(nf, fd.gd_cols, fd.gd.p, fd.gd_off.p, fd.rpos.p, fd.cpos.p,
                                    elem_ptr<double>(elem_cur), ws_M.p, ws_C.p, stream);
        } else {
            {
                Prof pr(this, "gather_f64", 0.0, (double)offM.back() * (8 + 16));
                table_gather<double>(tabs[cur], nf, fd.rows.p, fd.rows_off.p, fd.cols.p,
                                     fd.cols_off.p, fd.offMd.p, offM.back(), ws_M.p, stream);
            }
            {
                Prof pr(this, "gather_f64", 0.0, (double)offC.back() * (8 + 16));
                table_gather<double>(tabs[cur], nf, fd.cols.p, fd.cols_off.p, fd.cols.p,
                                     fd.cols_off.p, fd.offCd.p, offC.back(), ws_C.p, stream);
            }
        }
        timed(ST_GATHER, t0);
        ws_k.reserve(nf);
        ws_perm.reserve(F.cols_off.back());
        // Large tall faces take the Gram path (G = M^T M, pivoted Cholesky);
        // the rest, and every face when eps is below the Gram path's noise
        // floor, take the Householder column-pivoted QR.
        std::vector<unsigned char> mode(nf, 0);
        std::vector<int> gm, gn, gk;
        std::vector<long long> goffM, goffC;
        int max_n_house = 0, max_m_house = 0;
        double fl_h = 0, by_h = 0, fl_g = 0, by_g = 0, fl_c = 0;
        for (int f = 0; f < nf; ++f) {
            double m = F.m[f], n = F.n[f];
            bool gram = F.n[f] > gram_min_n && F.m[f] >= F.n[f] && eps >= gram_min_eps;
            mode[f] = gram ? 1 : 0;
            if (gram) {
                gm.push_back(F.n[f]);
                gn.push_back(F.n[f]);
                gk.push_back(F.m[f]);
                goffM.push_back(offM[f]);
                goffC.push_back(offC[f]);
                fl_g += 2 * m * n * n;
                by_g += 8 * (m * n + n * n);
                fl_c += n * n * n / 3;
            } else {
                max_n_house = std::max(max_n_house, F.n[f]);
                max_m_house = std::max(max_m_house, F.m[f]);
                double k = std::min(m, n);
                fl_h += 4 * m * n * k - 2 * k * k * (m + n) + 4.0 / 3.0 * k * k * k;
                by_h += 8 * (2 * m * n + n * n);
            }
        }
        ws_mode.upload(mode, stream);
        if (!gm.empty()) {
            int ng = (int)gm.size();
            ws_gm.upload(gm, stream);
            ws_gk.upload(gk, stream);
            ws_goffM.upload(goffM, stream);
            ws_goffC.upload(goffC, stream);
            // G = M^T M is symmetric: only the tiles on or below the diagonal.
            // A few large faces (the top levels) give too few output tiles
            // to fill the GPU over their long k: their k is then split into
            // slices of at least 256 rows, each product written to its own
            // partial block after the level's G, and the partials summed in
            // slice order (deterministic).
            long long lower_tiles = 0;
            for (int t = 0; t < ng; ++t) {
                const long long tm = (gm[t] + kMfmaTile - 1) / kMfmaTile;
                lower_tiles += tm * (tm + 1) / 2;
            }
            const int split = lower_tiles >= kGramSplitTiles
                                  ? 1
                                  : (int)std::min<long long>(
                                        8, (kGramSplitTiles + lower_tiles - 1) / lower_tiles);
            std::vector<GemmDesc> gd;
            std::vector<SliceSum> sums;
            long long part = offC.back();  // partial blocks follow the level's G
            for (int t = 0; t < ng; ++t) {
                const int sf = std::min(split, std::max(1, gk[t] / 256));
                const long long nn = (long long)gn[t] * gn[t];
                for (int q = 0; q < sf; ++q) {
                    const int k0 = (int)((long long)gk[t] * q / sf);
                    const int k1 = (int)((long long)gk[t] * (q + 1) / sf);
                    GemmDesc g{goffM[t] + k0, goffM[t] + k0, sf == 1 ? goffC[t] : part + q * nn,
                               gm[t], gn[t], k1 - k0, std::max(gk[t], 1), std::max(gk[t], 1),
                               std::max(gm[t], 1)};
                    g.lower = 1;
                    gd.push_back(g);
                }
                if (sf > 1) {
                    sums.push_back({goffC[t], part, gn[t], sf});
                    part += sf * nn;
                }
            }
            ws_G.reserve(part);
            ws_T.reserve(offC.back());
            {
                Prof pr(this, "gram_gemm", fl_g / 2, by_g);
                batched_gemm_desc<double>(true, gd, ws_M.p, ws_M.p, ws_G.p, 1.0, 0.0, gemm_ws,
                                          stream);
                sum_slices_lower(sums, d_gsum, ws_G.p, stream);
            }
            {
                Prof pr(this, "gram_pivchol", fl_c, 8.0 * offC.back());
                gram_pivchol(nf, F.n.data(), mode.data(), fd.n.p, ws_mode.p, ws_G.p, offC.data(),
                             fd.offCd.p, ws_T.p, eps, ws_perm.p, F.cols_off.data(),
                             fd.cols_off.p, ws_k.p, gram_ws, gemm_ws, stream);
            }
        }
        if (max_n_house > 0 || (int)gm.size() < nf) {
            Prof pr(this, "interp_decomp", fl_h, by_h);
            batched_interp_decomp(nf, fd.m.p, fd.n.p, fd.offMd.p, ws_M.p, eps, ws_perm.p,
                                  fd.cols_off.p, ws_k.p, ws_W.p, fd.offCd.p, max_n_house,
                                  max_m_house, ws_mode.p, stream);
        }
        timed(ST_ID, t0);
        // The level-0 faces are usually of full numerical rank: the host
        // plans the level-1 cells while the device runs their IDs instead of
        // after the synchronization below (kept only if the faces eliminate
        // nothing; element storage, whose level-0 faces read no position
        // table after their own plan).
        if (level == 0 && elem_on && L > 1 && ahead_option) {
            plan_cells_ahead(1);
            timed(ST_PLAN, t0);
        }
        std::vector<int> kk(nf), perm(F.cols_off.back());
        HIFDE_CHECK(hipMemcpyAsync(kk.data(), ws_k.p, nf * sizeof(int), hipMemcpyDeviceToHost,
                                   stream));
        HIFDE_CHECK(hipMemcpyAsync(perm.data(), ws_perm.p, perm.size() * sizeof(int),
                                   hipMemcpyDeviceToHost, stream));
        // the level's one synchronization: the skeleton sets plan what follows
        HIFDE_CHECK(hipStreamSynchronize(stream));
        stager.synced();
        timed(ST_DOWNLOAD, t0);
        if (!gm.empty()) {
            // Gram faces: W = R11^{-1} R12 now that the ranks are known
            std::vector<GramRDesc> rd;
            std::vector<long long> toff(1, 0), offW;
            double fl_w = 0;
            for (int f = 0; f < nf; ++f) {
                const int k = kk[f], n = F.n[f];
                if (!mode[f] || k == 0 || k == n) continue;
                rd.push_back({offC[f], toff.back(), k, n, F.cols_off[f]});
                toff.push_back(toff.back() + (long long)k * n);
                offW.push_back(offC[f]);
                fl_w += 2.0 * k * k * k + 2.0 * k * k * (n - k);
            }
            Prof pr(this, "gram_w", fl_w, 8.0 * toff.back() * 2);
            gram_w(rd, toff, ws_T.p, ws_perm.p, offW, ws_W.p, gram_ws, gj_ws, gemm_ws, stream);
        }
        // Every face kept all its columns and the next level's cells were
        // planned during the IDs: the face boxes (read by Apply and save
        // only) are built after the next level's cell launches are issued.
        bool none_red = true;
        for (int f = 0; f < nf && none_red; ++f) none_red = kk[f] == F.n[f];
        if (none_red && planned_ahead == level + 1) {
            deferred_faces = level;
            def_kk = std::move(kk);
            def_perm = std::move(perm);
            timed(ST_PLAN, t0);
            return;
        }
        face_boxes(level, kk, perm);
        // a face eliminates DOFs: the next level's cells see another mask
        if (B.tot_rr != 0 && planned_ahead == level + 1) planned_ahead = -1;
        timed(ST_PLAN, t0);
        if (B.tot_rr == 0) {
            // Every face keeps all its columns (full numerical rank, as every
            // level-0 face at P = 4 and eps 1e-4): nothing is eliminated,
            // SchurAfterID's Schur complement is zero and the level matrix
            // is unchanged (src/Factor.cpp:113-153, 262-311 with an empty
            // redundant set), so the transform, the Schur step and the
            // update are skipped.
            return;
        }
        T* A22 = a22_dest<T>(B);
        ws_a.reserve(B.tot_rs * sizeof(T));
        ws_b.reserve(B.tot_rs * sizeof(T));
        ws_c.reserve(B.tot_ss * sizeof(T));
        T* A21 = reinterpret_cast<T*>(ws_a.p);
        T* A12 = reinterpret_cast<T*>(ws_b.p);
        T* S = reinterpret_cast<T*>(ws_c.p);
        {
            double fl = 0, by = 0;
            for (int f = 0; f < nf; ++f) {
                double k = B.s[f], q = B.r[f], n = k + q;
                fl += 4 * k * k * q + 4 * k * q * q;
                by += 8 * (n * n + k * q) + sizeof(T) * (q * q + 2 * k * q);
            }
            Prof pr(this, "face_transform", fl, by);
            face_transform<T>(fd, B, A22, A21, A12);
        }
        // W (k x q) kept for Apply, packed at the r*s offsets.
        B.W.reserve(B.tot_rs * sizeof(T));
        pack_w<T>(B, fd);
        timed(ST_TRANSFORM, t0);
        schur_batch<T>(B, A21, A12, S, B.off_ss.data(), B.s.data(), false, t0);
        mark_eliminated(B);
        if (elem_on) {
            // A(skel_f, skel_f) += S_f, into the element of the face's own cell
            const BoxBatch& Cb = cells[level];
            std::vector<int> spos(B.skel_off.back());
            std::vector<FaceAddDesc> ad;
            ad.reserve(nf);
            for (int f = 0; f < nf; ++f) {
                const int* p = perm.data() + F.cols_off[f];
                for (int t = 0; t < B.s[f]; ++t)
                    spos[B.skel_off[f] + t] = fd.cpos1[F.cols_off[f] + p[t]];
                if (B.s[f] == 0) continue;
                const int c1 = F.c1[f];
                ad.push_back({Cb.off_ee[c1], B.off_ss[f], Cb.s[c1] + Cb.o[c1], B.s[f],
                              B.skel_off[f]});
            }
            d_spos.upload(spos, stream);
            Prof pr(this, fam<T>("face_add"), 0.0, (double)B.tot_ss * (sizeof(T) + 16));
            if constexpr (std::is_same<T, float>::value) {
                if (elem32)
                    face_add<float, float>(ad, d_fa, d_eoff, d_spos.p, S, elem_ptr<float>(elem_cur),
                                           sym_on ? 1 : 0, stream);
                else
                    face_add<float, double>(ad, d_fa, d_eoff, d_spos.p, S,
                                            elem_ptr<double>(elem_cur), sym_on ? 1 : 0, stream);
            } else {
                face_add<T, double>(ad, d_fa, d_eoff, d_spos.p, S, elem_ptr<double>(elem_cur),
                                    sym_on ? 1 : 0, stream);
            }
            timed(ST_SCATTER, t0);
        } else {
            scatter<T>(B, S, false, t0);
        }
    }

    // Blocks of Rot^T C Rot with Rot = I - E_{skel,red} W and C = A(F,F)
    // (SchurAfterID, src/Factor.cpp:113-153), without forming Rot:
    //   A12 = Tsr = C_sr - C_ss W                 (k x q)
    //   A22 = C_rr - C_rs W - W^T Tsr             (q x q)
    //   A21 = C_rs - W^T C_ss                     (q x k)
    // C is first permuted to skeleton-then-redundant order; all products are
    // FP64 batched GEMMs on sub-blocks of it; the blocks are then packed in T.
    template <class T>
    void face_transform(const FaceDev& fd, BoxBatch& B, T* A22, T* A21, T* A12) {
        const FacePlan& F = fd.F;
        const std::vector<long long>& offC = fd.offC;
        const int nf = B.nbox;
        ws_T.reserve(offC.back());
        permute_sym(nf, fd.offCd.p, fd.n.p, ws_perm.p, fd.cols_off.p, ws_C.p, ws_T.p, offC.back(),
                    stream);
        std::vector<GemmDesc> d1, d2, d3, d4;
        for (std::vector<GemmDesc>* v : {&d1, &d2, &d3, &d4}) v->reserve(nf);
        std::vector<CopyDesc> c22(nf), c21(nf), c12(nf);
        for (int f = 0; f < nf; ++f) {
            const int n = F.n[f], k = B.s[f], q = B.r[f];
            const long long base = offC[f], wb = offC[f], kn = (long long)k * n;
            c22[f] = {base, k, k, q, n};
            c21[f] = {base, k, 0, q, n};
            c12[f] = {base, 0, k, k, n};
            if (k == 0 || q == 0) continue;
            d1.push_back({base, wb, base + kn, k, q, k, n, k, n});          // Csr -= Css W
            d2.push_back({base + k, wb, base + k + kn, q, q, k, n, k, n});  // Crr -= Crs W
            d3.push_back({wb, base + kn, base + k + kn, q, q, k, k, n, n}); // Crr -= W^T Tsr
            d4.push_back({wb, base, base + k, q, k, k, k, n, n});           // Crs -= W^T Css
        }
        batched_gemm_desc<double>(false, d1, ws_T.p, ws_W.p, ws_T.p, -1.0, 1.0, gemm_ws, stream);
        batched_gemm_desc<double>(false, d2, ws_T.p, ws_W.p, ws_T.p, -1.0, 1.0, gemm_ws, stream);
        batched_gemm_desc<double>(true, d3, ws_W.p, ws_T.p, ws_T.p, -1.0, 1.0, gemm_ws, stream);
        batched_gemm_desc<double>(true, d4, ws_W.p, ws_T.p, ws_T.p, -1.0, 1.0, gemm_ws, stream);
        copy_blocks<T>(c22, B.off_rr, ws_T.p, A22, gemm_ws, stream);
        copy_blocks<T>(c21, B.off_rs, ws_T.p, A21, gemm_ws, stream);
        if (!sym_on) copy_blocks<T>(c12, B.off_rs, ws_T.p, A12, gemm_ws, stream);
    }

    template <class T>
    void pack_w(BoxBatch& B, const FaceDev& fd) {
        // Each face's W (k x q) from the n*n-strided ID workspace into the
        // packed r*s layout, converted to the factor precision.
        pack_boxes<T>(B.nbox, B.d_off_rs.p, fd.offCd.p, ws_W.p, reinterpret_cast<T*>(B.W.p),
                      B.tot_rs, stream);
    }

    // Mixed precision, by default: with an ID tolerance of at least 1e-5
    // every level is factored in FP32 (the level matrix, the ID and the PCG
    // stay FP64).  Measured at n=128, lognormal coefficients, eps 1e-4: 14
    // PCG iterations against 11 for the FP64 factor, and the fastest
    // factor + solve of the splits tried (FP64 on the top one or two levels:
    // 11 iterations, 7% slower in all).  Below that tolerance FP32 rounding
    // of the two top levels, the largest and worst-conditioned blocks, would
    // dominate the factor's error (at n=64, eps 1e-6: e_s 2e-4 against
    // 5e-6 in FP64), so they stay FP64.
    int fp32_levels_eff() const {
        if (precision != HIFDE_MIXED) return 0;
        if (fp32_levels >= 0) return std::min(fp32_levels, L);
        return eps >= 1e-5 ? L : std::max(L - 2, 0);
    }

    void run_factor() {
        if (!have_matrix) throw Error(HIFDE_ERR_STATE, "factor: no matrix set");
        // host planning threads of this call (several factor objects driven
        // from several host threads share the cores); the caller's setting
        // is restored on return
        struct OmpThreads {
            int prev;
            explicit OmpThreads(int n) : prev(omp_get_max_threads()) {
                if (n > 0) omp_set_num_threads(n);
            }
            ~OmpThreads() { omp_set_num_threads(prev); }
        } omp_threads(host_threads);
        clear_graphs();
        for (double& t : stage_t) t = 0.0;
        stage_lt.assign((size_t)L * ST_COUNT, 0.0);
        lu_info = 0;
        stager.stream = stream;
        struct StagerGuard {
            Stager* prev;
            explicit StagerGuard(Stager* s) : prev(active_stager()) { set_active_stager(s); }
            ~StagerGuard() { set_active_stager(prev); }
        } guard(&stager);
        TimerGuard tguard(this);
        // GEMM routing log of this factorization (the tuner's input)
        struct RecGuard {
            GemmRecorder* prev;
            RecGuard(GemmRecorder* r) : prev(active_gemm_recorder()) {
                if (r) set_active_gemm_recorder(r);
            }
            ~RecGuard() { set_active_gemm_recorder(prev); }
        } rguard(gemm_record ? &gemm_log : nullptr);
        gemm_log.log.clear();
        gemm_log.calls = 0;
        d_status.reserve(1);
        fill_int(d_status.p, 1, 0, stream);
        // Initialize (src/Factor.cpp:22-34)
        remaining.assign((size_t)NC * NC * NC, 0);
#pragma omp parallel for schedule(static)
        for (int k = 1; k < NC; ++k)
            for (int j = 1; j < NC; ++j)
                std::fill_n(remaining.begin() + lin(1, j, k), NC - 1, (unsigned char)1);
        cells.resize(L);
        faces.resize(L);
        for (std::vector<BoxBatch>* v : {&cells, &faces})
            for (BoxBatch& B : *v)
                for (int& x : B.rplan_R) x = 0;
        factored = false;
        planned_ahead = -1;
        deferred_faces = -1;
        sym_on = op_symmetric && sym_option != 0;
        elem_on = local_ok && elem_option != 0;
        elem32 = elem_on && elem_f32_option != 0 && fp32_levels_eff() >= L;
        if (!elem_on) plan0_valid = false;
        elem_cur = 0;
        if (!elem_on) {
            // Initialize(): the level matrix starts as the assembled operator.
            auto t0 = clk::now();
            long long nnz = nnz_op;
            cur = 0;
            tabs[0].sym = tabs[1].sym = sym_on;
            table_init(tabs[cur], (size_t)(sym_on ? (nnz + NC * NC * NC) / 2 : nnz), stream);
            Prof pr(this, "table_insert", 0.0, (double)nnz * (4 + 4 + 8 + 16));
            table_insert_coo(tabs[cur], d_rr.p, d_ci.p, d_v.p, nnz, stream);
            tab_inserted = (size_t)nnz;
            timed(ST_UPLOAD, t0);
        }
        for (int level = 0; level < L; ++level) {
            cur_level = level;
            if (level < fp32_levels_eff()) level_cells<float>(level);
            else level_cells<double>(level);
            finish_deferred_faces();
            if (level < L - 1) {
                if (level < fp32_levels_eff()) level_faces<float>(level);
                else level_faces<double>(level);
            }
        }
        finish_deferred_faces();
        // Every DOF has been eliminated: the level matrix is consumed.  Its
        // buffers and the workspaces are kept for the next factorization
        // unless keep_workspaces is off.
        HIFDE_CHECK(hipMemcpyAsync(&lu_info, d_status.p, sizeof(int), hipMemcpyDeviceToHost,
                                   stream));
        HIFDE_CHECK(hipStreamSynchronize(stream));
        stager.synced();
        resolve_profile();
        factored = true;
        a22_kept = keep_a22;
        if (!keep_a22)
            for (std::vector<BoxBatch>* v : {&cells, &faces})
                for (BoxBatch& B : *v) B.A22.release();
        if (!keep_workspaces) release_workspaces();
    }

    // The factorization's large workspaces (level-matrix elements, face and
    // Schur blocks, block-inverse and Gram panels); every user reserves them
    // again before use, so the next factorization reallocates them.
    void release_workspaces() {
        HIFDE_CHECK(hipStreamSynchronize(stream));
        for (DevBuf<double>& e : elem) e.release();
        for (DevBuf<unsigned char>* b : {&ws_a, &ws_b, &ws_c, &gj_ws.B, &gj_ws.D, &gj_ws.RO,
                                         &gj_ws.CW, &gj_ws.SW})
            b->release();
        for (DevBuf<double>* b : {&ws_M, &ws_C, &ws_W, &ws_T, &ws_G, &gram_ws.P, &gram_ws.RK,
                                  &gram_ws.dg})
            b->release();
    }

    // ------------------------------------------------------------------
    void upd(BoxBatch& B, const DevBuf<unsigned char>& Mbuf, int op, int pass, double alpha,
             double* u, double* side = nullptr) {
        if (B.esize == sizeof(float)) upd_t<float>(B, Mbuf, op, pass, alpha, u, side);
        else upd_t<double>(B, Mbuf, op, pass, alpha, u, side);
    }
    // op 0 of a cell batch, whose skeleton DOFs on shared faces get two
    // contributions: every output is formed once into the side buffers, then
    // the contributions are added in pass order (max faces, then min faces).
    // side: two NC^3 side buffers (one per pass flag)
    void upd_sides(BoxBatch& B, const DevBuf<unsigned char>& Mbuf, double alpha, double* u,
                   double* side) {
        if (B.nbox == 0 || B.tot_rs == 0) return;
        const long long n3 = (long long)NC * NC * NC;
        upd(B, Mbuf, 0, -1, alpha, u, side);
        side_add(B.skel_off.back(), B.d_skel_gid.p, B.d_skel_pass.p, side, n3, u, stream,
                 B.sides_paired);
    }
    void diag(BoxBatch& B, const DevBuf<unsigned char>& Mbuf, const double* u, double* out) {
        if (B.esize == sizeof(float)) diag_t<float>(B, Mbuf, u, out);
        else diag_t<double>(B, Mbuf, u, out);
    }
    template <class T>
    void upd_t(BoxBatch& B, const DevBuf<unsigned char>& Mbuf, int op, int pass, double alpha,
               double* u, double* side) {
        // with side, u_in is u and the outputs go to the side buffers
        double* uo = side ? side : u;
        const long long side_n = side ? (long long)NC * NC * NC : 0;
        const unsigned char* spp = (pass >= 0 || side) ? B.d_skel_pass.p : nullptr;
        // no red or no skeleton DOFs (faces of full rank): no contribution
        if (B.nbox == 0 || B.tot_rs == 0) return;
        const bool out_red = (op == 1 || op == 2);
        const bool in_red = (op == 0 || op == 3);
        const std::vector<int>& tl = out_red ? B.tiles_r : B.tiles_s;
        double by = (double)B.tot_rs * sizeof(T) +
                    16.0 * ((B.red_off.empty() ? 0 : B.red_off.back()) +
                            (B.skel_off.empty() ? 0 : B.skel_off.back()));
        Prof pr(this, fam<T>("apply_update"), 2.0 * B.tot_rs, by);
        if (B.max_r <= 64 && B.max_s <= 64) {
            const int nout = out_red ? B.red_off.back() : B.skel_off.back();
            apply_small<T>(B.nbox, nout, B.d_r.p, B.d_s.p, B.d_red_off.p, B.d_skel_off.p,
                           B.d_red_gid.p, B.d_skel_gid.p, spp, B.d_off_rs.p,
                           reinterpret_cast<const T*>(Mbuf.p), op, pass, alpha, u, uo,
                           2.0 * B.tot_rs, by, stream, side_n,
                           (op == 0 || op == 2) ? B.max_rs : 0);
            return;
        }
        if (wide_tiles(op, out_red ? B.max_r : B.max_s, out_red ? B.tiles_r2 : B.tiles_s2)) {
            const std::vector<int>& t2 = out_red ? B.tiles_r2 : B.tiles_s2;
            apply_tiles<T>(B.nbox, t2.back(), out_red ? B.d_tiles_r2.p : B.d_tiles_s2.p, B.d_r.p,
                           B.d_s.p, B.d_red_off.p, B.d_skel_off.p, B.d_red_gid.p, B.d_skel_gid.p,
                           spp, B.d_off_rs.p, reinterpret_cast<const T*>(Mbuf.p), op, pass,
                           alpha, u, uo, in_red ? B.max_r : B.max_s, 2.0 * B.tot_rs, by, stream,
                           side_n, 2);
            return;
        }
        apply_tiles<T>(B.nbox, tl.back(), out_red ? B.d_tiles_r.p : B.d_tiles_s.p, B.d_r.p,
                       B.d_s.p, B.d_red_off.p, B.d_skel_off.p, B.d_red_gid.p, B.d_skel_gid.p, spp,
                       B.d_off_rs.p, reinterpret_cast<const T*>(Mbuf.p), op, pass, alpha, u, uo,
                       in_red ? B.max_r : B.max_s, 2.0 * B.tot_rs, by, stream, side_n);
    }
    // Tiles of two rows per lane (128 outputs) for the non-transposed ops
    // (1, 3, 4): a wave's loads of one column then cover 512 contiguous
    // bytes instead of 256, which streams at 5.7-6.3 TB/s against 4.3 on
    // large column-major FP32 blocks (about 5.0 against 4.7 TB/s on the
    // level blocks).  Taken where outputs run past 64 and the launch still
    // has enough workgroups to fill the chip.
    static bool wide_tiles(int op, int max_out, const std::vector<int>& tiles2) {
        return (op == 1 || op == 3 || op == 4) && max_out > 64 && !tiles2.empty() &&
               tiles2.back() >= kApplyWideMinTiles;
    }
    template <class T>
    void diag_t(BoxBatch& B, const DevBuf<unsigned char>& Mbuf, const double* u, double* out) {
        if (B.nbox == 0 || B.tot_rr == 0) return;
        const double by =
            (double)B.tot_rr * sizeof(T) + 16.0 * (B.red_off.empty() ? 0 : B.red_off.back());
        Prof pr(this, fam<T>("apply_diag"), 2.0 * B.tot_rr, by);
        if (B.max_r <= 64) {
            apply_small<T>(B.nbox, B.red_off.back(), B.d_r.p, B.d_s.p, B.d_red_off.p,
                           B.d_skel_off.p, B.d_red_gid.p, B.d_skel_gid.p, nullptr, B.d_off_rr.p,
                           reinterpret_cast<const T*>(Mbuf.p), 4, -1, 1.0, u, out, 2.0 * B.tot_rr,
                           by, stream, 0, B.max_r * B.max_r);
            return;
        }
        const bool wide = wide_tiles(4, B.max_r, B.tiles_r2);
        apply_tiles<T>(B.nbox, wide ? B.tiles_r2.back() : B.tiles_r.back(),
                       wide ? B.d_tiles_r2.p : B.d_tiles_r.p, B.d_r.p, B.d_s.p, B.d_red_off.p,
                       B.d_skel_off.p, B.d_red_gid.p, B.d_skel_gid.p, nullptr, B.d_off_rr.p,
                       reinterpret_cast<const T*>(Mbuf.p), 4, -1, 1.0, u, out, B.max_r,
                       2.0 * B.tot_rr, by, stream, 0, wide ? 2 : 1);
    }
    // Apply tiling of a planned batch.
    void finalize_tiles(BoxBatch& B) {
        B.tiles_r.assign(B.nbox + 1, 0);
        B.tiles_s.assign(B.nbox + 1, 0);
        B.tiles_r2.assign(B.nbox + 1, 0);
        B.tiles_s2.assign(B.nbox + 1, 0);
        B.max_r = B.max_s = B.max_rs = 0;
        for (int b = 0; b < B.nbox; ++b) {
            B.tiles_r[b + 1] = B.tiles_r[b] + (B.r[b] + 63) / 64;
            B.tiles_s[b + 1] = B.tiles_s[b] + (B.s[b] + 63) / 64;
            B.tiles_r2[b + 1] = B.tiles_r2[b] + (B.r[b] + 127) / 128;
            B.tiles_s2[b + 1] = B.tiles_s2[b] + (B.s[b] + 127) / 128;
            B.max_r = std::max(B.max_r, B.r[b]);
            B.max_s = std::max(B.max_s, B.s[b]);
            B.max_rs = std::max(B.max_rs, B.r[b] * B.s[b]);
        }
        B.d_tiles_r.upload(B.tiles_r, stream);
        B.d_tiles_s.upload(B.tiles_s, stream);
        B.d_tiles_r2.upload(B.tiles_r2, stream);
        B.d_tiles_s2.upload(B.tiles_s2, stream);
    }

    // Apply (src/Factor.cpp:992-1081), one vector.  Each batch's blocks are
    // read in the precision they were factored in.  Scratch: tmp holds NC^3
    // doubles (the diagonal stage's output), side 2 NC^3 (the cell sweeps'
    // contributions by pass flag); both are written before they are read, so
    // they need no clearing.
    void apply_one(double* u, bool inv, double* tmp, double* side) {
        zero_boundary(u, NC, stream);
        const double sg = inv ? -1.0 : 1.0;
        for (int lv = 0; lv < L - 1; ++lv) {
            BoxBatch& C = cells[lv];
            if (inv) {
                upd_sides(C, C.X, -1.0, u, side);
            } else {
                upd(C, C.X, 1, -1, 1.0, u);
            }
            BoxBatch& F = faces[lv];
            if (F.nbox == 0) continue;
            if (inv) {
                upd(F, F.W, 2, -1, -1.0, u);
                upd(F, F.X, 0, -1, -1.0, u);
            } else {
                upd(F, F.W, 3, -1, 1.0, u);
                upd(F, F.X, 1, -1, 1.0, u);
            }
        }
        // Every interior DOF is redundant in exactly one box, so the diagonal
        // stage maps u to a fresh vector box by box and the result is copied back.
        vcopy(u, tmp, (long long)NC * NC * NC, stream);
        for (int lv = 0; lv < L; ++lv) {
            diag(cells[lv], inv ? cells[lv].A22inv : cells[lv].A22, u, tmp);
            diag(faces[lv], inv ? faces[lv].A22inv : faces[lv].A22, u, tmp);
        }
        vcopy(tmp, u, (long long)NC * NC * NC, stream);
        for (int lv = L - 2; lv >= 0; --lv) {
            BoxBatch& F = faces[lv];
            if (F.nbox) {
                if (inv) {
                    upd(F, F.X, 1, -1, -1.0, u);
                    upd(F, F.W, 3, -1, -1.0, u);
                } else {
                    upd(F, F.X, 0, -1, 1.0, u);
                    upd(F, F.W, 2, -1, 1.0, u);
                }
            }
            BoxBatch& C = cells[lv];
            if (inv) {
                upd(C, C.X, 1, -1, -1.0, u);
            } else {
                upd_sides(C, C.X, 1.0, u, side);
            }
        }
        (void)sg;
    }

    void check_apply(bool inv) const {
        if (!factored) throw Error(HIFDE_ERR_STATE, "apply: factor first");
        if (!inv && !a22_kept)
            throw Error(HIFDE_ERR_STATE,
                        "apply: the forward Apply F needs the A22 blocks (option keep_a22 was 0)");
    }
    void apply_dev(double* d_u, int nrhs, long long ldu, bool inv) {
        TimerGuard tguard(this);
        check_apply(inv);
        if (nrhs >= 1) apply_sweeps(d_u, nrhs, ldu, inv);
        resolve_profile();
    }

    // ------------------------------------------------------------------
    // The apply sweeps are a fixed sequence of some 40-100 launches for a
    // given factorization, vector and right-hand-side count.  The first call
    // with a given (vector, count, leading dimension, direction) runs
    // eagerly; it also sizes every workspace and builds every descriptor
    // table.  The second captures the sequence into a HIP graph, and later
    // calls replay it in one launch.  Graphs are dropped when the factor
    // changes.  Under full profiling (per-family events) the sweeps always
    // run eagerly; in single-kernel timing mode launches inside a capture
    // are not timed.
    struct ApplyGraph {
        double* u;
        int R;
        long long ld;
        bool inv;
        unsigned long long gen;  // apply_gen of the eager run that prepared it
        hipGraphExec_t exec;
    };
    std::vector<ApplyGraph> graphs;
    // Bumped whenever something a recorded sweep depends on changes: a
    // workspace moves or a multi-RHS descriptor table is rebuilt.  A graph is
    // captured and replayed only while it matches its eager run's value.
    unsigned long long apply_gen = 0;
    bool graph_option = true;
    static constexpr size_t kMaxGraphs = 8;
    void clear_graphs() {
        for (ApplyGraph& g : graphs)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
        graphs.clear();
    }
    void apply_eager(double* d_u, int R, long long ld, bool inv) {
        if (R > 1) {
            apply_multi(d_u, R, ld, inv);
        } else {
            const size_t n3 = (size_t)NC * NC * NC;
            apply_tmp.reserve(n3);
            apply_side.reserve(2 * n3);
            apply_one(d_u, inv, apply_tmp.p, apply_side.p);
        }
    }
    void apply_sweeps(double* d_u, int R, long long ld, bool inv) {
        const bool can_graph = graph_option && (!profiling || !profile_only.empty());
        if (!can_graph) {
            apply_eager(d_u, R, ld, inv);
            return;
        }
        ApplyGraph* g = nullptr;
        for (ApplyGraph& e : graphs)
            if (e.u == d_u && e.R == R && e.ld == ld && e.inv == inv) g = &e;
        if (!g || g->gen != apply_gen) {
            // first use, or the state changed since: eager, sizes the
            // workspaces and builds the descriptor tables
            const void* before[6] = {apply_tmp.p, apply_side.p, rhs_in.p,
                                     rhs_out.p,   rhs_side.p,   rhs_il.p};
            apply_eager(d_u, R, ld, inv);
            const void* after[6] = {apply_tmp.p, apply_side.p, rhs_in.p,
                                    rhs_out.p,   rhs_side.p,   rhs_il.p};
            for (int i = 0; i < 6; ++i)
                if (before[i] != after[i]) ++apply_gen;
            if (!g) {
                if (graphs.size() == kMaxGraphs) {
                    if (graphs.front().exec) (void)hipGraphExecDestroy(graphs.front().exec);
                    graphs.erase(graphs.begin());
                }
                graphs.push_back({d_u, R, ld, inv, apply_gen, nullptr});
            } else {
                if (g->exec) (void)hipGraphExecDestroy(g->exec);
                g->exec = nullptr;
                g->gen = apply_gen;
            }
            return;
        }
        if (!g->exec && !capture_graph(g, d_u, R, ld, inv)) {
            // the capture was refused or invalidated (for instance by a
            // device-wide synchronize on another host thread): none of the
            // captured launches ran, so the sweeps run eagerly instead, with
            // the same launches and the same bits
            apply_eager(d_u, R, ld, inv);
            return;
        }
        HIFDE_CHECK(hipGraphLaunch(g->exec, stream));
    }
    // Captures the sweeps into g->exec.  Returns false, with the stream out of
    // capture mode, the partial graph destroyed and the thread's HIP error
    // state cleared, when the capture fails for any reason (HIP refuses or
    // invalidates it); the object then stops capturing (graph_failures
    // counts these).  A failure that is not about the capture recurs in the
    // eager run that follows and is reported from there.
    int graph_failures = 0;
    // Packed batch (hifde_create_batch): nprob > 0 independent problems of
    // grid NC_sub packed as the octants of this object's grid; seg maps
    // every packed entry to its problem (255: none).
    int nprob = 0, NC_sub = 0;
    DevBuf<unsigned char> d_seg;
    bool gemm_record = false;  // log every GEMM descriptor of a factorization
    GemmRecorder gemm_log;
    bool gemm_table_on = true;  // route GEMMs by the tuning table where it has an entry
    bool capture_graph(ApplyGraph* g, double* d_u, int R, long long ld, bool inv) {
        if (graph_failures > 0) return false;
        hipGraph_t graph = nullptr;
        if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            (void)hipGetLastError();
            ++graph_failures;
            return false;
        }
        bool launched = true;
        try {
            apply_eager(d_u, R, ld, inv);
        } catch (const Error&) {
            launched = false;
        }
        const hipError_t ee = hipStreamEndCapture(stream, &graph);
        if (!launched || ee != hipSuccess || !graph) {
            if (graph) (void)hipGraphDestroy(graph);
            (void)hipGetLastError();  // clears the thread's error state
            ++graph_failures;
            return false;
        }
        const hipError_t e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            g->exec = nullptr;
            ++graph_failures;
            return false;
        }
        return true;
    }

    // ------------------------------------------------------------------
    // Several right-hand sides (columns of U, leading dimension ldu).  Every
    // box step of apply_one becomes gather -> batched GEMM -> scatter over
    // packed per-box blocks (multi_rhs.hip), so each factor block is read
    // once for all R vectors.  The GEMM descriptors of a batch are built once
    // per factorization and R.
    DevBuf<unsigned char> rhs_in, rhs_out;
    DevBuf<double> rhs_side, rhs_il;
    // batches whose boxes have at most this many inputs and outputs take the
    // fused kernel; larger ones the packed GEMM path
    int multi_fused_max = 512;

    static const int* ogid_of(const BoxBatch& B, bool red) {
        return red ? B.d_red_gid.p : B.d_skel_gid.p;
    }
    void upd_multi(BoxBatch& B, const DevBuf<unsigned char>& Mbuf, int op, double alpha,
                   double* U, RhsLayout ly, int R, bool side = false) {
        if (B.esize == sizeof(float)) upd_multi_t<float>(B, Mbuf, op, alpha, U, ly, R, side);
        else upd_multi_t<double>(B, Mbuf, op, alpha, U, ly, R, side);
    }
    // op as in apply_one: 0 X^T u_red -> skel, 1 X u_skel -> red, 2 W^T u_skel
    // -> red, 3 W u_red -> skel, 4 M u_red -> red (written, not added).  With
    // side, the outputs (op 0 of a cell batch) are parked by pass flag and
    // added in pass order.
    template <class T>
    void upd_multi_t(BoxBatch& B, const DevBuf<unsigned char>& Mbuf, int op, double alpha,
                     double* U, RhsLayout ly, int R, bool side) {
        if (B.nbox == 0) return;
        const bool in_red = (op == 0 || op == 3 || op == 4);
        const bool out_red = (op == 1 || op == 2 || op == 4);
        const std::vector<int>& ioff = in_red ? B.red_off : B.skel_off;
        const std::vector<int>& ooff = out_red ? B.red_off : B.skel_off;
        const long long nin = ioff.back(), nout = ooff.back();
        if (nout == 0) return;
        const long long n3 = (long long)NC * NC * NC;
        const int max_in = in_red ? B.max_r : B.max_s, max_out = out_red ? B.max_r : B.max_s;
        // few columns: the fused kernel for every size (the packed GEMMs'
        // tiles would be mostly idle); many: packed GEMMs beyond
        // multi_fused_max (measured at n=128: R=4 10.3 vs 12.5 ms per apply,
        // R=64 33.7 vs 42.5 ms with the threshold)
        const int fmax = R <= 16 && multi_fused_max > 0 ? (1 << 30) : multi_fused_max;
        const bool fused = op == 4 ? B.max_r <= 64 : (max_in <= fmax && max_out <= fmax);
        if (fused) {
            const double fl = 2.0 * (op == 4 ? B.tot_rr : B.tot_rs) * R;
            const double by = (double)(op == 4 ? B.tot_rr : B.tot_rs) * sizeof(T) +
                              (double)(nin + 2 * nout) * R * 8.0;
            Prof pr(this, fam<T>("apply_multi"), fl, by);
            if (side) rhs_side.reserve((size_t)2 * R * n3);
            const std::vector<int>& tl = out_red ? B.tiles_r : B.tiles_s;
            apply_multi_fused<T>(B.nbox, tl.back(), out_red ? B.d_tiles_r.p : B.d_tiles_s.p,
                                 B.d_r.p, B.d_s.p, B.d_red_off.p, B.d_skel_off.p, B.d_red_gid.p,
                                 B.d_skel_gid.p, B.d_skel_pass.p,
                                 op == 4 ? B.d_off_rr.p : B.d_off_rs.p,
                                 reinterpret_cast<const T*>(Mbuf.p), op, alpha, U, ly, R,
                                 side ? rhs_side.p : nullptr, side ? n3 * R : 0, fl, by, stream);
            if (side)
                rhs_side_add(nout, ogid_of(B, out_red), B.d_skel_pass.p, rhs_side.p, n3 * R, R, U,
                             ly, stream);
            return;
        }
        GemmPlan& pl = B.rplan[op];
        if (B.rplan_R[op] != R) {
            ++apply_gen;  // recorded sweeps read this table
            pl.clear();
            std::vector<GemmDesc> d;
            d.reserve(B.nbox);
            B.rplan_zero[op] = 0;
            for (int b = 0; b < B.nbox; ++b) {
                const int r = B.r[b], s = B.s[b];
                const int m = out_red ? r : s, k = in_red ? r : s;
                if (m == 0) continue;
                if (k == 0) {  // no inputs: the box adds nothing
                    B.rplan_zero[op] = 1;
                    continue;
                }
                GemmDesc g;
                g.offA = op == 4 ? B.off_rr[b] : B.off_rs[b];
                g.offB = (long long)ioff[b] * R;
                g.offC = (long long)ooff[b] * R;
                g.m = m;
                g.n = R;
                g.k = k;
                g.lda = (op == 0 || op == 1 || op == 4) ? r : s;
                g.ldb = k;
                g.ldc = m;
                d.push_back(g);
            }
            pl.esize = (int)sizeof(T);
            pl.use_table = gemm_table_on;
            pl.add(d, op == 0 || op == 2);
            pl.upload(stream);
            B.rplan_R[op] = R;
        }
        const double fl = 2.0 * (op == 4 ? B.tot_rr : B.tot_rs) * R;
        const double by = (double)(op == 4 ? B.tot_rr : B.tot_rs) * sizeof(T) +
                          (double)(nin + nout) * R * (sizeof(T) + 8.0) * 2;
        Prof pr(this, fam<T>("apply_multi"), fl, by);
        rhs_in.reserve((size_t)std::max(nin, 1LL) * R * sizeof(T));
        rhs_out.reserve((size_t)nout * R * sizeof(T));
        T* pin = reinterpret_cast<T*>(rhs_in.p);
        T* pout = reinterpret_cast<T*>(rhs_out.p);
        const int* d_ioff = in_red ? B.d_red_off.p : B.d_skel_off.p;
        const int* d_ooff = out_red ? B.d_red_off.p : B.d_skel_off.p;
        const int* igid = in_red ? B.d_red_gid.p : B.d_skel_gid.p;
        const int* ogid = out_red ? B.d_red_gid.p : B.d_skel_gid.p;
        rhs_gather<T>(B.nbox, nin, d_ioff, igid, U, ly, R, pin, stream);
        if (B.rplan_zero[op]) fill_zero<T>(pout, nout * R, stream);
        if (!pl.launches.empty())
            gemm_run<T, T>(pl, {0, (int)pl.launches.size()}, reinterpret_cast<const T*>(Mbuf.p),
                           pin, pout, (T)1, (T)0, stream);
        if (side) {
            rhs_side.reserve((size_t)2 * R * n3);
            rhs_scatter<T>(B.nbox, nout, d_ooff, ogid, pout, R, rhs_side.p, ly, alpha, 2,
                           B.d_skel_pass.p, n3 * R, stream);
            rhs_side_add(nout, ogid, B.d_skel_pass.p, rhs_side.p, n3 * R, R, U, ly, stream);
        } else {
            rhs_scatter<T>(B.nbox, nout, d_ooff, ogid, pout, R, U, ly, alpha, op == 4 ? 1 : 0,
                           nullptr, 0, stream);
        }
    }

    // apply_one for R columns.  The diagonal stage works in place: each box
    // reads and writes only its own redundant DOFs, every interior DOF is
    // redundant in exactly one box, and a box's inputs are gathered before
    // its outputs are written.
    void apply_multi(double* Ucm, int R, long long ldu, bool inv) {
        // the sweeps run on an interleaved copy (a DOF's R values contiguous)
        const long long n3 = (long long)NC * NC * NC;
        rhs_il.reserve((size_t)n3 * R);
        double* U = rhs_il.p;
        const RhsLayout ly{R, 1};
        to_interleaved(n3, R, Ucm, ldu, U, stream);
        zero_boundary_il(NC, R, U, stream);
        for (int lv = 0; lv < L - 1; ++lv) {
            BoxBatch& C = cells[lv];
            if (inv) upd_multi(C, C.X, 0, -1.0, U, ly, R, true);
            else upd_multi(C, C.X, 1, 1.0, U, ly, R);
            BoxBatch& F = faces[lv];
            if (F.nbox == 0) continue;
            if (inv) {
                upd_multi(F, F.W, 2, -1.0, U, ly, R);
                upd_multi(F, F.X, 0, -1.0, U, ly, R);
            } else {
                upd_multi(F, F.W, 3, 1.0, U, ly, R);
                upd_multi(F, F.X, 1, 1.0, U, ly, R);
            }
        }
        for (int lv = 0; lv < L; ++lv) {
            upd_multi(cells[lv], inv ? cells[lv].A22inv : cells[lv].A22, 4, 1.0, U, ly, R);
            upd_multi(faces[lv], inv ? faces[lv].A22inv : faces[lv].A22, 4, 1.0, U, ly, R);
        }
        for (int lv = L - 2; lv >= 0; --lv) {
            BoxBatch& F = faces[lv];
            if (F.nbox) {
                if (inv) {
                    upd_multi(F, F.X, 1, -1.0, U, ly, R);
                    upd_multi(F, F.W, 3, -1.0, U, ly, R);
                } else {
                    upd_multi(F, F.X, 0, 1.0, U, ly, R);
                    upd_multi(F, F.W, 2, 1.0, U, ly, R);
                }
            }
            BoxBatch& C = cells[lv];
            if (inv) upd_multi(C, C.X, 1, -1.0, U, ly, R);
            else upd_multi(C, C.X, 0, 1.0, U, ly, R, true);
        }
        from_interleaved(n3, R, U, Ucm, ldu, stream);
    }

    // ------------------------------------------------------------------
    double dot(const double* x, const double* y, long long n) {
        int nblk = dot_blocks(n);
        pcg_part.reserve(nblk);
        pcg_scalar.reserve(1);
        dot_partial(x, y, n, pcg_part.p, nblk, stream);
        reduce_partials(pcg_part.p, nblk, pcg_scalar.p, stream);
        double h = 0;
        HIFDE_CHECK(hipMemcpyAsync(&h, pcg_scalar.p, sizeof(double), hipMemcpyDeviceToHost,
                                   stream));
        HIFDE_CHECK(hipStreamSynchronize(stream));
        return h;
    }

    void spmv_dev(const double* x, double* y) {
        const long long n = (long long)NC * NC * NC;
        if (ell_width > 0) {
            Prof pr(this, "spmv", 2.0 * nnz_op, 12.0 * ell_width * n + 20.0 * n);
            ell_spmv((int)n, ell_width, d_rp.p, d_ell_v.p, d_ell_c.p, x, y, stream);
            return;
        }
        Prof pr(this, "spmv", 2.0 * nnz_op, 12.0 * nnz_op + 20.0 * n);
        csr_spmv((int)n, d_rp.p, d_ci.p, d_v.p, x, y, stream);
    }
    // The slot-major copy of the operator for the PCG's products, when its
    // rows hold at most kEllMaxWidth entries (every stencil operator: 7).
    void build_ell() {
        const long long n = (long long)NC * NC * NC;
        ell_width = 0;
        const int w = csr_max_row((int)n, d_rp.p, stream);
        if (w < 1 || w > kEllMaxWidth) return;
        d_ell_v.reserve((long long)w * n);
        d_ell_c.reserve((long long)w * n);
        csr_to_ell((int)n, w, d_rp.p, d_ci.p, d_v.p, d_ell_v.p, d_ell_c.p, stream);
        ell_width = w;
    }

    void pcg_dev(const double* b, double* x, double tol, int maxit, int* iters, double* relres) {
        TimerGuard tguard(this);
        if (!factored) throw Error(HIFDE_ERR_STATE, "pcg: factor first");
        const long long n = (long long)NC * NC * NC;
        pcg_buf.reserve(4 * n);
        double* r = pcg_buf.p;
        double* z = r + n;
        double* p = z + n;
        double* q = p + n;
        spmv_dev(x, q);
        vcopy(b, r, n, stream);
        axpy(-1.0, q, r, n, stream);
        zero_boundary(r, NC, stream);
        double bnorm = std::sqrt(dot(b, b, n));
        if (bnorm == 0.0) bnorm = 1.0;
        double rnorm = std::sqrt(dot(r, r, n));
        int it = 0;
        if (rnorm / bnorm > tol) {
            vcopy(r, z, n, stream);
            apply_dev(z, 1, n, true);
            vcopy(z, p, n, stream);
            double rz = dot(r, z, n);
            for (it = 1; it <= maxit; ++it) {
                spmv_dev(p, q);
                double pq = dot(p, q, n);
                double alpha = rz / pq;
                axpy(alpha, p, x, n, stream);
                axpy(-alpha, q, r, n, stream);
                rnorm = std::sqrt(dot(r, r, n));
                if (rnorm / bnorm <= tol) break;
                vcopy(r, z, n, stream);
                apply_dev(z, 1, n, true);
                double rz_new = dot(r, z, n);
                double beta = rz_new / rz;
                rz = rz_new;
                xpby(z, beta, p, n, stream);
            }
            if (it > maxit) it = maxit;
        }
        HIFDE_CHECK(hipStreamSynchronize(stream));
        resolve_profile();
        if (iters) *iters = it;
        if (relres) *relres = rnorm / bnorm;
    }

    // R independent PCG solves advanced together (one SpMV over R columns,
    // one preconditioner apply over R columns, R dots per reduction).  Each
    // column follows pcg_dev's recurrence; a converged column is frozen (its
    // step lengths are zero) while the others continue.
    DevBuf<double> mpcg_buf, mpcg_part, mpcg_scal;
    // F^{-1} on R columns (the single-vector sweeps for one column)
    void apply_cols(double* Z, int R, long long ld) { apply_sweeps(Z, R, ld, true); }
    // R independent PCG solves advanced together.  On a packed batch every
    // (column, problem) pair is its own solve: the operator and the factor
    // are block diagonal over the problems, so the recurrences separate once
    // the dots and step lengths are taken per problem (segment).  Results are
    // indexed [problem * R + column] there.
    void pcg_multi_dev(const double* B, long long ldb, double* X, long long ldx, int R,
                       double tol, int maxit, int* iters, double* relres) {
        TimerGuard tguard(this);
        if (!factored) throw Error(HIFDE_ERR_STATE, "pcg: factor first");
        const long long n = (long long)NC * NC * NC;
        const int S = nprob > 0 ? nprob : 1;
        const int RS = R * S;  // solves, [column * S + segment] internally
        mpcg_buf.reserve((size_t)4 * n * R);
        double* Rm = mpcg_buf.p;
        double* Z = Rm + n * R;
        double* Pm = Z + n * R;
        double* Q = Pm + n * R;
        const int nblk = std::max(1, std::min(256, 4096 / R));
        mpcg_part.reserve((size_t)nblk * RS);
        mpcg_scal.reserve((size_t)4 * RS);
        double* d_dot = mpcg_scal.p;       // RS dots
        double* d_a = d_dot + RS;          // RS step lengths
        double* d_ma = d_a + RS;           // their negatives
        double* d_b = d_ma + RS;           // RS betas
        std::vector<double> h(4 * RS);
        auto dots = [&](const double* x, const double* y) {
            if (nprob > 0)
                dot_cols_seg(n, x, n, y, n, R, d_seg.p, S, mpcg_part.p, nblk, d_dot, stream);
            else
                dot_cols(n, x, n, y, n, R, mpcg_part.p, nblk, d_dot, stream);
            HIFDE_CHECK(hipMemcpyAsync(h.data(), d_dot, RS * sizeof(double),
                                       hipMemcpyDeviceToHost, stream));
            HIFDE_CHECK(hipStreamSynchronize(stream));
            return std::vector<double>(h.begin(), h.begin() + RS);
        };
        auto upload_scalars = [&](const std::vector<double>& a, const std::vector<double>& b) {
            for (int c = 0; c < RS; ++c) {
                h[RS + c] = a[c];
                h[2 * RS + c] = -a[c];
                h[3 * RS + c] = b[c];
            }
            HIFDE_CHECK(hipMemcpyAsync(d_a, h.data() + RS, 3 * RS * sizeof(double),
                                       hipMemcpyHostToDevice, stream));
            HIFDE_CHECK(hipStreamSynchronize(stream));
        };
        auto axpy = [&](const double* a, const double* Xs, long long lx, double* Ys, long long ly,
                        bool xpby) {
            if (nprob > 0) axpy_cols_seg(n, a, Xs, lx, Ys, ly, R, xpby, d_seg.p, S, stream);
            else axpy_cols(n, a, Xs, lx, Ys, ly, R, xpby, stream);
        };
        // r = b - A x, boundary rows cleared
        {
            Prof pr(this, "spmv", 2.0 * nnz_op * R, (12.0 * nnz_op + 16.0 * n) * R);
            csr_spmm((int)n, d_rp.p, d_ci.p, d_v.p, X, ldx, Q, n, R, stream);
        }
        for (int c = 0; c < R; ++c) vcopy(B + c * ldb, Rm + c * n, n, stream);
        std::vector<double> ones(RS, -1.0), zeros(RS, 0.0);
        upload_scalars(ones, zeros);
        axpy(d_a, Q, n, Rm, n, false);
        for (int c = 0; c < R; ++c) zero_boundary(Rm + c * n, NC, stream);
        std::vector<double> bnorm(RS), rnorm(RS);
        {
            // ||b|| per solve (b has its own leading dimension)
            for (int c = 0; c < R; ++c) vcopy(B + c * ldb, Z + c * n, n, stream);
            std::vector<double> bb = dots(Z, Z);
            for (int c = 0; c < RS; ++c) {
                bnorm[c] = std::sqrt(bb[c]);
                if (bnorm[c] == 0.0) bnorm[c] = 1.0;
            }
        }
        std::vector<double> rr = dots(Rm, Rm);
        std::vector<int> it(RS, 0);
        std::vector<char> done(RS, 0);
        int active = 0;
        for (int c = 0; c < RS; ++c) {
            rnorm[c] = std::sqrt(rr[c]);
            done[c] = rnorm[c] / bnorm[c] <= tol;
            active += !done[c];
        }
        if (active > 0) {
            vcopy(Rm, Z, n * R, stream);
            apply_cols(Z, R, n);
            vcopy(Z, Pm, n * R, stream);
            std::vector<double> rz = dots(Rm, Z);
            std::vector<double> alpha(RS), beta(RS);
            for (int k = 1; k <= maxit && active > 0; ++k) {
                {
                    Prof pr(this, "spmv", 2.0 * nnz_op * R, (12.0 * nnz_op + 16.0 * n) * R);
                    csr_spmm((int)n, d_rp.p, d_ci.p, d_v.p, Pm, n, Q, n, R, stream);
                }
                std::vector<double> pq = dots(Pm, Q);
                for (int c = 0; c < RS; ++c) alpha[c] = done[c] ? 0.0 : rz[c] / pq[c];
                upload_scalars(alpha, zeros);
                axpy(d_a, Pm, n, X, ldx, false);
                axpy(d_ma, Q, n, Rm, n, false);
                rr = dots(Rm, Rm);
                for (int c = 0; c < RS; ++c) {
                    if (done[c]) continue;
                    it[c] = k;
                    rnorm[c] = std::sqrt(rr[c]);
                    if (rnorm[c] / bnorm[c] <= tol) {
                        done[c] = 1;
                        --active;
                    }
                }
                if (active == 0) break;
                vcopy(Rm, Z, n * R, stream);
                apply_cols(Z, R, n);
                std::vector<double> rz_new = dots(Rm, Z);
                for (int c = 0; c < RS; ++c) {
                    beta[c] = done[c] ? 0.0 : rz_new[c] / rz[c];
                    if (!done[c]) rz[c] = rz_new[c];
                }
                upload_scalars(zeros, beta);
                axpy(d_b, Z, n, Pm, n, true);
            }
        }
        HIFDE_CHECK(hipStreamSynchronize(stream));
        resolve_profile();
        for (int c = 0; c < R; ++c)
            for (int sg = 0; sg < S; ++sg) {
                const int q = c * S + sg, o = sg * R + c;
                if (iters) iters[o] = it[q];
                if (relres) relres[o] = rnorm[q] / bnorm[q];
            }
    }

    // ------------------------------------------------------------------
    // Factor persistence: the operator, every box's index and its factor
    // blocks, so that a later process can apply and solve without factoring.
    // Binary, native byte order; the header names the grid, tolerance and
    // precision, and a load checks them against the object's own.
    static constexpr char kFileMagic[8] = {'H', 'I', 'F', 'D', 'E', 'F', '0', '2'};
    static constexpr char kFileEnd[8] = {'H', 'I', 'F', 'D', 'E', 'E', 'N', 'D'};
    struct File {
        FILE* fp = nullptr;
        bool writing = false;
        File(const char* path, bool w) : writing(w) {
            fp = std::fopen(path, w ? "wb" : "rb");
            if (!fp) throw Error(HIFDE_ERR_ARG, std::string("cannot open ") + path);
        }
        ~File() {
            if (fp) std::fclose(fp);
        }
        void put(const void* p, size_t n) {
            if (n && std::fwrite(p, 1, n, fp) != n) throw Error(HIFDE_ERR_STATE, "write failed");
        }
        void get(void* p, size_t n) {
            if (n && std::fread(p, 1, n, fp) != n)
                throw Error(HIFDE_ERR_STATE, "factor file truncated or unreadable");
        }
        template <class V>
        void put_vec(const std::vector<V>& v) {
            const long long n = (long long)v.size();
            put(&n, sizeof n);
            put(v.data(), v.size() * sizeof(V));
        }
        template <class V>
        void get_vec(std::vector<V>& v, long long max_n) {
            long long n = 0;
            get(&n, sizeof n);
            if (n < 0 || n > max_n) throw Error(HIFDE_ERR_STATE, "factor file corrupt");
            v.resize((size_t)n);
            get(v.data(), (size_t)n * sizeof(V));
        }
    };
    // Device bytes to the file through a bounded host buffer.
    void put_dev(File& fl, const unsigned char* d, long long bytes) {
        std::vector<unsigned char> h((size_t)std::min<long long>(bytes, 64LL << 20));
        for (long long o = 0; o < bytes; o += (long long)h.size()) {
            const size_t n = (size_t)std::min<long long>((long long)h.size(), bytes - o);
            HIFDE_CHECK(hipMemcpyAsync(h.data(), d + o, n, hipMemcpyDeviceToHost, stream));
            HIFDE_CHECK(hipStreamSynchronize(stream));
            fl.put(h.data(), n);
        }
    }
    void get_dev(File& fl, DevBuf<unsigned char>& d, long long bytes) {
        d.reserve((size_t)std::max<long long>(bytes, 1));
        std::vector<unsigned char> h((size_t)std::min<long long>(std::max<long long>(bytes, 1), 64LL << 20));
        for (long long o = 0; o < bytes; o += (long long)h.size()) {
            const size_t n = (size_t)std::min<long long>((long long)h.size(), bytes - o);
            fl.get(h.data(), n);
            HIFDE_CHECK(hipMemcpyAsync(d.p + o, h.data(), n, hipMemcpyHostToDevice, stream));
            HIFDE_CHECK(hipStreamSynchronize(stream));
        }
    }

    void save(const char* path) {
        if (!factored) throw Error(HIFDE_ERR_STATE, "save: factor first");
        HIFDE_CHECK(hipStreamSynchronize(stream));
        File fl(path, true);
        fl.put(kFileMagic, 8);
        const int hdr[7] = {N, P, precision, L, quirk ? 1 : 0, lu_info, a22_kept ? 1 : 0};
        fl.put(hdr, sizeof hdr);
        fl.put(&eps, sizeof eps);
        // operator, as COO in CSR order
        ensure_host_csr();
        std::vector<int> rows(h_ci.size());
        for (size_t i = 0; i + 1 < h_rp.size(); ++i)
            for (int q = h_rp[i]; q < h_rp[i + 1]; ++q) rows[q] = (int)i;
        fl.put_vec(rows);
        fl.put_vec(h_ci);
        fl.put_vec(h_v);
        for (int lv = 0; lv < L; ++lv)
            for (int kind = 0; kind < 2; ++kind) {
                const BoxBatch& B = kind == 0 ? cells[lv] : faces[lv];
                const int bh[2] = {B.nbox, B.esize};
                fl.put(bh, sizeof bh);
                fl.put_vec(B.r);
                fl.put_vec(B.s);
                fl.put_vec(B.red_gid);
                fl.put_vec(B.skel_gid);
                fl.put_vec(B.skel_pass);
                if (B.nbox == 0) continue;
                if (a22_kept) put_dev(fl, B.A22.p, B.tot_rr * B.esize);
                put_dev(fl, B.A22inv.p, B.tot_rr * B.esize);
                put_dev(fl, B.X.p, B.tot_rs * B.esize);
                if (kind == 1) put_dev(fl, B.W.p, B.tot_rs * B.esize);
            }
        fl.put(kFileEnd, 8);
    }

    void load(const char* path) {
        File fl(path, false);
        char magic[8];
        fl.get(magic, 8);
        if (std::memcmp(magic, kFileMagic, 8) != 0) throw Error(HIFDE_ERR_ARG, "not a factor file");
        int hdr[7];
        double feps = 0;
        fl.get(hdr, sizeof hdr);
        fl.get(&feps, sizeof feps);
        if (hdr[0] != N || hdr[1] != P || hdr[2] != precision || hdr[3] != L || feps != eps ||
            hdr[4] != (quirk ? 1 : 0))
            throw Error(HIFDE_ERR_ARG, "load: the file holds another grid, tolerance or precision");
        if (hdr[6] != 0 && hdr[6] != 1) throw Error(HIFDE_ERR_STATE, "factor file corrupt");
        const bool file_a22 = hdr[6] == 1;
        const long long nd = (long long)NC * NC * NC;
        std::vector<int> rows, cols;
        std::vector<double> vals;
        fl.get_vec(rows, 64 * nd);
        fl.get_vec(cols, 64 * nd);
        fl.get_vec(vals, 64 * nd);
        if (rows.size() != cols.size() || rows.size() != vals.size())
            throw Error(HIFDE_ERR_STATE, "factor file corrupt");
        set_coo((long long)rows.size(), rows.data(), cols.data(), vals.data());
        factored = false;
        clear_graphs();
        cells.resize(L);
        faces.resize(L);
        for (int lv = 0; lv < L; ++lv)
            for (int kind = 0; kind < 2; ++kind) {
                BoxBatch& B = kind == 0 ? cells[lv] : faces[lv];
                B.reset_host();
                int bh[2];
                fl.get(bh, sizeof bh);
                if (bh[0] < 0 || bh[0] > nd || (bh[1] != 4 && bh[1] != 8))
                    throw Error(HIFDE_ERR_STATE, "factor file corrupt");
                B.nbox = bh[0];
                B.esize = bh[1];
                fl.get_vec(B.r, nd);
                fl.get_vec(B.s, nd);
                fl.get_vec(B.red_gid, nd);
                fl.get_vec(B.skel_gid, nd);
                fl.get_vec(B.skel_pass, nd);
                B.sides_paired =
                    !B.skel_pass.empty() &&
                    2 * (long long)std::count(B.skel_pass.begin(), B.skel_pass.end(), 1) ==
                        (long long)B.skel_pass.size();
                if ((int)B.r.size() != B.nbox || (int)B.s.size() != B.nbox)
                    throw Error(HIFDE_ERR_STATE, "factor file corrupt");
                // every size within the grid, every prefix sum within int
                // range, before any offset is formed
                long long sr = 0, ss = 0;
                for (int b = 0; b < B.nbox; ++b) {
                    if (B.r[b] < 0 || B.r[b] > nd || B.s[b] < 0 || B.s[b] > nd)
                        throw Error(HIFDE_ERR_STATE, "factor file corrupt");
                    sr += B.r[b];
                    ss += B.s[b];
                }
                if (sr > nd || ss > nd) throw Error(HIFDE_ERR_STATE, "factor file corrupt");
                B.red_off = prefix_i(B.r);
                B.skel_off = prefix_i(B.s);
                if ((long long)B.red_gid.size() != B.red_off.back() ||
                    (long long)B.skel_gid.size() != B.skel_off.back() ||
                    (!B.skel_pass.empty() && B.skel_pass.size() != B.skel_gid.size()))
                    throw Error(HIFDE_ERR_STATE, "factor file corrupt");
                for (int g : B.red_gid)
                    if (g < 0 || g >= nd) throw Error(HIFDE_ERR_STATE, "factor file corrupt");
                for (int g : B.skel_gid)
                    if (g < 0 || g >= nd) throw Error(HIFDE_ERR_STATE, "factor file corrupt");
                B.off_rr = prefix_ll(B.r, B.r);
                B.off_rs = prefix_ll(B.r, B.s);
                B.off_ss = prefix_ll(B.s, B.s);
                B.tot_rr = B.off_rr.back();
                B.tot_rs = B.off_rs.back();
                B.tot_ss = B.off_ss.back();
                if (B.nbox == 0) continue;
                upload_batch_index(B);
                if (file_a22) get_dev(fl, B.A22, B.tot_rr * B.esize);
                else B.A22.release();
                get_dev(fl, B.A22inv, B.tot_rr * B.esize);
                get_dev(fl, B.X, B.tot_rs * B.esize);
                if (kind == 1) get_dev(fl, B.W, B.tot_rs * B.esize);
            }
        fl.get(magic, 8);
        if (std::memcmp(magic, kFileEnd, 8) != 0) throw Error(HIFDE_ERR_STATE, "factor file corrupt");
        HIFDE_CHECK(hipStreamSynchronize(stream));
        lu_info = hdr[5];
        a22_kept = file_a22;
        plan0_valid = false;  // the level-0 index is rebuilt by the next factorization
        factored = true;
    }

    // ------------------------------------------------------------------
    // Operator assembly: rows/cols/vals with Sparse::Add semantics.
    void set_coo(long long nnz, const int* rows, const int* cols, const double* vals) {
        const long long nd = (long long)NC * NC * NC;
        for (long long t = 0; t < nnz; ++t)
            if (rows[t] < 0 || rows[t] >= nd || cols[t] < 0 || cols[t] >= nd)
                throw Error(HIFDE_ERR_ARG, "set_matrix: index out of range");
        // CSR with entries in the order given, duplicates merged in order.
        std::vector<int> cnt(nd + 1, 0);
        for (long long t = 0; t < nnz; ++t) cnt[rows[t] + 1]++;
        std::vector<int> rp(nd + 1, 0);
        for (long long i = 0; i < nd; ++i) rp[i + 1] = rp[i] + cnt[i + 1];
        std::vector<int> ci(nnz);
        std::vector<double> v(nnz);
        std::vector<int> fill(rp.begin(), rp.end() - 1);
        for (long long t = 0; t < nnz; ++t) {
            int p = fill[rows[t]]++;
            ci[p] = cols[t];
            v[p] = vals[t];
        }
        h_rp.assign(nd + 1, 0);
        h_ci.clear();
        h_v.clear();
        h_ci.reserve(nnz);
        h_v.reserve(nnz);
        for (long long i = 0; i < nd; ++i) {
            size_t start = h_ci.size();
            for (int p = rp[i]; p < rp[i + 1]; ++p) {
                size_t q = start;
                while (q < h_ci.size() && h_ci[q] != ci[p]) ++q;
                if (q < h_ci.size()) h_v[q] += v[p];
                else { h_ci.push_back(ci[p]); h_v.push_back(v[p]); }
            }
            h_rp[i + 1] = (int)h_ci.size();
        }
        // bitwise symmetry of the assembled operator
        op_symmetric = true;
        for (long long i = 0; i < nd && op_symmetric; ++i)
            for (int p = h_rp[i]; p < h_rp[i + 1] && op_symmetric; ++p) {
                const int j = h_ci[p];
                bool found = false;
                for (int q = h_rp[j]; q < h_rp[j + 1]; ++q)
                    if (h_ci[q] == (int)i) {
                        found = h_v[q] == h_v[p];
                        break;
                    }
                op_symmetric = found;
            }
        nnz_op = (long long)h_ci.size();
        host_csr_valid = true;
        d_rp.upload(h_rp, stream);
        d_ci.upload(h_ci, stream);
        d_v.upload(h_v, stream);
        std::vector<int> rr(h_ci.size());
        for (long long i = 0; i < nd; ++i)
            for (int p = h_rp[i]; p < h_rp[i + 1]; ++p) rr[p] = (int)i;
        d_rr.upload(rr, stream);
        build_ell();
        // Element storage needs every coupling between two unknowns to lie
        // inside one level-0 cell closure: per axis the coordinates must
        // share a cell (a coordinate on a cell plane belongs to both sides).
        {
            const int w = P, c = NC / P;
            auto cells_of = [&](int x, int& lo, int& hi) {
                if (x % w == 0) {
                    lo = std::max(x / w - 1, 0);
                    hi = std::min(x / w, c - 1);
                } else {
                    lo = hi = x / w;
                }
            };
            auto interior = [&](int g) {
                const int x = g % NC, y = (g / NC) % NC, z = g / (NC * NC);
                return x >= 1 && x <= N && y >= 1 && y <= N && z >= 1 && z <= N;
            };
            int ok = 1;
#pragma omp parallel for schedule(static) reduction(min : ok)
            for (long long i = 0; i < nd; ++i) {
                if (!interior((int)i)) continue;
                const int ic[3] = {(int)(i % NC), (int)((i / NC) % NC), (int)(i / ((long long)NC * NC))};
                for (int p = h_rp[i]; p < h_rp[i + 1]; ++p) {
                    const int j = h_ci[p];
                    if (!interior(j)) continue;
                    const int jc[3] = {j % NC, (j / NC) % NC, j / (NC * NC)};
                    for (int a = 0; a < 3; ++a) {
                        int l1, h1, l2, h2;
                        cells_of(ic[a], l1, h1);
                        cells_of(jc[a], l2, h2);
                        if (std::max(l1, l2) > std::min(h1, h2)) ok = 0;
                    }
                }
            }
            local_ok = ok != 0;
        }
        HIFDE_CHECK(hipStreamSynchronize(stream));
        have_matrix = true;
        factored = false;
    }

    void set_stencil(const double* a, const double* V, int symmetric) {
        // SetupStencil (include/setup_stencil.hpp:18-74)
        const int m = NC + 1;
        auto T = [&](const double* t, int i, int j, int k) {
            return t[i + (size_t)m * (j + (size_t)m * k)];
        };
        const double h = 1.0 / NC, hh = h * h;
        const long long cap = 7LL * N * N * N;
        std::vector<int> rows(cap), cols(cap);
        std::vector<double> vals(cap);
        // 7 slots per interior point, in the reference's push order
        // (diagonal, i-1, i+1, j-1, j+1, k-1, k+1); absent neighbours are
        // compacted out afterwards.
#pragma omp parallel for schedule(static)
        for (int i = 1; i <= N; ++i)
            for (int j = 1; j <= N; ++j)
                for (int k = 1; k <= N; ++k) {
                    long long base = 7LL * (((long long)(i - 1) * N + (j - 1)) * N + (k - 1));
                    const int nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k},
                                          {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
                    double ap = T(a, i, j, k);
                    double d = 0.0;
                    if (!symmetric) {
                        d += (T(a, i - 1, j, k) + T(a, i + 1, j, k)) / hh;
                        d += (T(a, i, j - 1, k) + T(a, i, j + 1, k)) / hh;
                        d += (T(a, i, j, k - 1) + T(a, i, j, k + 1)) / hh;
                    } else {
                        for (int t = 0; t < 6; ++t)
                            d += 0.5 * (ap + T(a, nb[t][0], nb[t][1], nb[t][2])) / hh;
                    }
                    d += T(V, i, j, k);
                    int p = lin(i, j, k);
                    rows[base] = p;
                    cols[base] = p;
                    vals[base] = d;
                    for (int t = 0; t < 6; ++t) {
                        int x = nb[t][0], y = nb[t][1], z = nb[t][2];
                        bool in = x >= 1 && x <= N && y >= 1 && y <= N && z >= 1 && z <= N;
                        rows[base + 1 + t] = in ? p : -1;
                        cols[base + 1 + t] = in ? lin(x, y, z) : -1;
                        double aq = T(a, x, y, z);
                        vals[base + 1 + t] = symmetric ? -0.5 * (ap + aq) / hh : -aq / hh;
                    }
                }
        long long w = 0;
        for (long long t = 0; t < cap; ++t)
            if (rows[t] >= 0) {
                rows[w] = rows[t];
                cols[w] = cols[t];
                vals[w] = vals[t];
                ++w;
            }
        set_coo(w, rows.data(), cols.data(), vals.data());
    }

    // SetupStencil from device nodal tensors (include/setup_stencil.hpp:6-76):
    // the CSR is built in HBM, with the same rows, entry order and values as
    // set_stencil.  Nearest-neighbour couplings always share a level-0 cell
    // closure, so the element storage applies.
    void set_stencil_device(const double* d_a, const double* d_V, int symmetric) {
        const long long nd = (long long)NC * NC * NC;
        nnz_op = stencil_csr(NC, d_a, d_V, symmetric, d_rp, d_ci, d_v, d_rr, stream);
        host_csr_valid = false;
        op_symmetric = csr_symmetric(nd, d_rp.p, d_ci.p, d_v.p, stream);
        build_ell();
        local_ok = true;
        HIFDE_CHECK(hipStreamSynchronize(stream));
        have_matrix = true;
        factored = false;
    }

    // The nprob problems' stencils assembled into the packed grid: each
    // problem's rows are its own operator's (as set_stencil_device would
    // build them) relabelled into its octant; the separating planes and any
    // unused octant get identity rows.
    void set_stencil_batch_device(const double* const* a, const double* const* V, int symmetric) {
        const long long nd = (long long)NC * NC * NC;
        DevBuf<const double*> ptrs;
        std::vector<const double*> h(2 * nprob);
        for (int o = 0; o < nprob; ++o) {
            h[o] = a[o];
            h[nprob + o] = V[o];
        }
        ptrs.upload(h, stream);
        nnz_op = pack_stencil_csr(NC_sub, nprob, ptrs.p, ptrs.p + nprob, symmetric, d_rp, d_ci,
                                  d_v, d_rr, d_seg, stream);
        host_csr_valid = false;
        op_symmetric = csr_symmetric(nd, d_rp.p, d_ci.p, d_v.p, stream);
        build_ell();
        local_ok = true;
        HIFDE_CHECK(hipStreamSynchronize(stream));
        have_matrix = true;
        factored = false;
    }
};

// ===========================================================================
// C ABI
// ===========================================================================
namespace {

template <class F>
int guarded(F&& fn) {
    try {
        return fn();
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return HIFDE_ERR_HIP;
    }
}

// Householder column-pivoted QR of a batch of host matrices on the GPU (the
// kernel of the factorization's ID).  M_packed is overwritten with the QR
// factors.  With k_out/perm_out/W_out the ID (rank rule with eps, skeleton
// then redundant columns, W) is returned; with jpvt_out/tau_out the pivots
// and reflector scalars.
void qr_batch(int nbatch, const int* m, const int* n, double* M_packed, double eps, int* k_out,
              int* perm_out, double* W_out, int* jpvt_out, double* tau_out) {
    hipStream_t s = aux_stream();
    OwnerGuard og(s);
    std::vector<int> hm(m, m + nbatch), hn(n, n + nbatch);
    std::vector<long long> offM = prefix_ll(hm, hn), offW = prefix_ll(hn, hn);
    std::vector<int> poff = prefix_i(hn);
    int max_n = 0, max_m = 0;
    for (int b = 0; b < nbatch; ++b) {
        if (hm[b] < 0 || hn[b] < 0) throw Error(HIFDE_ERR_ARG, "negative size");
        max_n = std::max(max_n, hn[b]);
        max_m = std::max(max_m, hm[b]);
    }
    DevBuf<int> dm, dn, dperm, dpoff, dk, djp;
    DevBuf<long long> doffM, doffW;
    DevBuf<double> dM, dW, dtau;
    dm.upload(hm, s);
    dn.upload(hn, s);
    doffM.upload(offM, s);
    doffW.upload(offW, s);
    dpoff.upload(poff, s);
    dM.upload(M_packed, offM.back(), s);
    dW.reserve(offW.back());
    dperm.reserve(poff.back());
    dk.reserve(nbatch);
    if (jpvt_out) djp.reserve(poff.back());
    if (tau_out) dtau.reserve(poff.back());
    batched_interp_decomp(nbatch, dm.p, dn.p, doffM.p, dM.p, eps, dperm.p, dpoff.p, dk.p, dW.p,
                          doffW.p, max_n, max_m, nullptr, s, jpvt_out ? djp.p : nullptr,
                          tau_out ? dtau.p : nullptr);
    auto down = [&](void* dst, const void* src, size_t bytes) {
        if (bytes) HIFDE_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    };
    down(M_packed, dM.p, offM.back() * sizeof(double));
    if (k_out) down(k_out, dk.p, nbatch * sizeof(int));
    if (perm_out) down(perm_out, dperm.p, poff.back() * sizeof(int));
    if (W_out) down(W_out, dW.p, offW.back() * sizeof(double));
    if (jpvt_out) down(jpvt_out, djp.p, poff.back() * sizeof(int));
    std::vector<double> tau_all(tau_out ? poff.back() : 0);
    if (tau_out) down(tau_all.data(), dtau.p, poff.back() * sizeof(double));
    HIFDE_CHECK(hipStreamSynchronize(s));
    if (tau_out) {
        // min(m, n) reflector scalars per matrix, packed one after another
        long long o = 0;
        for (int b = 0; b < nbatch; ++b) {
            const int mn = std::min(hm[b], hn[b]);
            std::copy(tau_all.begin() + poff[b], tau_all.begin() + poff[b] + mn, tau_out + o);
            o += mn;
        }
    }
}

template <class T>
void read_dev(const DevBuf<unsigned char>& buf, long long off, long long cnt, double* out,
              hipStream_t s) {
    if (!out || cnt <= 0) return;
    std::vector<T> tmp(cnt);
    HIFDE_CHECK(hipMemcpyAsync(tmp.data(), reinterpret_cast<const T*>(buf.p) + off,
                               cnt * sizeof(T), hipMemcpyDeviceToHost, s));
    HIFDE_CHECK(hipStreamSynchronize(s));
    for (long long i = 0; i < cnt; ++i) out[i] = (double)tmp[i];
}

}  // namespace

extern "C" {

const char* hifde_last_error(void) { return g_last_error.c_str(); }

namespace {
// Parses a routing table written by the tuner: one entry per line,
// "esize transA lower kbin mnbin nbbin kind", '#' starts a comment.
int load_tune_table(const char* path) {
    FILE* fp = fopen(path, "r");
    if (!fp) return -1;
    std::vector<signed char> t(kTuneEntries, -1);
    char line[512];
    int n = 0;
    while (fgets(line, sizeof line, fp)) {
        if (char* c = strchr(line, '#')) *c = 0;
        int v[7];
        if (sscanf(line, "%d %d %d %d %d %d %d", v, v + 1, v + 2, v + 3, v + 4, v + 5, v + 6) !=
            7)
            continue;
        if ((v[0] != 4 && v[0] != 8) || v[1] < 0 || v[1] > 1 || v[2] < 0 || v[2] > 1 ||
            v[3] < 0 || v[3] >= kTuneKBins || v[4] < 0 || v[4] >= kTuneMNBins || v[5] < 0 ||
            v[5] >= kTuneNBBins || v[6] < 0 || v[6] > 2) {
            fclose(fp);
            throw Error(HIFDE_ERR_ARG, std::string("gemm table: bad entry in ") + path);
        }
        const int key =
            (((((v[0] == 4 ? 0 : 1) * 2 + v[1]) * 2 + v[2]) * kTuneKBins + v[3]) * kTuneMNBins +
             v[4]) * kTuneNBBins + v[5];
        t[key] = (signed char)v[6];
        ++n;
    }
    fclose(fp);
    gemm_tune_install(t.data());
    return n;
}
// The table shipped next to the library (gemm_tune.txt), loaded once when
// the first factor object is created.
std::once_flag g_table_once;
std::string default_table_path() {
    Dl_info info;
    if (!dladdr(reinterpret_cast<void*>(&load_tune_table), &info) || !info.dli_fname) return "";
    std::string p = info.dli_fname;
    const size_t slash = p.rfind('/');
    return (slash == std::string::npos ? std::string(".") : p.substr(0, slash)) +
           "/gemm_tune.txt";
}
}  // namespace

int hifde_gemm_table_load(const char* path) {
    return guarded([&]() -> int {
        const std::string p = path ? std::string(path) : default_table_path();
        const int n = load_tune_table(p.c_str());
        if (n < 0) throw Error(HIFDE_ERR_ARG, "gemm_table_load: cannot read " + p);
        return n;
    });
}

int hifde_gemm_table_clear(void) {
    std::vector<signed char> none(kTuneEntries, -1);
    gemm_tune_install(none.data());
    return HIFDE_OK;
}

long long hifde_gemm_log(const hifde_factor* f, int* out, long long cap) {
    if (!f) return HIFDE_ERR_ARG;
    const std::vector<int>& v = f->gemm_log.log;
    if (out) std::copy(v.begin(), v.begin() + std::min<long long>(cap, (long long)v.size()), out);
    return (long long)v.size();
}

int hifde_gemm_time(int precision, int transA, int lower, int nbatch, const int* m, const int* n,
                    const int* k, int kind, int reps, double* avg_seconds) {
    return guarded([&]() -> int {
        if (nbatch <= 0 || !m || !n || !k || kind < 0 || kind > 2 || reps < 1 || !avg_seconds)
            throw Error(HIFDE_ERR_ARG, "gemm_time: bad arguments");
        if (precision != HIFDE_FP64 && precision != HIFDE_MIXED)
            throw Error(HIFDE_ERR_ARG, "gemm_time: precision");
        hipStream_t s = aux_stream();
        OwnerGuard og(s);
        std::vector<GemmDesc> d(nbatch);
        long long oa = 0, ob = 0, oc = 0;
        for (int b = 0; b < nbatch; ++b) {
            if (m[b] < 1 || n[b] < 1 || k[b] < 1) throw Error(HIFDE_ERR_ARG, "gemm_time: bad sizes");
            GemmDesc& g = d[b];
            g.offA = oa;
            g.offB = ob;
            g.offC = oc;
            g.m = m[b];
            g.n = n[b];
            g.k = k[b];
            g.lda = transA ? k[b] : m[b];
            g.ldb = k[b];
            g.ldc = m[b];
            g.lower = lower ? 1 : 0;
            oa += (long long)m[b] * k[b];
            ob += (long long)k[b] * n[b];
            oc += (long long)m[b] * n[b];
        }
        const size_t es = precision == HIFDE_MIXED ? 4 : 8;
        DevBuf<unsigned char> A, B, C;
        A.reserve(oa * es);
        B.reserve(ob * es);
        C.reserve(oc * es);
        // operands of values in [-1, 1] (timing does not depend on them)
        if (es == 8) {
            gen_uniform(reinterpret_cast<double*>(A.p), oa, 11, -1.0, 1.0, false, s);
            gen_uniform(reinterpret_cast<double*>(B.p), ob, 12, -1.0, 1.0, false, s);
        } else {
            DevBuf<double> tmp;
            tmp.reserve(std::max(oa, ob));
            gen_uniform(tmp.p, oa, 11, -1.0, 1.0, false, s);
            convert_d2f(tmp.p, reinterpret_cast<float*>(A.p), oa, s);
            gen_uniform(tmp.p, ob, 12, -1.0, 1.0, false, s);
            convert_d2f(tmp.p, reinterpret_cast<float*>(B.p), ob, s);
            HIFDE_CHECK(hipStreamSynchronize(s));
        }
        GemmWork ws;
        ws.plan.force_kind = kind;
        ws.plan.use_table = false;
        hipEvent_t e0, e1;
        HIFDE_CHECK(hipEventCreate(&e0));
        HIFDE_CHECK(hipEventCreate(&e1));
        auto run = [&]() {
            if (es == 4)
                batched_gemm_desc<float>(transA != 0, d, reinterpret_cast<const float*>(A.p),
                                         reinterpret_cast<const float*>(B.p),
                                         reinterpret_cast<float*>(C.p), 1.0f, 0.0f, ws, s);
            else
                batched_gemm_desc<double>(transA != 0, d, reinterpret_cast<const double*>(A.p),
                                          reinterpret_cast<const double*>(B.p),
                                          reinterpret_cast<double*>(C.p), 1.0, 0.0, ws, s);
        };
        float ms = 0.f;
        try {
            run();  // warm-up (and plan upload)
            HIFDE_CHECK(hipEventRecord(e0, s));
            for (int r = 0; r < reps; ++r) run();
            HIFDE_CHECK(hipEventRecord(e1, s));
            HIFDE_CHECK(hipEventSynchronize(e1));
            HIFDE_CHECK(hipEventElapsedTime(&ms, e0, e1));
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *avg_seconds = ms * 1e-3 / reps;
        return HIFDE_OK;
    });
}

int hifde_create(int N, int P, double epsilon, int precision, int flags, hifde_factor** out) {
    return guarded([&]() -> int {
        if (!out || N < 1 || P < 2 || !(epsilon > 0.0))
            throw Error(HIFDE_ERR_ARG, "create: invalid N, P or epsilon");
        if (precision != HIFDE_FP64 && precision != HIFDE_MIXED)
            throw Error(HIFDE_ERR_ARG, "create: unknown precision");
        int NC = N + 1;
        int q = NC / P;
        if (q < 1 || (q & (q - 1)) != 0 || NC % P != 0)
            throw Error(HIFDE_ERR_ARG, "create: (N+1)/P must be a power of two");
        if (NC >= 1024) throw Error(HIFDE_ERR_SIZE, "create: N + 1 must be below 1024");
        std::call_once(g_table_once, [] {
            const std::string p = default_table_path();
            if (!p.empty()) (void)load_tune_table(p.c_str());
        });
        std::unique_ptr<hifde_factor> f(new hifde_factor());
        f->N = N;
        f->P = P;
        f->NC = NC;
        f->L = (int)std::lround(std::log2((double)(NC / P))) + 1;  // src/Factor.cpp:36
        f->eps = epsilon;
        f->precision = precision;
        f->quirk = (flags & HIFDE_FLAG_FRONT_QUIRK) != 0;
        // 64 x 64 matrix-core tiles where the 128 x 128 ones would waste
        // over a quarter of their output (option "gemm_tile64")
        f->gemm_ws.plan.tile64 = f->gj_ws.plan.tile64 = 1;
        HIFDE_CHECK(hipGetDevice(&f->device));
        HIFDE_CHECK(hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking));
        HIFDE_CHECK(hipEventCreateWithFlags(&f->entry_ev, hipEventDisableTiming));
        f->scratch_dev = resolve_scratch(f->device);
        *out = f.release();
        return HIFDE_OK;
    });
}

int hifde_set_stream(hifde_factor* f, void* stream) {
    if (!f) return HIFDE_ERR_ARG;
    f->caller_stream = static_cast<hipStream_t>(stream);
    return HIFDE_OK;
}

int hifde_create_batch(int N, int P, double epsilon, int precision, int flags, int nprob,
                       hifde_factor** out) {
    return guarded([&]() -> int {
        if (!out || nprob < 1 || nprob > 8 || N < 1)
            throw Error(HIFDE_ERR_ARG, "create_batch: 1 to 8 problems");
        hifde_factor* f = nullptr;
        const int rc = hifde_create(2 * (N + 1) - 1, P, epsilon, precision, flags, &f);
        if (rc < 0) throw Error(rc, g_last_error);
        f->nprob = nprob;
        f->NC_sub = N + 1;
        *out = f;
        return HIFDE_OK;
    });
}

int hifde_set_stencil_batch_device(hifde_factor* f, const double* const* d_a,
                                   const double* const* d_v, int symmetric) {
    return guarded([&]() -> int {
        OwnerGuard og(f ? f->stream : nullptr);
        if (!f || !d_a || !d_v) throw Error(HIFDE_ERR_ARG, "set_stencil_batch: null argument");
        if (f->nprob < 1) throw Error(HIFDE_ERR_STATE, "set_stencil_batch: not a packed batch");
        for (int o = 0; o < f->nprob; ++o)
            if (!d_a[o] || !d_v[o]) throw Error(HIFDE_ERR_ARG, "set_stencil_batch: null tensor");
        f->after_caller_stream();
        f->set_stencil_batch_device(d_a, d_v, symmetric);
        return HIFDE_OK;
    });
}

int hifde_batch_pack(hifde_factor* f, double* d_sub, long long ld_sub, double* d_packed,
                     long long ld_packed, int nrhs, int unpack) {
    return guarded([&]() -> int {
        if (!f || !d_sub || !d_packed || nrhs < 0) throw Error(HIFDE_ERR_ARG, "batch_pack: bad args");
        if (f->nprob < 1) throw Error(HIFDE_ERR_STATE, "batch_pack: not a packed batch");
        const long long ns = (long long)f->NC_sub * f->NC_sub * f->NC_sub;
        const long long np = (long long)f->NC * f->NC * f->NC;
        if ((nrhs * f->nprob > 1 && ld_sub < ns) || (nrhs > 1 && ld_packed < np))
            throw Error(HIFDE_ERR_ARG, "batch_pack: leading dimension too small");
        if (nrhs == 0) return HIFDE_OK;
        OwnerGuard og(f->stream);
        f->after_caller_stream();
        pack_vectors(f->NC_sub, f->nprob, nrhs, d_sub, std::max(ld_sub, ns), d_packed,
                     std::max(ld_packed, np), unpack != 0, f->stream);
        HIFDE_CHECK(hipStreamSynchronize(f->stream));
        return HIFDE_OK;
    });
}

void hifde_destroy(hifde_factor* f) {
    if (!f) return;
    hipStream_t s = f->stream;
    delete f;
    if (s) (void)hipStreamDestroy(s);
}

int hifde_set_matrix_coo(hifde_factor* f, long long nnz, const int* rows, const int* cols,
                         const double* vals) {
    return guarded([&]() -> int {
        OwnerGuard og(f ? f->stream : nullptr);
        if (!f || nnz < 0 || (nnz && (!rows || !cols || !vals)))
            throw Error(HIFDE_ERR_ARG, "set_matrix_coo: bad arguments");
        f->set_coo(nnz, rows, cols, vals);
        return HIFDE_OK;
    });
}

int hifde_set_stencil(hifde_factor* f, const double* a, const double* v, int symmetric) {
    return guarded([&]() -> int {
        OwnerGuard og(f ? f->stream : nullptr);
        if (!f || !a || !v) throw Error(HIFDE_ERR_ARG, "set_stencil: null argument");
        f->set_stencil(a, v, symmetric);
        return HIFDE_OK;
    });
}

int hifde_set_stencil_device(hifde_factor* f, const double* d_a, const double* d_v,
                             int symmetric) {
    return guarded([&]() -> int {
        OwnerGuard og(f ? f->stream : nullptr);
        if (!f || !d_a || !d_v) throw Error(HIFDE_ERR_ARG, "set_stencil_device: null argument");
        f->after_caller_stream();
        f->set_stencil_device(d_a, d_v, symmetric);
        return HIFDE_OK;
    });
}

long long hifde_operator_csr(hifde_factor* f, int* rowptr, int* colind, double* vals) {
    long long nnz = 0;
    const int rc = guarded([&]() -> int {
        if (!f) throw Error(HIFDE_ERR_ARG, "operator_csr: null");
        if (!f->have_matrix) throw Error(HIFDE_ERR_STATE, "operator_csr: no matrix");
        OwnerGuard og(f->stream);
        f->ensure_host_csr();
        if (rowptr) std::copy(f->h_rp.begin(), f->h_rp.end(), rowptr);
        if (colind) std::copy(f->h_ci.begin(), f->h_ci.end(), colind);
        if (vals) std::copy(f->h_v.begin(), f->h_v.end(), vals);
        nnz = f->nnz_op;
        return HIFDE_OK;
    });
    return rc < 0 ? (long long)rc : nnz;
}

int hifde_factor_run(hifde_factor* f) {
    return guarded([&]() -> int {
        if (!f) throw Error(HIFDE_ERR_ARG, "factor: null");
        f->run_factor();
        HIFDE_CHECK(hipStreamSynchronize(f->stream));
        return f->lu_info;
    });
}

int hifde_apply_device(hifde_factor* f, double* d_u, int nrhs, long long ldu, int inverse) {
    return guarded([&]() -> int {
        if (!f || !d_u || nrhs < 0) throw Error(HIFDE_ERR_ARG, "apply: bad arguments");
        long long n = (long long)f->NC * f->NC * f->NC;
        if (nrhs > 1 && ldu < n) throw Error(HIFDE_ERR_ARG, "apply: ldu too small");
        f->after_caller_stream();
        f->apply_dev(d_u, nrhs, ldu, inverse != 0);
        HIFDE_CHECK(hipStreamSynchronize(f->stream));
        return HIFDE_OK;
    });
}

int hifde_apply_device_noalloc(hifde_factor* f, double* d_u, int inverse) {
    return guarded([&]() -> int {
        if (!f || !d_u) throw Error(HIFDE_ERR_ARG, "apply_noalloc: bad arguments");
        const long long n = (long long)f->NC * f->NC * f->NC;
        if (n > kNoAllocMaxCells)
            throw Error(HIFDE_ERR_SIZE, "apply_noalloc: (N+1)^3 = " + std::to_string(n) +
                                            " exceeds the scratch limit 128^3");
        if (!f->factored) throw Error(HIFDE_ERR_STATE, "apply_noalloc: factor first");
        f->check_apply(inverse != 0);
        if (!f->scratch_dev) throw Error(HIFDE_ERR_STATE, "apply_noalloc: no scratch");
        OwnerGuard og(f->stream);
        struct NoAlloc {
            hifde_factor* f;
            explicit NoAlloc(hifde_factor* x) : f(x) { f->no_alloc_call = true; }
            ~NoAlloc() { f->no_alloc_call = false; }
        } na(f);
        KernelTimer* prev = active_timer();
        set_active_timer(nullptr);
        try {
            f->after_caller_stream();
            f->apply_one(d_u, inverse != 0, f->scratch_dev, f->scratch_dev + n);
            HIFDE_CHECK(hipStreamSynchronize(f->stream));
        } catch (...) {
            set_active_timer(prev);
            throw;
        }
        set_active_timer(prev);
        return HIFDE_OK;
    });
}

int hifde_apply(hifde_factor* f, double* u, int inverse) {
    return guarded([&]() -> int {
        OwnerGuard og(f ? f->stream : nullptr);
        if (!f || !u) throw Error(HIFDE_ERR_ARG, "apply: bad arguments");
        long long n = (long long)f->NC * f->NC * f->NC;
        DevBuf<double> d;
        d.upload(u, n, f->stream);
        f->apply_dev(d.p, 1, n, inverse != 0);
        HIFDE_CHECK(hipMemcpyAsync(u, d.p, n * sizeof(double), hipMemcpyDeviceToHost, f->stream));
        HIFDE_CHECK(hipStreamSynchronize(f->stream));
        return HIFDE_OK;
    });
}

int hifde_spmv(hifde_factor* f, const double* x, double* y, int transpose) {
    return guarded([&]() -> int {
        OwnerGuard og(f ? f->stream : nullptr);
        if (!f || !x || !y || !f->have_matrix) throw Error(HIFDE_ERR_ARG, "spmv: bad arguments");
        long long n = (long long)f->NC * f->NC * f->NC;
        if (transpose) {
            // y = A^T x (the reference's column iteration, src/Factor.cpp:1099-1113)
            f->ensure_host_csr();
            std::fill(y, y + n, 0.0);
            for (long long i = 0; i < n; ++i)
                for (int p = f->h_rp[i]; p < f->h_rp[i + 1]; ++p) y[f->h_ci[p]] += f->h_v[p] * x[i];
        } else {
            DevBuf<double> dx, dy;
            dx.upload(x, n, f->stream);
            dy.reserve(n);
            f->spmv_dev(dx.p, dy.p);
            HIFDE_CHECK(hipMemcpyAsync(y, dy.p, n * sizeof(double), hipMemcpyDeviceToHost,
                                       f->stream));
            HIFDE_CHECK(hipStreamSynchronize(f->stream));
        }
        return HIFDE_OK;
    });
}

int hifde_pcg_device(hifde_factor* f, const double* d_b, double* d_x, double tol, int maxit,
                     int* iters, double* relres) {
    return guarded([&]() -> int {
        if (!f || !d_b || !d_x || maxit < 0) throw Error(HIFDE_ERR_ARG, "pcg: bad arguments");
        if (f->nprob > 0)
            throw Error(HIFDE_ERR_STATE, "pcg: a packed batch solves with hifde_pcg_multi_device");
        f->after_caller_stream();
        f->pcg_dev(d_b, d_x, tol, maxit, iters, relres);
        return HIFDE_OK;
    });
}

int hifde_pcg_multi_device(hifde_factor* f, const double* d_b, long long ldb, double* d_x,
                           long long ldx, int nrhs, double tol, int maxit, int* iters,
                           double* relres) {
    return guarded([&]() -> int {
        if (!f || !d_b || !d_x || maxit < 0 || nrhs < 0)
            throw Error(HIFDE_ERR_ARG, "pcg_multi: bad arguments");
        const long long n = (long long)f->NC * f->NC * f->NC;
        if (nrhs > 1 && (ldb < n || ldx < n))
            throw Error(HIFDE_ERR_ARG, "pcg_multi: leading dimension too small");
        if (nrhs == 0) return HIFDE_OK;
        f->after_caller_stream();
        f->pcg_multi_dev(d_b, ldb, d_x, ldx, nrhs, tol, maxit, iters, relres);
        return HIFDE_OK;
    });
}

int hifde_pcg(hifde_factor* f, const double* b, double* x, double tol, int maxit, int* iters,
              double* relres) {
    return guarded([&]() -> int {
        OwnerGuard og(f ? f->stream : nullptr);
        if (!f || !b || !x) throw Error(HIFDE_ERR_ARG, "pcg: bad arguments");
        if (f->nprob > 0)
            throw Error(HIFDE_ERR_STATE, "pcg: a packed batch solves with hifde_pcg_multi_device");
        long long n = (long long)f->NC * f->NC * f->NC;
        DevBuf<double> db, dx;
        db.upload(b, n, f->stream);
        dx.upload(x, n, f->stream);
        f->pcg_dev(db.p, dx.p, tol, maxit, iters, relres);
        HIFDE_CHECK(hipMemcpyAsync(x, dx.p, n * sizeof(double), hipMemcpyDeviceToHost,
                                   f->stream));
        HIFDE_CHECK(hipStreamSynchronize(f->stream));
        return HIFDE_OK;
    });
}

int hifde_factor_save(hifde_factor* f, const char* path) {
    return guarded([&]() -> int {
        OwnerGuard og(f ? f->stream : nullptr);
        if (!f || !path) throw Error(HIFDE_ERR_ARG, "save: bad arguments");
        f->save(path);
        return HIFDE_OK;
    });
}

int hifde_factor_load(hifde_factor* f, const char* path) {
    return guarded([&]() -> int {
        OwnerGuard og(f ? f->stream : nullptr);
        if (!f || !path) throw Error(HIFDE_ERR_ARG, "load: bad arguments");
        f->load(path);
        return f->lu_info;
    });
}

int hifde_num_levels(const hifde_factor* f) { return f ? f->L : HIFDE_ERR_ARG; }

int hifde_level_counts(const hifde_factor* f, int level, int* out) {
    if (!f || !out || level < 0 || level >= f->L || !f->factored) return HIFDE_ERR_ARG;
    const BoxBatch& C = f->cells[level];
    const BoxBatch& F = f->faces[level];
    out[0] = C.nbox;
    out[1] = C.red_off.empty() ? 0 : C.red_off.back();
    out[2] = F.nbox;
    out[3] = F.red_off.empty() ? 0 : F.red_off.back();
    out[4] = F.skel_off.empty() ? 0 : F.skel_off.back();
    return HIFDE_OK;
}

int hifde_box_count(const hifde_factor* f, int level, int kind) {
    if (!f || level < 0 || level >= f->L || !f->factored) return HIFDE_ERR_ARG;
    return kind == 0 ? f->cells[level].nbox : f->faces[level].nbox;
}

int hifde_box_sizes(const hifde_factor* f, int level, int kind, int box, int* out) {
    if (!f || !out || level < 0 || level >= f->L || !f->factored) return HIFDE_ERR_ARG;
    const BoxBatch& B = kind == 0 ? f->cells[level] : f->faces[level];
    if (box < 0 || box >= B.nbox) return HIFDE_ERR_ARG;
    out[0] = B.r[box];
    out[1] = B.s[box];
    return HIFDE_OK;
}

int hifde_box_data(const hifde_factor* f, int level, int kind, int box, int* red_gid,
                   int* skel_gid, double* A22, double* A22inv, double* X, double* W) {
    return guarded([&]() -> int {
        if (!f || level < 0 || level >= f->L || !f->factored)
            throw Error(HIFDE_ERR_ARG, "box_data: bad arguments");
        const BoxBatch& B = kind == 0 ? f->cells[level] : f->faces[level];
        if (box < 0 || box >= B.nbox) throw Error(HIFDE_ERR_ARG, "box_data: bad box");
        int r = B.r[box], s = B.s[box];
        if (red_gid) std::copy_n(B.red_gid.data() + B.red_off[box], r, red_gid);
        if (skel_gid) std::copy_n(B.skel_gid.data() + B.skel_off[box], s, skel_gid);
        auto rd = [&](const DevBuf<unsigned char>& buf, long long off, long long cnt, double* o) {
            if (B.esize == sizeof(float)) read_dev<float>(buf, off, cnt, o, f->stream);
            else read_dev<double>(buf, off, cnt, o, f->stream);
        };
        if (A22 && !f->a22_kept)
            throw Error(HIFDE_ERR_STATE, "box_data: A22 was not kept (option keep_a22)");
        rd(B.A22, B.off_rr[box], (long long)r * r, A22);
        rd(B.A22inv, B.off_rr[box], (long long)r * r, A22inv);
        rd(B.X, B.off_rs[box], (long long)r * s, X);
        if (kind == 1) rd(B.W, B.off_rs[box], (long long)r * s, W);
        return HIFDE_OK;
    });
}

int hifde_stage_times(const hifde_factor* f, double* out, int nout) {
    if (!f || !out) return HIFDE_ERR_ARG;
    int n = std::min(nout, (int)ST_COUNT);
    for (int i = 0; i < n; ++i) out[i] = f->stage_t[i];
    return n;
}

int hifde_level_stage_times(const hifde_factor* f, int level, double* out, int nout) {
    if (!f || !out || nout < 0) return HIFDE_ERR_ARG;
    if (level < 0 || level >= f->L) return HIFDE_ERR_ARG;
    const int n = std::min(nout, (int)ST_COUNT);
    for (int i = 0; i < n; ++i)
        out[i] = (size_t)(level * ST_COUNT + i) < f->stage_lt.size()
                     ? f->stage_lt[level * ST_COUNT + i]
                     : 0.0;
    return n;
}

int hifde_set_option(hifde_factor* f, const char* name, double value) {
    return guarded([&]() -> int {
        OwnerGuard og(f ? f->stream : nullptr);
        if (!f || !name) throw Error(HIFDE_ERR_ARG, "set_option: null");
        std::string n(name);
        if (n == "gram_min_n") f->gram_min_n = (int)value;
        else if (n == "gram_min_eps") f->gram_min_eps = value;
        else if (n == "fp32_levels") f->fp32_levels = (int)value;
        else if (n == "symmetric") f->sym_option = (int)value;
        else if (n == "elements") f->elem_option = (int)value;
        else if (n == "elements_fp32") f->elem_f32_option = (int)value;
        else if (n == "stage_sync") f->stage_sync = value != 0;
        else if (n == "gram_panel_reg") f->gram_ws.reg_panel = value != 0;
        else if (n == "symmetric_inverse") f->sym_gj = value != 0;
        else if (n == "bf16x3") f->split_schur = value != 0;
        else if (n == "gemm_tile64") {
            if (value < 0 || value > 2) throw Error(HIFDE_ERR_ARG, "gemm_tile64: 0, 1 or 2");
            f->gemm_ws.plan.tile64 = f->gj_ws.plan.tile64 = (int)value;
        }
        else if (n == "gj_window") {
            const int w = (int)value;
            if (w != 0 && (w < 64 || w > 4096 || w % 64 != 0))
                throw Error(HIFDE_ERR_ARG, "gj_window: 0 or a multiple of 64 in [64, 4096]");
            f->gj_ws.panel = w;
        }
        else if (n == "host_threads") f->host_threads = std::max(0, (int)value);
        else if (n == "gemm_record") f->gemm_record = value != 0;
        else if (n == "gemm_table") {
            f->gemm_table_on = value != 0;
            f->gemm_ws.plan.use_table = f->gj_ws.plan.use_table = f->gram_ws.plan.use_table =
                f->gemm_table_on;
            for (auto* v : {&f->cells, &f->faces})
                for (BoxBatch& B : *v)
                    for (int op = 0; op < 5; ++op) B.rplan_R[op] = 0;  // re-planned
            f->clear_graphs();
        }
        else if (n == "keep_a22") f->keep_a22 = value != 0;
        else if (n == "plan_ahead") f->ahead_option = value != 0;
        else if (n == "gj_fused_steps") f->gj_ws.fused_steps = value != 0;
        else if (n == "keep_workspaces") {
            f->keep_workspaces = value != 0;
            if (!f->keep_workspaces && f->factored) f->release_workspaces();
        }
        else if (n == "multi_fused_max") {
            f->multi_fused_max = (int)value;
            f->clear_graphs();
        } else if (n == "graphs") {
            f->graph_option = value != 0;
            f->graph_failures = 0;
            f->clear_graphs();
        }
        else throw Error(HIFDE_ERR_ARG, "set_option: unknown option " + n);
        return HIFDE_OK;
    });
}

int hifde_set_profiling(hifde_factor* f, int on) {
    if (!f) return HIFDE_ERR_ARG;
    f->profiling = on != 0;
    return HIFDE_OK;
}

int hifde_set_profile_kernel(hifde_factor* f, const char* kernel) {
    if (!f) return HIFDE_ERR_ARG;
    f->profile_only = kernel ? kernel : "";
    return HIFDE_OK;
}

int hifde_reset_kernel_stats(hifde_factor* f) {
    return guarded([&]() -> int {
        if (!f) throw Error(HIFDE_ERR_ARG, "null");
        f->resolve_profile();
        for (auto& k : f->fam_stats) k = hifde_factor::KStat();
        return HIFDE_OK;
    });
}

int hifde_kernel_stats(const hifde_factor* f, int idx, char* name, int name_len, double* out) {
    if (!f) return HIFDE_ERR_ARG;
    int n = (int)f->fam_names.size();
    if (idx < 0 || idx >= n) return n;
    if (name && name_len > 0) {
        std::strncpy(name, f->fam_names[idx].c_str(), name_len - 1);
        name[name_len - 1] = 0;
    }
    if (out) {
        const auto& k = f->fam_stats[idx];
        out[0] = (double)k.launches;
        out[1] = k.sec;
        out[2] = k.flops;
        out[3] = k.bytes;
    }
    return n;
}

const char* hifde_stage_name(int stage) {
    return (stage >= 0 && stage < ST_COUNT) ? kStageNames[stage] : "";
}

long long hifde_factor_bytes(const hifde_factor* f) {
    if (!f) return HIFDE_ERR_ARG;
    long long tot = 0;
    for (int lv = 0; lv < (int)f->cells.size(); ++lv) {
        for (const BoxBatch* B : {&f->cells[lv], &f->faces[lv]}) {
            tot += (long long)((f->a22_kept ? 2 : 1) * B->tot_rr + B->tot_rs) *
                   (long long)B->esize;
            if (B == &f->faces[lv]) tot += (long long)B->tot_rs * (long long)B->esize;
        }
    }
    return tot;
}

int hifde_interp_decomp_batch(int nbatch, const int* m, const int* n, double* M_packed,
                              double eps, int* k_out, int* perm_out, double* W_out) {
    return guarded([&]() -> int {
        if (nbatch < 0 || (nbatch && (!m || !n || !M_packed || !k_out || !perm_out || !W_out)))
            throw Error(HIFDE_ERR_ARG, "interp_decomp_batch: bad arguments");
        if (nbatch == 0) return HIFDE_OK;
        qr_batch(nbatch, m, n, M_packed, eps, k_out, perm_out, W_out, nullptr, nullptr);
        return HIFDE_OK;
    });
}

int hifde_gen_uniform(double* d_out, long long n, unsigned long long seed, double lo, double hi,
                      int pow10, void* stream) {
    return guarded([&]() -> int {
        if (n < 0 || (n && !d_out)) throw Error(HIFDE_ERR_ARG, "gen_uniform: bad arguments");
        gen_uniform(d_out, n, seed, lo, hi, pow10 != 0, static_cast<hipStream_t>(stream));
        return HIFDE_OK;
    });
}

int hifde_gen_normal(double* d_out, long long n, unsigned long long seed, void* stream) {
    return guarded([&]() -> int {
        if (n < 0 || (n && !d_out)) throw Error(HIFDE_ERR_ARG, "gen_normal: bad arguments");
        gen_normal(d_out, n, seed, static_cast<hipStream_t>(stream));
        return HIFDE_OK;
    });
}

int hifde_gen_rhs_normal(double* d_out, int NC, int nrhs, long long ld, unsigned long long seed,
                         void* stream) {
    return guarded([&]() -> int {
        if (NC < 2 || nrhs < 0 || (nrhs && !d_out) ||
            (nrhs > 1 && ld < (long long)NC * NC * NC))
            throw Error(HIFDE_ERR_ARG, "gen_rhs_normal: bad arguments");
        gen_rhs(d_out, NC, nrhs, nrhs > 1 ? ld : (long long)NC * NC * NC, seed,
                static_cast<hipStream_t>(stream));
        return HIFDE_OK;
    });
}

int hifde_gen_gaussian_field(double* d_out, int NC, unsigned long long seed, double corr,
                             int lognormal, void* stream) {
    return guarded([&]() -> int {
        if (NC < 2 || !d_out || !(corr > 0.0))
            throw Error(HIFDE_ERR_ARG, "gen_gaussian_field: bad arguments");
        hipStream_t s = static_cast<hipStream_t>(stream);
        if (!s) {
            // the caller's stream is the legacy default stream: start after
            // its work, then run on this thread's own stream (the
            // temporaries are grown and freed in stream order)
            HIFDE_CHECK(hipStreamSynchronize(nullptr));
            s = aux_stream();
        }
        OwnerGuard og(s);
        gen_gaussian_field(d_out, NC, seed, corr, lognormal != 0, s);
        return HIFDE_OK;
    });
}

int hifde_pivoted_qr_batch(int nbatch, const int* m, const int* n, double* A_packed,
                           int* jpvt_out, double* tau_out) {
    return guarded([&]() -> int {
        if (nbatch < 0 || (nbatch && (!m || !n || !A_packed || !jpvt_out)))
            throw Error(HIFDE_ERR_ARG, "pivoted_qr_batch: bad arguments");
        if (nbatch == 0) return HIFDE_OK;
        qr_batch(nbatch, m, n, A_packed, 0.0, nullptr, nullptr, nullptr, jpvt_out, tau_out);
        return HIFDE_OK;
    });
}

int hifde_triangular_solve_batch(int nbatch, const int* m, const int* n, const double* R_packed,
                                 double* B_packed) {
    return guarded([&]() -> int {
        if (nbatch < 0 || (nbatch && (!m || !n || !R_packed || !B_packed)))
            throw Error(HIFDE_ERR_ARG, "triangular_solve_batch: bad arguments");
        if (nbatch == 0) return HIFDE_OK;
        hipStream_t s = aux_stream();
        OwnerGuard og(s);
        std::vector<int> hm(m, m + nbatch), hn(n, n + nbatch);
        int max_n = 0;
        for (int b = 0; b < nbatch; ++b) {
            if (hm[b] < 0 || hn[b] < 0) throw Error(HIFDE_ERR_ARG, "negative size");
            max_n = std::max(max_n, hn[b]);
        }
        std::vector<long long> offR = prefix_ll(hm, hm), offB = prefix_ll(hm, hn);
        DevBuf<int> dm, dn;
        DevBuf<long long> doffR, doffB;
        DevBuf<double> dR, dB;
        dm.upload(hm, s);
        dn.upload(hn, s);
        doffR.upload(offR, s);
        doffB.upload(offB, s);
        dR.upload(R_packed, offR.back(), s);
        dB.upload(B_packed, offB.back(), s);
        batched_trsm_upper(nbatch, dm.p, dn.p, doffR.p, dR.p, doffB.p, dB.p, max_n, s);
        if (offB.back())
            HIFDE_CHECK(hipMemcpyAsync(B_packed, dB.p, offB.back() * sizeof(double),
                                       hipMemcpyDeviceToHost, s));
        HIFDE_CHECK(hipStreamSynchronize(s));
        return HIFDE_OK;
    });
}

}  // extern "C"

namespace {
template <class T>
int schur_batch_impl(int nb, const int* r, const int* s, const double* A22, const double* A21,
                     const double* A12, double* A22inv, double* X, double* S, int* info,
                     bool sym, bool split) {
    hipStream_t st = aux_stream();
    OwnerGuard og(st);
    std::vector<int> hr(r, r + nb), hs(s, s + nb);
    std::vector<long long> orr = prefix_ll(hr, hr), ors = prefix_ll(hr, hs), oss = prefix_ll(hs, hs);
    int max_r = 0;
    for (int x : hr) max_r = std::max(max_r, x);
    auto up = [&](const double* src, long long cnt, DevBuf<T>& d) {
        std::vector<T> tmp(src, src + cnt);
        d.upload(tmp, st);
    };
    DevBuf<T> dA, dA21, dA12, dX, dS;
    DevBuf<int> dr, ds, dinfo, tiles;
    DevBuf<long long> drr, drs, dss;
    up(A22, orr.back(), dA);
    up(A21, ors.back(), dA21);
    up(A12, ors.back(), dA12);
    dr.upload(hr, st);
    ds.upload(hs, st);
    drr.upload(orr, st);
    drs.upload(ors, st);
    dss.upload(oss, st);
    dX.reserve(ors.back());
    dS.reserve(oss.back());
    dinfo.reserve(nb);
    HIFDE_CHECK(hipMemsetAsync(dinfo.p, 0, nb * sizeof(int), st));
    GJWork gw;
    GemmWork gws;
    batched_gj_inverse<T>(nb, hr.data(), orr.data(), dr.p, drr.p, dA.p, dinfo.p, gw, gws, st,
                          sym);
    gws.plan.split = split;
    batched_gemm<T>(false, nb, hr.data(), hs.data(), hr.data(), orr.data(), ors.data(), ors.data(),
                    dA.p, dA21.p, dX.p, (T)1, gws, st);
    batched_gemm<T>(false, nb, hs.data(), hs.data(), hr.data(), ors.data(), ors.data(), oss.data(),
                    dA12.p, dX.p, dS.p, (T)-1, gws, st);
    auto down = [&](DevBuf<T>& d, long long cnt, double* dst) {
        std::vector<T> tmp(cnt);
        if (cnt)
            HIFDE_CHECK(hipMemcpyAsync(tmp.data(), d.p, cnt * sizeof(T), hipMemcpyDeviceToHost, st));
        HIFDE_CHECK(hipStreamSynchronize(st));
        for (long long i = 0; i < cnt; ++i) dst[i] = (double)tmp[i];
    };
    down(dA, orr.back(), A22inv);
    down(dX, ors.back(), X);
    down(dS, oss.back(), S);
    HIFDE_CHECK(hipMemcpyAsync(info, dinfo.p, nb * sizeof(int), hipMemcpyDeviceToHost, st));
    HIFDE_CHECK(hipStreamSynchronize(st));
    return HIFDE_OK;
}
}  // namespace

namespace {
template <class T>
int gemm_batch_impl(int transA, int nb, const int* m, const int* n, const int* k, const double* A,
                    const double* B, double* C, bool split, int tile64, int force,
                    bool table, bool lower) {
    hipStream_t st = aux_stream();
    OwnerGuard og(st);
    std::vector<long long> oA(nb + 1, 0), oB(nb + 1, 0), oC(nb + 1, 0);
    for (int b = 0; b < nb; ++b) {
        if (m[b] < 0 || n[b] < 0 || k[b] < 0) throw Error(HIFDE_ERR_ARG, "negative size");
        oA[b + 1] = oA[b] + (long long)m[b] * k[b];
        oB[b + 1] = oB[b] + (long long)k[b] * n[b];
        oC[b + 1] = oC[b] + (long long)m[b] * n[b];
    }
    auto up = [&](const double* src, long long cnt, DevBuf<T>& d) {
        std::vector<T> tmp(src, src + cnt);
        d.upload(tmp, st);
    };
    DevBuf<T> dA, dB, dC;
    up(A, oA[nb], dA);
    up(B, oB[nb], dB);
    dC.reserve(oC[nb]);
    GemmWork gws;
    gws.plan.split = split;
    gws.plan.tile64 = tile64;
    gws.plan.force_kind = force;
    gws.plan.use_table = table;
    batched_gemm<T>(transA != 0, nb, m, n, k, oA.data(), oB.data(), oC.data(), dA.p, dB.p, dC.p,
                    (T)1, gws, st, lower);
    std::vector<T> tmp(oC[nb]);
    if (oC[nb])
        HIFDE_CHECK(
            hipMemcpyAsync(tmp.data(), dC.p, oC[nb] * sizeof(T), hipMemcpyDeviceToHost, st));
    HIFDE_CHECK(hipStreamSynchronize(st));
    for (long long i = 0; i < oC[nb]; ++i) C[i] = (double)tmp[i];
    return HIFDE_OK;
}
}  // namespace

extern "C" {

int hifde_gemm_batch(int precision, int transA, int nbatch, const int* m, const int* n,
                     const int* k, const double* A, const double* B, double* C) {
    return guarded([&]() -> int {
        if (nbatch < 0 || (nbatch && (!m || !n || !k || !A || !B || !C)))
            throw Error(HIFDE_ERR_ARG, "gemm_batch: bad arguments");
        if (nbatch == 0) return HIFDE_OK;
        const bool split = (precision & HIFDE_BF16X3) != 0;
        const int tile64 = (precision & HIFDE_GEMM_TILE64) ? 3 : 0;
        const int force = (precision & HIFDE_GEMM_FORCE_FMA)      ? 0
                          : (precision & HIFDE_GEMM_FORCE_CORE128) ? 1
                          : (precision & HIFDE_GEMM_FORCE_CORE64)  ? 2
                                                                   : -1;
        const bool table = (precision & HIFDE_GEMM_RULE) == 0;
        const bool lower = (precision & HIFDE_GEMM_LOWER) != 0;
        const int prec = precision & ~(HIFDE_BF16X3 | HIFDE_GEMM_TILE64 | HIFDE_GEMM_FORCE_FMA |
                                       HIFDE_GEMM_FORCE_CORE128 | HIFDE_GEMM_FORCE_CORE64 |
                                       HIFDE_GEMM_RULE | HIFDE_GEMM_LOWER);
        if (split && prec != HIFDE_MIXED)
            throw Error(HIFDE_ERR_ARG, "gemm_batch: HIFDE_BF16X3 needs HIFDE_MIXED");
        if (prec == HIFDE_MIXED)
            return gemm_batch_impl<float>(transA, nbatch, m, n, k, A, B, C, split, tile64, force,
                                          table, lower);
        return gemm_batch_impl<double>(transA, nbatch, m, n, k, A, B, C, false, tile64, force,
                                       table, lower);
    });
}

int hifde_schur_batch(int precision, int nbatch, const int* r, const int* s, const double* A22,
                      const double* A21, const double* A12, double* A22inv, double* X, double* S,
                      int* info) {
    return guarded([&]() -> int {
        if (nbatch < 0 || (nbatch && (!r || !s || !A22 || !A21 || !A12 || !A22inv || !X || !S ||
                                      !info)))
            throw Error(HIFDE_ERR_ARG, "schur_batch: bad arguments");
        if (nbatch == 0) return HIFDE_OK;
        const bool sym = (precision & HIFDE_SCHUR_SYMMETRIC) != 0;
        const bool split = (precision & HIFDE_BF16X3) != 0;
        const int prec = precision & ~(HIFDE_SCHUR_SYMMETRIC | HIFDE_BF16X3);
        if (split && prec != HIFDE_MIXED)
            throw Error(HIFDE_ERR_ARG, "schur_batch: HIFDE_BF16X3 needs HIFDE_MIXED");
        if (prec == HIFDE_MIXED)
            return schur_batch_impl<float>(nbatch, r, s, A22, A21, A12, A22inv, X, S, info, sym,
                                           split);
        return schur_batch_impl<double>(nbatch, r, s, A22, A21, A12, A22inv, X, S, info, sym,
                                        false);
    });
}

}  // extern "C"
