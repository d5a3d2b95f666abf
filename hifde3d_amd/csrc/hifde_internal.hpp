// Internal declarations shared by the HIP translation units of libhifde_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace hifde {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIFDE_CHECK(x)                                                                   \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess)                                                            \
            throw ::hifde::Error(-2, std::string(#x) + ": " + hipGetErrorString(e_));    \
    } while (0)

#define HIFDE_LAUNCH_CHECK() HIFDE_CHECK(hipGetLastError())

// Pinned staging arena for host-to-device uploads.  hipMemcpyAsync from
// pageable memory returns only once the data is copied; staged through
// pinned memory an upload is queued on the stream like a kernel and the host
// runs ahead.  The arena is reused once the stream has been synchronized
// (synced()), and synchronizes the stream itself when it runs out of room.
struct Stager {
    unsigned char* buf = nullptr;
    size_t cap = 0, used = 0;
    hipStream_t stream = nullptr;
    // buffers outgrown since the last synchronization (copies from them may
    // still be in flight); freed by synced()
    std::vector<unsigned char*> retired;
    Stager() = default;
    Stager(const Stager&) = delete;
    Stager& operator=(const Stager&) = delete;
    ~Stager() {
        for (unsigned char* p : retired) (void)hipHostFree(p);
        if (buf) (void)hipHostFree(buf);
    }
    const void* stage(const void* src, size_t bytes);
    // the stream has been synchronized: every staged copy has completed
    void synced() {
        used = 0;
        for (unsigned char* p : retired) (void)hipHostFree(p);
        retired.clear();
    }
};
// Stager of the calling thread's current factorization (null: plain copies).
Stager* active_stager();
void set_active_stager(Stager* st);
inline void upload_bytes(void* dst, const void* src, size_t bytes, hipStream_t s) {
    Stager* st = active_stager();
    if (st && st->stream == s) src = st->stage(src, bytes);
    HIFDE_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
}

// Per-launch device timing of individual kernels: HIP events recorded on the
// launch stream around each launch, with the launch's algorithmic flops and
// bytes.  Active while a profiled factorization or solve runs.
struct KernelTimer {
    virtual ~KernelTimer() = default;
    // brackets one launch of `kernel` on stream s
    virtual void begin(hipStream_t s, const char* kernel) = 0;
    virtual void end(hipStream_t s, double flops, double bytes) = 0;
};
KernelTimer* active_timer();
void set_active_timer(KernelTimer* t);

// Stream of the factor object the calling thread is working for (null
// outside its calls): the only stream that uses that object's buffers.
hipStream_t owner_stream();
void set_owner_stream(hipStream_t s);

// Owning device buffer.  Grows, never shrinks, unless release() is called.
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void reserve(size_t cnt) {
        if (cnt <= n && p) return;
        // work queued earlier may still read the old buffer: wait for the
        // owning call's stream.  Never the whole device or the legacy null
        // stream: a device-wide synchronize waits for other objects' streams
        // and invalidates their graph captures.  Every entry point that grows
        // a buffer names its stream (OwnerGuard / TimerGuard).
        if (p) {
            const hipStream_t os = owner_stream();
            if (!os) throw Error(-4, "internal: device buffer grown outside an owning call");
            HIFDE_CHECK(hipStreamSynchronize(os));
        }
        release();
        HIFDE_CHECK(hipMalloc(reinterpret_cast<void**>(&p), (cnt ? cnt : 1) * sizeof(T)));
        n = cnt;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void upload(const T* h, size_t cnt, hipStream_t s) {
        reserve(cnt);
        if (cnt) upload_bytes(p, h, cnt * sizeof(T), s);
    }
    void upload(const std::vector<T>& h, hipStream_t s) { upload(h.data(), h.size(), s); }
};

// Level-matrix slot: key = row << 32 | col, EMPTY when unused.
struct __align__(16) Slot {
    unsigned long long key;
    double val;
};
constexpr unsigned long long kEmptyKey = ~0ull;

// Device hash table holding the level matrix A(i, j) for live DOFs.
struct LevelTable {
    DevBuf<Slot> slots;
    size_t cap = 0;    // power of two
    bool sym = false;  // one entry per unordered pair, keyed (min, max)
};

// Output tile of one MFMA GEMM workgroup (128 x 128).
constexpr int kMfmaTile = 128;

struct GemmDesc {
    long long offA, offB, offC;
    int m, n, k, lda, ldb, ldc;
    int lower = 0;  // 1: only output tiles on or below the block diagonal
    int flag = -1;  // >= 0: skipped when flags[flag] == 0 (device flags)
};
struct CopyDesc {
    long long off;  // source block base
    int r0, c0, rows, ld;
};
// Descriptors of many batched GEMM launches, built on the host, uploaded in
// one copy each and then launched one by one (gemm_run).
struct GemmPlan {
    struct Launch {
        int mfma;  // kernel: 0 FMA tiles, 1 matrix cores (128 x 128), 2 (64 x 64)
        int transA, nb, d0, t0, total;
        double flops, ab_elems, cells;  // algorithmic work of the launch
        int split;  // matrix-core launch with bf16x3 split products (FP32 only)
        int contig = 0;  // matrix-core launch: contiguous tile range per XCD
    };
    // Launches added while set use the split (bf16x3) matrix-core products;
    // clear() leaves it as it is.
    bool split = false;
    // 64 x 64 matrix-core tiles (kind 2): 0 never; 1 for products of depth
    // above 64 whose 128 x 128 tiles would compute over a quarter more
    // output; 2 also for products of depth 64 or less with both output sides
    // of at least 32; 3 every product with both output sides of at least 32.
    // clear() leaves it as it is.
    int tile64 = 0;
    // Bytes per operand element of the products added (4 or 8): part of the
    // tuning-table key; set by the caller before add().
    int esize = 8;
    // Route by the process-wide tuning table (gemm_table_*) where it has an
    // entry for a descriptor's class; otherwise by the fixed rule.
    bool use_table = true;
    // >= 0: every descriptor that kernel kind can run goes to it (timing
    // replays of the tuner); -1: normal routing.
    int force_kind = -1;
    std::vector<GemmDesc> desc;
    std::vector<int> tiles;  // per launch: nb + 1 local tile offsets
    std::vector<Launch> launches;
    DevBuf<GemmDesc> d_desc;
    DevBuf<int> d_tiles;
    void clear() {
        desc.clear();
        tiles.clear();
        launches.clear();
    }
    // Appends one batched GEMM (outputs narrower than half an MFMA tile on
    // the FMA kernel, the rest on the MFMA kernel); returns its launches
    // [first, second).  one_launch: a latency-bound call of few descriptors
    // whose descriptors would go to several kernels goes to one of them.
    std::pair<int, int> add(const std::vector<GemmDesc>& d, bool transA,
                            bool one_launch = false);
    void upload(hipStream_t s) {
        d_desc.upload(desc, s);
        d_tiles.upload(tiles, s);
    }
};

// One batch of boxes of a half level (cells of level l, or faces of level l).
// Dense blocks are column-major and packed box after box; offsets are in
// elements.
struct BoxBatch {
    int nbox = 0;
    int esize = 8;  // bytes per element of the stored factor blocks
    std::vector<int> r, s;                      // |red|, |skel| per box
    std::vector<long long> off_rr, off_rs, off_ss;  // prefix sums of r*r, r*s, s*s (nbox+1)
    std::vector<int> red_off, skel_off;         // prefix sums of r, s (nbox+1)
    std::vector<int> red_gid, skel_gid;         // global DOF ids, reference order
    std::vector<unsigned char> skel_pass;       // cells: 1 if the DOF lies on a min face
    DevBuf<int> d_r, d_s, d_red_off, d_skel_off, d_red_gid, d_skel_gid;
    DevBuf<unsigned char> d_skel_pass;
    DevBuf<long long> d_off_rr, d_off_rs, d_off_ss;
    // Factor data in the factor precision (float or double), as raw bytes.
    DevBuf<unsigned char> A22, A22inv, X, W;
    long long tot_rr = 0, tot_rs = 0, tot_ss = 0;
    // apply tiling: prefix sums of ceil(r/64) and ceil(s/64), and of
    // ceil(r/128) and ceil(s/128) (tiles of two rows per lane)
    std::vector<int> tiles_r, tiles_s, tiles_r2, tiles_s2;
    DevBuf<int> d_tiles_r, d_tiles_s, d_tiles_r2, d_tiles_s2;
    int max_r = 0, max_s = 0, max_rs = 0;  // largest r, s and r s of a box
    // Cells: as many skeleton entries on min faces as on max faces (each
    // shared face DOF has one of each), so the two contributions a DOF gets
    // from the side buffers can be added by one thread.
    bool sides_paired = false;
    // Faces only: per-face box geometry.
    std::vector<int> face_kind;
    // Cells, element storage of the level matrix: the live edge and corner
    // DOFs of the closure ("other"), and the element block over
    // [skel | other] kept after the cell's Schur complement.
    std::vector<int> o, other_off, other_gid;
    std::vector<long long> off_ee;
    long long tot_ee = 0;
    // Multi-RHS apply: per op, the batched GEMM of every box for rplan_R
    // right-hand sides (0: not built for the current factorization).
    GemmPlan rplan[5];
    int rplan_R[5] = {0, 0, 0, 0, 0};
    int rplan_zero[5] = {0, 0, 0, 0, 0};  // some box has outputs but no inputs
    // Clears the index description; device buffers are kept for reuse.
    void reset_host() {
        for (int& x : rplan_R) x = 0;
        nbox = 0;
        r.clear(); s.clear(); o.clear();
        off_rr.clear(); off_rs.clear(); off_ss.clear(); off_ee.clear();
        red_off.clear(); skel_off.clear(); other_off.clear();
        red_gid.clear(); skel_gid.clear(); other_gid.clear(); skel_pass.clear();
        tot_rr = tot_rs = tot_ss = tot_ee = 0;
        tiles_r.clear(); tiles_s.clear(); tiles_r2.clear(); tiles_s2.clear();
        max_r = max_s = max_rs = 0;
        sides_paired = false;
        face_kind.clear();
    }
};

// Element storage descriptors (level_elem.hip).
struct InitDesc {            // one level-0 cell
    long long rr, rs, ee;    // offsets of its A22, A21/A12 and element blocks
    long long tab;           // offset of its position table
    int gid;                 // offset of its DOF list [red | skel | other]
    int r, s, o;
    int lo[3], ext[3], cc[3];  // closure box origin, extent, cell coordinates
};
struct ExtendDesc {          // one (parent, child) pair of an extend-add pass
    long long src;           // child element offset
    long long rr, rs, ee;    // parent block offsets
    int e;                   // child element size (its leading dimension)
    int map;                 // offset of the child's parent-position map
    int r, s, ep;            // parent red / skel counts and element size
};
constexpr unsigned kNoPos = 0xFFFFu;  // absent position in a packed pair
struct FaceGatherDesc {      // one face
    long long e1, e2;        // elements of the two adjacent cells
    long long offM, offC;
    int ld1, ld2, m, n;
    int rows, cols;          // offsets of the row / column position arrays
};
struct FaceAddDesc {         // one face's Schur complement into its owner's element
    long long e1, offS;
    int ld1, k, sp;
};
struct MirrorDesc {
    long long ee;
    int s, ld;
};

}  // namespace hifde

// ---------------------------------------------------------------------------
// Kernel launchers (defined in the .hip translation units)
// ---------------------------------------------------------------------------
namespace hifde {

// level table
void table_init(LevelTable& t, size_t min_entries, hipStream_t s);
void table_insert_coo(LevelTable& t, const int* d_rows, const int* d_cols, const double* d_vals,
                      long long nnz, hipStream_t s);
// out(box) = A(rows(box), cols(box)), column-major, converted to T.
template <class T>
void table_gather(const LevelTable& t, int nbox, const int* d_rows, const int* d_row_off,
                  const int* d_cols, const int* d_col_off, const long long* d_out_off,
                  long long total, T* d_out, hipStream_t s);
// A(skel_i, skel_j) += S(i, j) for pass-selected pairs (pass < 0: all pairs).
template <class T>
void table_scatter(LevelTable& t, int nbox, const int* d_s, const int* d_skel_off,
                   const int* d_skel_gid, const unsigned char* d_skel_pass,
                   const long long* d_off_ss, long long total, const T* d_S, int pass,
                   hipStream_t s);
// dst = entries of src whose row and column are alive, in a table sized to
// also hold `extra` further inserts.  dst's buffer only ever grows.
// Returns the number of live entries copied.
size_t table_compact(const LevelTable& src, LevelTable& dst, const unsigned char* d_alive,
                     size_t extra, DevBuf<unsigned long long>& counter, hipStream_t s);
size_t table_count(const LevelTable& t, hipStream_t s);

// dense batched kernels
// Workspace of the blocked inverse (pivots, pivot rows, tile list).
// One strided block copy: rows x cols from src (leading dim lds) to dst
// (leading dim ldd); sel picks one of two source buffers.
struct Copy2D {
    long long src, dst;
    int rows, cols, lds, ldd, sel;
};


template <class T, class TC = T>
void gemm_run(const GemmPlan& p, std::pair<int, int> h, const T* d_A, const T* d_B, TC* d_C,
              T alpha, T beta, hipStream_t s, const int* d_flags = nullptr);

// GEMM routing table written by the project's tuner (hifde3d_amd/tune.py):
// per descriptor class -- element size, transA, lower, depth bin, bin of the
// smaller output side, bin of the call's descriptor count -- the kernel kind
// that ran its recorded products fastest (0 FMA 64 x 64, 1 matrix-core
// 128 x 128, 2 matrix-core 64 x 64), or -1 (no entry: fixed rule).
constexpr int kTuneKBins = 8, kTuneMNBins = 6, kTuneNBBins = 5;
constexpr int kTuneEntries = 2 * 2 * 2 * kTuneKBins * kTuneMNBins * kTuneNBBins;
int gemm_tune_key(int esize, int transA, int lower, int m, int n, int k, int nb);
int gemm_tune_lookup(int key);  // -1 when absent
void gemm_tune_install(const signed char* kinds);  // kTuneEntries kinds
// Every descriptor added while a recorder is active on this thread is
// logged: call id, esize, transA, lower, m, n, k, descriptors in the call,
// routed kind.
struct GemmRecorder {
    std::vector<int> log;  // 9 ints per descriptor
    int calls = 0;
};
GemmRecorder* active_gemm_recorder();
void set_active_gemm_recorder(GemmRecorder* r);

struct GemmWork {
    GemmPlan plan;
    DevBuf<CopyDesc> cdesc;
    DevBuf<long long> off;
};
// dst(i, j) = sign(j) src(i, j), or src(j, i) when trans, for i < rows and
// j < cols (only i < j when upper == 1, only i > j when upper == 2);
// sign(j) = -s0 for j >= neg_from, s0 before; avg 1: dst(i, j) becomes the
// mean of its value and that; avg 2 (dst == src, trans, upper == 1): the
// block is made symmetric, both triangles taking (a_ij + a_ji) / 2.
// sel: 0 the matrices, 1 the symmetric-path workspace.
struct SymCopy {
    long long dst, src;
    int rows, cols, ldd, lds;
    int trans, upper, neg_from, s0, dsel, ssel;
    int avg;
};
struct GJWork {
    DevBuf<int> wn, wld;              // per outer step and box: window size, leading dim
    DevBuf<long long> woff, b_off;    // window offsets; inner R offsets
    DevBuf<unsigned char> B, D;       // inner R, inner diagonal inverses
    DevBuf<unsigned char> RO, CW;     // outer R (kGJPanel x n), outer column panel (n x kGJPanel)
    DevBuf<Copy2D> cdesc;
    DevBuf<int> bix, info;            // blocked-path boxes and their status
    GemmPlan plan;                    // every GEMM of one inversion
    DevBuf<unsigned char> SW;         // symmetric path: V, W (n x kGJPanel), V^T per box
    DevBuf<SymCopy> sdesc;
    DevBuf<long long> stile;          // per copy launch: tile prefix of its descriptors
    // outer window width (a multiple of 64); 0: the default (512)
    int panel = 0;
    // each inner step's column kernel shares a launch with the next step's
    // diagonal block (k_bgj_col_diag); off: separate launches
    bool fused_steps = true;
};

// C = alpha op(A) B + beta C per descriptor (op(A) = A^T when transA).  C
// may be FP64 for FP32 products (accumulated in FP32, added in FP64).
template <class T, class TC = T>
void batched_gemm_desc(bool transA, const std::vector<GemmDesc>& desc, const T* d_A, const T* d_B,
                       TC* d_C, T alpha, T beta, GemmWork& ws, hipStream_t s,
                       const int* d_flags = nullptr);
// Packed boxes (ld = rows): C = alpha op(A) B with host offsets.  lower:
// square boxes compute only their output tiles on or below the diagonal.
template <class T>
void batched_gemm(bool transA, int nbox, const int* h_m, const int* h_n, const int* h_k,
                  const long long* h_offA, const long long* h_offB, const long long* h_offC,
                  const T* d_A, const T* d_B, T* d_C, T alpha, GemmWork& ws, hipStream_t s,
                  bool lower = false);
template <class D>
void copy_blocks(const std::vector<CopyDesc>& desc, const std::vector<long long>& dst_off,
                 const double* d_src, D* d_dst, GemmWork& ws, hipStream_t s);
// In-place inverses of the boxes (n x n, ld n).  sym: every box is symmetric;
// boxes wider than one window then read only their lower triangle and the
// outer level updates only lower triangles (half the flops).
template <class T>
void batched_gj_inverse(int nbox, const int* h_n, const long long* h_off, const int* d_n,
                        const long long* d_off, T* d_A, int* d_info, GJWork& ws, GemmWork& gws,
                        hipStream_t s, bool sym = false);
void permute_sym(int nbox, const long long* d_off, const int* d_n, const int* d_perm,
                 const int* d_perm_off, const double* d_C, double* d_Cp, long long total,
                 hipStream_t s);
// mode[f] != 0 selects the Gram path for face f; the Householder kernel skips it.
// d_M is overwritten with the QR factors (R on and above the diagonal, the
// Householder vectors below it, dgeqp3's layout).  d_jpvt / d_tau, when not
// null, receive each face's column pivots (0-based, at d_perm_off[f]) and
// reflector scalars (min(m, n) of them, at d_perm_off[f]).
void batched_interp_decomp(int nface, const int* d_m, const int* d_n, const long long* d_offM,
                           double* d_M, double eps, int* d_perm, const int* d_perm_off,
                           int* d_k, double* d_W, const long long* d_offW, int max_n,
                           int max_m, const unsigned char* d_mode, hipStream_t s,
                           int* d_jpvt = nullptr, double* d_tau = nullptr);
// Device problem assembly and input generators (assemble.hip).
void gen_uniform(double* d_out, long long n, unsigned long long seed, double lo, double hi,
                 bool pow10, hipStream_t s);
void gen_normal(double* d_out, long long n, unsigned long long seed, hipStream_t s);
void gen_rhs(double* d_out, int NC, int nrhs, long long ld, unsigned long long seed,
             hipStream_t s);
void gen_gaussian_field(double* d_out, int NC, unsigned long long seed, double corr,
                        bool lognormal, hipStream_t s);
long long stencil_csr(int NC, const double* d_a, const double* d_V, int symmetric,
                      DevBuf<int>& rp, DevBuf<int>& ci, DevBuf<double>& val, DevBuf<int>& rr,
                      hipStream_t s);
bool csr_symmetric(long long nd, const int* d_rp, const int* d_ci, const double* d_val,
                   hipStream_t s);
// B_b := R_b^{-1} B_b, R_b upper triangular m_b x m_b, B_b m_b x n_b, both
// packed column-major at their offsets.
void batched_trsm_upper(int nbox, const int* d_m, const int* d_n, const long long* d_offR,
                        const double* d_R, const long long* d_offB, double* d_B, int max_n,
                        hipStream_t s);
// Gram-path ID (faces with mode != 0): blocked pivoted Cholesky of the
// lower triangle of G (overwritten), rows of R, pivot order and rank; then,
// once the ranks are known on the host, W = R11^{-1} R12.
constexpr int kGramNB = 32;
struct GramRDesc {
    long long r, out;  // R of the face (rows at r + i n), [R11 | R12] output
    int k, n, perm;    // rank, columns, offset of the face's pivot order
};
struct GramWork {
    GemmPlan plan;
    DevBuf<double> P, dg, tol, RK;
    DevBuf<long long> offP, roff, ko;
    DevBuf<int> active, count, kd, info;
    DevBuf<GramRDesc> rdesc;
    std::vector<int> host_active;
    // panel rows in registers (faces up to 1536 columns); false takes the
    // LDS/global-memory panel for every size
    bool reg_panel = true;
};
void gram_pivchol(int nface, const int* h_n, const unsigned char* h_mode, const int* d_n,
                  const unsigned char* d_mode, double* d_G, const long long* h_offG,
                  const long long* d_offG, double* d_R, double eps, int* d_perm,
                  const int* h_perm_off, const int* d_perm_off, int* d_k, GramWork& ws,
                  GemmWork& gws, hipStream_t s);
// Lower triangle of an n x n block as the sum of `slices` partial blocks
// (split-k Gram products), added in slice order.
struct SliceSum {
    long long dst, src;  // block offsets; slice s at src + s n^2
    int n, slices;
};
void sum_slices_lower(const std::vector<SliceSum>& desc, DevBuf<SliceSum>& d_desc, double* d_G,
                      hipStream_t s);
void gram_w(const std::vector<GramRDesc>& desc, const std::vector<long long>& toff,
            const double* d_R, const int* d_perm, const std::vector<long long>& offW,
            double* d_W, GramWork& ws, GJWork& gj, GemmWork& gws, hipStream_t s);
void fold_active(int n, const int* d_active, int* d_count, hipStream_t s);
void convert_d2f(const double* src, float* dst, long long n, hipStream_t s);
// *d_status = info[b] of the first box b with a nonzero status, unless
// *d_status is already nonzero.
void fold_info(int nbox, const int* d_info, int* d_status, hipStream_t s);
void fill_int(int* p, int n, int v, hipStream_t s);
template <class D>
void pack_boxes(int nbox, const long long* d_dst_off, const long long* d_src_off,
                const double* d_src, D* d_dst, long long total, hipStream_t s);

// element storage of the level matrix (level_elem.hip); elements are FP64,
// or FP32 (TE) when every level is factored in FP32; the frontal blocks A22 /
// A21 / A12 are in the level's factor precision T
template <class T>
void fill_zero(T* p, long long n, hipStream_t s);
template <class T, class TE>
void frontal_init(int ncell, const InitDesc* d_desc, const int* d_tab, const int* d_gids,
                  const int* d_rp, const int* d_ci, const double* d_v, int NC, int w, int c,
                  T* A22, T* A21, T* A12, TE* E, hipStream_t s);
// One launch per pass (child position); each pass adds one child per parent.
template <class Td, class TE>
void extend_add(const std::vector<std::vector<ExtendDesc>>& passes, DevBuf<ExtendDesc>& dbuf,
                DevBuf<long long>& obuf, const int* d_maps, const TE* src, Td* A22, Td* A21,
                Td* A12, TE* E, hipStream_t s);
// Uploads the face descriptors and their column prefix; returns the number of
// face columns (the launch size of face_gather).
long long face_gather_prepare(const std::vector<FaceGatherDesc>& desc,
                              DevBuf<FaceGatherDesc>& dbuf, DevBuf<long long>& obuf,
                              hipStream_t s);
template <class TE>
void face_gather(int nface, long long total_cols, const FaceGatherDesc* d_desc,
                 const long long* d_off, const unsigned* d_rpos, const unsigned* d_cpos,
                 const TE* E, double* M, double* C, hipStream_t s);
template <class T, class TE>
void face_add(const std::vector<FaceAddDesc>& desc, DevBuf<FaceAddDesc>& dbuf,
              DevBuf<long long>& obuf, const int* d_spos, const T* S, TE* E, int sym,
              hipStream_t s);
template <class TE>
void mirror_lower(const std::vector<MirrorDesc>& desc, DevBuf<MirrorDesc>& dbuf,
                  DevBuf<long long>& obuf, TE* E, hipStream_t s);

// apply sweeps: one tile of 64 (or 128) outputs of one box per workgroup
// (see apply_pcg.hip).  Fewest 128-output tiles for which a launch of the non-transposed apply ops
// uses two rows per lane.
constexpr int kApplyWideMinTiles = 512;
template <class T>
void apply_tiles(int nbox, int ntiles, const int* d_tile_off, const int* d_r, const int* d_s,
                 const int* d_red_off, const int* d_skel_off, const int* d_red_gid,
                 const int* d_skel_gid, const unsigned char* d_skel_pass, const long long* d_off,
                 const T* d_M, int op, int pass, double alpha, const double* d_in,
                 double* d_out, int max_in, double flops, double bytes, hipStream_t s,
                 long long side_n = 0, int rows_per_lane = 1);
// Batches whose boxes have at most 64 inputs and outputs: one thread per output.
// max_block: the largest box block (elements); boxes are staged in LDS when
// four such blocks fit in 64 KB (the callers pass it for ops 0 and 2, whose
// lanes read columns, and for op 4; 0: read in place).
template <class T>
void apply_small(int nbox, int nout_total, const int* d_r, const int* d_s, const int* d_red_off,
                 const int* d_skel_off, const int* d_red_gid, const int* d_skel_gid,
                 const unsigned char* d_skel_pass, const long long* d_off, const T* d_M, int op,
                 int pass, double alpha, const double* d_in, double* d_out, double flops,
                 double bytes, hipStream_t s, long long side_n = 0, int max_block = 0);
// side_n > 0 (op 0): each output goes to d_out[pass flag * side_n + DOF]
// instead of being added to the DOF; side_add then adds the pass-0 and then
// the pass-1 contributions of a batch's skeleton DOFs to u.  paired: every
// DOF has exactly one entry of each pass flag; one launch then adds both
// contributions, in pass order, from the pass-1 entry.
void side_add(int n, const int* d_gid, const unsigned char* d_sp, const double* d_side,
              long long side_n, double* d_u, hipStream_t s, bool paired = false);
void zero_boundary(double* d_u, int NC, hipStream_t s);

// several right-hand sides (multi_rhs.hip).  A block of R vectors: entry
// (dof g, column c) at U[g sg + c sc]; column-major is {1, ld}, interleaved
// (a DOF's R values contiguous) is {R, 1}.
struct RhsLayout {
    long long sg, sc;
};
template <class T>
void rhs_gather(int nbox, long long nentries, const int* d_off, const int* d_gid, const double* U,
                RhsLayout ly, int R, T* pack, hipStream_t s);
// mode 0: add alpha * pack; 1: write pack; 2: per-pass side buffers (laid
// out like U, pstride apart)
template <class T>
void rhs_scatter(int nbox, long long nentries, const int* d_off, const int* d_gid, const T* pack,
                 int R, double* U, RhsLayout ly, double alpha, int mode, const unsigned char* d_sp,
                 long long pstride, hipStream_t s);
void to_interleaved(long long n, int R, const double* U, long long ld, double* W, hipStream_t s);
void from_interleaved(long long n, int R, const double* W, double* U, long long ld,
                      hipStream_t s);
void zero_boundary_il(int NC, int R, double* W, hipStream_t s);
// gather -> product -> scatter in one kernel (boxes of moderate size)
template <class T>
void apply_multi_fused(int nbox, int ntiles, const int* d_tile_off, const int* d_r, const int* d_s,
                       const int* d_red_off, const int* d_skel_off, const int* d_red_gid,
                       const int* d_skel_gid, const unsigned char* d_skel_pass,
                       const long long* d_off, const T* d_M, int op, double alpha, double* U,
                       RhsLayout ly, int R, double* side, long long pstride, double flops,
                       double bytes, hipStream_t s);
void rhs_side_add(long long nentries, const int* d_gid, const unsigned char* d_sp,
                  const double* side, long long pstride, int R, double* U, RhsLayout ly,
                  hipStream_t s);
void csr_spmm(int n, const int* d_rp, const int* d_ci, const double* d_v, const double* X,
              long long ldx, double* Y, long long ldy, int R, hipStream_t s);
void dot_cols(long long n, const double* X, long long ldx, const double* Y, long long ldy, int R,
              double* d_part, int nblk, double* d_out, hipStream_t s);
void axpy_cols(long long n, const double* d_a, const double* X, long long ldx, double* Y,
               long long ldy, int R, bool xpby, hipStream_t s);
// The same per column and segment (d_seg: segment of each entry, 255 for
// none; 1 <= nseg <= 8): d_out[c nseg + s]; scalars d_a[c nseg + s].
void dot_cols_seg(long long n, const double* X, long long ldx, const double* Y, long long ldy,
                  int R, const unsigned char* d_seg, int nseg, double* d_part, int nblk,
                  double* d_out, hipStream_t s);
void axpy_cols_seg(long long n, const double* d_a, const double* X, long long ldx, double* Y,
                   long long ldy, int R, bool xpby, const unsigned char* d_seg, int nseg,
                   hipStream_t s);
// Packed batches (assemble.hip): nprob problems of grid NC as the octants of
// a grid 2 NC wide; CSR with identity rows on the separating planes and on
// unused octants, and the segment (problem) of every packed entry.
long long pack_stencil_csr(int NC, int nprob, const double* const* d_a_ptrs,
                           const double* const* d_V_ptrs, int symmetric, DevBuf<int>& rp,
                           DevBuf<int>& ci, DevBuf<double>& val, DevBuf<int>& rr,
                           DevBuf<unsigned char>& seg, hipStream_t s);
void pack_vectors(int NC, int nprob, int R, double* sub, long long ld_sub, double* packed,
                  long long ld_packed, bool unpack, hipStream_t s);

// vectors / CSR SpMV for PCG
void csr_spmv(int n, const int* d_rowptr, const int* d_colind, const double* d_val,
              const double* d_x, double* d_y, hipStream_t s);
// Slot-major copy of a CSR whose rows hold at most kEllMaxWidth entries:
// ell_v[q n + i] and ell_c[q n + i] are row i's q-th entry (in CSR order).
// The SpMV then reads entries coalesced and sums each row in CSR order, so
// its results equal csr_spmv's bit for bit.
constexpr int kEllMaxWidth = 8;
int csr_max_row(int n, const int* d_rowptr, hipStream_t s);
void csr_to_ell(int n, int width, const int* d_rowptr, const int* d_colind, const double* d_val,
                double* d_ell_v, int* d_ell_c, hipStream_t s);
void ell_spmv(int n, int width, const int* d_rowptr, const double* d_ell_v, const int* d_ell_c,
              const double* d_x, double* d_y, hipStream_t s);
void dot_partial(const double* x, const double* y, long long n, double* d_part, int nblk,
                 hipStream_t s);
void reduce_partials(const double* d_part, int nblk, double* d_out, hipStream_t s);
void axpy(double a, const double* x, double* y, long long n, hipStream_t s);       // y += a x
void xpby(const double* x, double b, double* y, long long n, hipStream_t s);       // y = x + b y
void vcopy(const double* x, double* y, long long n, hipStream_t s);

int dot_blocks(long long n);

}  // namespace hifde
