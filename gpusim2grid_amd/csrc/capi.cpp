// C ABI of the batched Newton-Raphson power flow (see include/gpusim2grid_amd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/gpusim2grid_amd.h"
#include "capi_common.hpp"
#include "device_common.hpp"
#include "plan.hpp"

namespace g2g {

// kernels (kernels.hip)
template <typename Real, bool PerRowY, bool WithJ>
__global__ void k_mismatch(MismatchArgs<Real> a);
template <typename Real, bool WithJ>
__global__ void k_mis_status(MismatchArgs<Real> a);
template <typename Real, int W, bool SPLIT, int RG, int MODE>
__global__ void k_factor(FactorArgs<Real> a);
template <typename Real, int W, bool SPLIT>
__global__ void k_backward(BackwardArgs<Real> a);
template <typename Real>
__global__ void k_update(UpdateArgs<Real> a);
template <typename Real>
__global__ void k_tiles_to_rows_f64(const Real* in, int rows, int m, double* out, long long ld);
__global__ void k_slots_range_to_rows(const double2* in, int row0, int rows, int m, double2* out, long long ld);
template <typename Real>
__global__ void k_ext_pre(ExtArgs<Real> a);
template <typename Real>
__global__ void k_ext_post(ExtArgs<Real> a);
template <typename Real>
__global__ void k_ext_update(ExtArgs<Real> a);
template <typename Real>
__global__ void k_ext_init(ExtArgs<Real> a);
template <typename Real>
__global__ void k_lin_residual(RefineArgs<Real> a);
template <typename Real, bool PerRowY>
__global__ void k_ir_residual(IrArgs<Real> a);
template <typename Real>
__global__ void k_axpy_tiles(const Real* Xs, Real* X, const int* tile_active, int d);
template <typename T>
__global__ void k_rows_to_tiles(const T* in, long long ld, int rows, int m, int row_offset, T* out, int broadcast);
template <typename T>
__global__ void k_tiles_to_rows(const T* in, int rows, int m, T* out, long long ld);
__global__ void k_set_row_edits(const int* row_ptr, const int* ent, const double2* delta, int R, const int* slots,
                                int cap, int* n_ed, int* ed_ent, double2* ed_val);
__global__ void k_edit_flags(const int* n_ed, const int* ed_ent, int cap, int rows, const int* entry_row, int n,
                             int* flag);
template <typename T>
__global__ void k_regrow_edits(const T* src, int cap_old, T* dst, int cap_new, int tiles);
template <typename T>
__global__ void k_rows_to_slots(const T* src, long long ld, int R, int m, const int* slots, T* dst);
template <typename T>
__global__ void k_slots_to_rows(const T* src, int R, int m, const int* slots, T* dst, long long ld);
__global__ void k_set_slots(const int* slots, int R, int* v, int value);
__global__ void k_mask_v_nan(const unsigned char* mask, long long total, double2* V);
__global__ void k_restart_v(const int* status, const unsigned char* warm, int n, const double2* V0, int v0_row,
                            double2* V);
__global__ void k_v_uniform(const double2* V, int rows, int n, int* differs);
__global__ void k_restart_rows(int* status, unsigned char* warm, const int* it0, int* budget, int rows, int iter,
                               int max_iter, int* requeued, int* tile_restart);
template <typename Real>
__global__ void k_gather_astream(const double* vals, long long ldv, const double* rhs, long long ldb, int rows,
                                 int nnz, int d, const int* a_pos, const int* f_pos, int a_len, Real* A);
__global__ void k_row_absmax(const double* vals, long long ldv, int rows, int nnz, double* out);
__global__ void k_fill_int(int* p, int v, long long n);
__global__ void k_add_int(int* p, int v);
__global__ void k_count_tiles(const int* status, int rows_pad, int* tile_active);
__global__ void k_branch_flows(const double2* V, int n, const int* from, const int* to, const double2* y4,
                               const double* base_or, const double* base_ex, int m, int T, double* or_t,
                               double* ex_t);
__global__ void k_limit_violations(ViolationArgs a);
size_t limit_violations_lds(int K);
template <typename T>
__global__ void k_move_rows(const T* src, T* dst, const int* map, int R, int m, int gather);
template <typename T>
__global__ void k_move_vals(const T* src, T* dst, const int* map, int R, int gather, int add);

// learned warm start (mlp_capi.cpp)
void mlp_warm_start_tiles(g2g_mlp_s* p, const double2* S_tiles, int M, int n, const unsigned char* d_mask,
                          int n_mask, double2* V_tiles, hipStream_t st);

}  // namespace g2g

using namespace g2g;


// a segment table as the kernels read it: 9 ints per walker
inline std::vector<int> seg_words(const std::vector<g2g::SymbolicLU::Seg>& segs)
{
    std::vector<int> w;
    for (const auto& g : segs)
        w.insert(w.end(), {g.prog_f_off, g.prog_f_len, g.prog_b_off, g.prog_b_len, g.a_off, g.a_len, g.u_off, g.u_len,
                           g.n_piv});
    return w;
}

struct g2g_plan_s {
    bool grid = false;
    NrPlan nr;           // grid plans
    SymbolicLU raw;      // raw matrix plans
    int nnz = 0;
    int device = 0;
    int* d_prog_f = nullptr;
    int* d_prog_b = nullptr;
    int* d_seg = nullptr;      // per walker segments (SymbolicLU::Seg, 9 ints each)
    int* d_bseg = nullptr;     // per backward walker segments
    int* d_k1_bus = nullptr;
    int* d_k1_ent = nullptr;
    double2* d_Y = nullptr;
    int* d_perm = nullptr;
    int* d_apos = nullptr;
    int* d_fpos = nullptr;
    int* d_upd = nullptr;     // per bus {bus, theta unknown, |V| unknown}
    int n_upd = 0;
    int* d_colmap = nullptr;  // [n][2]: each bus's theta and |V| unknown, -1 for none
    int* d_entry_row = nullptr;  // [nY]: the Ybus row of each entry
    int* d_k1_ent_t = nullptr;  // k1_ent with the A positions of the transposed entries (adjoint)
    std::vector<int> raw_p, raw_i;  // a raw plan's matrix pattern (CSR)
    std::vector<int> raw_order;     // and its pivot order (empty: minimum degree)
    // variants with fewer walkers, built when a precision's walker kernels do
    // not fit the LDS
    std::vector<std::unique_ptr<g2g_plan_s>> fitted;
    int* d_Jp = nullptr;      // the matrix CSR in unknown labels (refinement residuals)
    int* d_Ji = nullptr;
    // split walk: programs, segments, and the mismatch tables of its layout
    int* d_prog_fs = nullptr;
    int* d_prog_bs = nullptr;
    int* d_split_seg = nullptr;
    int* d_split_bseg = nullptr;
    int* d_k1_bus_s = nullptr;
    int* d_k1_ent_s = nullptr;
    int prog_fs_stride = 0, prog_bs_stride = 0;
    bool split() const { return lu().n_split > 0 && (!grid || nr.split_ok()); }
    // extension rows (nr.xt): device copies of its tables
    struct {
        int *inj_ptr = nullptr, *inj = nullptr, *jrec = nullptr, *frec = nullptr, *xcol = nullptr;
        int *h_bus1 = nullptr, *h_bus2 = nullptr, *h_status = nullptr, *vc_reg = nullptr, *vc_ptr = nullptr;
        int *sh_c1 = nullptr, *sh_ck = nullptr;
        double *ds_w = nullptr, *jconst = nullptr, *h_par = nullptr, *vc_vset = nullptr, *vc_slope = nullptr,
               *vc_w = nullptr;
        void free_all()
        {
            for (void* q : {static_cast<void*>(inj_ptr), static_cast<void*>(inj), static_cast<void*>(jrec),
                            static_cast<void*>(frec), static_cast<void*>(xcol), static_cast<void*>(h_bus1),
                            static_cast<void*>(h_bus2), static_cast<void*>(h_status), static_cast<void*>(vc_reg),
                            static_cast<void*>(vc_ptr), static_cast<void*>(sh_c1), static_cast<void*>(sh_ck),
                            static_cast<void*>(ds_w), static_cast<void*>(jconst), static_cast<void*>(h_par),
                            static_cast<void*>(vc_vset), static_cast<void*>(vc_slope), static_cast<void*>(vc_w)})
                dfree(q);
        }
    } dx;
    bool ext() const { return grid && nr.xt.on(); }
    int prog_f_stride = 0, prog_b_stride = 0;
    bool uploaded = false;
    const SymbolicLU& lu() const { return grid ? nr.lu : raw; }
    ~g2g_plan_s()
    {
        dfree(d_prog_f); dfree(d_prog_b); dfree(d_seg); dfree(d_bseg); dfree(d_k1_bus); dfree(d_k1_ent);
        dfree(d_Y); dfree(d_perm); dfree(d_apos); dfree(d_fpos); dfree(d_upd); dfree(d_Jp); dfree(d_Ji);
        dfree(d_colmap); dfree(d_entry_row);
        dfree(d_k1_ent_t);
        dx.free_all();
        dfree(d_prog_fs); dfree(d_prog_bs); dfree(d_split_seg); dfree(d_split_bseg); dfree(d_k1_bus_s);
        dfree(d_k1_ent_s);
    }
    // the mismatch program writing J transposed: entry (r, c) of J goes to the
    // A position of (c, r) (the pattern is structurally symmetric)
    void upload_transposed()
    {
        if (d_k1_ent_t) return;
        const NrPlan& P = nr;
        std::vector<int> tpos(static_cast<size_t>(P.nJ), -1);
        for (int r = 0; r < P.d; ++r)
            for (int e = P.Jp[r]; e < P.Jp[r + 1]; ++e) {
                const int c = P.Ji[e];
                const auto b0 = P.Ji.begin() + P.Jp[c], b1 = P.Ji.begin() + P.Jp[c + 1];
                const auto it = std::lower_bound(b0, b1, r);
                if (it == b1 || *it != r) throw std::runtime_error("the Jacobian pattern is not structurally symmetric");
                tpos[e] = nr.lu.a_pos_of_entry[static_cast<size_t>(it - P.Ji.begin())];
            }
        // a k1 entry's a11..a22 are a_pos_of_entry[map]: replace them by the
        // transposed entries' positions
        std::vector<int> ent = P.k1_ent;
        std::vector<int> inv(static_cast<size_t>(nr.lu.a_len + 1), -1);
        for (int e = 0; e < P.nJ; ++e)
            if (nr.lu.a_pos_of_entry[e] >= 0) inv[nr.lu.a_pos_of_entry[e]] = e;
        for (size_t q = 0; q < ent.size(); q += 6)
            for (int b = 2; b < 6; ++b)
                if (ent[q + b] >= 0) ent[q + b] = tpos[inv[ent[q + b]]];
        d_k1_ent_t = dupload(ent);
    }
    // device copies of the tables, made on first use (plans can be built and
    // inspected on a host without a GPU)
    void upload()
    {
        if (uploaded) return;
        HIPCHK(hipGetDevice(&device));
        const SymbolicLU& s = lu();
        if (s.max_c > kMaxC)
            throw std::runtime_error("pivot column of " + std::to_string(s.max_c) + " entries exceeds the kernels' " +
                                     std::to_string(kMaxC));
        // programs padded to whole prefetch chunks
        auto padded = [](const std::vector<int>& v, int& stride) {
            stride = static_cast<int>((v.size() + kProgChunk - 1) / kProgChunk * kProgChunk);
            std::vector<int> out(v);
            out.resize(static_cast<size_t>(stride), 0);
            return out;
        };
        d_prog_f = dupload(padded(s.prog_f, prog_f_stride));
        d_prog_b = dupload(padded(s.prog_b, prog_b_stride));
        d_seg = dupload(seg_words(s.seg));
        d_bseg = dupload(seg_words(s.bseg));
        d_perm = dupload(s.perm);
        d_apos = dupload(s.a_pos_of_entry);
        d_fpos = dupload(s.f_pos);
        d_Jp = dupload(grid ? nr.Jp : raw_p);
        d_Ji = dupload(grid ? nr.Ji : raw_i);
        if (split()) {
            d_prog_fs = dupload(padded(s.prog_fs, prog_fs_stride));
            d_prog_bs = dupload(padded(s.prog_bs, prog_bs_stride));
            d_split_seg = dupload(seg_words(s.split_seg));
            d_split_bseg = dupload(seg_words(s.split_bseg));
            if (grid) {
                d_k1_bus_s = dupload(nr.k1_bus_s);
                d_k1_ent_s = dupload(nr.k1_ent_s);
            }
        }
        if (grid) {
            d_k1_bus = dupload(nr.k1_bus);
            d_k1_ent = dupload(nr.k1_ent);
            std::vector<double2> y(static_cast<size_t>(nr.nY));
            for (int k = 0; k < nr.nY; ++k) y[k] = make_double2(nr.Yx[2 * k], nr.Yx[2 * k + 1]);
            d_Y = dupload(y);
            std::vector<int> upd;
            for (int i = 0; i < nr.n; ++i) {
                if (nr.col_th[i] < 0) continue;  // no angle unknown: a fixed bus (or the slack's P row)
                upd.insert(upd.end(), {i, nr.col_th[i], nr.row_q[i]});
            }
            n_upd = static_cast<int>(upd.size() / 3);
            d_upd = dupload(upd);
            std::vector<int> colmap(2 * static_cast<size_t>(nr.n));
            for (int i = 0; i < nr.n; ++i) {
                colmap[2 * i] = nr.col_th[i];
                colmap[2 * i + 1] = nr.row_q[i];
            }
            d_colmap = dupload(colmap);
            std::vector<int> entry_row(static_cast<size_t>(nr.nY));
            for (int i = 0; i < nr.n; ++i)
                for (int q = nr.Yp[i]; q < nr.Yp[i + 1]; ++q) entry_row[q] = i;
            d_entry_row = dupload(entry_row);
            if (nr.xt.on()) {
                const ExtTables& x = nr.xt;
                dx.inj_ptr = dupload(x.inj_ptr);
                dx.inj = dupload(x.inj);
                dx.jrec = dupload(x.jrec);
                dx.frec = dupload(x.frec);
                dx.xcol = dupload(x.xcol);
                dx.h_bus1 = dupload(x.h_bus1);
                dx.h_bus2 = dupload(x.h_bus2);
                dx.h_status = dupload(x.h_status);
                dx.vc_reg = dupload(x.vc_reg);
                dx.vc_ptr = dupload(x.vc_ptr);
                dx.sh_c1 = dupload(x.sh_c1);
                dx.sh_ck = dupload(x.sh_ck);
                dx.ds_w = dupload(x.ds_w);
                dx.jconst = dupload(x.jconst);
                dx.h_par = dupload(x.h_par);
                dx.vc_vset = dupload(x.vc_vset);
                dx.vc_slope = dupload(x.vc_slope);
                dx.vc_w = dupload(x.vc_w);
            }
        }
        uploaded = true;
    }
};

// staging buffers of the streamed-row calls
enum StageBuf { kStSlots, kStS, kStV, kStEdPtr, kStEdEnt, kStEdDelta, kStMask, kStSg, kStVg, kStageBufs };

struct g2g_batch_s {
    g2g_plan_s* plan = nullptr;
    int rows = 0, rows_pad = 0, T = 0, precision = G2G_FP64, per_row_y = 0, active_rows = 0;
    hipStream_t stream = nullptr;
    double2* V = nullptr;
    double2* S = nullptr;
    // outage batches: each row's Ybus edits (RowEdits), capacity ed_cap per row
    int ed_cap = 0;
    int* ed_n = nullptr;
    int* ed_ent = nullptr;
    double2* ed_val = nullptr;
    int* ed_flag = nullptr;  // [T][n]
    void* A = nullptr;
    // mixed precision with FP64-residual refinement: the factorization's L
    // multipliers and the refinement's right-hand sides, in the A stream's
    // layout (allocated on first use)
    void* Lm = nullptr;
    void* U = nullptr;
    void* X = nullptr;         // solved Newton step [T][d][64]
    int a_stride = 0, u_stride = 0;
    double* res = nullptr;
    double* jmax = nullptr;
    // workgroups per tile of the mismatch and update kernels on a full
    // batch, and the mismatch's per-split residual partials
    // [2][splits][tiles x 64], room for part_cap (split, tile) pairs (a
    // launch over fewer tiles may split each into more workgroups)
    int mis_splits = 1;
    int part_cap = 1;
    double* mis_part = nullptr;
    int* status = nullptr;
    int* iters = nullptr;
    int* tile_active = nullptr;
    int* n_active = nullptr;  // rows and tiles still iterating (2 ints)
    int* clamped = nullptr;
    int* error = nullptr;     // stream kernels' give-up flag
    int* h_active = nullptr;  // pinned: rows and tiles still iterating, rows restarted
    // warm solves: the caller's start voltages (the fallback start)
    double2* V0 = nullptr;
    bool v0_row = false;  // V0 holds one row [n] (every row's fallback start is the same)
    // V as last set: every row the same vector, kept in v_row [n] (the
    // calls that change V clear the flag)
    bool v_bcast = false;
    double2* v_row = nullptr;
    int* tile_restart = nullptr;  // [T] tiles with a row restarted at this iteration
    int* differs = nullptr;
    // step control (settings of the next solve)
    int ir_steps = 0;                 // refinement steps of every Newton step
    double max_dva = 0.0, max_dvm = 0.0;  // step scaling bounds (max_dva <= 0: off)
    void* Bs = nullptr;               // refinement: saved rhs and step [T][d][64]
    void* Xs = nullptr;
    // tail compaction: the full batch's V, Sbus, per-row Ybus, statuses,
    // iteration counts and residuals while the few rows left iterate in
    // dense tiles, and the row map (compacted row -> batch row)
    double2* Vk = nullptr;
    double2* Sk = nullptr;
    int* edk_n = nullptr;  // the edits' compaction copies
    int* edk_ent = nullptr;
    double2* edk_val = nullptr;
    int* edk_flag = nullptr;
    int* stk = nullptr;
    int* itk = nullptr;
    double* resk = nullptr;
    int* mapk = nullptr;
    int compact_R = 0, compact_saved = 0;
    // streamed rows (continuous batching): per-row start iteration and the
    // streamed rounds: the iteration clock on the device (the kernels read
    // it, so a round is the same launches every time), and the round
    // captured as a HIP graph (per round shape), replayed
    int* clock_dev = nullptr;
    bool stream_graph = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    // where the graph is captured and replayed: the batch's stream, or for a
    // batch on the null stream (which cannot be captured) a stream of its
    // own, ordered after the batch's earlier work and before its later work
    // by two events
    hipStream_t graph_stream = nullptr;
    bool own_graph_stream = false;
    hipEvent_t graph_ev[2] = {nullptr, nullptr};
    long long graph_key[6] = {0, 0, 0, 0, 0, 0};
    long long graph_replays = 0;
    // batch's iteration clock
    int* it0 = nullptr;
    int clock = 0;
    bool streaming = false;
    // learned starts with in-place fallback: per-row update budgets, the rows
    // still on their learned start, the restart count, and compaction copies
    bool restart_on = false;
    int restart_max_iter = 0;
    double contraction = 0.0;  // the quadratic-contraction rejection of learned starts (0: off)
    bool ir_fp64 = false;      // refinement residuals in FP64 from V (else the solve's precision)
    int n_cu = 256;            // the device's compute units
    // the warm start's resident outage mask [rows][wmask_n] (nullable)
    unsigned char* wmask = nullptr;
    int wmask_n = 0;
    int* budget = nullptr;
    unsigned char* warm = nullptr;
    int* requeued = nullptr;
    double2* V0k = nullptr;
    int *it0k = nullptr, *budgetk = nullptr;
    unsigned char* warmk = nullptr;
    // split walk: the pending buffer [T][pend_len][64], whether the A / U
    // streams currently hold the split layout, and the tiles a launch covers
    void* pend = nullptr;
    bool split_layout = false;
    int split_T = 0;
    // extension rows: per-row state [T][n_xs][64] (sa, Q of each controller),
    // its compaction copy, and the extension buses' residual terms [T][n_xb][64]
    double* xs = nullptr;
    double* xsk = nullptr;
    double2* sx = nullptr;
    // islanded buses per row [T][n][64] (nullable), and its compaction copy
    unsigned char* mask = nullptr;
    unsigned char* maskk = nullptr;
    // grow-only device staging of the streamed-row calls (kept between calls,
    // so a refill or a take allocates and frees nothing once warm)
    void* stage[kStageBufs] = {};
    size_t stage_cap[kStageBufs] = {};
    std::vector<hipEvent_t> ev;
    ~g2g_batch_s()
    {
        dfree(V); dfree(S); dfree(ed_n); dfree(ed_ent); dfree(ed_val); dfree(ed_flag); dfree(A); dfree(U); dfree(X); dfree(res); dfree(jmax); dfree(mis_part);
        dfree(status); dfree(iters); dfree(tile_active); dfree(n_active); dfree(clamped); dfree(error);
        dfree(V0); dfree(Bs); dfree(Xs); dfree(Lm); dfree(wmask); dfree(differs); dfree(tile_restart); dfree(v_row);
        dfree(clock_dev);
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (own_graph_stream) (void)hipStreamDestroy(graph_stream);
        for (hipEvent_t e : graph_ev)
            if (e) (void)hipEventDestroy(e);
        dfree(Vk); dfree(Sk); dfree(edk_n); dfree(edk_ent); dfree(edk_val); dfree(edk_flag); dfree(stk); dfree(itk); dfree(resk); dfree(mapk); dfree(it0);
        dfree(mask); dfree(maskk); dfree(xs); dfree(xsk); dfree(sx); dfree(pend);
        dfree(budget); dfree(warm); dfree(requeued); dfree(V0k); dfree(it0k); dfree(budgetk); dfree(warmk);
        for (void*& q : stage) dfree(q);
        if (h_active) (void)hipHostFree(h_active);
        for (auto e : ev) (void)hipEventDestroy(e);
    }
    size_t real_size() const { return precision == G2G_MIXED ? sizeof(float) : sizeof(double); }
};

namespace {

template <typename Real>
constexpr int chunk_of() { return sizeof(Real) == sizeof(double) ? kChunkF64 : kChunkF32; }

inline int round_up(int v, int m) { return (v + m - 1) / m * m; }

// U positions one tile holds: the stream in whole chunks, reaching far enough
// that every producer, which streams its segment in whole chunks from the
// segment's own start (a backward set need not start on a chunk boundary),
// stays inside the tile's range; the last tile's reads then stay inside the
// buffer
int u_stride_of(const SymbolicLU& s, int real_size)
{
    const int ch = real_size == 8 ? kChunkF64 : kChunkF32;
    int end = s.u_len;
    for (const auto& g : s.seg) end = std::max(end, g.u_off + round_up(g.u_len, value_chunk(real_size, s.walkers)));
    for (const auto& g : s.bseg) end = std::max(end, g.u_off + round_up(g.u_len, value_chunk(real_size, s.bwalkers)));
    for (const auto& g : s.split_seg) end = std::max(end, g.u_off + round_up(g.u_len, value_chunk(real_size, kSplitRingF)));
    for (const auto& g : s.split_bseg) end = std::max(end, g.u_off + round_up(g.u_len, value_chunk(real_size, kSplitRingB)));
    return round_up(end, ch);
}

// A positions one tile holds: the main stream, and the split walk's, each
// read in whole chunks from every segment's start
int a_stride_of(const SymbolicLU& s, int real_size)
{
    const int ch = real_size == 8 ? kChunkF64 : kChunkF32;
    int end = s.a_len;
    for (const auto& g : s.seg) end = std::max(end, g.a_off + round_up(g.a_len, value_chunk(real_size, s.walkers)));
    for (const auto& g : s.split_seg) end = std::max(end, g.a_off + round_up(g.a_len, value_chunk(real_size, kSplitRingF)));
    return round_up(end, ch);
}

template <typename Real>
constexpr int ring_slots_of() { return sizeof(Real) == sizeof(double) ? kRingSlotsF64 : kRingSlotsF32; }

// walker control areas + flags + value and program rings, then the slots
// (plus the zero and trash slots) and each walker's long-column scratch rows
// (kernels.hip's LDS layout)
template <typename Real>
size_t factor_lds(const SymbolicLU& s, int rg = 0)
{
    const int scratch = s.max_c > kFastC ? 2 * s.max_c * s.walkers : 0;
    return static_cast<size_t>(walk_ring_bytes(sizeof(Real), s.walkers, rg > 0 ? rg : s.walkers)) +
           static_cast<size_t>(s.slots_f + 2 + scratch) * kLanes * sizeof(Real);
}

// The factorization's ring geometry: its own walkers' (0), or, for two FP32
// walkers whose slots do not fit beside full rings, half-size chunks
// (G2G_HALF_RINGS=1 takes them whenever they fit).
template <typename Real>
int factor_ring_geometry(const SymbolicLU& s, size_t limit)
{
    if (sizeof(Real) != 4 || s.walkers != 2) return 0;
    static const int force = [] {
        const char* e = std::getenv("G2G_HALF_RINGS");
        return e ? std::atoi(e) : -1;
    }();
    if (force == 0) return 0;
    const bool half_fits = factor_lds<Real>(s, kHalfRingGeometry) <= limit;
    if (force == 1) return half_fits ? kHalfRingGeometry : 0;
    return factor_lds<Real>(s) > limit && half_fits ? kHalfRingGeometry : 0;
}

template <typename Real>
size_t factor_lds_fit(const SymbolicLU& s, size_t limit)
{
    return factor_lds<Real>(s, factor_ring_geometry<Real>(s, limit));
}

template <typename Real>
size_t backward_lds(const SymbolicLU& s)
{
    return static_cast<size_t>(walk_ring_bytes(sizeof(Real), s.bwalkers)) +
           static_cast<size_t>(s.slots_b + 2) * kLanes * sizeof(Real);
}

size_t lds_limit()
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    int v = 0;
    HIPCHK(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    return static_cast<size_t>(v);
}

// The plan plan_fitting(P, fp64) returns, if it has been built already
// (nullptr otherwise): a lookup that builds and allocates nothing.
const g2g_plan_s* plan_fitted(const g2g_plan_s* P, bool fp64)
{
    const size_t limit = lds_limit();
    auto fits = [&](const SymbolicLU& s) {
        const size_t f = fp64 ? factor_lds_fit<double>(s, limit) : factor_lds_fit<float>(s, limit);
        const size_t b = fp64 ? backward_lds<double>(s) : backward_lds<float>(s);
        return f <= limit && b <= limit;
    };
    if (fits(P->lu())) return P;
    const int wb = P->lu().bwalkers;
    for (const auto& [wf, wbk] : {std::pair<int, int>{1, wb}, std::pair<int, int>{1, 1}}) {
        const g2g_plan_s* q = nullptr;
        for (const auto& c : P->fitted)
            if (c->lu().walkers == wf && c->lu().bwalkers == wbk) q = c.get();
        if (!q) return nullptr;
        if (fits(q->lu())) return q;
    }
    return P->fitted.empty() ? nullptr : P->fitted.back().get();
}

// The plan whose walker kernels fit the LDS at this precision: P itself, or
// its one-walker variant (built once, owned by P).
g2g_plan_s* plan_fitting(g2g_plan_s* P, bool fp64)
{
    const size_t limit = lds_limit();
    auto fits = [&](const SymbolicLU& s) {
        const size_t f = fp64 ? factor_lds_fit<double>(s, limit) : factor_lds_fit<float>(s, limit);
        const size_t b = fp64 ? backward_lds<double>(s) : backward_lds<float>(s);
        return f <= limit && b <= limit;
    };
    if (fits(P->lu())) return P;
    // the same pattern with fewer walkers: one factorization walker (the
    // walker split duplicates the slot working set), then one for both
    const int wb = P->lu().bwalkers;
    for (const auto& [wf, wbk] : {std::pair<int, int>{1, wb}, std::pair<int, int>{1, 1}}) {
        g2g_plan_s* q = nullptr;
        for (const auto& c : P->fitted)
            if (c->lu().walkers == wf && c->lu().bwalkers == wbk) q = c.get();
        if (!q) {
            auto c = std::make_unique<g2g_plan_s>();
            c->grid = P->grid;
            c->nnz = P->nnz;
            if (P->grid) {
                const NrPlan& nr = P->nr;
                std::vector<int> pv;
                std::vector<char> is_pq(static_cast<size_t>(nr.n), 0);
                for (int b : nr.pq) is_pq[b] = 1;
                for (int b : nr.pvpq)
                    if (!is_pq[b]) pv.push_back(b);
                c->nr = build_nr_plan(nr.n, nr.Yp.data(), nr.Yi.data(), nr.Yx.empty() ? nullptr : nr.Yx.data(),
                                      static_cast<int>(pv.size()), pv.data(), nr.npq, nr.pq.data(), nr.slack, wf,
                                      wbk, &nr.ext);
            } else {
                c->raw = build_symbolic(P->raw.d, P->raw_p.data(), P->raw_i.data(), P->raw_order, wf, wbk);
                c->raw_p = P->raw_p;
                c->raw_i = P->raw_i;
                c->raw_order = P->raw_order;
            }
            P->fitted.push_back(std::move(c));
            q = P->fitted.back().get();
        }
        if (fits(q->lu())) return q;
    }
    return P->fitted.back().get();  // one walker each: the launch reports what does not fit
}

template <typename Real, int MODE>
void launch_factor_mode(const SymbolicLU& s, int rg, size_t lds, const FactorArgs<Real>& a, int T, hipStream_t st)
{
    const void* fn = nullptr;
    if constexpr (sizeof(Real) == 4) {
        if (rg == kHalfRingGeometry)
            fn = reinterpret_cast<const void*>(&k_factor<Real, 2, false, kHalfRingGeometry, MODE>);
    }
    if (!fn)
        fn = s.walkers == 2 ? reinterpret_cast<const void*>(&k_factor<Real, 2, false, 0, MODE>)
                            : reinterpret_cast<const void*>(&k_factor<Real, 1, false, 0, MODE>);
    HIPCHK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    if constexpr (sizeof(Real) == 4) {
        if (rg == kHalfRingGeometry) {
            k_factor<Real, 2, false, kHalfRingGeometry, MODE><<<T, kLanes * factor_walk_waves(2, MODE), lds, st>>>(a);
            HIPCHK(hipGetLastError());
            return;
        }
    }
    if (s.walkers == 2) k_factor<Real, 2, false, 0, MODE><<<T, kLanes * factor_walk_waves(2, MODE), lds, st>>>(a);
    else k_factor<Real, 1, false, 0, MODE><<<T, kLanes * factor_walk_waves(1, MODE), lds, st>>>(a);
    HIPCHK(hipGetLastError());
}

// The main-layout walk over the A stream: the factorization (mode
// kFactorFull), the factorization storing its L multipliers in the A stream
// (kFactorStoreL), or the forward substitution alone over such a stream
// (kForwardOnly: y into the U stream, nothing else written).  The stored
// modes are FP32 only.
template <typename Real>
void launch_factor(const g2g_plan_s* P, const Real* A, Real* U, int a_stride, int u_stride, const int* tile_active,
                   const double* jmax, int* clamped, int* error, int rows, int T, double eps, hipStream_t st,
                   int mode = kFactorFull, Real* Lw = nullptr)
{
    const SymbolicLU& s = P->lu();
    const size_t limit = lds_limit();
    const int rg = factor_ring_geometry<Real>(s, limit);
    const size_t lds = factor_lds<Real>(s, rg);
    if (lds > limit) throw std::runtime_error("factorization needs " + std::to_string(lds) + " bytes of LDS");
    FactorArgs<Real> a{A,     U,    P->d_prog_f, tile_active, jmax,    clamped, error,  s.d, a_stride, u_stride,
                       P->prog_f_stride, s.slots_f, rows, s.max_c, eps, P->d_seg, T, 0, nullptr, 0};
    if constexpr (sizeof(Real) == 4) {
        if (mode == kFactorStoreL) {
            if (!Lw) throw std::logic_error("a multiplier-storing factorization needs its stream");
            a.Lw = Lw;
            return launch_factor_mode<Real, kFactorStoreL>(s, rg, lds, a, T, st);
        }
        if (mode == kForwardOnly) {
            a.clamped = nullptr;  // no pivots: the factorization's counts stay
            return launch_factor_mode<Real, kForwardOnly>(s, rg, lds, a, T, st);
        }
    } else {
        if (mode != kFactorFull) throw std::logic_error("stored-multiplier walks are FP32 only");
    }
    launch_factor_mode<Real, kFactorFull>(s, rg, lds, a, T, st);
}

template <typename Real>
void launch_backward(const g2g_plan_s* P, const Real* U, int u_stride, const int* tile_active, Real* X, int* error,
                     int rows, int T, hipStream_t st)
{
    const SymbolicLU& s = P->lu();
    const size_t lds = backward_lds<Real>(s);
    if (lds > lds_limit()) throw std::runtime_error("backward substitution needs too much LDS");
    BackwardArgs<Real> a{U, P->d_prog_b, tile_active, X, error, s.d, s.u_len, u_stride, P->prog_b_stride, rows,
                         s.max_c, P->d_bseg, s.n_top, T, 0};
    const void* fn = s.bwalkers == 4   ? reinterpret_cast<const void*>(&k_backward<Real, 4, false>)
                     : s.bwalkers == 2 ? reinterpret_cast<const void*>(&k_backward<Real, 2, false>)
                                       : reinterpret_cast<const void*>(&k_backward<Real, 1, false>);
    HIPCHK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    if (s.bwalkers == 4) k_backward<Real, 4, false><<<T, kLanes * walk_waves(4), lds, st>>>(a);
    else if (s.bwalkers == 2) k_backward<Real, 2, false><<<T, kLanes * walk_waves(2), lds, st>>>(a);
    else k_backward<Real, 1, false><<<T, kLanes * walk_waves(1), lds, st>>>(a);
    HIPCHK(hipGetLastError());
}

// LDS of a split-walk workgroup: its rings, then its slots (+ zero and trash)
// and long-column scratch rows
template <typename Real>
size_t split_factor_lds(const SymbolicLU& s)
{
    const int scratch = s.max_c > kFastC ? 2 * s.max_c : 0;
    return static_cast<size_t>(split_ring_bytes(sizeof(Real), kSplitRingF)) +
           static_cast<size_t>(s.slots_fs + 2 + scratch) * kLanes * sizeof(Real);
}

template <typename Real>
size_t split_backward_lds(const SymbolicLU& s)
{
    return static_cast<size_t>(split_ring_bytes(sizeof(Real), kSplitRingB)) +
           static_cast<size_t>(s.slots_bs + 2) * kLanes * sizeof(Real);
}

// The split walk's factorization: the fine sets, then the middle parts, then
// the top part, one launch each over the first T tiles.  clamped counts are
// added per row, so they start at zero.
template <typename Real>
void launch_factor_split(const g2g_plan_s* P, const Real* A, Real* U, int a_stride, int u_stride,
                         const int* tile_active, const double* jmax, int* clamped, int* error, int rows, int T,
                         double eps, Real* pend, hipStream_t st)
{
    const SymbolicLU& s = P->lu();
    const size_t lds = split_factor_lds<Real>(s);
    if (lds > lds_limit()) throw std::runtime_error("split factorization needs " + std::to_string(lds) + " bytes of LDS");
    if (clamped) {
        k_fill_int<<<blocks_for(static_cast<long long>(T) * kLanes, 256), 256, 0, st>>>(clamped, 0,
                                                                                       static_cast<long long>(T) * kLanes);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_factor<Real, 1, true, 0, kFactorFull>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    const int levels[3][2] = {{0, s.split_fine}, {s.split_fine, s.split_mid}, {s.split_fine + s.split_mid, 1}};
    for (const auto& lv : levels) {
        if (lv[1] <= 0) continue;
        FactorArgs<Real> a{A,   U,        P->d_prog_fs, tile_active, jmax, clamped, error, s.d, a_stride,
                           u_stride, P->prog_fs_stride, s.slots_fs, rows, s.max_c, eps, P->d_split_seg,
                           T,   lv[0],    pend,         s.pend_len};
        k_factor<Real, 1, true, 0, kFactorFull><<<T * lv[1], kLanes * (2 + kSplitProducers), lds, st>>>(a);
        HIPCHK(hipGetLastError());
    }
}

// The split walk's backward substitution: the top part, then the middle
// parts, then the fine sets.
template <typename Real>
void launch_backward_split(const g2g_plan_s* P, const Real* U, int u_stride, const int* tile_active, Real* X,
                           int* error, int rows, int T, hipStream_t st)
{
    const SymbolicLU& s = P->lu();
    const size_t lds = split_backward_lds<Real>(s);
    if (lds > lds_limit()) throw std::runtime_error("split backward substitution needs too much LDS");
    HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_backward<Real, 1, true>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    const int levels[3][2] = {{s.split_fine + s.split_mid, 1}, {s.split_fine, s.split_mid}, {0, s.split_fine}};
    for (const auto& lv : levels) {
        if (lv[1] <= 0) continue;
        BackwardArgs<Real> a{U, P->d_prog_bs, tile_active, X, error, s.d, s.split_u_len, u_stride,
                             P->prog_bs_stride, rows, s.max_c, P->d_split_bseg, 0, T, lv[0]};
        k_backward<Real, 1, true><<<T * lv[1], kLanes * (2 + kSplitProducers), lds, st>>>(a);
        HIPCHK(hipGetLastError());
    }
}

// `steps` rounds of x += (LU)^-1 (b - A x) on the tiles' A streams: the
// residual overwrites the A stream's right-hand side, and the factors are
// recomputed from the unchanged matrix values (bitwise the same factors).
template <typename Real>
void refine(const g2g_plan_s* P, Real* A, Real* U, Real* X, Real* Bs, Real* Xs, int a_stride, int u_stride,
            const int* tile_active, const double* jmax, int* clamped, int* error, int rows, int T, double eps,
            int steps, hipStream_t st)
{
    const SymbolicLU& s = P->lu();
    for (int k = 0; k < steps; ++k) {
        RefineArgs<Real> ra{A, X, Bs, Xs, P->d_Jp, P->d_Ji, P->d_apos, P->d_fpos, tile_active, s.d, a_stride,
                            k == 0 ? 1 : 0};
        k_lin_residual<Real><<<T, kLanes * kMisWaves, 0, st>>>(ra);
        HIPCHK(hipGetLastError());
        launch_factor<Real>(P, A, U, a_stride, u_stride, tile_active, jmax, clamped, error, rows, T, eps, st);
        launch_backward<Real>(P, U, u_stride, tile_active, X, error, rows, T, st);
        const int per_tile = s.d * kLanes;
        k_axpy_tiles<Real><<<dim3(std::min(64, (per_tile + 255) / 256), T), 256, 0, st>>>(Xs, X, tile_active, s.d);
        HIPCHK(hipGetLastError());
    }
}

// The batch's row edits as the kernels take them (all null for a batch
// sharing one Ybus).
RowEdits row_edits(const g2g_batch_s* b)
{
    return RowEdits{b->ed_n, b->ed_ent, b->ed_val, b->ed_flag, b->ed_cap};
}

// flags of the Ybus rows that rows [0, rows) edit, afresh
void edit_flags(g2g_batch_s* b, const int* n_ed, const int* ent, int rows, int* flag)
{
    const NrPlan& nr = b->plan->nr;
    const long long nf = static_cast<long long>(b->T) * nr.n;
    k_fill_int<<<blocks_for(nf, 256), 256, 0, b->stream>>>(flag, 0, nf);
    if (rows > 0)
        k_edit_flags<<<blocks_for(rows, 256), 256, 0, b->stream>>>(n_ed, ent, b->ed_cap, rows, b->plan->d_entry_row,
                                                                   nr.n, flag);
    HIPCHK(hipGetLastError());
}

// CSR edit lists with each row's repeated entries merged (their deltas
// summed), in place of (row_ptr, ent, delta of interleaved pairs)
struct EditLists {
    std::vector<int> rp, ent;
    std::vector<double2> val;
    int most = 0;
    EditLists(const int32_t* row_ptr, const int32_t* entry, const double* delta, int rows)
    {
        rp.assign(1, 0);
        for (int r = 0; r < rows; ++r) {
            const size_t start = ent.size();
            for (int q = row_ptr[r]; q < row_ptr[r + 1]; ++q) {
                const double2 d = make_double2(delta[2 * q], delta[2 * q + 1]);
                size_t k = start;
                while (k < ent.size() && ent[k] != entry[q]) ++k;
                if (k < ent.size()) {
                    val[k].x += d.x;
                    val[k].y += d.y;
                } else {
                    ent.push_back(entry[q]);
                    val.push_back(d);
                }
            }
            most = std::max(most, static_cast<int>(ent.size() - start));
            rp.push_back(static_cast<int>(ent.size()));
        }
    }
};

// room for cap edits per row (the lists kept when keep is set)
void ensure_edit_cap(g2g_batch_s* b, int cap, bool keep)
{
    if (cap <= b->ed_cap) return;
    cap = (cap + 7) / 8 * 8;
    const size_t n = static_cast<size_t>(b->rows_pad) * cap;
    int* ent = dalloc<int>(n);
    double2* val = dalloc<double2>(n);
    if (keep && b->ed_cap > 0) {
        const long long m = static_cast<long long>(b->T) * b->ed_cap * kLanes;
        k_regrow_edits<int><<<blocks_for(m, 256), 256, 0, b->stream>>>(b->ed_ent, b->ed_cap, ent, cap, b->T);
        k_regrow_edits<double2><<<blocks_for(m, 256), 256, 0, b->stream>>>(b->ed_val, b->ed_cap, val, cap, b->T);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(b->stream));
    }
    DeviceScratch::free_one(b->ed_ent);
    DeviceScratch::free_one(b->ed_val);
    DeviceScratch::free_one(b->edk_ent);  // compaction copies follow the capacity
    DeviceScratch::free_one(b->edk_val);
    b->ed_ent = ent;
    b->ed_val = val;
    b->ed_cap = cap;
}

// The walk mode of a Newton step's main-layout factorization.  With
// FP64-residual refinement in mixed precision it stores its L multipliers in
// the A stream, so each refinement round solves with a forward-only walk
// instead of a second factorization (G2G_IR_REFACTOR=1: the second
// factorization, for comparison).
template <typename Real>
int factor_walk_mode(const g2g_batch_s* b)
{
    if (sizeof(Real) != 4 || b->ir_steps <= 0 || !b->ir_fp64) return kFactorFull;
    const char* e = std::getenv("G2G_IR_REFACTOR");
    return e && std::atoi(e) == 1 ? kFactorFull : kFactorStoreL;
}

// the multipliers' stream of kFactorStoreL (zeroed once: positions no walk
// writes are read and dropped by the forward walk)
template <typename Real>
Real* multiplier_stream(g2g_batch_s* b)
{
    if (factor_walk_mode<Real>(b) != kFactorStoreL) return nullptr;
    if (!b->Lm) {
        const size_t bytes = static_cast<size_t>(b->rows_pad) * b->a_stride * sizeof(Real);
        b->Lm = dalloc<unsigned char>(bytes);
        HIPCHK(hipMemsetAsync(b->Lm, 0, bytes, b->stream));
    }
    return static_cast<Real*>(b->Lm);
}

// Refinement rounds of a Newton step with the residual in FP64 from V
// (k_ir_residual), against the solve's factors: r = F(V) - J(V) x over the
// A stream's rhs, its forward substitution (a forward-only walk over the
// stored multipliers, or the factorization again: bitwise the same factors),
// the backward substitution, x += the correction.
template <typename Real>
void refine_fp64(g2g_batch_s* b, double eps, int T)
{
    const g2g_plan_s* P = b->plan;
    const NrPlan& nr = P->nr;
    Real* A = static_cast<Real*>(b->A);
    Real* X = static_cast<Real*>(b->X);
    Real* Xs = static_cast<Real*>(b->Xs);
    const bool sp = b->split_layout;  // the tail's split walk: its layout's tables and launches
    // workgroups per tile: the mismatch's setting, doubled (up to 32) while
    // the launch stays within one round of resident workgroups
    int splits = std::max(1, b->mis_splits);
    while (splits < 32 && 2 * splits * T <= 512) splits *= 2;
    // the forward walk's stream: the rhs goes into the multipliers' stream
    Real* L = sp ? nullptr : (factor_walk_mode<Real>(b) == kFactorStoreL ? static_cast<Real*>(b->Lm) : nullptr);
    for (int k = 0; k < b->ir_steps; ++k) {
        IrArgs<Real> a{sp ? P->d_k1_bus_s : P->d_k1_bus, sp ? P->d_k1_ent_s : P->d_k1_ent, nr.k1_nbus, P->d_Y,
                       row_edits(b), b->V, b->S, b->mask, P->d_colmap, L ? L : A, X, Xs, b->tile_active, nr.n, nr.nY,
                       nr.d, b->a_stride, splits};
        if (b->per_row_y) k_ir_residual<Real, true><<<T * splits, kLanes * kMismatchWaves, 0, b->stream>>>(a);
        else k_ir_residual<Real, false><<<T * splits, kLanes * kMismatchWaves, 0, b->stream>>>(a);
        HIPCHK(hipGetLastError());
        if (sp) {
            launch_factor_split<Real>(P, A, static_cast<Real*>(b->U), b->a_stride, b->u_stride, b->tile_active,
                                      b->jmax, b->clamped, b->error, b->active_rows, T, eps,
                                      static_cast<Real*>(b->pend), b->stream);
            launch_backward_split<Real>(P, static_cast<const Real*>(b->U), b->u_stride, b->tile_active, X,
                                        b->error, b->active_rows, T, b->stream);
        } else {
            launch_factor<Real>(P, L ? L : A, static_cast<Real*>(b->U), b->a_stride, b->u_stride, b->tile_active,
                                b->jmax, b->clamped, b->error, b->active_rows, T, eps, b->stream,
                                L ? kForwardOnly : kFactorFull);
            launch_backward<Real>(P, static_cast<const Real*>(b->U), b->u_stride, b->tile_active, X, b->error,
                                  b->active_rows, T, b->stream);
        }
        if (k + 1 < b->ir_steps) {  // the last correction is added by the update
            const int per_tile = nr.d * kLanes;
            k_axpy_tiles<Real><<<dim3(std::min(64, (per_tile + 255) / 256), T), 256, 0, b->stream>>>(
                Xs, X, b->tile_active, nr.d);
            HIPCHK(hipGetLastError());
        }
    }
}

// A Newton step's refinement: FP64 residual or the solve's own (the
// reference's, BatchedLu::refine, batched_lu.cu:391-411).
template <typename Real>
void refine_step(g2g_batch_s* b, double eps, int T)
{
    if (b->ir_fp64) {
        refine_fp64<Real>(b, eps, T);
        return;
    }
    refine<Real>(b->plan, static_cast<Real*>(b->A), static_cast<Real*>(b->U), static_cast<Real*>(b->X),
                 static_cast<Real*>(b->Bs), static_cast<Real*>(b->Xs), b->a_stride, b->u_stride, b->tile_active,
                 b->jmax, b->clamped, b->error, b->active_rows, T, eps, b->ir_steps, b->stream);
}

// Workgroups per tile of the mismatch and update kernels (G2G_MIS_SPLITS:
// 1, 2 or 4; default 4).  A tile's kernels are latency-bound chains over its
// buses, so splitting the buses over more workgroups shortens a launch on a
// full batch (two workgroups resident per CU) and on a batch's last few tiles.
int default_mis_splits()
{
    if (const char* e = std::getenv("G2G_MIS_SPLITS")) {
        const int v = std::atoi(e);
        if (v == 1 || v == 2 || v == 4) return v;
    }
    return 4;
}

constexpr int kRestartChunks = 16;  // workgroups per tile of a restart copy


// The tiles that can hold iterating rows: the compacted rows' tiles, or all.
int live_tiles(const g2g_batch_s* b) { return b->compact_R > 0 ? (b->compact_R + kLanes - 1) / kLanes : b->T; }

// Workgroups per tile of a mismatch or update launch over `tiles` tiles: the
// batch's setting, doubled (up to 32) while the launch stays within one
// round of resident workgroups and the partials within their buffer.  A
// tile's buses form a latency-bound chain per wave, so a launch over a few
// tiles runs as many times faster as it has workgroups per tile.
int launch_splits(const g2g_batch_s* b, int tiles)
{
    constexpr int kMaxSplits = 32, kResidentWgs = 512;
    int s = b->mis_splits;
    while (s < kMaxSplits && 2 * s * tiles <= kResidentWgs && 2 * s * tiles <= b->part_cap) s *= 2;
    return s;
}

template <typename Real>
void launch_update(g2g_batch_s* b)
{
    const g2g_plan_s* P = b->plan;
    const int T = live_tiles(b);
    // step scaling: the row's maxima in one workgroup
    const int splits = b->max_dva > 0.0 ? 1 : launch_splits(b, T);
    // FP64-residual refinement leaves its last correction in X and the step
    // before it in Xs: the update adds them
    const Real* add = b->ir_steps > 0 && b->ir_fp64 ? static_cast<const Real*>(b->Xs) : nullptr;
    UpdateArgs<Real> a{static_cast<const Real*>(b->X), P->d_upd, P->n_upd, b->tile_active, b->status, b->V,
                       P->nr.d, P->nr.n, b->active_rows, b->max_dva, b->max_dvm, splits, add};
    k_update<Real><<<T * splits, kLanes * kMisWaves, 0, b->stream>>>(a);
    HIPCHK(hipGetLastError());
}

template <typename Real>
ExtArgs<Real> ext_args(g2g_batch_s* b)
{
    const g2g_plan_s* P = b->plan;
    const ExtTables& x = P->nr.xt;
    ExtArgs<Real> a{};
    a.V = b->V;
    a.S = b->S;
    a.xs = b->xs;
    a.sx = b->sx;
    a.A = static_cast<Real*>(b->A);
    a.X = static_cast<const Real*>(b->X);
    a.tile_active = b->tile_active;
    a.status = b->status;
    a.rows = b->active_rows;
    a.n = P->nr.n;
    a.d = P->nr.d;
    a.a_stride = b->a_stride;
    a.n_xs = x.n_xs;
    a.n_xb = static_cast<int>(x.xb_bus.size());
    a.sa = x.sa;
    a.q0 = x.q0;
    a.inj_ptr = P->dx.inj_ptr;
    a.inj = P->dx.inj;
    a.ds_w = P->dx.ds_w;
    a.n_jrec = static_cast<int>(x.jrec.size() / 4);
    a.jrec = P->dx.jrec;
    a.jconst = P->dx.jconst;
    a.n_frec = static_cast<int>(x.frec.size() / 3);
    a.frec = P->dx.frec;
    a.xcol = P->dx.xcol;
    a.h_bus1 = P->dx.h_bus1;
    a.h_bus2 = P->dx.h_bus2;
    a.h_status = P->dx.h_status;
    a.h_par = P->dx.h_par;
    a.vc_reg = P->dx.vc_reg;
    a.vc_ptr = P->dx.vc_ptr;
    a.vc_vset = P->dx.vc_vset;
    a.vc_slope = P->dx.vc_slope;
    a.vc_w = P->dx.vc_w;
    a.sh_c1 = P->dx.sh_c1;
    a.sh_ck = P->dx.sh_ck;
    a.part = b->mis_part;  // launch_mismatch sets the split count and the partials' stride
    a.part_stride = b->rows_pad;
    a.splits = 1;
    return a;
}

// the extension state's start: sa = sum Re(Sbus) per row, Q(c) = 0
template <typename Real>
void launch_ext_init(g2g_batch_s* b)
{
    if (!b->plan->ext()) return;
    k_ext_init<Real><<<b->T, kLanes, 0, b->stream>>>(ext_args<Real>(b));
    HIPCHK(hipGetLastError());
}

template <typename Real>
void launch_ext_update(g2g_batch_s* b)
{
    if (!b->plan->ext() || b->plan->nr.xt.n_xs == 0) return;
    k_ext_update<Real><<<live_tiles(b), kLanes, 0, b->stream>>>(ext_args<Real>(b));
    HIPCHK(hipGetLastError());
}

// Whether a mismatch launch groups each tile's split workgroups on one XCD
// (G2G_MIS_XCD=0: split-major instead).
bool mis_xcd_group()
{
    static const bool on = [] {
        const char* e = std::getenv("G2G_MIS_XCD");
        return !(e && std::atoi(e) == 0);
    }();
    return on;
}

// only (nullable): evaluate only the tiles with only[tile] > 0; their
// counts go to a spare counter (the caller keeps the full pass's).
// clock (nullable): the iteration is *clock + iter.
template <typename Real>
void launch_mismatch(g2g_batch_s* b, bool with_j, int iter, int max_iter, double tol, int freeze,
                     double growth = 0.0, bool transposed = false, const int* only = nullptr,
                     const int* clock = nullptr)
{
    const NrPlan& P = b->plan->nr;
    const bool ext = b->plan->ext();
    const int T = live_tiles(b);
    const int splits = launch_splits(b, T);
    const int part_stride = T * kLanes;
    if (ext) {
        k_ext_pre<Real><<<T, kLanes, 0, b->stream>>>(ext_args<Real>(b));
        HIPCHK(hipGetLastError());
    }
    const int* k1b = b->split_layout ? b->plan->d_k1_bus_s : b->plan->d_k1_bus;
    const int* k1e = b->split_layout ? b->plan->d_k1_ent_s : transposed ? b->plan->d_k1_ent_t : b->plan->d_k1_ent;
    MismatchArgs<Real> a{k1b, k1e, P.k1_nbus, b->plan->d_Y, row_edits(b), b->V, b->S,
                         static_cast<Real*>(b->A), P.n, P.nY, b->a_stride, b->res, b->jmax, b->status, b->iters,
                         b->tile_active, only ? b->n_active + 2 : b->n_active, b->active_rows, iter, max_iter, tol,
                         freeze, growth,
                         growth > 0.0 ? b->contraction : 0.0,
                         (b->streaming || b->restart_on) ? b->it0 : nullptr, b->mask,
                         b->restart_on ? b->budget : nullptr, b->restart_on ? b->warm : nullptr,
                         splits, (splits > 1 || ext) ? b->mis_part : nullptr, part_stride,
                         ext ? b->sx : nullptr, static_cast<int>(P.xt.xb_bus.size()), T,
                         splits > 1 && mis_xcd_group() ? 1 : 0, only, clock};
    const int grid_tiles = a.xcd_group ? (T + kXcds - 1) / kXcds * kXcds : T;
    const dim3 grid(grid_tiles * splits), block(kLanes * kMismatchWaves);
    if (b->per_row_y) {
        if (with_j) k_mismatch<Real, true, true><<<grid, block, 0, b->stream>>>(a);
        else k_mismatch<Real, true, false><<<grid, block, 0, b->stream>>>(a);
    } else {
        if (with_j) k_mismatch<Real, false, true><<<grid, block, 0, b->stream>>>(a);
        else k_mismatch<Real, false, false><<<grid, block, 0, b->stream>>>(a);
    }
    HIPCHK(hipGetLastError());
    if (ext) {
        ExtArgs<Real> ea = ext_args<Real>(b);
        if (!with_j) ea.n_jrec = 0;  // an F-only pass leaves the J values alone
        ea.splits = splits;
        ea.part_stride = part_stride;
        k_ext_post<Real><<<T, kLanes, 0, b->stream>>>(ea);
        HIPCHK(hipGetLastError());
    }
    if (a.part) {
        if (with_j) k_mis_status<Real, true><<<T, kLanes, 0, b->stream>>>(a);
        else k_mis_status<Real, false><<<T, kLanes, 0, b->stream>>>(a);
        HIPCHK(hipGetLastError());
    }
}

void reset_rows(g2g_batch_s* b)
{
    // status: active for [0, active_rows), not simulated beyond
    std::vector<int> st(static_cast<size_t>(b->rows_pad), kNotSimulated);
    std::fill(st.begin(), st.begin() + b->active_rows, kActive);
    HIPCHK(hipMemcpyAsync(b->status, st.data(), st.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
    k_fill_int<<<blocks_for(b->rows_pad, 256), 256, 0, b->stream>>>(b->iters, 0, b->rows_pad);
    k_count_tiles<<<b->T, kLanes, 0, b->stream>>>(b->status, b->rows_pad, b->tile_active);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(b->stream));  // st is a host temporary
}

// per-solve event bookkeeping (phase times summed over every launch)
struct SolveTimer {
    g2g_batch_s* b;
    int nev = 0;
    std::vector<std::pair<int, int>> ph[4];  // mismatch, factor, backward + update, warm start
    std::vector<char> factor_main;          // per factorization: a main-layout walk (not the split walk)
    long long rows_main = 0;                // rows of the main-layout factorizations
    explicit SolveTimer(g2g_batch_s* bb, int max_launches) : b(bb)
    {
        while (static_cast<int>(b->ev.size()) < 2 * max_launches + 4) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            b->ev.push_back(e);
        }
    }
    int mark()
    {
        if (nev >= static_cast<int>(b->ev.size())) throw std::runtime_error("solve event pool exhausted");
        HIPCHK(hipEventRecord(b->ev[nev], b->stream));
        return nev++;
    }
    double ms(int k, bool main_only = false) const
    {
        double t = 0.0;
        for (size_t q = 0; q < ph[k].size(); ++q) {
            if (main_only && !(q < factor_main.size() && factor_main[q])) continue;
            float m = 0.f;
            HIPCHK(hipEventElapsedTime(&m, b->ev[ph[k][q].first], b->ev[ph[k][q].second]));
            t += m;
        }
        return t;
    }
};

// Tail compaction.  The LU walk of a tile costs the same whatever number of
// its rows are active, but the mismatch and update kernels cost in
// proportion to the active tiles: once few rows are left, they are gathered
// into dense tiles (V, Sbus, per-row Ybus, iteration counts, residuals), the
// iterations continue there, and the results are scattered back.
void compact_rows(g2g_batch_s* b)
{
    const NrPlan& nr = b->plan->nr;
    const size_t rp = static_cast<size_t>(b->rows_pad);
    hipStream_t s = b->stream;
    if (!b->Vk) {
        b->Vk = dalloc<double2>(rp * nr.n);
        b->Sk = dalloc<double2>(rp * nr.n);
        b->stk = dalloc<int>(rp);
        b->itk = dalloc<int>(rp);
        b->resk = dalloc<double>(rp);
        b->mapk = dalloc<int>(rp);
        // lanes beyond the compacted rows stay finite (|V| = 1)
        HIPCHK(hipMemcpyAsync(b->Vk, b->V, rp * nr.n * sizeof(double2), hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemsetAsync(b->Sk, 0, rp * nr.n * sizeof(double2), s));
    }
    if (b->per_row_y && !b->edk_ent) {
        const size_t m = rp * b->ed_cap;
        if (!b->edk_n) b->edk_n = dalloc<int>(rp);
        if (!b->edk_flag) b->edk_flag = dalloc<int>(static_cast<size_t>(b->T) * nr.n);
        b->edk_ent = dalloc<int>(m);
        b->edk_val = dalloc<double2>(m);
    }
    const int rows = b->active_rows;
    std::vector<int> st(static_cast<size_t>(rows));
    std::vector<unsigned char> wm(b->restart_on ? static_cast<size_t>(rows) : 0);
    HIPCHK(hipMemcpyAsync(st.data(), b->status, st.size() * sizeof(int), hipMemcpyDeviceToHost, s));
    if (b->restart_on) HIPCHK(hipMemcpyAsync(wm.data(), b->warm, wm.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    std::vector<int> map;
    // the rows still iterating, and those about to restart from their fallback start
    for (int r = 0; r < rows; ++r)
        if (st[r] == kActive || (b->restart_on && st[r] == kRejected && wm[r])) map.push_back(r);
    const int R = static_cast<int>(map.size());
    HIPCHK(hipMemcpyAsync(b->mapk, map.data(), map.size() * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(b->stk, b->status, rp * sizeof(int), hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(b->itk, b->iters, rp * sizeof(int), hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(b->resk, b->res, rp * sizeof(double), hipMemcpyDeviceToDevice, s));
    const unsigned gv = blocks_for(static_cast<long long>(R) * nr.n, 256);
    k_move_rows<double2><<<gv, 256, 0, s>>>(b->V, b->Vk, b->mapk, R, nr.n, 1);
    k_move_rows<double2><<<gv, 256, 0, s>>>(b->S, b->Sk, b->mapk, R, nr.n, 1);
    if (b->per_row_y) {  // the rows' edits move with them (lanes beyond: none)
        k_fill_int<<<blocks_for(rp, 256), 256, 0, s>>>(b->edk_n, 0, rp);
        k_move_vals<int><<<blocks_for(R, 256), 256, 0, s>>>(b->ed_n, b->edk_n, b->mapk, R, 1, 0);
        const unsigned ge = blocks_for(static_cast<long long>(R) * b->ed_cap, 256);
        k_move_rows<int><<<ge, 256, 0, s>>>(b->ed_ent, b->edk_ent, b->mapk, R, b->ed_cap, 1);
        k_move_rows<double2><<<ge, 256, 0, s>>>(b->ed_val, b->edk_val, b->mapk, R, b->ed_cap, 1);
        HIPCHK(hipGetLastError());
        edit_flags(b, b->edk_n, b->edk_ent, R, b->edk_flag);
    }
    if (b->mask) {
        if (!b->maskk) b->maskk = dalloc<unsigned char>(rp * nr.n);
        HIPCHK(hipMemsetAsync(b->maskk, 0, rp * nr.n, s));
        k_move_rows<unsigned char><<<gv, 256, 0, s>>>(b->mask, b->maskk, b->mapk, R, nr.n, 1);
        std::swap(b->mask, b->maskk);
    }
    k_fill_int<<<blocks_for(rp, 256), 256, 0, s>>>(b->status, kNotSimulated, rp);
    const unsigned gr = blocks_for(R, 256);
    k_move_vals<int><<<gr, 256, 0, s>>>(b->stk, b->status, b->mapk, R, 1, 0);
    k_move_vals<int><<<gr, 256, 0, s>>>(b->itk, b->iters, b->mapk, R, 1, 0);
    k_move_vals<double><<<gr, 256, 0, s>>>(b->resk, b->res, b->mapk, R, 1, 0);
    if (b->restart_on) {
        if (!b->V0k) {
            b->V0k = dalloc<double2>(rp * nr.n);
            b->it0k = dalloc<int>(rp);
            b->budgetk = dalloc<int>(rp);
            b->warmk = dalloc<unsigned char>(rp);
        }
        if (!b->v0_row) k_move_rows<double2><<<gv, 256, 0, s>>>(b->V0, b->V0k, b->mapk, R, nr.n, 1);
        k_move_vals<int><<<gr, 256, 0, s>>>(b->it0, b->it0k, b->mapk, R, 1, 0);
        k_move_vals<int><<<gr, 256, 0, s>>>(b->budget, b->budgetk, b->mapk, R, 1, 0);
        k_move_vals<unsigned char><<<gr, 256, 0, s>>>(b->warm, b->warmk, b->mapk, R, 1, 0);
        if (!b->v0_row) std::swap(b->V0, b->V0k);
        std::swap(b->it0, b->it0k);
        std::swap(b->budget, b->budgetk);
        std::swap(b->warm, b->warmk);
    }
    if (b->xs) {  // extension state moves with its rows
        const int nxs = b->plan->nr.xt.n_xs;
        if (!b->xsk) {
            b->xsk = dalloc<double>(rp * nxs);
            HIPCHK(hipMemsetAsync(b->xsk, 0, rp * nxs * sizeof(double), s));
        }
        k_move_rows<double><<<blocks_for(static_cast<long long>(R) * nxs, 256), 256, 0, s>>>(b->xs, b->xsk, b->mapk, R,
                                                                                              nxs, 1);
        std::swap(b->xs, b->xsk);
    }
    HIPCHK(hipGetLastError());
    std::swap(b->V, b->Vk);
    std::swap(b->S, b->Sk);
    if (b->per_row_y) {
        std::swap(b->ed_n, b->edk_n);
        std::swap(b->ed_ent, b->edk_ent);
        std::swap(b->ed_val, b->edk_val);
        std::swap(b->ed_flag, b->edk_flag);
    }
    b->compact_saved = rows;
    b->compact_R = R;
    b->active_rows = R;
    k_count_tiles<<<b->T, kLanes, 0, s>>>(b->status, b->rows_pad, b->tile_active);
    HIPCHK(hipGetLastError());
}

void uncompact_rows(g2g_batch_s* b)
{
    const NrPlan& nr = b->plan->nr;
    const size_t rp = static_cast<size_t>(b->rows_pad);
    hipStream_t s = b->stream;
    const int R = b->compact_R;
    k_move_rows<double2><<<blocks_for(static_cast<long long>(R) * nr.n, 256), 256, 0, s>>>(b->V, b->Vk, b->mapk, R,
                                                                                          nr.n, 0);
    if (b->xs) {
        const int nxs = nr.xt.n_xs;
        k_move_rows<double><<<blocks_for(static_cast<long long>(R) * nxs, 256), 256, 0, s>>>(b->xs, b->xsk, b->mapk, R,
                                                                                              nxs, 0);
        std::swap(b->xs, b->xsk);
    }
    std::swap(b->V, b->Vk);
    std::swap(b->S, b->Sk);
    if (b->per_row_y) {
        std::swap(b->ed_n, b->edk_n);
        std::swap(b->ed_ent, b->edk_ent);
        std::swap(b->ed_val, b->edk_val);
        std::swap(b->ed_flag, b->edk_flag);
    }
    if (b->mask) std::swap(b->mask, b->maskk);
    if (b->restart_on) {  // the fallback start and first counts go back unread
        if (!b->v0_row) std::swap(b->V0, b->V0k);
        std::swap(b->it0, b->it0k);
        std::swap(b->budget, b->budgetk);
        std::swap(b->warm, b->warmk);
        // restarts while compacted changed the budgets and the warm flags
        const unsigned g = blocks_for(R, 256);
        k_move_vals<int><<<g, 256, 0, s>>>(b->budgetk, b->budget, b->mapk, R, 0, 0);
        k_move_vals<unsigned char><<<g, 256, 0, s>>>(b->warmk, b->warm, b->mapk, R, 0, 0);
    }
    const unsigned gr = blocks_for(R, 256);
    k_move_vals<int><<<gr, 256, 0, s>>>(b->status, b->stk, b->mapk, R, 0, 0);
    k_move_vals<int><<<gr, 256, 0, s>>>(b->iters, b->itk, b->mapk, R, 0, 0);
    k_move_vals<double><<<gr, 256, 0, s>>>(b->res, b->resk, b->mapk, R, 0, 0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(b->status, b->stk, rp * sizeof(int), hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(b->iters, b->itk, rp * sizeof(int), hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(b->res, b->resk, rp * sizeof(double), hipMemcpyDeviceToDevice, s));
    b->active_rows = b->compact_saved;
    b->compact_R = 0;
    k_count_tiles<<<b->T, kLanes, 0, s>>>(b->status, b->rows_pad, b->tile_active);
    HIPCHK(hipGetLastError());
}

// Whether the rest of a solve takes the split walk: when few tiles are left
// (a launch over a tile costs the whole walk of its rows, the split walk cuts
// it into dozens of independent programs); G2G_SPLIT=0 / 1 forces it off / on.
int split_env()
{
    if (const char* e = std::getenv("G2G_SPLIT")) return std::atoi(e) ? 1 : 0;
    return -1;
}

constexpr int kSplitMaxTiles = 128;  // active tiles up to which the split walk is faster

bool use_split(const g2g_batch_s* b, int active_tiles)
{
    // refinement from the A stream's J reads the main layout's positions
    if (!b->plan->split() || !b->pend || (b->ir_steps > 0 && !b->ir_fp64)) return false;
    const int env = split_env();
    if (env >= 0) return env == 1;
    return active_tiles <= kSplitMaxTiles;
}

// tiles a split launch covers: the compacted rows' tiles, or all
int split_tiles(const g2g_batch_s* b) { return live_tiles(b); }

// Whether to (re)compact: the active rows, packed, would free at least
// kCompactMinTiles tiles and a quarter of the active ones, and the LU walk
// gains from it.  A tile's walk is one workgroup, a latency-bound chain, so
// a launch costs about one walk per round of workgroups over the CUs: fewer
// tiles help once they drop a round, or fit the split walk.  (The mismatch
// and update kernels skip inactive tiles; their saving alone does not repay
// the row moves, which for per-row Ybus batches carry the whole Ybus.)
// A batch in its own row order about to take the split walk is packed
// whatever the counts: that walk covers every live tile, which is all of
// such a batch's tiles, and its layout switch zeroes their A stream (twice:
// on and off), a cost in proportion to the batch, not to the few rows left.
// G2G_COMPACT=0 / 1: never / whenever packing frees a quarter of the active
// tiles (or precedes the split walk).
constexpr int kCompactMinTiles = 16;
bool worth_compacting(const g2g_batch_s* b, int active_rows, int active_tiles)
{
    const int packed = (active_rows + kLanes - 1) / kLanes;
    const char* e = std::getenv("G2G_COMPACT");
    if (e && std::atoi(e) == 0) return false;
    if (b->compact_R == 0 && !b->split_layout && packed < b->T && use_split(b, active_tiles)) return true;
    if (4 * packed > 3 * active_tiles) return false;
    if (e) return packed < active_tiles;
    if (active_tiles - packed < kCompactMinTiles) return false;
    const int rounds_now = (active_tiles + b->n_cu - 1) / b->n_cu;
    const int rounds_packed = (packed + b->n_cu - 1) / b->n_cu;
    return rounds_packed < rounds_now || (b->plan->split() && packed <= kSplitMaxTiles);
}

// The A stream of tiles [0, T) to zero.  Each layout leaves its fill
// positions (factor entries with no J value) at zero, and the two layouts'
// fill positions differ, so a switch of layout starts from zeros.
void zero_a_tiles(g2g_batch_s* b, int T)
{
    const long long words = static_cast<long long>(T) * b->a_stride * kLanes * static_cast<long long>(b->real_size()) / 4;
    k_fill_int<<<blocks_for(words, 256), 256, 0, b->stream>>>(static_cast<int*>(b->A), 0, words);
    HIPCHK(hipGetLastError());
    b->split_T = std::max(b->split_T, T);
}

template <typename Real>
int nr_loop_body(g2g_batch_s* b, int max_iter, double tol, int freeze, double growth, SolveTimer& tm,
                 long long& rows_factored, bool& compacted)
{
    const double eps = b->precision == G2G_MIXED ? 1e-5 : 1e-12;  // own_lu/batched_lu.cuh:50-56
    if (b->ir_steps > 0 && !b->Bs) {
        const size_t bytes = static_cast<size_t>(b->rows_pad) * b->plan->nr.d * b->real_size();
        b->Bs = dalloc<unsigned char>(bytes);
        b->Xs = dalloc<unsigned char>(bytes);
    }
    int it = 0;
    int restarts = 0;
    for (;; ++it) {
        k_fill_int<<<1, 64, 0, b->stream>>>(b->n_active, 0, 2);
        int e0 = tm.mark();
        launch_mismatch<Real>(b, true, it, max_iter, tol, freeze, growth);
        tm.ph[0].push_back({e0, tm.mark()});
        if (it >= max_iter) break;
        if (b->restart_on) {
            // rows whose learned start failed at this mismatch take their
            // fallback start now, with max_iter more updates
            const NrPlan& nr = b->plan->nr;
            k_restart_v<<<dim3(live_tiles(b), kRestartChunks), 256, 0, b->stream>>>(b->status, b->warm, nr.n, b->V0,
                                                                                  b->v0_row ? 1 : 0, b->V);
            k_fill_int<<<blocks_for(live_tiles(b), 256), 256, 0, b->stream>>>(b->tile_restart, 0, live_tiles(b));
            k_restart_rows<<<blocks_for(b->active_rows, 256), 256, 0, b->stream>>>(
                b->status, b->warm, b->it0, b->budget, b->active_rows, it, b->restart_max_iter, b->requeued,
                b->tile_restart);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(b->h_active + 2, b->requeued, sizeof(int), hipMemcpyDeviceToHost, b->stream));
        }
        HIPCHK(hipMemcpyAsync(b->h_active, b->n_active, 2 * sizeof(int), hipMemcpyDeviceToHost, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
        if (b->restart_on && b->h_active[2] > restarts) {
            // evaluate the tiles of the restarted rows again, so their J and
            // F at the new start are in the A stream; the other rows of those
            // tiles come out as before (no rejection test this time), and the
            // counts stay the first pass's (a row marked for a restart counted
            // as iterating)
            restarts = b->h_active[2];
            e0 = tm.mark();
            launch_mismatch<Real>(b, true, it, max_iter, tol, freeze, 0.0, false, b->tile_restart);
            tm.ph[0].push_back({e0, tm.mark()});
        }
        if (*b->h_active == 0) break;
        if (worth_compacting(b, b->h_active[0], b->h_active[1])) {
            // (again: the rows go back to their batch places first)
            if (compacted) uncompact_rows(b);
            compact_rows(b);
            compacted = true;
            b->h_active[1] = (b->h_active[0] + kLanes - 1) / kLanes;  // the packed rows' tiles
            // the A stream was written in the old row order: evaluate again
            // (same V, same decisions; no rejection test) in the compacted one
            k_fill_int<<<1, 64, 0, b->stream>>>(b->n_active, 0, 2);
            e0 = tm.mark();
            launch_mismatch<Real>(b, true, it, max_iter, tol, freeze, 0.0);
            tm.ph[0].push_back({e0, tm.mark()});
        }
        if (!b->split_layout && use_split(b, b->h_active[1])) {
            // the rest of the solve takes the split walk: the A stream again,
            // in its layout (same V, same decisions), over zeros (its fill
            // positions are not the main layout's)
            zero_a_tiles(b, split_tiles(b));
            b->split_layout = true;
            k_fill_int<<<1, 64, 0, b->stream>>>(b->n_active, 0, 2);
            e0 = tm.mark();
            launch_mismatch<Real>(b, true, it, max_iter, tol, freeze, 0.0);
            tm.ph[0].push_back({e0, tm.mark()});
        }
        e0 = tm.mark();
        if (b->split_layout)
            launch_factor_split<Real>(b->plan, static_cast<const Real*>(b->A), static_cast<Real*>(b->U), b->a_stride,
                                      b->u_stride, b->tile_active, b->jmax, b->clamped, b->error, b->active_rows,
                                      split_tiles(b), eps, static_cast<Real*>(b->pend), b->stream);
        else
            launch_factor<Real>(b->plan, static_cast<const Real*>(b->A), static_cast<Real*>(b->U), b->a_stride,
                                b->u_stride, b->tile_active, b->jmax, b->clamped, b->error, b->active_rows, b->T,
                                eps, b->stream, factor_walk_mode<Real>(b), multiplier_stream<Real>(b));
        tm.ph[1].push_back({e0, tm.mark()});
        tm.factor_main.push_back(b->split_layout ? 0 : 1);
        if (!b->split_layout) tm.rows_main += *b->h_active;
        rows_factored += *b->h_active;
        e0 = tm.mark();
        if (b->split_layout)
            launch_backward_split<Real>(b->plan, static_cast<const Real*>(b->U), b->u_stride, b->tile_active,
                                        static_cast<Real*>(b->X), b->error, b->active_rows, split_tiles(b), b->stream);
        else
            launch_backward<Real>(b->plan, static_cast<const Real*>(b->U), b->u_stride, b->tile_active,
                                  static_cast<Real*>(b->X), b->error, b->active_rows, b->T, b->stream);
        if (b->ir_steps > 0) refine_step<Real>(b, eps, b->split_layout ? split_tiles(b) : b->T);
        launch_update<Real>(b);
        launch_ext_update<Real>(b);
        tm.ph[2].push_back({e0, tm.mark()});
    }
    return it;
}

// Newton iterations on the batch's active rows until none is left or the
// budget is spent; returns the iterations launched.  When no more than a
// quarter of the rows (and at most all but four tiles' worth) are still
// iterating, they are compacted into dense tiles for the rest of the loop.
template <typename Real>
int nr_loop(g2g_batch_s* b, int max_iter, double tol, int freeze, double growth, SolveTimer& tm,
            long long& rows_factored)
{
    bool compacted = false;
    b->split_layout = false;
    b->split_T = 0;
    // back to the main layout: the tiles the split walk used go back to zeros
    auto leave_split = [&] {
        if (b->split_layout) zero_a_tiles(b, b->split_T);
        b->split_layout = false;
        b->split_T = 0;
    };
    try {
        const int it = nr_loop_body<Real>(b, max_iter, tol, freeze, growth, tm, rows_factored, compacted);
        leave_split();
        if (compacted) uncompact_rows(b);
        return it;
    } catch (...) {
        leave_split();
        if (compacted) uncompact_rows(b);
        throw;
    }
}

// One round of a streamed batch: k Newton iterations on its active slots,
// continuous rule, each row against its own iteration count.  A row whose
// last update brought it below tol is found by the next round's first
// mismatch (its count is then the number of updates it had).
template <typename Real>
void stream_round_launches(g2g_batch_s* b, int k, int max_iter, double tol)
{
    const double eps = b->precision == G2G_MIXED ? 1e-5 : 1e-12;  // own_lu/batched_lu.cuh:50-56
    for (int q = 0; q < k; ++q) {
        k_fill_int<<<1, 64, 0, b->stream>>>(b->n_active, 0, 2);
        launch_mismatch<Real>(b, true, q, max_iter, tol, 1, 0.0, false, nullptr, b->clock_dev);
        launch_factor<Real>(b->plan, static_cast<const Real*>(b->A), static_cast<Real*>(b->U), b->a_stride,
                            b->u_stride, b->tile_active, b->jmax, b->clamped, b->error, b->active_rows, b->T, eps,
                            b->stream, factor_walk_mode<Real>(b), multiplier_stream<Real>(b));
        launch_backward<Real>(b->plan, static_cast<const Real*>(b->U), b->u_stride, b->tile_active,
                              static_cast<Real*>(b->X), b->error, b->active_rows, b->T, b->stream);
        if (b->ir_steps > 0) refine_step<Real>(b, eps, b->T);
        launch_update<Real>(b);
        launch_ext_update<Real>(b);
    }
    k_add_int<<<1, 64, 0, b->stream>>>(b->clock_dev, k);
    HIPCHK(hipGetLastError());
}

// A round's launches depend only on (k, max_iter, tol, refinement steps and
// kind, step scaling): with graphs on, they are captured once per such shape
// and replayed.
template <typename Real>
void stream_round(g2g_batch_s* b, int k, int max_iter, double tol)
{
    if (b->ir_steps > 0 && !b->Bs) {
        const size_t bytes = static_cast<size_t>(b->rows_pad) * b->plan->nr.d * b->real_size();
        b->Bs = dalloc<unsigned char>(bytes);
        b->Xs = dalloc<unsigned char>(bytes);
    }
    multiplier_stream<Real>(b);  // allocated (and cleared) outside any graph capture
    if (!b->stream_graph) {
        stream_round_launches<Real>(b, k, max_iter, tol);
    } else {
        auto bits = [](double x) {
            long long v = 0;
            std::memcpy(&v, &x, sizeof(x));
            return v;
        };
        // everything a round's launches take that can change between rounds
        const long long refine = 4LL * b->ir_steps + (b->ir_fp64 ? 2 : 0) + (factor_walk_mode<Real>(b) == kFactorStoreL);
        const long long key[6] = {k, max_iter, bits(tol), refine, bits(b->max_dva), bits(b->max_dvm)};
        if (!b->graph_exec || !std::equal(key, key + 6, b->graph_key)) {
            if (b->graph_exec) HIPCHK(hipGraphExecDestroy(b->graph_exec));
            if (b->graph) HIPCHK(hipGraphDestroy(b->graph));
            b->graph_exec = nullptr;
            b->graph = nullptr;
            // function attributes (dynamic LDS) are set by an eager round
            // first, so the capture records kernel launches only
            stream_round_launches<Real>(b, k, max_iter, tol);
            b->clock += k;
            if (!b->graph_stream) {
                if (b->stream) {
                    b->graph_stream = b->stream;
                } else {
                    HIPCHK(hipStreamCreate(&b->graph_stream));  // default flags and priority
                    b->own_graph_stream = true;
                    for (hipEvent_t& e : b->graph_ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
                }
            }
            hipStream_t keep = b->stream;
            b->stream = b->graph_stream;  // the launches below are recorded on it
            HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
            try {
                stream_round_launches<Real>(b, k, max_iter, tol);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(b->stream, &g);
                if (g) (void)hipGraphDestroy(g);
                b->stream = keep;
                throw;
            }
            const hipError_t e = hipStreamEndCapture(b->stream, &b->graph);
            b->stream = keep;
            HIPCHK(e);
            HIPCHK(hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
            std::copy(key, key + 6, b->graph_key);
            return;
        }
        if (b->own_graph_stream) {  // after the batch's earlier work on its stream
            HIPCHK(hipEventRecord(b->graph_ev[0], b->stream));
            HIPCHK(hipStreamWaitEvent(b->graph_stream, b->graph_ev[0], 0));
        }
        HIPCHK(hipGraphLaunch(b->graph_exec, b->graph_stream));
        if (b->own_graph_stream) {  // and before its later work
            HIPCHK(hipEventRecord(b->graph_ev[1], b->graph_stream));
            HIPCHK(hipStreamWaitEvent(b->stream, b->graph_ev[1], 0));
        }
        ++b->graph_replays;
    }
    b->clock += k;
}

// islanded buses of the solved rows to NaN
void mask_islanded_v(g2g_batch_s* b)
{
    if (!b->mask) return;
    const long long total = static_cast<long long>(b->rows_pad) * b->plan->nr.n;
    k_mask_v_nan<<<blocks_for(total, 256), 256, 0, b->stream>>>(b->mask, total, b->V);
    HIPCHK(hipGetLastError());
}

void check_stream_error(g2g_batch_s* b)
{
    int err = 0;
    HIPCHK(hipMemcpyAsync(&err, b->error, sizeof(int), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    if (err) throw std::runtime_error("a stream wavefront gave up waiting for its data (results invalid)");
}

void fill_stats(g2g_batch_s* b, const SolveTimer& tm, int e_begin, int e_end, int iterations,
                long long rows_factored, g2g_solve_stats* stats)
{
    if (!stats) return;
    std::memset(stats, 0, sizeof(*stats));
    stats->iterations = iterations;
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, b->ev[e_begin], b->ev[e_end]));
    stats->ms_total = ms;
    stats->ms_mismatch = tm.ms(0);
    stats->ms_factor = tm.ms(1);
    stats->ms_backward = tm.ms(2);
    stats->ms_warm_start = tm.ms(3);
    const SymbolicLU& s = b->plan->lu();
    stats->bytes_factor = static_cast<int64_t>(s.a_len + s.u_len) * static_cast<int64_t>(b->real_size()) *
                          static_cast<int64_t>(b->active_rows);
    stats->rows_factored = rows_factored;
    stats->n_factor_launches = static_cast<int32_t>(tm.ph[1].size());
    stats->ms_factor_main = tm.ms(1, true);
    stats->rows_factored_main = tm.rows_main;
    for (char m : tm.factor_main) stats->n_factor_main += m ? 1 : 0;
    std::vector<int> st(static_cast<size_t>(b->active_rows));
    if (!st.empty()) HIPCHK(hipMemcpy(st.data(), b->status, st.size() * sizeof(int), hipMemcpyDeviceToHost));
    for (int v : st) {
        stats->n_converged += v == kConverged;
        stats->n_maxiter += v == kMaxIter;
        stats->n_diverged += v == kDiverged;
    }
}

template <typename Real>
void solve_impl(g2g_batch_s* b, int max_iter, double tol, int sched, g2g_solve_stats* stats)
{
    const int freeze = sched == G2G_SCHED_CONTINUOUS ? 1 : 0;
    b->streaming = false;
    reset_rows(b);
    launch_ext_init<Real>(b);
    SolveTimer tm(b, 3 * (max_iter + 1) + 2);
    const int e_begin = tm.mark();
    long long rows_factored = 0;
    const int it = nr_loop<Real>(b, max_iter, tol, freeze, 0.0, tm, rows_factored);
    mask_islanded_v(b);
    const int e_end = tm.mark();
    check_stream_error(b);
    fill_stats(b, tm, e_begin, e_end, it, rows_factored, stats);
}

// Learned start with in-place fallback: every row starts from the MLP's
// prediction with a budget of k_warm updates; a row whose mismatch grows
// more than `growth`, turns non-finite or is not converged within k_warm
// takes the caller's start voltages (the V the batch held before the call)
// at once and gets the full budget of max_iter updates on top, while the
// other rows carry on, so a failure costs its own iterations only.  A row's
// iteration count is the sum of its updates from both starts.
// The warm start's outage mask on the device: the caller's host mask for
// this call (uploaded into `tmp`), else the batch's resident one, else none.
unsigned char* warm_mask(g2g_batch_s* b, const unsigned char* mask, int& n_mask, DeviceScratch& tmp)
{
    if (n_mask > 0) {
        unsigned char* d = tmp.get<unsigned char>(static_cast<size_t>(b->active_rows) * n_mask);
        HIPCHK(hipMemcpyAsync(d, mask, static_cast<size_t>(b->active_rows) * n_mask, hipMemcpyHostToDevice,
                              b->stream));
        return d;
    }
    n_mask = b->wmask ? b->wmask_n : 0;
    return b->wmask;
}

template <typename Real>
void solve_warm_impl(g2g_batch_s* b, g2g_mlp_s* mlp, const unsigned char* mask, int n_mask, int k_warm,
                     double growth, int max_iter, double tol, g2g_solve_stats* stats)
{
    const NrPlan& nr = b->plan->nr;
    const size_t rp = static_cast<size_t>(b->rows_pad);
    const size_t nv = rp * nr.n;
    b->streaming = false;
    if (!b->V0) b->V0 = dalloc<double2>(nv);
    if (!b->budget) {
        b->budget = dalloc<int>(rp);
        b->warm = dalloc<unsigned char>(rp);
        b->requeued = dalloc<int>(1);
        b->tile_restart = dalloc<int>(b->T);
    }
    if (!b->it0) b->it0 = dalloc<int>(rp);
    DeviceScratch tmp;
    unsigned char* d_mask = warm_mask(b, mask, n_mask, tmp);
    reset_rows(b);
    SolveTimer tm(b, 3 * (k_warm + max_iter + 2) + 8);
    const int e_begin = tm.mark();
    hipStream_t s = b->stream;
    // the fallback start: one row when every row's V is the same (a
    // broadcast start, known from g2g_batch_set_v or checked), else the whole V
    if (b->v_bcast) {
        b->v0_row = true;
        HIPCHK(hipMemcpyAsync(b->V0, b->v_row, static_cast<size_t>(nr.n) * sizeof(double2), hipMemcpyDeviceToDevice,
                              s));
    } else {
    if (!b->differs) b->differs = dalloc<int>(1);
    k_fill_int<<<1, 64, 0, s>>>(b->differs, 0, 1);
    const long long vt = (b->active_rows + kLanes - 1) / kLanes;
    if (vt > 0)
        k_v_uniform<<<blocks_for(vt * nr.n * kLanes, 256), 256, 0, s>>>(b->V, b->active_rows, nr.n, b->differs);
    HIPCHK(hipGetLastError());
    int differs = 0;
    HIPCHK(hipMemcpyAsync(&differs, b->differs, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    b->v0_row = differs == 0;
    if (b->v0_row) k_tiles_to_rows<double2><<<blocks_for(nr.n, 256), 256, 0, s>>>(b->V, 1, nr.n, b->V0, nr.n);
    else HIPCHK(hipMemcpyAsync(b->V0, b->V, nv * sizeof(double2), hipMemcpyDeviceToDevice, s));
    }
    b->v_bcast = false;  // the solve changes V
    mlp_warm_start_tiles(mlp, b->S, b->active_rows, nr.n, d_mask, n_mask, b->V, s);
    tm.ph[3].push_back({e_begin, tm.mark()});
    k_fill_int<<<blocks_for(rp, 256), 256, 0, s>>>(b->it0, 0, rp);
    k_fill_int<<<blocks_for(rp, 256), 256, 0, s>>>(b->budget, k_warm, rp);
    HIPCHK(hipMemsetAsync(b->warm, 1, rp, s));
    k_fill_int<<<1, 64, 0, s>>>(b->requeued, 0, 1);
    HIPCHK(hipGetLastError());
    long long rows_factored = 0;
    b->restart_on = true;
    b->restart_max_iter = max_iter;
    int it = 0;
    try {
        // a row fails by update k_warm at the latest and then has max_iter more
        it = nr_loop<Real>(b, k_warm + max_iter, tol, 1, growth, tm, rows_factored);
    } catch (...) {
        b->restart_on = false;
        throw;
    }
    b->restart_on = false;
    mask_islanded_v(b);
    const int e_end = tm.mark();
    int R = 0;
    HIPCHK(hipMemcpyAsync(&R, b->requeued, sizeof(int), hipMemcpyDeviceToHost, s));
    check_stream_error(b);
    fill_stats(b, tm, e_begin, e_end, it, rows_factored, stats);
    if (stats) stats->n_requeued = R;
}

// rows x m complex (row stride ld complex; ld = 0 broadcasts) -> tiled [T][m][64]
void to_tiles(const double* src, long long ld, int rows, int m, int row0, int on_device, double2* dst,
              hipStream_t st)
{
    const double2* dsrc = reinterpret_cast<const double2*>(src);
    double2* tmp = nullptr;
    if (!on_device) {
        const size_t cnt = ld == 0 ? static_cast<size_t>(m) : static_cast<size_t>(rows - 1) * ld + m;
        tmp = dalloc<double2>(cnt);
        HIPCHK(hipMemcpyAsync(tmp, src, cnt * sizeof(double2), hipMemcpyHostToDevice, st));
        dsrc = tmp;
    }
    const long long total = static_cast<long long>(rows) * m;
    k_rows_to_tiles<double2><<<blocks_for(total, 256), 256, 0, st>>>(dsrc, ld, rows, m, row0, dst, ld == 0 ? 1 : 0);
    HIPCHK(hipGetLastError());
    if (tmp) {
        HIPCHK(hipStreamSynchronize(st));
        dfree(tmp);
    }
}

void check_rows(const g2g_batch_s* b, int rows, int row0)
{
    if (rows < 0 || row0 < 0 || row0 + rows > b->rows) throw std::invalid_argument("row range outside the batch");
}

// Raw batched solve of `rows` matrices on a plan's pattern (host arrays).
template <typename Real>
void raw_solve(g2g_plan_s* P, int rows, const double* vals, const double* rhs, double* x, double pivot_eps,
               int ir_steps, int32_t* clamped, void* stream)
{
    hipStream_t st = static_cast<hipStream_t>(stream);
    P = plan_fitting(P, sizeof(Real) == sizeof(double));
    P->upload();
    const SymbolicLU& s = P->lu();
    const int T = (rows + kLanes - 1) / kLanes;
    const size_t rp = static_cast<size_t>(T) * kLanes;
    const int nnz = P->nnz, d = s.d;
    const int a_stride = a_stride_of(s, sizeof(Real)), u_stride = u_stride_of(s, sizeof(Real));
    DeviceScratch B;
    double* d_vals = B.get<double>(static_cast<size_t>(rows) * nnz);
    double* d_rhs = B.get<double>(static_cast<size_t>(rows) * d);
    Real* A = B.get<Real>(rp * a_stride);
    Real* U = B.get<Real>(rp * u_stride);
    Real* X = B.get<Real>(rp * d);
    double* jm = B.get<double>(rp);
    int* ta = B.get<int>(T);
    int* cl = B.get<int>(rp);
    int* err = B.get<int>(1);
    try {
        HIPCHK(hipMemsetAsync(err, 0, sizeof(int), st));
        HIPCHK(hipMemcpyAsync(d_vals, vals, static_cast<size_t>(rows) * nnz * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d_rhs, rhs, static_cast<size_t>(rows) * d * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemsetAsync(A, 0, rp * a_stride * sizeof(Real), st));
        HIPCHK(hipMemsetAsync(jm, 0, rp * sizeof(double), st));
        k_fill_int<<<blocks_for(T, 256), 256, 0, st>>>(ta, kLanes, T);
        k_row_absmax<<<rows, 256, 0, st>>>(d_vals, nnz, rows, nnz, jm);
        const long long total = static_cast<long long>(rows) * (nnz + d);
        k_gather_astream<Real><<<blocks_for(total, 256), 256, 0, st>>>(d_vals, nnz, d_rhs, d, rows, nnz, d,
                                                                        P->d_apos, P->d_fpos, a_stride, A);
        HIPCHK(hipGetLastError());
        launch_factor<Real>(P, A, U, a_stride, u_stride, ta, jm, cl, err, rows, T, pivot_eps, st);
        launch_backward<Real>(P, U, u_stride, ta, X, err, rows, T, st);
        if (ir_steps > 0) {
            Real* Bs = B.get<Real>(rp * d);
            Real* Xs = B.get<Real>(rp * d);
            refine<Real>(P, A, U, X, Bs, Xs, a_stride, u_stride, ta, jm, cl, err, rows, T, pivot_eps, ir_steps, st);
        }
        int herr = 0;
        HIPCHK(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, st));
        std::vector<Real> hx(rp * d);
        HIPCHK(hipMemcpyAsync(hx.data(), X, hx.size() * sizeof(Real), hipMemcpyDeviceToHost, st));
        if (clamped) HIPCHK(hipMemcpyAsync(clamped, cl, rows * sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (herr) throw std::runtime_error("a stream wavefront gave up waiting for its data (results invalid)");
        for (int r = 0; r < rows; ++r)
            for (int e = 0; e < d; ++e)
                x[static_cast<size_t>(r) * d + e] = hx[((static_cast<size_t>(r / kLanes) * d) + e) * kLanes + (r % kLanes)];
    } catch (...) {
        (void)hipStreamSynchronize(st);
        throw;
    }
}

// ---- allocation-free raw solve ------------------------------------------
// Scratch of g2g_lu_solve_batch_noalloc: one array per device, held until the
// process ends.  Host code never passes `scratch` itself to a kernel (on the
// host it names a stand-in, not device memory): scratch_base() resolves the
// current device's copy once with hipGetSymbolAddress and keeps it.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

unsigned char* scratch_base()
{
    static std::mutex mu;
    static std::vector<unsigned char*> base;  // per device ordinal
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if (static_cast<size_t>(dev) >= base.size()) base.resize(static_cast<size_t>(dev) + 1, nullptr);
    if (!base[dev]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw std::runtime_error("hipGetSymbolAddress returned no address for the scratch array");
        base[dev] = static_cast<unsigned char*>(p);
    }
    return base[dev];
}

// Bump allocation inside the scratch array (256-byte aligned pieces).
struct ScratchCarve {
    unsigned char* base;
    size_t used = 0;
    template <typename T>
    T* take(size_t count)
    {
        T* p = reinterpret_cast<T*>(base + used);
        used += (count * sizeof(T) + 255) / 256 * 256;
        return p;
    }
};

// Scratch bytes of a raw solve of `rows` rows (A, U and X streams, the clamp
// scales, tile counts, the error flag, and the refinement copies).
template <typename Real>
size_t raw_scratch_bytes(const SymbolicLU& s, int rows, int ir_steps)
{
    ScratchCarve c{nullptr};
    const int T = (rows + kLanes - 1) / kLanes;
    const size_t rp = static_cast<size_t>(T) * kLanes;
    c.take<Real>(rp * a_stride_of(s, sizeof(Real)));
    c.take<Real>(rp * u_stride_of(s, sizeof(Real)));
    c.take<Real>(rp * s.d);
    c.take<double>(rp);
    c.take<int>(static_cast<size_t>(T));
    c.take<int>(1);
    if (ir_steps > 0) {
        c.take<Real>(rp * s.d);
        c.take<Real>(rp * s.d);
    }
    return c.used;
}

// The raw solve on device arrays with every intermediate in the scratch
// array: no allocation.  Work is enqueued on `stream`; the call waits for it
// only to read the walkers' error flag.
template <typename Real>
void raw_solve_noalloc(const g2g_plan_s* P0, int rows, const double* d_vals, const double* d_rhs, double* d_x,
                       double pivot_eps, int ir_steps, int32_t* d_clamped, void* stream)
{
    hipStream_t st = static_cast<hipStream_t>(stream);
    const g2g_plan_s* P = plan_fitted(P0, sizeof(Real) == sizeof(double));
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (!P || !P->uploaded || P->device != dev)
        throw std::invalid_argument("the plan's device tables for this precision are not resident on this device: "
                                    "run g2g_lu_solve_batch once with this plan and precision first");
    const SymbolicLU& s = P->lu();
    const size_t need = raw_scratch_bytes<Real>(s, rows, ir_steps);
    if (need > kNoAllocScratchBytes) {
        const size_t per = raw_scratch_bytes<Real>(s, kLanes, ir_steps);
        throw std::invalid_argument(std::to_string(rows) + " rows need " + std::to_string(need >> 20) +
                                    " MiB of scratch, more than its " + std::to_string(kNoAllocScratchBytes >> 20) +
                                    " MiB; at most " + std::to_string(kNoAllocScratchBytes / per * kLanes) +
                                    " rows for this plan and precision");
    }
    const int T = (rows + kLanes - 1) / kLanes;
    const size_t rp = static_cast<size_t>(T) * kLanes;
    const int nnz = P->nnz, d = s.d;
    const int a_stride = a_stride_of(s, sizeof(Real)), u_stride = u_stride_of(s, sizeof(Real));
    ScratchCarve c{scratch_base()};
    Real* A = c.take<Real>(rp * a_stride);
    Real* U = c.take<Real>(rp * u_stride);
    Real* X = c.take<Real>(rp * d);
    double* jm = c.take<double>(rp);
    int* ta = c.take<int>(static_cast<size_t>(T));
    int* err = c.take<int>(1);
    // the scratch keeps the previous call's contents: clear what the kernels
    // accumulate into or only partly write (fill kernels, no memset nodes)
    const long long a_words = static_cast<long long>(rp * a_stride * sizeof(Real) / 4);
    k_fill_int<<<blocks_for(a_words, 256), 256, 0, st>>>(reinterpret_cast<int*>(A), 0, a_words);
    const long long j_words = static_cast<long long>(rp * 2);
    k_fill_int<<<blocks_for(j_words, 256), 256, 0, st>>>(reinterpret_cast<int*>(jm), 0, j_words);
    k_fill_int<<<1, 64, 0, st>>>(err, 0, 1);
    k_fill_int<<<blocks_for(T, 256), 256, 0, st>>>(ta, kLanes, T);
    k_row_absmax<<<rows, 256, 0, st>>>(d_vals, nnz, rows, nnz, jm);
    const long long total = static_cast<long long>(rows) * (nnz + d);
    k_gather_astream<Real><<<blocks_for(total, 256), 256, 0, st>>>(d_vals, nnz, d_rhs, d, rows, nnz, d, P->d_apos,
                                                                    P->d_fpos, a_stride, A);
    HIPCHK(hipGetLastError());
    launch_factor<Real>(P, A, U, a_stride, u_stride, ta, jm, d_clamped, err, rows, T, pivot_eps, st);
    launch_backward<Real>(P, U, u_stride, ta, X, err, rows, T, st);
    if (ir_steps > 0) {
        Real* Bs = c.take<Real>(rp * d);
        Real* Xs = c.take<Real>(rp * d);
        refine<Real>(P, A, U, X, Bs, Xs, a_stride, u_stride, ta, jm, d_clamped, err, rows, T, pivot_eps, ir_steps,
                     st);
    }
    const long long xs = static_cast<long long>(rows) * d;
    k_tiles_to_rows_f64<Real><<<blocks_for(xs, 256), 256, 0, st>>>(X, rows, d, d_x, d);
    HIPCHK(hipGetLastError());
    int herr = 0;
    HIPCHK(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (herr) throw std::runtime_error("a stream wavefront gave up waiting for its data (results invalid)");
}

// J^T x = rhs for the rows [0, active_rows) at the batch's V: the mismatch
// program fills the A stream with J transposed, rhs replaces its right-hand
// side, and the factorization / backward programs of the (symmetric) pattern
// solve it.  The solve's statuses and iteration counts are kept.
template <typename Real>
void adjoint_impl(g2g_batch_s* b, const double* rhs, double* x)
{
    const NrPlan& nr = b->plan->nr;
    const size_t rp = static_cast<size_t>(b->rows_pad);
    const int rows = b->active_rows;
    b->plan->upload_transposed();
    DeviceScratch tmp;
    int* saved = tmp.get<int>(rp);
    int* saved_it = tmp.get<int>(rp);
    double* saved_res = tmp.get<double>(rp);
    double* d_rhs = tmp.get<double>(static_cast<size_t>(rows) * nr.d);
    hipStream_t st = b->stream;
    HIPCHK(hipMemcpyAsync(saved, b->status, rp * sizeof(int), hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(saved_it, b->iters, rp * sizeof(int), hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(saved_res, b->res, rp * sizeof(double), hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(d_rhs, rhs, static_cast<size_t>(rows) * nr.d * sizeof(double), hipMemcpyHostToDevice, st));
    const bool streaming = b->streaming;
    b->streaming = false;
    try {
        reset_rows(b);
        k_fill_int<<<1, 64, 0, st>>>(b->n_active, 0, 2);
        launch_mismatch<Real>(b, true, 0, 1 << 30, 1e-300, 0, 0.0, true);
        k_gather_astream<Real><<<blocks_for(static_cast<long long>(rows) * nr.d, 256), 256, 0, st>>>(
            nullptr, 0, d_rhs, nr.d, rows, 0, nr.d, b->plan->d_apos, b->plan->d_fpos, b->a_stride,
            static_cast<Real*>(b->A));
        HIPCHK(hipGetLastError());
        // every row of the batch is solved; the clamp scale is |J|'s, which
        // the transposed fill computed
        const double eps = b->precision == G2G_MIXED ? 1e-5 : 1e-12;
        k_count_tiles<<<b->T, kLanes, 0, st>>>(b->status, b->rows_pad, b->tile_active);
        launch_factor<Real>(b->plan, static_cast<const Real*>(b->A), static_cast<Real*>(b->U), b->a_stride,
                            b->u_stride, b->tile_active, b->jmax, b->clamped, b->error, rows, b->T, eps, st);
        launch_backward<Real>(b->plan, static_cast<const Real*>(b->U), b->u_stride, b->tile_active,
                              static_cast<Real*>(b->X), b->error, rows, b->T, st);
        std::vector<Real> hx(rp * nr.d);
        HIPCHK(hipMemcpyAsync(hx.data(), b->X, hx.size() * sizeof(Real), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(b->status, saved, rp * sizeof(int), hipMemcpyDeviceToDevice, st));
        HIPCHK(hipMemcpyAsync(b->iters, saved_it, rp * sizeof(int), hipMemcpyDeviceToDevice, st));
        HIPCHK(hipMemcpyAsync(b->res, saved_res, rp * sizeof(double), hipMemcpyDeviceToDevice, st));
        k_count_tiles<<<b->T, kLanes, 0, st>>>(b->status, b->rows_pad, b->tile_active);
        HIPCHK(hipStreamSynchronize(st));
        b->streaming = streaming;
        check_stream_error(b);
        for (int r = 0; r < rows; ++r)
            for (int e = 0; e < nr.d; ++e)
                x[static_cast<size_t>(r) * nr.d + e] =
                    hx[((static_cast<size_t>(r / kLanes) * nr.d) + e) * kLanes + (r % kLanes)];
    } catch (...) {
        (void)hipStreamSynchronize(st);
        b->streaming = streaming;
        throw;
    }
}


}  // namespace

extern "C" {

const char* g2g_last_error(void) { return last_error().c_str(); }

int g2g_version(void) { return 100; }

int g2g_set_device(int32_t device)
{
    return guarded([&] {
        int count = 0;
        HIPCHK(hipGetDeviceCount(&count));
        if (device >= count) throw std::invalid_argument("device ordinal out of range");
        if (device >= 0) HIPCHK(hipSetDevice(device));
    });
}

int g2g_plan_create(int32_t n, const int32_t* Yp, const int32_t* Yi, const double* Yx, int32_t npv,
                    const int32_t* pv, int32_t npq, const int32_t* pq, int32_t slack, g2g_plan_t* out)
{
    return guarded([&] {
        if (!out || !Yp || !Yi || !Yx || (npv && !pv) || (npq && !pq)) throw std::invalid_argument("null argument");
        auto p = std::make_unique<g2g_plan_s>();
        p->grid = true;
        p->nr = build_nr_plan(n, Yp, Yi, Yx, npv, pv, npq, pq, slack);
        p->nnz = p->nr.nJ;
        *out = p.release();
    });
}

int g2g_plan_create_ext(int32_t n, const int32_t* Yp, const int32_t* Yi, const double* Yx, int32_t npv,
                        const int32_t* pv, int32_t npq, const int32_t* pq, int32_t slack, const g2g_extensions* ext,
                        g2g_plan_t* out)
{
    return guarded([&] {
        if (!out || !Yp || !Yi || !Yx || (npv && !pv) || (npq && !pq)) throw std::invalid_argument("null argument");
        ExtSpec x;
        if (ext) {
            auto vec = [](const auto* p, int k) {
                using T = std::remove_cv_t<std::remove_pointer_t<decltype(p)>>;
                if (k > 0 && !p) throw std::invalid_argument("null extension array");
                return k > 0 ? std::vector<T>(p, p + k) : std::vector<T>();
            };
            if (ext->slack_weights) x.slack_w = vec(ext->slack_weights, n);
            const int nh = ext->n_hvdc;
            if (nh < 0 || ext->n_vc_groups < 0) throw std::invalid_argument("negative extension count");
            x.h_bus1 = vec(ext->hvdc_bus1, nh);
            x.h_bus2 = vec(ext->hvdc_bus2, nh);
            x.h_status = vec(ext->hvdc_status, nh);
            x.h_p0 = vec(ext->hvdc_p0, nh);
            x.h_k = vec(ext->hvdc_k, nh);
            x.h_lf1 = vec(ext->hvdc_lf1, nh);
            x.h_lf2 = vec(ext->hvdc_lf2, nh);
            x.h_r = vec(ext->hvdc_r, nh);
            x.h_pmax12 = vec(ext->hvdc_pmax12, nh);
            x.h_pmax21 = vec(ext->hvdc_pmax21, nh);
            const int ng = ext->n_vc_groups;
            if (ng > 0) {
                if (!ext->vc_ptr) throw std::invalid_argument("null extension array");
                x.vc_ptr = vec(ext->vc_ptr, ng + 1);
                const int nc = x.vc_ptr[ng];
                if (nc < 0) throw std::invalid_argument("bad voltage-control group pointers");
                x.vc_reg = vec(ext->vc_reg_bus, ng);
                x.vc_vset = vec(ext->vc_v_set, ng);
                x.vc_bus = vec(ext->vc_bus, nc);
                x.vc_slope = vec(ext->vc_slope, nc);
                x.vc_w = vec(ext->vc_weight, nc);
            }
        }
        auto p = std::make_unique<g2g_plan_s>();
        p->grid = true;
        p->nr = build_nr_plan(n, Yp, Yi, Yx, npv, pv, npq, pq, slack, 0, 0, &x);
        p->nnz = p->nr.nJ;
        *out = p.release();
    });
}

int g2g_lu_plan_create(int32_t d, const int32_t* Ap, const int32_t* Ai, const int32_t* order, g2g_plan_t* out)
{
    return guarded([&] {
        if (!out || !Ap || !Ai) throw std::invalid_argument("null argument");
        auto p = std::make_unique<g2g_plan_s>();
        p->grid = false;
        std::vector<int> ord;
        if (order) ord.assign(order, order + d);
        p->raw = build_symbolic(d, Ap, Ai, ord, default_walkers(d), default_bwalkers(d));
        p->raw_order = ord;
        p->nnz = Ap[d];
        p->raw_p.assign(Ap, Ap + d + 1);
        p->raw_i.assign(Ai, Ai + Ap[d]);
        *out = p.release();
    });
}

void g2g_plan_destroy(g2g_plan_t plan) { delete plan; }

int g2g_plan_get_info(g2g_plan_t P, g2g_plan_info* info)
{
    return guarded([&] {
        if (!P || !info) throw std::invalid_argument("null argument");
        const SymbolicLU& s = P->lu();
        std::memset(info, 0, sizeof(*info));
        info->d = s.d;
        info->nnz = P->nnz;
        info->slack = -1;
        if (P->grid) {
            info->n = P->nr.n;
            info->nY = P->nr.nY;
            info->nJ = P->nr.nJ;
            info->npvpq = P->nr.npvpq;
            info->npq = P->nr.npq;
            info->slack = P->nr.slack;
        }
        info->nLU = s.nLU;
        info->n_updates = s.n_updates;
        info->n_steps = s.n_steps;
        info->walkers = s.walkers;
        info->bwalkers = s.bwalkers;
        info->n_top = s.n_top;
        info->max_c = s.max_c;
        info->height = s.height;
        info->slots_f = s.slots_f;
        info->slots_b = s.slots_b;
        info->a_len = s.a_len;
        info->u_len = s.u_len;
        info->u_stride_f64 = u_stride_of(s, 8);
        info->u_stride_f32 = u_stride_of(s, 4);
        info->n_split = s.n_split;
        info->slots_fs = s.slots_fs;
        info->slots_bs = s.slots_bs;
        info->pend_len = s.pend_len;
        info->split_a_len = s.split_a_len;
        info->split_u_len = s.split_u_len;
        info->a_stride_f64 = a_stride_of(s, 8);
        info->a_stride_f32 = a_stride_of(s, 4);
        if (P->grid) {
            info->n_ext = P->nr.n_ext;
            info->n_xs = P->nr.xt.n_xs;
            info->sa_col = P->nr.xt.sa_col;
        } else {
            info->sa_col = -1;
        }
    });
}

int g2g_plan_get_array(g2g_plan_t P, int32_t which, int32_t* out, int64_t capacity, int64_t* length)
{
    return guarded([&] {
        if (!P) throw std::invalid_argument("null plan");
        const SymbolicLU& s = P->lu();
        std::vector<int> tmp;
        const std::vector<int>* v = nullptr;
        switch (which) {
        case G2G_ARR_JP: if (!P->grid) throw std::invalid_argument("not a grid plan"); v = &P->nr.Jp; break;
        case G2G_ARR_JI: if (!P->grid) throw std::invalid_argument("not a grid plan"); v = &P->nr.Ji; break;
        case G2G_ARR_MAPS:
            if (!P->grid) throw std::invalid_argument("not a grid plan");
            for (int b = 0; b < 4; ++b) tmp.insert(tmp.end(), P->nr.map[b].begin(), P->nr.map[b].end());
            v = &tmp;
            break;
        case G2G_ARR_PVPQ: if (!P->grid) throw std::invalid_argument("not a grid plan"); v = &P->nr.pvpq; break;
        case G2G_ARR_PERM: v = &s.perm; break;
        case G2G_ARR_COLP: v = &s.colp; break;
        case G2G_ARR_COLI: v = &s.coli; break;
        case G2G_ARR_PROG_F: v = &s.prog_f; break;
        case G2G_ARR_PROG_B: v = &s.prog_b; break;
        case G2G_ARR_APOS: v = &s.a_pos_of_entry; break;
        case G2G_ARR_FPOS: v = &s.f_pos; break;
        case G2G_ARR_UOFF: v = &s.u_off; break;
        case G2G_ARR_SEGS:
            tmp = seg_words(s.seg);
            v = &tmp;
            break;
        case G2G_ARR_BSEGS:
            tmp = seg_words(s.bseg);
            v = &tmp;
            break;
        case G2G_ARR_PROG_FS: v = &s.prog_fs; break;
        case G2G_ARR_PROG_BS: v = &s.prog_bs; break;
        case G2G_ARR_SPLIT_SEGS:
            tmp = seg_words(s.split_seg);
            v = &tmp;
            break;
        case G2G_ARR_SPLIT_BSEGS:
            tmp = seg_words(s.split_bseg);
            v = &tmp;
            break;
        case G2G_ARR_SPLIT_APOS: v = &s.split_apos; break;
        case G2G_ARR_SPLIT_FPOS: v = &s.split_fpos; break;
        default: throw std::invalid_argument("unknown array id");
        }
        if (length) *length = static_cast<int64_t>(v->size());
        if (out) {
            if (capacity < static_cast<int64_t>(v->size())) throw std::invalid_argument("output buffer too small");
            std::copy(v->begin(), v->end(), out);
        }
    });
}

int g2g_batch_create(g2g_plan_t P, int32_t rows, int32_t precision, int32_t per_row_ybus, void* stream,
                     g2g_batch_t* out)
{
    return guarded([&] {
        if (!P || !out) throw std::invalid_argument("null argument");
        if (!P->grid) throw std::invalid_argument("a batch needs a grid plan");
        if (rows <= 0) throw std::invalid_argument("rows must be positive");
        if (precision != G2G_FP64 && precision != G2G_MIXED) throw std::invalid_argument("unknown precision");
        P = plan_fitting(P, precision == G2G_FP64);
        P->upload();
        auto b = std::make_unique<g2g_batch_s>();
        b->plan = P;
        b->rows = b->active_rows = rows;
        b->T = (rows + kLanes - 1) / kLanes;
        b->rows_pad = b->T * kLanes;
        b->precision = precision;
        b->per_row_y = per_row_ybus ? 1 : 0;
        b->stream = static_cast<hipStream_t>(stream);
        const NrPlan& nr = P->nr;
        const size_t rp = static_cast<size_t>(b->rows_pad);
        b->V = dalloc<double2>(rp * nr.n);
        b->S = dalloc<double2>(rp * nr.n);
        HIPCHK(hipMemsetAsync(b->S, 0, rp * nr.n * sizeof(double2), b->stream));
        // padding lanes: |V| = 1 keeps every kernel finite on them
        {
            std::vector<double2> ones(static_cast<size_t>(nr.n), make_double2(1.0, 0.0));
            double2* d1 = dupload(ones);
            k_rows_to_tiles<double2><<<blocks_for(static_cast<long long>(b->rows_pad) * nr.n, 256), 256, 0,
                                       b->stream>>>(d1, 0, b->rows_pad, nr.n, 0, b->V, 1);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(b->stream));
            dfree(d1);
        }
        if (b->per_row_y) {  // no edits yet: every row on the shared Ybus
            b->ed_n = dalloc<int>(rp);
            k_fill_int<<<blocks_for(rp, 256), 256, 0, b->stream>>>(b->ed_n, 0, rp);
            b->ed_flag = dalloc<int>(static_cast<size_t>(b->T) * nr.n);
            ensure_edit_cap(b.get(), 16, false);
            edit_flags(b.get(), b->ed_n, b->ed_ent, 0, b->ed_flag);
        }
        b->a_stride = a_stride_of(nr.lu, static_cast<int>(b->real_size()));
        b->u_stride = u_stride_of(nr.lu, static_cast<int>(b->real_size()));
        b->A = dalloc<unsigned char>(rp * b->a_stride * b->real_size());
        b->U = dalloc<unsigned char>(rp * b->u_stride * b->real_size());
        b->X = dalloc<unsigned char>(rp * nr.d * b->real_size());
        HIPCHK(hipMemsetAsync(b->A, 0, rp * b->a_stride * b->real_size(), b->stream));
        b->res = dalloc<double>(rp);
        b->jmax = dalloc<double>(rp);
        HIPCHK(hipMemsetAsync(b->res, 0, rp * sizeof(double), b->stream));  // rows never simulated read 0
        HIPCHK(hipMemsetAsync(b->jmax, 0, rp * sizeof(double), b->stream));
        {
            int dev = 0, cu = 0;
            HIPCHK(hipGetDevice(&dev));
            HIPCHK(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev));
            if (cu > 0) b->n_cu = cu;
        }
        b->mis_splits = default_mis_splits();
        b->part_cap = std::max(std::max(b->mis_splits, 4) * b->T, 128);
        b->mis_part = dalloc<double>(2 * static_cast<size_t>(b->part_cap) * kLanes);
        if (P->split())
            b->pend = dalloc<unsigned char>(rp * static_cast<size_t>(std::max(1, nr.lu.pend_len)) * b->real_size());
        if (P->ext()) {
            const size_t nxs = static_cast<size_t>(std::max(1, nr.xt.n_xs));
            b->xs = dalloc<double>(rp * nxs);
            HIPCHK(hipMemsetAsync(b->xs, 0, rp * nxs * sizeof(double), b->stream));
            b->sx = dalloc<double2>(rp * std::max<size_t>(1, nr.xt.xb_bus.size()));
        }
        b->status = dalloc<int>(rp);
        b->iters = dalloc<int>(rp);
        b->clamped = dalloc<int>(rp);
        b->tile_active = dalloc<int>(b->T);
        b->n_active = dalloc<int>(4);  // rows, tiles; a partial pass's spare pair
        b->error = dalloc<int>(1);
        HIPCHK(hipMemsetAsync(b->error, 0, sizeof(int), b->stream));
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&b->h_active), 4 * sizeof(int), hipHostMallocDefault));
        k_fill_int<<<blocks_for(rp, 256), 256, 0, b->stream>>>(b->status, kNotSimulated, rp);
        k_fill_int<<<blocks_for(rp, 256), 256, 0, b->stream>>>(b->iters, 0, rp);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(b->stream));
        *out = b.release();
    });
}

void g2g_batch_destroy(g2g_batch_t b) { delete b; }

int g2g_batch_set_v(g2g_batch_t b, const double* V, int64_t ld, int32_t rows, int32_t row0, int32_t on_device)
{
    return guarded([&] {
        if (!b || !V) throw std::invalid_argument("null argument");
        check_rows(b, rows, row0);
        const int n = b->plan->nr.n;
        to_tiles(V, ld, rows, n, row0, on_device, b->V, b->stream);
        b->v_bcast = ld == 0 && row0 == 0 && rows == b->rows;
        if (b->v_bcast) {  // keep the one row (a later warm solve's fallback start)
            if (!b->v_row) b->v_row = dalloc<double2>(static_cast<size_t>(n));
            HIPCHK(hipMemcpyAsync(b->v_row, V, static_cast<size_t>(n) * sizeof(double2),
                                  on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, b->stream));
            if (!on_device) HIPCHK(hipStreamSynchronize(b->stream));
        }
    });
}

int g2g_batch_set_sbus(g2g_batch_t b, const double* S, int64_t ld, int32_t rows, int32_t row0, int32_t on_device)
{
    return guarded([&] {
        if (!b || !S) throw std::invalid_argument("null argument");
        check_rows(b, rows, row0);
        to_tiles(S, ld, rows, b->plan->nr.n, row0, on_device, b->S, b->stream);
    });
}

int g2g_batch_set_ybus_edits(g2g_batch_t b, const int32_t* row_ptr, const int32_t* entry, const double* delta,
                             int32_t rows)
{
    return guarded([&] {
        if (!b || !row_ptr) throw std::invalid_argument("null argument");
        if (!b->per_row_y) throw std::invalid_argument("the batch was created without per-row Ybus values");
        check_rows(b, rows, 0);
        const NrPlan& nr = b->plan->nr;
        const int ne = row_ptr[rows];
        for (int r = 0; r < rows; ++r)
            if (row_ptr[r + 1] < row_ptr[r]) throw std::invalid_argument("row_ptr is not monotone");
        for (int q = 0; q < ne; ++q)
            if (entry[q] < 0 || entry[q] >= nr.nY) throw std::invalid_argument("edit entry out of range");
        const EditLists el(row_ptr, entry, delta, rows);
        ensure_edit_cap(b, el.most, false);
        k_fill_int<<<blocks_for(b->rows_pad, 256), 256, 0, b->stream>>>(b->ed_n, 0, b->rows_pad);
        if (ne > 0) {
            int* d_rp = dupload(el.rp);
            int* d_en = dupload(el.ent);
            double2* d_de = dupload(el.val);
            k_set_row_edits<<<blocks_for(rows, 256), 256, 0, b->stream>>>(d_rp, d_en, d_de, rows, nullptr, b->ed_cap,
                                                                          b->ed_n, b->ed_ent, b->ed_val);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(b->stream));
            dfree(d_rp);
            dfree(d_en);
            dfree(d_de);
        }
        edit_flags(b, b->ed_n, b->ed_ent, b->rows_pad, b->ed_flag);
        HIPCHK(hipStreamSynchronize(b->stream));
    });
}

int g2g_batch_set_active_rows(g2g_batch_t b, int32_t active_rows)
{
    return guarded([&] {
        if (!b) throw std::invalid_argument("null batch");
        if (active_rows < 0 || active_rows > b->rows) throw std::invalid_argument("active_rows outside the batch");
        b->active_rows = active_rows;
    });
}

int g2g_batch_set_step_control(g2g_batch_t b, int32_t ir_steps, int32_t scaling, double max_dva, double max_dvm)
{
    return guarded([&] {
        if (!b) throw std::invalid_argument("null batch");
        if (ir_steps < 0) throw std::invalid_argument("ir_steps must be >= 0");
        if (scaling && !(max_dva > 0.0 && max_dvm > 0.0))
            throw std::invalid_argument("step scaling needs positive max_dva and max_dvm");
        if (scaling && b->plan->ext()) throw std::invalid_argument("step scaling is not available with extension rows");
        b->ir_steps = ir_steps;
        b->max_dva = scaling ? max_dva : 0.0;
        b->max_dvm = scaling ? max_dvm : 0.0;
    });
}

int g2g_batch_solve(g2g_batch_t b, int32_t max_iter, double tol, int32_t scheduling, g2g_solve_stats* stats)
{
    return guarded([&] {
        if (b) b->v_bcast = false;  // V changes (or is handed out)
        if (!b) throw std::invalid_argument("null batch");
        if (max_iter < 0) throw std::invalid_argument("max_iter must be >= 0");
        if (!(tol > 0.0)) throw std::invalid_argument("tol must be positive");
        if (scheduling != G2G_SCHED_CHUNKED && scheduling != G2G_SCHED_CONTINUOUS)
            throw std::invalid_argument("unknown scheduling");
        if (b->precision == G2G_MIXED) solve_impl<float>(b, max_iter, tol, scheduling, stats);
        else solve_impl<double>(b, max_iter, tol, scheduling, stats);
    });
}

int g2g_batch_solve_warm(g2g_batch_t b, g2g_mlp_t mlp, const uint8_t* mask, int32_t n_mask, int32_t k_warm,
                         double growth, int32_t max_iter, double tol, g2g_solve_stats* stats)
{
    return guarded([&] {
        if (!b || !mlp) throw std::invalid_argument("null argument");
        if (n_mask < 0 || (n_mask > 0 && !mask)) throw std::invalid_argument("mask missing");
        if (k_warm < 0 || max_iter < 0) throw std::invalid_argument("iteration budgets must be >= 0");
        if (!(tol > 0.0)) throw std::invalid_argument("tol must be positive");
        if (growth < 0.0 || growth != growth) throw std::invalid_argument("growth must be >= 0 (0: off)");
        if (b->plan->ext()) throw std::invalid_argument("learned warm starts are not available with extension rows");
        if (b->precision == G2G_MIXED)
            solve_warm_impl<float>(b, mlp, mask, n_mask, k_warm, growth, max_iter, tol, stats);
        else solve_warm_impl<double>(b, mlp, mask, n_mask, k_warm, growth, max_iter, tol, stats);
    });
}

int g2g_batch_set_refinement_residual(g2g_batch_t b, int32_t fp64)
{
    return guarded([&] {
        if (!b) throw std::invalid_argument("null batch");
        if (fp64 && b->plan->ext())
            throw std::invalid_argument("the FP64 refinement residual is not available with extension rows");
        b->ir_fp64 = fp64 != 0;
    });
}

int g2g_batch_set_warm_mask(g2g_batch_t b, const uint8_t* mask, int32_t n_mask)
{
    return guarded([&] {
        if (!b) throw std::invalid_argument("null batch");
        if (n_mask < 0 || (n_mask > 0 && !mask)) throw std::invalid_argument("mask missing");
        HIPCHK(hipStreamSynchronize(b->stream));
        DeviceScratch::free_one(b->wmask);
        b->wmask_n = 0;
        if (n_mask == 0) return;
        const size_t bytes = static_cast<size_t>(b->rows) * n_mask;
        b->wmask = dalloc<unsigned char>(bytes);
        HIPCHK(hipMemcpy(b->wmask, mask, bytes, hipMemcpyHostToDevice));
        b->wmask_n = n_mask;
    });
}

int g2g_batch_set_warm_contraction(g2g_batch_t b, double q)
{
    return guarded([&] {
        if (!b) throw std::invalid_argument("null batch");
        if (!(q >= 0.0)) throw std::invalid_argument("the contraction bound must be >= 0 (0: off)");
        b->contraction = q;
    });
}

int g2g_batch_warm_start(g2g_batch_t b, g2g_mlp_t mlp, const uint8_t* mask, int32_t n_mask)
{
    return guarded([&] {
        if (b) b->v_bcast = false;  // V changes (or is handed out)
        if (!b || !mlp) throw std::invalid_argument("null argument");
        if (n_mask < 0 || (n_mask > 0 && !mask)) throw std::invalid_argument("mask missing");
        DeviceScratch tmp;
        unsigned char* d_mask = warm_mask(b, mask, n_mask, tmp);
        mlp_warm_start_tiles(mlp, b->S, b->active_rows, b->plan->nr.n, d_mask, n_mask, b->V, b->stream);
        HIPCHK(hipStreamSynchronize(b->stream));
    });
}

int g2g_batch_branch_flows(g2g_batch_t b, int32_t m, const int32_t* from, const int32_t* to, const double* y4,
                           const double* base_or, const double* base_ex, const int32_t* trip_ptr,
                           const int32_t* trip_branch, double* or_amps, double* ex_amps)
{
    return guarded([&] {
        if (!b || !from || !to || !y4 || !base_or || !base_ex || !or_amps || !ex_amps)
            throw std::invalid_argument("null argument");
        if (m <= 0) throw std::invalid_argument("m must be positive");
        const int n = b->plan->nr.n;
        for (int l = 0; l < m; ++l)
            if (from[l] < -1 || from[l] >= n || to[l] < -1 || to[l] >= n)
                throw std::invalid_argument("branch bus out of range");
        const int rows = b->active_rows;
        const int T = (rows + kLanes - 1) / kLanes;
        DeviceScratch s;
        int* d_f = s.get<int>(m);
        int* d_t = s.get<int>(m);
        double2* d_y = s.get<double2>(4 * static_cast<size_t>(m));
        double* d_bo = s.get<double>(m);
        double* d_be = s.get<double>(m);
        const size_t cnt = static_cast<size_t>(T) * m * kLanes;
        double* d_or = s.get<double>(cnt);
        double* d_ex = s.get<double>(cnt);
        hipStream_t st = b->stream;
        HIPCHK(hipMemcpyAsync(d_f, from, m * sizeof(int), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d_t, to, m * sizeof(int), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d_y, y4, 4 * static_cast<size_t>(m) * sizeof(double2), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d_bo, base_or, m * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(d_be, base_ex, m * sizeof(double), hipMemcpyHostToDevice, st));
        k_branch_flows<<<blocks_for(static_cast<long long>(cnt), 256), 256, 0, st>>>(
            b->V, n, d_f, d_t, d_y, d_bo, d_be, m, T, d_or, d_ex);
        HIPCHK(hipGetLastError());
        std::vector<double> ho(cnt), he(cnt);
        HIPCHK(hipMemcpyAsync(ho.data(), d_or, cnt * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(he.data(), d_ex, cnt * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        for (int r = 0; r < rows; ++r)
            for (int l = 0; l < m; ++l) {
                const size_t i = (static_cast<size_t>(r / kLanes) * m + l) * kLanes + (r % kLanes);
                or_amps[static_cast<size_t>(r) * m + l] = ho[i];
                ex_amps[static_cast<size_t>(r) * m + l] = he[i];
            }
        // tripped branches carry no current
        if (trip_ptr && trip_branch)
            for (int r = 0; r < rows; ++r)
                for (int q = trip_ptr[r]; q < trip_ptr[r + 1]; ++q) {
                    const int l = trip_branch[q];
                    if (l < 0 || l >= m) throw std::invalid_argument("tripped branch out of range");
                    or_amps[static_cast<size_t>(r) * m + l] = 0.0;
                    ex_amps[static_cast<size_t>(r) * m + l] = 0.0;
                }
    });
}

int g2g_batch_limit_violations(g2g_batch_t b, const g2g_limit_table* t, const int32_t* trip_ptr,
                               const int32_t* trip_branch, double tol, double rel_tol, int32_t capacity,
                               g2g_violations* out)
{
    return guarded([&] {
        if (!b || !t || !out) throw std::invalid_argument("null argument");
        if (!t->from || !t->to || !t->y4 || !t->base_or || !t->base_ex || !t->limit1_ka || !t->limit2_ka ||
            !t->vn_kv || !t->vmin_kv || !t->vmax_kv)
            throw std::invalid_argument("null limit table array");
        if (!out->count || !out->truncated || !out->n_low || !out->n_high || !out->n_current || !out->element_type ||
            !out->element_id || !out->side || !out->type || !out->value || !out->limit)
            throw std::invalid_argument("null output array");
        if (capacity < 1 || capacity > G2G_MAX_VIOLATION_CAPACITY)
            throw std::invalid_argument("violation capacity must be in [1, " +
                                        std::to_string(G2G_MAX_VIOLATION_CAPACITY) + "]");
        if (!(rel_tol >= 0.0 && rel_tol < 1.0)) throw std::invalid_argument("rel_tol must be in [0, 1[");
        const int m = t->m, n = b->plan->nr.n;
        if (m < 0) throw std::invalid_argument("m must be >= 0");
        if (t->n_lines < 0 || t->n_lines > m) throw std::invalid_argument("n_lines must be in [0, m]");
        for (int l = 0; l < m; ++l)
            if (t->from[l] < -1 || t->from[l] >= n || t->to[l] < -1 || t->to[l] >= n)
                throw std::invalid_argument("branch bus out of range");
        const int rows = b->active_rows;
        if (rows <= 0) return;
        if (trip_ptr) {
            if (!trip_branch && trip_ptr[rows] > trip_ptr[0]) throw std::invalid_argument("null trip_branch");
            if (trip_ptr[0] != 0) throw std::invalid_argument("trip_ptr must start at 0");
            for (int r = 0; r < rows; ++r) {
                if (trip_ptr[r + 1] < trip_ptr[r]) throw std::invalid_argument("trip_ptr is not monotone");
                for (int q = trip_ptr[r]; q < trip_ptr[r + 1]; ++q)
                    if (trip_branch[q] < 0 || trip_branch[q] >= m)
                        throw std::invalid_argument("tripped branch out of range");
            }
        }
        const int K = capacity;
        const size_t R = static_cast<size_t>(rows);
        const size_t nrec = R * 3 * K;
        const int ntrip = trip_ptr ? trip_ptr[rows] : 0;
        DeviceScratch s;
        hipStream_t st = b->stream;
        auto up = [&](auto* host, size_t cnt) {
            using T = std::remove_const_t<std::remove_pointer_t<decltype(host)>>;
            T* d = s.get<T>(std::max<size_t>(cnt, 1));
            if (cnt) HIPCHK(hipMemcpyAsync(d, host, cnt * sizeof(T), hipMemcpyHostToDevice, st));
            return d;
        };
        ViolationArgs a{};
        a.V = b->V;
        a.n = n;
        a.rows = rows;
        a.status = b->status;
        a.res = b->res;
        a.tol = tol;
        a.up = 1.0 + rel_tol;
        a.dn = 1.0 - rel_tol;
        a.m = m;
        a.n_lines = t->n_lines;
        a.from = up(t->from, m);
        a.to = up(t->to, m);
        a.y4 = reinterpret_cast<const double2*>(up(t->y4, 8 * static_cast<size_t>(m)));
        a.base_or = up(t->base_or, m);
        a.base_ex = up(t->base_ex, m);
        a.lim1 = up(t->limit1_ka, m);
        a.lim2 = up(t->limit2_ka, m);
        a.vn_kv = up(t->vn_kv, n);
        a.vmin = up(t->vmin_kv, n);
        a.vmax = up(t->vmax_kv, n);
        if (trip_ptr) {
            a.trip_ptr = up(trip_ptr, R + 1);
            a.trip = up(trip_branch ? trip_branch : trip_ptr, static_cast<size_t>(ntrip));
        }
        a.K = K;
        a.count = s.get<int>(R);
        a.truncated = s.get<int>(R);
        a.n_type = s.get<int>(3 * R);
        a.etype = s.get<int>(nrec);
        a.eid = s.get<int>(nrec);
        a.side = s.get<int>(nrec);
        a.vtype = s.get<int>(nrec);
        a.value = s.get<double>(nrec);
        a.limit = s.get<double>(nrec);
        const size_t lds = limit_violations_lds(K);
        HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_limit_violations),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
        k_limit_violations<<<rows, kLanes, lds, st>>>(a);
        HIPCHK(hipGetLastError());
        auto down = [&](auto* host, const auto* dev, size_t cnt) {
            HIPCHK(hipMemcpyAsync(host, dev, cnt * sizeof(*host), hipMemcpyDeviceToHost, st));
        };
        std::vector<int> nt(3 * R);
        down(out->count, a.count, R);
        down(out->truncated, a.truncated, R);
        down(nt.data(), a.n_type, 3 * R);
        down(out->element_type, a.etype, nrec);
        down(out->element_id, a.eid, nrec);
        down(out->side, a.side, nrec);
        down(out->type, a.vtype, nrec);
        down(out->value, a.value, nrec);
        down(out->limit, a.limit, nrec);
        HIPCHK(hipStreamSynchronize(st));
        for (size_t r = 0; r < R; ++r) {
            out->n_low[r] = nt[3 * r];
            out->n_high[r] = nt[3 * r + 1];
            out->n_current[r] = nt[3 * r + 2];
        }
    });
}

int g2g_batch_get_v(g2g_batch_t b, double* V, int64_t ld, int32_t on_device)
{
    return guarded([&] {
        if (!b || !V) throw std::invalid_argument("null argument");
        const int n = b->plan->nr.n;
        if (ld < n) throw std::invalid_argument("ld smaller than n");
        const long long total = static_cast<long long>(b->rows) * n;
        if (on_device) {
            k_tiles_to_rows<double2><<<blocks_for(total, 256), 256, 0, b->stream>>>(
                b->V, b->rows, n, reinterpret_cast<double2*>(V), ld);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(b->stream));
        } else {
            double2* tmp = dalloc<double2>(static_cast<size_t>(b->rows) * ld);
            k_tiles_to_rows<double2><<<blocks_for(total, 256), 256, 0, b->stream>>>(b->V, b->rows, n, tmp, ld);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(V, tmp, static_cast<size_t>(b->rows) * ld * sizeof(double2),
                                  hipMemcpyDeviceToHost, b->stream));
            HIPCHK(hipStreamSynchronize(b->stream));
            dfree(tmp);
        }
    });
}

int g2g_batch_get_v_rows(g2g_batch_t b, double* V, int64_t ld, int32_t rows, int32_t row0, int32_t on_device)
{
    return guarded([&] {
        if (!b || !V) throw std::invalid_argument("null argument");
        check_rows(b, rows, row0);
        const int n = b->plan->nr.n;
        if (ld < n) throw std::invalid_argument("ld smaller than n");
        if (rows == 0) return;
        const long long total = static_cast<long long>(rows) * n;
        double2* dst = reinterpret_cast<double2*>(V);
        DeviceScratch tmp;
        if (!on_device) dst = tmp.get<double2>(static_cast<size_t>(rows) * ld);
        k_slots_range_to_rows<<<blocks_for(total, 256), 256, 0, b->stream>>>(b->V, row0, rows, n, dst, ld);
        HIPCHK(hipGetLastError());
        if (!on_device)
            HIPCHK(hipMemcpyAsync(V, dst, static_cast<size_t>(rows) * ld * sizeof(double2), hipMemcpyDeviceToHost,
                                  b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
    });
}

int g2g_batch_get_results(g2g_batch_t b, int32_t* status, int32_t* iters, double* residual)
{
    return guarded([&] {
        if (!b) throw std::invalid_argument("null batch");
        HIPCHK(hipStreamSynchronize(b->stream));
        const size_t r = static_cast<size_t>(b->rows);
        if (status) HIPCHK(hipMemcpy(status, b->status, r * sizeof(int), hipMemcpyDeviceToHost));
        if (iters) HIPCHK(hipMemcpy(iters, b->iters, r * sizeof(int), hipMemcpyDeviceToHost));
        if (residual) HIPCHK(hipMemcpy(residual, b->res, r * sizeof(double), hipMemcpyDeviceToHost));
    });
}

int g2g_batch_eval(g2g_batch_t b, double* F, double* J, double* residual)
{
    return guarded([&] {
        if (!b) throw std::invalid_argument("null batch");
        const NrPlan& nr = b->plan->nr;
        const size_t rp = static_cast<size_t>(b->rows_pad);
        // evaluate every row as active without touching the solve's statuses
        int* saved = dalloc<int>(rp);
        int* saved_it = dalloc<int>(rp);
        HIPCHK(hipMemcpyAsync(saved, b->status, rp * sizeof(int), hipMemcpyDeviceToDevice, b->stream));
        HIPCHK(hipMemcpyAsync(saved_it, b->iters, rp * sizeof(int), hipMemcpyDeviceToDevice, b->stream));
        reset_rows(b);
        k_fill_int<<<1, 64, 0, b->stream>>>(b->n_active, 0, 2);
        if (b->precision == G2G_MIXED) launch_mismatch<float>(b, J != nullptr, 0, 1 << 30, 1e-300, 0);
        else launch_mismatch<double>(b, J != nullptr, 0, 1 << 30, 1e-300, 0);
        const size_t rs = b->real_size();
        std::vector<unsigned char> host(rp * b->a_stride * rs);
        HIPCHK(hipMemcpyAsync(host.data(), b->A, host.size(), hipMemcpyDeviceToHost, b->stream));
        if (residual) HIPCHK(hipMemcpyAsync(residual, b->res, b->rows * sizeof(double), hipMemcpyDeviceToHost, b->stream));
        HIPCHK(hipMemcpyAsync(b->status, saved, rp * sizeof(int), hipMemcpyDeviceToDevice, b->stream));
        HIPCHK(hipMemcpyAsync(b->iters, saved_it, rp * sizeof(int), hipMemcpyDeviceToDevice, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
        dfree(saved);
        dfree(saved_it);
        auto at = [&](int r, int pos) -> double {
            const size_t i = ((static_cast<size_t>(r / kLanes) * b->a_stride) + pos) * kLanes + (r % kLanes);
            if (rs == sizeof(float)) return reinterpret_cast<const float*>(host.data())[i];
            return reinterpret_cast<const double*>(host.data())[i];
        };
        for (int r = 0; r < b->rows; ++r) {
            if (F)
                for (int e = 0; e < nr.d; ++e) F[static_cast<size_t>(r) * nr.d + e] = at(r, nr.lu.f_pos[e]);
            if (J)
                for (int q = 0; q < nr.nJ; ++q) J[static_cast<size_t>(r) * nr.nJ + q] = at(r, nr.lu.a_pos_of_entry[q]);
        }
    });
}

int g2g_batch_get_ext_state(g2g_batch_t b, double* out)
{
    return guarded([&] {
        if (!b || !out) throw std::invalid_argument("null argument");
        const int nxs = b->plan->nr.xt.n_xs;
        if (!b->xs || nxs == 0) return;
        std::vector<double> h(static_cast<size_t>(b->rows_pad) * nxs);
        HIPCHK(hipMemcpyAsync(h.data(), b->xs, h.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
        for (int r = 0; r < b->rows; ++r)
            for (int q = 0; q < nxs; ++q)
                out[static_cast<size_t>(r) * nxs + q] = h[(static_cast<size_t>(r / kLanes) * nxs + q) * kLanes + r % kLanes];
    });
}

int g2g_batch_set_ext_state(g2g_batch_t b, const double* in)
{
    return guarded([&] {
        if (!b || !in) throw std::invalid_argument("null argument");
        const int nxs = b->plan->nr.xt.n_xs;
        if (!b->xs || nxs == 0) return;
        std::vector<double> h(static_cast<size_t>(b->rows_pad) * nxs, 0.0);
        for (int r = 0; r < b->rows; ++r)
            for (int q = 0; q < nxs; ++q)
                h[(static_cast<size_t>(r / kLanes) * nxs + q) * kLanes + r % kLanes] = in[static_cast<size_t>(r) * nxs + q];
        HIPCHK(hipMemcpyAsync(b->xs, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
    });
}

int g2g_batch_solve_adjoint(g2g_batch_t b, const double* rhs, double* x)
{
    return guarded([&] {
        if (!b || !rhs || !x) throw std::invalid_argument("null argument");
        if (b->plan->ext()) throw std::invalid_argument("the adjoint is not available with extension rows");
        if (b->precision == G2G_MIXED) adjoint_impl<float>(b, rhs, x);
        else adjoint_impl<double>(b, rhs, x);
    });
}

int g2g_lu_solve_batch(g2g_plan_t P, int32_t rows, int32_t precision, const double* vals, const double* rhs,
                       double* x, double pivot_eps, int32_t ir_steps, int32_t* clamped, void* stream)
{
    return guarded([&] {
        if (!P || !vals || !rhs || !x) throw std::invalid_argument("null argument");
        if (rows <= 0) throw std::invalid_argument("rows must be positive");
        if (ir_steps < 0) throw std::invalid_argument("ir_steps must be >= 0");
        if (precision == G2G_MIXED) raw_solve<float>(P, rows, vals, rhs, x, pivot_eps, ir_steps, clamped, stream);
        else if (precision == G2G_FP64) raw_solve<double>(P, rows, vals, rhs, x, pivot_eps, ir_steps, clamped, stream);
        else throw std::invalid_argument("unknown precision");
    });
}

int g2g_lu_solve_batch_noalloc(g2g_plan_t P, int32_t rows, int32_t precision, const double* d_vals,
                               const double* d_rhs, double* d_x, double pivot_eps, int32_t ir_steps,
                               int32_t* d_clamped, void* stream)
{
    return guarded([&] {
        if (!P || !d_vals || !d_rhs || !d_x) throw std::invalid_argument("null argument");
        if (rows <= 0) throw std::invalid_argument("rows must be positive");
        if (ir_steps < 0) throw std::invalid_argument("ir_steps must be >= 0");
        if (P->grid) throw std::invalid_argument("a raw matrix plan (g2g_lu_plan_create) is needed");
        if (precision == G2G_MIXED)
            raw_solve_noalloc<float>(P, rows, d_vals, d_rhs, d_x, pivot_eps, ir_steps, d_clamped, stream);
        else if (precision == G2G_FP64)
            raw_solve_noalloc<double>(P, rows, d_vals, d_rhs, d_x, pivot_eps, ir_steps, d_clamped, stream);
        else throw std::invalid_argument("unknown precision");
    });
}

int g2g_batch_set_bus_mask(g2g_batch_t b, const int32_t* row_ptr, const int32_t* buses, int32_t rows)
{
    return guarded([&] {
        if (!b) throw std::invalid_argument("null batch");
        const NrPlan& nr = b->plan->nr;
        const size_t rp = static_cast<size_t>(b->rows_pad);
        if (!row_ptr) {  // masking off
            dfree(b->mask);
            b->mask = nullptr;
            return;
        }
        if (b->streaming) throw std::invalid_argument("streamed rows take no bus mask");
        if (b->plan->ext()) throw std::invalid_argument("bus masks are not available with extension rows");
        if (!buses && row_ptr[rows] > row_ptr[0]) throw std::invalid_argument("null buses");
        check_rows(b, rows, 0);
        std::vector<unsigned char> h(rp * nr.n, 0);
        for (int r = 0; r < rows; ++r) {
            if (row_ptr[r + 1] < row_ptr[r]) throw std::invalid_argument("row_ptr is not monotone");
            for (int q = row_ptr[r]; q < row_ptr[r + 1]; ++q) {
                const int i = buses[q];
                if (i < 0 || i >= nr.n) throw std::invalid_argument("masked bus out of range");
                if (i == nr.slack) throw std::invalid_argument("the slack bus cannot be masked");
                h[((static_cast<size_t>(r / kLanes) * nr.n) + i) * kLanes + (r % kLanes)] = 1;
            }
        }
        if (!b->mask) b->mask = dalloc<unsigned char>(rp * nr.n);
        HIPCHK(hipMemcpyAsync(b->mask, h.data(), h.size(), hipMemcpyHostToDevice, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
    });
}

int g2g_batch_stream_begin(g2g_batch_t b)
{
    return guarded([&] {
        if (b) b->v_bcast = false;  // V changes (or is handed out)
        if (!b) throw std::invalid_argument("null batch");
        if (b->mask) throw std::invalid_argument("streamed rows take no bus mask: clear it first");
        if (b->plan->ext()) throw std::invalid_argument("streamed rows are not available with extension rows");
        const size_t rp = static_cast<size_t>(b->rows_pad);
        if (!b->it0) b->it0 = dalloc<int>(rp);
        k_fill_int<<<blocks_for(rp, 256), 256, 0, b->stream>>>(b->it0, 0, rp);
        k_fill_int<<<blocks_for(rp, 256), 256, 0, b->stream>>>(b->status, kNotSimulated, rp);
        k_fill_int<<<blocks_for(rp, 256), 256, 0, b->stream>>>(b->iters, 0, rp);
        k_count_tiles<<<b->T, kLanes, 0, b->stream>>>(b->status, b->rows_pad, b->tile_active);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(b->stream));
        if (!b->clock_dev) b->clock_dev = dalloc<int>(1);
        k_fill_int<<<1, 64, 0, b->stream>>>(b->clock_dev, 0, 1);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(b->stream));
        b->clock = 0;
        b->active_rows = b->rows;
        b->streaming = true;
    });
}

}  // extern "C"

namespace {
// staging buffer i of the batch with room for n elements of T (grown, after
// the stream drained, when too small)
template <typename T>
T* stage_buf(g2g_batch_s* b, int i, size_t n)
{
    const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    if (bytes > b->stage_cap[i]) {
        const size_t cap = std::max(bytes, 2 * b->stage_cap[i]);
        HIPCHK(hipStreamSynchronize(b->stream));
        dfree(b->stage[i]);
        b->stage[i] = nullptr;
        b->stage_cap[i] = 0;
        b->stage[i] = dalloc<unsigned char>(cap);
        b->stage_cap[i] = cap;
    }
    return static_cast<T*>(b->stage[i]);
}
}  // namespace

extern "C" {

int g2g_batch_stream_fill(g2g_batch_t b, int32_t n, const int32_t* slots, const double* S, int64_t lds,
                          const double* V0, int64_t ldv, int32_t on_device, const int32_t* edit_ptr,
                          const int32_t* edit_ent, const double* edit_delta)
{
    return guarded([&] {
        if (b) b->v_bcast = false;  // V changes (or is handed out)
        if (!b || (n > 0 && (!slots || !S || !V0))) throw std::invalid_argument("null argument");
        if (!b->streaming) throw std::invalid_argument("call g2g_batch_stream_begin first");
        if (n <= 0) return;
        const NrPlan& nr = b->plan->nr;
        for (int r = 0; r < n; ++r)
            if (slots[r] < 0 || slots[r] >= b->rows) throw std::invalid_argument("slot outside the batch");
        if (edit_ptr && !b->per_row_y) throw std::invalid_argument("the batch was created without per-row Ybus values");
        hipStream_t st = b->stream;
        int* d_slots = stage_buf<int>(b, kStSlots, n);
        HIPCHK(hipMemcpyAsync(d_slots, slots, n * sizeof(int), hipMemcpyHostToDevice, st));
        auto rows_src = [&](const double* src, long long ld, int buf) -> const double2* {
            if (on_device) return reinterpret_cast<const double2*>(src);
            const size_t cnt = ld == 0 ? static_cast<size_t>(nr.n) : static_cast<size_t>(n - 1) * ld + nr.n;
            double2* d = stage_buf<double2>(b, buf, cnt);
            HIPCHK(hipMemcpyAsync(d, src, cnt * sizeof(double2), hipMemcpyHostToDevice, st));
            return d;
        };
        const unsigned g = blocks_for(static_cast<long long>(n) * nr.n, 256);
        k_rows_to_slots<double2><<<g, 256, 0, st>>>(rows_src(S, lds, kStS), lds, n, nr.n, d_slots, b->S);
        k_rows_to_slots<double2><<<g, 256, 0, st>>>(rows_src(V0, ldv, kStV), ldv, n, nr.n, d_slots, b->V);
        if (b->per_row_y) {  // the slots' edits (none without edit lists), then the tiles' flags
            const int* d_rp = nullptr;
            const int* d_en = nullptr;
            const double2* d_de = nullptr;
            if (edit_ptr && edit_ptr[n] > 0) {
                for (int r = 0; r < n; ++r)
                    if (edit_ptr[r + 1] < edit_ptr[r]) throw std::invalid_argument("edit_ptr is not monotone");
                for (int q = 0; q < edit_ptr[n]; ++q)
                    if (edit_ent[q] < 0 || edit_ent[q] >= nr.nY) throw std::invalid_argument("edit entry out of range");
                const EditLists el(edit_ptr, edit_ent, edit_delta, n);
                const int ne = el.rp[n];
                ensure_edit_cap(b, el.most, true);
                int* rp_ = stage_buf<int>(b, kStEdPtr, n + 1);
                int* en_ = stage_buf<int>(b, kStEdEnt, ne);
                double2* de_ = stage_buf<double2>(b, kStEdDelta, ne);
                HIPCHK(hipMemcpyAsync(rp_, el.rp.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice, st));
                HIPCHK(hipMemcpyAsync(en_, el.ent.data(), ne * sizeof(int), hipMemcpyHostToDevice, st));
                HIPCHK(hipMemcpyAsync(de_, el.val.data(), ne * sizeof(double2), hipMemcpyHostToDevice, st));
                d_rp = rp_;
                d_en = en_;
                d_de = de_;
            }
            k_set_row_edits<<<blocks_for(n, 256), 256, 0, st>>>(d_rp, d_en, d_de, n, d_slots, b->ed_cap, b->ed_n,
                                                                b->ed_ent, b->ed_val);
            HIPCHK(hipGetLastError());
            edit_flags(b, b->ed_n, b->ed_ent, b->rows_pad, b->ed_flag);
        }
        const unsigned gs = blocks_for(n, 256);
        k_set_slots<<<gs, 256, 0, st>>>(d_slots, n, b->status, kActive);
        k_set_slots<<<gs, 256, 0, st>>>(d_slots, n, b->it0, b->clock);
        k_set_slots<<<gs, 256, 0, st>>>(d_slots, n, b->iters, 0);
        k_count_tiles<<<b->T, kLanes, 0, st>>>(b->status, b->rows_pad, b->tile_active);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));  // the caller's host rows may be reused on return
    });
}

int g2g_batch_stream_round(g2g_batch_t b, int32_t k, int32_t max_iter, double tol)
{
    return guarded([&] {
        if (b) b->v_bcast = false;  // V changes (or is handed out)
        if (!b) throw std::invalid_argument("null batch");
        if (!b->streaming) throw std::invalid_argument("call g2g_batch_stream_begin first");
        if (k <= 0 || max_iter < 0) throw std::invalid_argument("k must be positive and max_iter >= 0");
        if (b->precision == G2G_MIXED) stream_round<float>(b, k, max_iter, tol);
        else stream_round<double>(b, k, max_iter, tol);
        check_stream_error(b);
    });
}

int g2g_batch_set_stream_graph(g2g_batch_t b, int32_t on)
{
    return guarded([&] {
        if (!b) throw std::invalid_argument("null batch");
        b->stream_graph = on != 0;
    });
}

int g2g_batch_stream_graph_info(g2g_batch_t b, int32_t* on, int64_t* replays)
{
    return guarded([&] {
        if (!b) throw std::invalid_argument("null batch");
        if (on) *on = b->stream_graph ? 1 : 0;
        if (replays) *replays = b->graph_replays;
    });
}

int g2g_batch_stream_warm(g2g_batch_t b, g2g_mlp_t mlp, int32_t n, const int32_t* slots, const uint8_t* mask,
                          int32_t n_mask)
{
    return guarded([&] {
        if (b) b->v_bcast = false;  // V changes (or is handed out)
        if (!b || !mlp || (n > 0 && !slots)) throw std::invalid_argument("null argument");
        if (!b->streaming) throw std::invalid_argument("call g2g_batch_stream_begin first");
        if (n_mask < 0 || (n_mask > 0 && !mask)) throw std::invalid_argument("mask missing");
        if (n <= 0) return;
        const int N = b->plan->nr.n;
        for (int r = 0; r < n; ++r)
            if (slots[r] < 0 || slots[r] >= b->rows) throw std::invalid_argument("slot outside the batch");
        hipStream_t st = b->stream;
        const size_t kp = static_cast<size_t>(n + kLanes - 1) / kLanes * kLanes;
        int* d_slots = stage_buf<int>(b, kStSlots, n);
        double2* Sg = stage_buf<double2>(b, kStSg, kp * N);
        double2* Vg = stage_buf<double2>(b, kStVg, kp * N);
        HIPCHK(hipMemcpyAsync(d_slots, slots, n * sizeof(int), hipMemcpyHostToDevice, st));
        unsigned char* d_mask = nullptr;
        if (n_mask > 0) {
            d_mask = stage_buf<unsigned char>(b, kStMask, static_cast<size_t>(n) * n_mask);
            HIPCHK(hipMemcpyAsync(d_mask, mask, static_cast<size_t>(n) * n_mask, hipMemcpyHostToDevice, st));
        }
        // the slots' injections into dense tiles, the MLP on them, the start
        // voltages back into the slots
        const unsigned g = blocks_for(static_cast<long long>(n) * N, 256);
        k_move_rows<double2><<<g, 256, 0, st>>>(b->S, Sg, d_slots, n, N, 1);
        // the slots' start V: the fixed buses keep theirs
        k_move_rows<double2><<<g, 256, 0, st>>>(b->V, Vg, d_slots, n, N, 1);
        HIPCHK(hipGetLastError());
        mlp_warm_start_tiles(mlp, Sg, n, N, d_mask, n_mask, Vg, st);
        k_move_rows<double2><<<g, 256, 0, st>>>(Vg, b->V, d_slots, n, N, 0);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
    });
}

int g2g_batch_stream_take(g2g_batch_t b, int32_t n, const int32_t* slots, double* V, int64_t ldv, int32_t on_device)
{
    return guarded([&] {
        if (!b || (n > 0 && !slots)) throw std::invalid_argument("null argument");
        if (!b->streaming) throw std::invalid_argument("call g2g_batch_stream_begin first");
        if (n <= 0) return;
        const NrPlan& nr = b->plan->nr;
        for (int r = 0; r < n; ++r)
            if (slots[r] < 0 || slots[r] >= b->rows) throw std::invalid_argument("slot outside the batch");
        hipStream_t st = b->stream;
        int* d_slots = stage_buf<int>(b, kStSlots, n);
        HIPCHK(hipMemcpyAsync(d_slots, slots, n * sizeof(int), hipMemcpyHostToDevice, st));
        if (V) {
            double2* dst = on_device ? reinterpret_cast<double2*>(V)
                                     : stage_buf<double2>(b, kStV, static_cast<size_t>(n) * nr.n);
            const long long ld = on_device ? ldv : nr.n;
            k_slots_to_rows<double2><<<blocks_for(static_cast<long long>(n) * nr.n, 256), 256, 0, st>>>(
                b->V, n, nr.n, d_slots, dst, ld);
            HIPCHK(hipGetLastError());
            if (!on_device) {
                if (ldv == nr.n) {  // contiguous host rows: one copy (direct DMA when they are pinned)
                    HIPCHK(hipMemcpyAsync(V, dst, static_cast<size_t>(n) * nr.n * sizeof(double2),
                                          hipMemcpyDeviceToHost, st));
                    HIPCHK(hipStreamSynchronize(st));
                } else {
                    std::vector<double2> h(static_cast<size_t>(n) * nr.n);
                    HIPCHK(hipMemcpyAsync(h.data(), dst, h.size() * sizeof(double2), hipMemcpyDeviceToHost, st));
                    HIPCHK(hipStreamSynchronize(st));
                    for (int r = 0; r < n; ++r)
                        std::memcpy(V + 2 * static_cast<size_t>(r) * ldv, h.data() + static_cast<size_t>(r) * nr.n,
                                    nr.n * sizeof(double2));
                }
            }
        }
        k_set_slots<<<blocks_for(n, 256), 256, 0, st>>>(d_slots, n, b->status, kNotSimulated);
        k_count_tiles<<<b->T, kLanes, 0, st>>>(b->status, b->rows_pad, b->tile_active);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
    });
}

int g2g_batch_device_arrays(g2g_batch_t b, void** V_tiles, void** S_tiles)
{
    return guarded([&] {
        if (b) b->v_bcast = false;  // V changes (or is handed out)
        if (!b) throw std::invalid_argument("null batch");
        if (V_tiles) *V_tiles = b->V;
        if (S_tiles) *S_tiles = b->S;
    });
}

}  // extern "C"
