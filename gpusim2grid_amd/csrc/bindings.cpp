// _gpusim2grid: the reference's pybind11 module names (src/_cpp/python_bindings.cpp)
// as a C++ layer over the C ABI of include/gpusim2grid_amd.h.
//
// Sessions own a plan (the grid's symbolic layer) and device batches; every
// solve runs the HIP kernels behind g2g_batch_solve.  The array ("tuple")
// mode of the reference is what is mirrored: the bare [pvpq | pq] system
// (README.md:51-58), no lightsim2grid bridge.
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <limits>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "gpusim2grid_amd.h"

namespace py = pybind11;
using cplx = std::complex<double>;
using CArr = py::array_t<cplx, py::array::c_style | py::array::forcecast>;
using DArr = py::array_t<double, py::array::c_style | py::array::forcecast>;
using IArr = py::array_t<int32_t, py::array::c_style | py::array::forcecast>;

namespace {

void check(int rc)
{
    if (rc == G2G_OK) return;
    if (rc == G2G_ERR_INVALID) throw py::value_error(g2g_last_error());
    throw std::runtime_error(g2g_last_error());
}

void hipchk(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// A device buffer shared by a session and the DLPack tensors aliasing it: a
// session that needs a larger one allocates anew, so an exported tensor
// never points to freed memory (it sees no further runs, though).
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    int device = 0;
    explicit DevBuf(size_t b) : bytes(b)
    {
        hipchk(hipGetDevice(&device), "hipGetDevice");
        hipchk(hipMalloc(&p, b ? b : 1), "hipMalloc");
    }
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    cplx* c() const { return static_cast<cplx*>(p); }
};
using DevBufPtr = std::shared_ptr<DevBuf>;

void ensure(DevBufPtr& b, size_t bytes)
{
    if (!b || b->bytes < bytes) b = std::make_shared<DevBuf>(bytes);
}

void upload(DevBufPtr& b, const void* host, size_t bytes)
{
    ensure(b, bytes);
    if (bytes) hipchk(hipMemcpy(b->p, host, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
}

// pinned host staging of downloads (grow-only)
struct Pinned {
    void* p = nullptr;
    size_t bytes = 0;
    Pinned() = default;
    Pinned(const Pinned&) = delete;
    Pinned& operator=(const Pinned&) = delete;
    ~Pinned()
    {
        if (p) (void)hipHostFree(p);
    }
    void* get(size_t b)
    {
        if (b > bytes) {
            if (p) (void)hipHostFree(p);
            p = nullptr;
            bytes = 0;
            hipchk(hipHostMalloc(&p, b, hipHostMallocDefault), "hipHostMalloc");
            bytes = b;
        }
        return p;
    }
};

// DLPack (the dlpack.h v0.8 ABI): a managed tensor of complex128 values on a
// ROCm device aliasing a session buffer
struct DLDevice_ {
    int32_t device_type, device_id;
};
struct DLDataType_ {
    uint8_t code, bits;
    uint16_t lanes;
};
struct DLTensor_ {
    void* data;
    DLDevice_ device;
    int32_t ndim;
    DLDataType_ dtype;
    int64_t* shape;
    int64_t* strides;
    uint64_t byte_offset;
};
struct DLManagedTensor_ {
    DLTensor_ dl_tensor;
    void* manager_ctx;
    void (*deleter)(DLManagedTensor_*);
};
constexpr int32_t kDLROCM_ = 10;
constexpr uint8_t kDLComplex_ = 5;

struct AliasCtx {
    DevBufPtr buf;
    int64_t shape[2];
};

// a "dltensor" capsule aliasing buf as (rows, n) (ndim 2) or (n,) (ndim 1);
// the tensor keeps the buffer alive, and the session's next run overwrites it
py::object alias_capsule(const DevBufPtr& buf, int64_t rows, int64_t n, int ndim)
{
    if (static_cast<size_t>(rows * n) * sizeof(cplx) > buf->bytes) throw std::logic_error("capsule beyond its buffer");
    auto* ctx = new AliasCtx{buf, {rows, n}};
    auto* mt = new DLManagedTensor_{};
    mt->dl_tensor.data = buf->p;
    mt->dl_tensor.device = DLDevice_{kDLROCM_, buf->device};
    mt->dl_tensor.ndim = ndim;
    mt->dl_tensor.dtype = DLDataType_{kDLComplex_, 128, 1};
    mt->dl_tensor.shape = ndim == 2 ? ctx->shape : ctx->shape + 1;
    mt->dl_tensor.strides = nullptr;
    mt->dl_tensor.byte_offset = 0;
    mt->manager_ctx = ctx;
    mt->deleter = [](DLManagedTensor_* self) {
        delete static_cast<AliasCtx*>(self->manager_ctx);
        delete self;
    };
    PyObject* cap = PyCapsule_New(mt, "dltensor", [](PyObject* c) {
        // an importer renames the capsule and owns the tensor; otherwise it is ours
        if (PyCapsule_IsValid(c, "dltensor")) {
            auto* m = static_cast<DLManagedTensor_*>(PyCapsule_GetPointer(c, "dltensor"));
            if (m && m->deleter) m->deleter(m);
        }
    });
    if (!cap) {
        mt->deleter(mt);
        throw py::error_already_set();
    }
    return py::reinterpret_steal<py::object>(cap);
}

double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ---- enums (python_bindings.cpp:1294-1361,1576-1654; row_status.hpp:28-43) ----
enum class RowStatus : int { Converged = 0, MaxIter = 1, Diverged = 2, NotSimulated = 3 };
enum class BatchScheduling : int { Chunked = 0, Continuous = 1 };
enum class LinearSolver : int { Cudss = 0, OwnLu = 1 };
enum class ReorderingAlg : int { Default = 0, BtfColamd, Colamd, Amd, NestedDissection, None };
enum class MatchingAlg : int { None = 0, MaxDiagCount, MaxMinDiag, MaxMinDiagAlt, MaxDiagSum, MaxDiagProduct, Auto };
enum class PivotEpsilonAlg : int { Default = 0, Scaled, Static };
enum class ContingencySolverType : int { DirectRefactorEvery = 0, DirectIter0Only, DirectRefactorEveryN };
enum class Precision : int { Fp64 = G2G_FP64, Mixed = G2G_MIXED };

// ---- timings (timing_utils.hpp:38-53,131-171,348-400) ------------------------
struct TimingEntry {
    double gpu_ms = 0.0;
    double wall_ms = 0.0;
};

struct AcPfTimings {
    double t_init_ms = 0.0;     // plan build + device upload
    double t_analyze_ms = 0.0;  // symbolic analysis (ordering, programs)
    TimingEntry t_mismatch;     // fused SpMV + F + J fill + residual
    TimingEntry t_refactorize;  // factorization + forward substitution
    TimingEntry t_solve;        // backward substitution + V update
    double t_total_ms = 0.0;
    double t_copy_v_to_host_ms = 0.0;
    int nb_iter = 0;
    int n_refactorize = 0;
    bool converged = false;
    double residual = 0.0;
};

struct BatchTimings {
    double t_base_case_ms = 0.0;
    double t_analysis_ms = 0.0;
    double t_alloc_ms = 0.0;
    double t_preprocess_ms = 0.0;   // outage edits, connectivity filter
    TimingEntry t_mismatch;
    TimingEntry t_refactorize;
    TimingEntry t_solve;
    TimingEntry t_warm_start;
    double t_run_ms = 0.0;          // wall time of the last run()
    double t_copy_V_to_host_ms = 0.0;
    double t_violations_ms = 0.0;   // limit checks (compute_limit_violations)
    int n_chunks = 0;
    int n_rows = 0;
    int n_converged = 0;
    int n_requeued = 0;
    long long rows_factored = 0;
    double total_ms() const { return t_run_ms; }
};

// ---- grid core ------------------------------------------------------------
struct Csr {
    int n = 0;
    std::vector<int32_t> p, i;
    std::vector<double> x;  // interleaved complex
};

Csr csr_from(py::object Y)
{
    py::object sp = py::module_::import("scipy.sparse");
    py::object C = sp.attr("csr_matrix")(Y);
    C.attr("sum_duplicates")();
    C.attr("sort_indices")();
    py::tuple shape = C.attr("shape");
    const int n = shape[0].cast<int>();
    if (shape[1].cast<int>() != n) throw py::value_error("Ybus must be square");
    Csr c;
    c.n = n;
    IArr ip = C.attr("indptr").cast<IArr>();
    IArr ii = C.attr("indices").cast<IArr>();
    CArr xx = C.attr("data").cast<CArr>();
    c.p.assign(ip.data(), ip.data() + ip.size());
    c.i.assign(ii.data(), ii.data() + ii.size());
    c.x.resize(2 * static_cast<size_t>(xx.size()));
    std::memcpy(c.x.data(), xx.data(), c.x.size() * sizeof(double));
    return c;
}

std::vector<int32_t> ivec(const IArr& a) { return std::vector<int32_t>(a.data(), a.data() + a.size()); }

struct Grid {
    Csr Y;
    std::vector<int32_t> pv, pq;
    std::vector<int32_t> fixed;  // buses outside pv and pq (no unknowns)
    int slack = -1;
    g2g_plan_t plan = nullptr;
    g2g_plan_info info{};
    double t_analyze_ms = 0.0;

    int n_xs = 0;  // extension state values per row

    Grid(py::object Ybus, const IArr& slack_ids, const IArr& pv_, const IArr& pq_, py::object ext = py::none())
    {
        Y = csr_from(Ybus);
        pv = ivec(pv_);
        pq = ivec(pq_);
        // array mode (acpf_nr.cu:644-647,1732-1733): the solve reads no slack
        // list; every bus in neither pv nor pq keeps its start V.  The angle
        // anchor is the first slack id when it is such a bus, else the first one.
        std::vector<char> has_unknowns(static_cast<size_t>(Y.n), 0);
        for (const auto* v : {&pv, &pq})
            for (int32_t b : *v)
                if (b >= 0 && b < Y.n) has_unknowns[b] = 1;
        for (int b = 0; b < Y.n; ++b)
            if (!has_unknowns[b]) fixed.push_back(b);
        if (fixed.empty()) throw py::value_error("every bus is pv or pq: the system has no fixed (slack) bus");
        slack = fixed[0];
        if (slack_ids.size() > 0) {
            const int s0 = slack_ids.data()[0];
            if (s0 >= 0 && s0 < Y.n && !has_unknowns[s0]) slack = s0;
        }
        const double t0 = now_ms();
        if (ext.is_none()) {
            check(g2g_plan_create(Y.n, Y.p.data(), Y.i.data(), Y.x.data(), static_cast<int32_t>(pv.size()),
                                  pv.data(), static_cast<int32_t>(pq.size()), pq.data(), slack, &plan));
        } else {
            // extension rows: a dict {"slack_weights": [n] | None, "hvdc": {bus1, bus2, status, p0, k, lf1,
            // lf2, r, pmax12, pmax21} | None, "vc": {reg_bus, v_set, ptr, bus, slope, weight} | None}
            py::dict d = ext.cast<py::dict>();
            std::vector<DArr> keep_d;
            std::vector<IArr> keep_i;
            auto darr = [&](py::handle h) { keep_d.push_back(h.cast<DArr>()); return keep_d.back().data(); };
            auto iarr = [&](py::handle h) { keep_i.push_back(h.cast<IArr>()); return keep_i.back().data(); };
            g2g_extensions x{};
            if (d.contains("slack_weights") && !d["slack_weights"].is_none()) {
                x.slack_weights = darr(d["slack_weights"]);
                if (keep_d.back().size() != Y.n) throw py::value_error("slack_weights must have n_bus entries");
            }
            if (d.contains("hvdc") && !d["hvdc"].is_none()) {
                py::dict h = d["hvdc"].cast<py::dict>();
                x.hvdc_bus1 = iarr(h["bus1"]);
                x.n_hvdc = static_cast<int32_t>(keep_i.back().size());
                auto hi = [&](const char* k) {
                    const int32_t* q = iarr(h[k]);
                    if (keep_i.back().size() != x.n_hvdc) throw py::value_error("hvdc arrays differ in length");
                    return q;
                };
                auto hd = [&](const char* k) {
                    const double* q = darr(h[k]);
                    if (keep_d.back().size() != x.n_hvdc) throw py::value_error("hvdc arrays differ in length");
                    return q;
                };
                x.hvdc_bus2 = hi("bus2");
                x.hvdc_status = hi("status");
                x.hvdc_p0 = hd("p0");
                x.hvdc_k = hd("k");
                x.hvdc_lf1 = hd("lf1");
                x.hvdc_lf2 = hd("lf2");
                x.hvdc_r = hd("r");
                x.hvdc_pmax12 = hd("pmax12");
                x.hvdc_pmax21 = hd("pmax21");
            }
            if (d.contains("vc") && !d["vc"].is_none()) {
                py::dict v = d["vc"].cast<py::dict>();
                x.vc_reg_bus = iarr(v["reg_bus"]);
                x.n_vc_groups = static_cast<int32_t>(keep_i.back().size());
                x.vc_v_set = darr(v["v_set"]);
                if (keep_d.back().size() != x.n_vc_groups) throw py::value_error("v_set needs one entry per group");
                x.vc_ptr = iarr(v["ptr"]);
                if (keep_i.back().size() != x.n_vc_groups + 1) throw py::value_error("ptr needs groups + 1 entries");
                const int nc = x.vc_ptr[x.n_vc_groups];
                x.vc_bus = iarr(v["bus"]);
                if (keep_i.back().size() != nc) throw py::value_error("bus needs one entry per controller");
                x.vc_slope = darr(v["slope"]);
                if (keep_d.back().size() != nc) throw py::value_error("slope needs one entry per controller");
                x.vc_weight = darr(v["weight"]);
                if (keep_d.back().size() != nc) throw py::value_error("weight needs one entry per controller");
            }
            check(g2g_plan_create_ext(Y.n, Y.p.data(), Y.i.data(), Y.x.data(), static_cast<int32_t>(pv.size()),
                                      pv.data(), static_cast<int32_t>(pq.size()), pq.data(), slack, &x, &plan));
        }
        t_analyze_ms = now_ms() - t0;
        check(g2g_plan_get_info(plan, &info));
        n_xs = info.n_xs;
    }
    ~Grid() { g2g_plan_destroy(plan); }
    Grid(const Grid&) = delete;
    Grid& operator=(const Grid&) = delete;

    int n() const { return Y.n; }
    std::vector<int32_t> array(int which) const
    {
        int64_t len = 0;
        check(g2g_plan_get_array(plan, which, nullptr, 0, &len));
        std::vector<int32_t> out(static_cast<size_t>(len));
        check(g2g_plan_get_array(plan, which, out.data(), len, &len));
        return out;
    }
    // CSR position of Ybus entry (r, c), -1 if absent
    int64_t pos(int r, int c) const
    {
        const auto b = Y.i.begin() + Y.p[r], e = Y.i.begin() + Y.p[r + 1];
        const auto it = std::lower_bound(b, e, c);
        return (it != e && *it == c) ? static_cast<int64_t>(it - Y.i.begin()) : -1;
    }
};

// device batch owned by a session
struct Batch {
    g2g_batch_t h = nullptr;
    int rows = 0;
    int precision = G2G_FP64;
    int per_row_y = 0;
    Batch(const Grid& g, int rows_, int precision_, int per_row_y_)
        : rows(rows_), precision(precision_), per_row_y(per_row_y_)
    {
        check(g2g_batch_create(g.plan, rows, precision, per_row_y, nullptr, &h));
    }
    ~Batch() { g2g_batch_destroy(h); }
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;
};

void add(TimingEntry& t, double gpu_ms) { t.gpu_ms += gpu_ms; t.wall_ms += gpu_ms; }

// refinement rounds and step scaling of a session's solves (ir_n_steps,
// scaling_max_voltage_change, max_dVa, max_dVm: python_bindings.cpp:1369-1560)
struct StepControl {
    int ir_steps = 0;
    bool scaling = false;
    double max_dva = 0.5, max_dvm = 0.1;
    StepControl() = default;
    StepControl(int ir, bool sc, double dva, double dvm) : ir_steps(ir), scaling(sc), max_dva(dva), max_dvm(dvm)
    {
        if (ir < 0) throw py::value_error("ir_n_steps must be >= 0");
        if (sc && !(dva > 0.0 && dvm > 0.0)) throw py::value_error("max_dVa and max_dVm must be positive");
    }
    void apply(g2g_batch_t h) const { check(g2g_batch_set_step_control(h, ir_steps, scaling ? 1 : 0, max_dva, max_dvm)); }
};

// ---- limit violations of the solved rows (compute_limit_violations) ------
struct ViolJob {
    g2g_limit_table table{};
    const std::vector<std::vector<int>>* trips = nullptr;  // tripped branches per row (nullable)
    double tol = 1e-6;
    double rel_tol = 1e-9;
    int K = 16;
};

// per row: count (-1 not simulated), truncated, true counts per type; records [rows][3K]
struct ViolResults {
    int K = 0;
    std::vector<int32_t> count, truncated, n_low, n_high, n_current, etype, eid, side, type;
    std::vector<double> value, limit;
    void reset(int rows, int k)
    {
        K = k;
        const size_t R = static_cast<size_t>(rows), nrec = R * 3 * k;
        count.assign(R, -1);
        truncated.assign(R, 0);
        n_low.assign(R, -1);
        n_high.assign(R, -1);
        n_current.assign(R, -1);
        etype.assign(nrec, 0);
        eid.assign(nrec, 0);
        side.assign(nrec, 0);
        type.assign(nrec, 0);
        value.assign(nrec, std::nan(""));
        limit.assign(nrec, std::nan(""));
    }
};

// ---- the solve of many rows, chunked by batch_size -----------------------
struct RowJob {
    const cplx* S = nullptr;       // host: rows x n, or one shared row (S_stride 0)
    int64_t S_stride = 0;
    const cplx* V0 = nullptr;      // host: rows x n, or one shared row
    int64_t V_stride = 0;
    const cplx* dS = nullptr;      // the same on the device (filled by the session)
    const cplx* dV0 = nullptr;
    cplx* dV = nullptr;            // device results: rows x n
    // per-row Ybus edits (CSR-like, nullable)
    const std::vector<int32_t>* edit_ptr = nullptr;
    const std::vector<int32_t>* edit_ent = nullptr;
    const std::vector<cplx>* edit_delta = nullptr;
    std::vector<char> skip;        // rows not simulated (islanding trips)
    // per-row islanded buses (CSR-like, nullable; handle_disconnected_grid)
    const std::vector<int32_t>* mask_ptr = nullptr;
    const std::vector<int32_t>* mask_bus = nullptr;
    int max_iter = 10;
    double tol = 1e-8;
    int sched = G2G_SCHED_CONTINUOUS;
    uintptr_t mlp = 0;             // warm start (0: none)
    int k_warm = 4;
    double growth = 2.0;
    StepControl step;
    const ViolJob* viol = nullptr;  // limit checks after each chunk (nullable)
};

struct RowResults {
    std::vector<int32_t> status, iters;
    std::vector<double> res;
    std::vector<double> xs;  // extension state, rows x n_xs
    ViolResults viol;
};

// contiguous runs [first row, length, chunk offset] of the chunk rows live[c0 .. c0 + m)
std::vector<std::array<int, 3>> row_runs(const std::vector<int>& live, size_t c0, int m)
{
    std::vector<std::array<int, 3>> runs;
    for (int k = 0; k < m; ++k) {
        const int r = live[c0 + k];
        if (!runs.empty() && runs.back()[0] + runs.back()[1] == r) ++runs.back()[1];
        else runs.push_back({r, 1, k});
    }
    return runs;
}

// the limit checks of chunk rows live[c0 .. c0 + m), scattered to their rows
void check_violations(g2g_batch_t h, const ViolJob& vj, const std::vector<int>& live, size_t c0, int m,
                      ViolResults& out)
{
    const int K = vj.K;
    const size_t nrec = static_cast<size_t>(m) * 3 * K;
    std::vector<int32_t> tp(1, 0), tb;
    for (int k = 0; k < m; ++k) {
        const int r = live[c0 + k];
        if (vj.trips && static_cast<size_t>(r) < vj.trips->size())
            for (int l : (*vj.trips)[r]) tb.push_back(l);
        tp.push_back(static_cast<int32_t>(tb.size()));
    }
    std::vector<int32_t> cnt(m), tr(m), nl(m), nh(m), nc(m), et(nrec), ei(nrec), sd(nrec), ty(nrec);
    std::vector<double> va(nrec), li(nrec);
    g2g_violations o{cnt.data(), tr.data(), nl.data(), nh.data(), nc.data(), et.data(), ei.data(), sd.data(),
                     ty.data(), va.data(), li.data()};
    check(g2g_batch_limit_violations(h, &vj.table, tp.data(), tb.empty() ? nullptr : tb.data(), vj.tol, vj.rel_tol,
                                     K, &o));
    const size_t w = 3 * static_cast<size_t>(K);
    for (int k = 0; k < m; ++k) {
        const size_t r = static_cast<size_t>(live[c0 + k]);
        out.count[r] = cnt[k];
        out.truncated[r] = tr[k];
        out.n_low[r] = nl[k];
        out.n_high[r] = nh[k];
        out.n_current[r] = nc[k];
        const size_t a = k * w, b = r * w;
        std::copy_n(et.begin() + a, w, out.etype.begin() + b);
        std::copy_n(ei.begin() + a, w, out.eid.begin() + b);
        std::copy_n(sd.begin() + a, w, out.side.begin() + b);
        std::copy_n(ty.begin() + a, w, out.type.begin() + b);
        std::copy_n(va.begin() + a, w, out.value.begin() + b);
        std::copy_n(li.begin() + a, w, out.limit.begin() + b);
    }
}

// The rows' solve, chunked by batch_size, device-resident: every chunk's
// injections and start voltages are gathered on the device from the
// session's device rows (job.dS / job.dV0) and its voltages go straight to the
// session's device results (job.dV); only statuses, counts and residuals
// (and the limit-check records) come to the host.
void run_rows(const Grid& g, std::unique_ptr<Batch>& batch, int precision, int batch_size, int n_rows,
              const RowJob& job, RowResults& out, BatchTimings& tm)
{
    const int n = g.n();
    out.status.assign(n_rows, static_cast<int32_t>(RowStatus::NotSimulated));
    out.iters.assign(n_rows, 0);
    out.res.assign(n_rows, std::nan(""));
    out.xs.assign(static_cast<size_t>(n_rows) * g.n_xs, std::nan(""));
    if (job.viol) out.viol.reset(n_rows, job.viol->K);
    else out.viol = ViolResults{};
    // simulated rows, in order
    std::vector<int> live;
    for (int r = 0; r < n_rows; ++r)
        if (job.skip.empty() || !job.skip[r]) live.push_back(r);
    if (live.empty()) return;
    const int cap = std::max(1, std::min<int>(batch_size > 0 ? batch_size : static_cast<int>(live.size()),
                                              static_cast<int>(live.size())));
    const int per_row_y = job.edit_ptr ? 1 : 0;
    if (!batch || batch->rows < cap || batch->precision != precision || batch->per_row_y != per_row_y) {
        const double t0 = now_ms();
        batch.reset();
        batch = std::make_unique<Batch>(g, cap, precision, per_row_y);
        tm.t_alloc_ms += now_ms() - t0;
    }
    const int B = batch->rows;
    std::vector<int32_t> st(B), it(B);
    std::vector<double> rs(B), xs(static_cast<size_t>(B) * g.n_xs);
    std::vector<int32_t> eptr, eent, mptr, mbus;
    std::vector<cplx> edel;
    if (!job.mask_ptr) check(g2g_batch_set_bus_mask(batch->h, nullptr, nullptr, 0));
    auto dptr = [](const cplx* p) { return reinterpret_cast<const double*>(p); };
    for (size_t c0 = 0; c0 < live.size(); c0 += B) {
        const int m = static_cast<int>(std::min<size_t>(B, live.size() - c0));
        const auto runs = row_runs(live, c0, m);
        check(g2g_batch_set_active_rows(batch->h, m));
        job.step.apply(batch->h);
        if (job.S_stride == 0) {
            check(g2g_batch_set_sbus(batch->h, dptr(job.dS), 0, m, 0, 1));
        } else {
            for (const auto& rn : runs)
                check(g2g_batch_set_sbus(batch->h, dptr(job.dS + rn[0] * job.S_stride), job.S_stride, rn[1], rn[2], 1));
        }
        if (job.V_stride == 0) {
            check(g2g_batch_set_v(batch->h, dptr(job.dV0), 0, m, 0, 1));
        } else {
            for (const auto& rn : runs)
                check(g2g_batch_set_v(batch->h, dptr(job.dV0 + rn[0] * job.V_stride), job.V_stride, rn[1], rn[2], 1));
        }
        if (per_row_y) {
            eptr.assign(1, 0);
            eent.clear();
            edel.clear();
            for (int k = 0; k < m; ++k) {
                const int r = live[c0 + k];
                for (int q = (*job.edit_ptr)[r]; q < (*job.edit_ptr)[r + 1]; ++q) {
                    eent.push_back((*job.edit_ent)[q]);
                    edel.push_back((*job.edit_delta)[q]);
                }
                eptr.push_back(static_cast<int32_t>(eent.size()));
            }
            check(g2g_batch_set_ybus_edits(batch->h, eptr.data(), eent.data(),
                                           reinterpret_cast<const double*>(edel.data()), m));
        }
        if (job.mask_ptr) {
            mptr.assign(1, 0);
            mbus.clear();
            for (int k = 0; k < m; ++k) {
                const int r = live[c0 + k];
                for (int q = (*job.mask_ptr)[r]; q < (*job.mask_ptr)[r + 1]; ++q) mbus.push_back((*job.mask_bus)[q]);
                mptr.push_back(static_cast<int32_t>(mbus.size()));
            }
            check(g2g_batch_set_bus_mask(batch->h, mptr.data(), mbus.data(), m));
        }
        g2g_solve_stats s{};
        if (job.mlp)
            check(g2g_batch_solve_warm(batch->h, reinterpret_cast<g2g_mlp_t>(job.mlp), nullptr, 0, job.k_warm,
                                       job.growth, job.max_iter, job.tol, &s));
        else check(g2g_batch_solve(batch->h, job.max_iter, job.tol, job.sched, &s));
        add(tm.t_mismatch, s.ms_mismatch);
        add(tm.t_refactorize, s.ms_factor);
        add(tm.t_solve, s.ms_backward);
        add(tm.t_warm_start, s.ms_warm_start);
        tm.rows_factored += s.rows_factored;
        tm.n_requeued += s.n_requeued;
        ++tm.n_chunks;
        const double t0 = now_ms();
        for (const auto& rn : runs)
            check(g2g_batch_get_v_rows(batch->h, reinterpret_cast<double*>(job.dV + static_cast<size_t>(rn[0]) * n), n,
                                       rn[1], rn[2], 1));
        tm.t_copy_V_to_host_ms += now_ms() - t0;
        check(g2g_batch_get_results(batch->h, st.data(), it.data(), rs.data()));
        if (g.n_xs) check(g2g_batch_get_ext_state(batch->h, xs.data()));
        for (int k = 0; k < m; ++k) {
            const int r = live[c0 + k];
            out.status[r] = st[k];
            out.iters[r] = it[k];
            out.res[r] = rs[k];
            std::copy_n(xs.begin() + static_cast<size_t>(k) * g.n_xs, g.n_xs,
                        out.xs.begin() + static_cast<size_t>(r) * g.n_xs);
        }
        if (job.viol) {
            const double tv = now_ms();
            check_violations(batch->h, *job.viol, live, c0, m, out.viol);
            tm.t_violations_ms += now_ms() - tv;
        }
    }
    tm.n_rows = n_rows;
    tm.n_converged = static_cast<int>(std::count(out.status.begin(), out.status.end(), 0));
}

template <typename T>
py::array_t<T> to_array(const std::vector<T>& v)
{
    py::array_t<T> a(static_cast<py::ssize_t>(v.size()));
    std::copy(v.begin(), v.end(), a.mutable_data());
    return a;
}


void set_device(int device)
{
    if (device >= 0) check(g2g_set_device(device));
}

// ---- AcPfNrSession (python_bindings.cpp:1369-1560) -------------------------
class AcPfNrSession {
public:
    AcPfNrSession(py::object Ybus, CArr Vinit, CArr Sbus, IArr slack_ids, DArr /*slack_weights*/, IArr pv, IArr pq,
                  int max_iter, double tol, int device, bool presolved_v, StepControl step, Precision precision,
                  py::object extensions = py::none())
        : max_iter_(max_iter), tol_(tol), precision_(static_cast<int>(precision))
    {
        set_device(device);
        const double t0 = now_ms();
        grid_ = std::make_unique<Grid>(Ybus, slack_ids, pv, pq, extensions);
        const int n = grid_->n();
        if (Vinit.size() != n || Sbus.size() != n) throw py::value_error("Vinit / Sbus must have n_bus entries");
        S_.assign(Sbus.data(), Sbus.data() + n);
        V_.assign(Vinit.data(), Vinit.data() + n);
        // |V| set-points of the buses without a |V| unknown (PV and fixed),
        // from the first start: no Newton update moves them
        const auto qrow = q_row_of_bus();
        for (int b = 0; b < n; ++b)
            if (qrow.data()[b] < 0) vm_set_.emplace_back(b, std::abs(V_[b]));
        batch_ = std::make_unique<Batch>(*grid_, 1, precision_, 0);
        step.apply(batch_->h);
        timings_.t_analyze_ms = grid_->t_analyze_ms;
        timings_.t_init_ms = now_ms() - t0;
        if (presolved_v) {
            upload();
            double r = 0.0;
            check(g2g_batch_eval(batch_->h, nullptr, nullptr, &r));
            if (!(r < tol_))
                throw std::runtime_error("presolved_v: ||F(Vinit)||inf = " + std::to_string(r) +
                                         " is not below tol");
            timings_.converged = true;
            timings_.residual = r;
            keep_v();
        } else {
            solve(max_iter_, tol_);
        }
    }

    AcPfTimings resolve(CArr Vinit, py::object Sbus, int max_iter, double tol)
    {
        const int n = grid_->n();
        if (Vinit.size() != n) throw py::value_error("Vinit must have n_bus entries");
        V_.assign(Vinit.data(), Vinit.data() + n);
        if (!Sbus.is_none()) {
            CArr s = Sbus.cast<CArr>();
            if (s.size() != n) throw py::value_error("Sbus must have n_bus entries");
            S_.assign(s.data(), s.data() + n);
        }
        py::gil_scoped_release nogil;
        solve(max_iter < 0 ? max_iter_ : max_iter, tol < 0 ? tol_ : tol);
        return timings_;
    }

    // LSGrid.ac_pf's contract: a start whose |V| at the PV and fixed buses is
    // replaced by their set-points (angles kept), solved; true if converged
    bool solve_at_setpoints(CArr Vinit, int max_iter, double tol)
    {
        const int n = grid_->n();
        if (Vinit.size() != n) throw py::value_error("Vinit must have n_bus entries");
        V_.assign(Vinit.data(), Vinit.data() + n);
        for (const auto& [b, vm] : vm_set_) V_[b] = std::polar(vm, std::arg(V_[b]));
        py::gil_scoped_release nogil;
        solve(max_iter, tol);
        return timings_.converged;
    }

    CArr get_v()
    {
        const double t0 = now_ms();
        CArr out(grid_->n());
        hipchk(hipMemcpy(out.mutable_data(), dV_->p, grid_->n() * sizeof(cplx), hipMemcpyDeviceToHost),
               "hipMemcpy D2H");
        timings_.t_copy_v_to_host_ms = now_ms() - t0;
        return out;
    }
    // a DLPack (n_bus,) complex128 tensor aliasing the session's device V:
    // resolve() overwrites it (acpf_nr/gpu_facade.py's v_dlpack contract)
    py::object v_dlpack() { return alias_capsule(dV_, 1, grid_->n(), 1); }
    py::array_t<double> get_ext_state() const
    {
        std::vector<double> xs(static_cast<size_t>(grid_->n_xs));
        if (!xs.empty()) check(g2g_batch_get_ext_state(batch_->h, xs.data()));
        return to_array(xs);
    }

    DArr get_F()
    {
        DArr F(grid_->info.d);
        check(g2g_batch_eval(batch_->h, F.mutable_data(), nullptr, nullptr));
        return F;
    }

    // J^T x = rhs at the solved V (python_bindings.cpp:1525, solve_JT_dlpack);
    // a host array here, the facade wraps it as a device capsule
    DArr solve_JT(DArr rhs)
    {
        const int d = grid_->info.d;
        if (rhs.size() != d) throw py::value_error("rhs must have dim_J entries");
        DArr x(d);
        {
            py::gil_scoped_release nogil;
            check(g2g_batch_solve_adjoint(batch_->h, rhs.data(), x.mutable_data()));
        }
        return x;
    }

    py::tuple get_J()
    {
        std::vector<double> F(grid_->info.d), J(grid_->info.nJ);
        check(g2g_batch_eval(batch_->h, F.data(), J.data(), nullptr));
        return py::make_tuple(to_array(grid_->array(G2G_ARR_JP)), to_array(grid_->array(G2G_ARR_JI)), to_array(J));
    }

    const AcPfTimings& timings() const { return timings_; }
    int n_pvpq() const { return grid_->info.npvpq; }
    int n_pq() const { return grid_->info.npq; }
    int dim_J() const { return grid_->info.d; }
    py::array_t<int32_t> pvpq() const { return to_array(grid_->array(G2G_ARR_PVPQ)); }
    py::array_t<int32_t> pq() const { return to_array(grid_->pq); }
    py::array_t<int32_t> p_row_of_bus() const
    {
        std::vector<int32_t> r(grid_->n(), -1);
        const auto pvpq = grid_->array(G2G_ARR_PVPQ);
        for (size_t k = 0; k < pvpq.size(); ++k) r[pvpq[k]] = static_cast<int32_t>(k);
        return to_array(r);
    }
    py::array_t<int32_t> q_row_of_bus() const
    {
        std::vector<int32_t> r(grid_->n(), -1);
        for (size_t k = 0; k < grid_->pq.size(); ++k) r[grid_->pq[k]] = grid_->info.npvpq + static_cast<int32_t>(k);
        return to_array(r);
    }

private:
    void upload()
    {
        const int n = grid_->n();
        check(g2g_batch_set_sbus(batch_->h, reinterpret_cast<const double*>(S_.data()), n, 1, 0, 0));
        check(g2g_batch_set_v(batch_->h, reinterpret_cast<const double*>(V_.data()), n, 1, 0, 0));
    }
    void solve(int max_iter, double tol)
    {
        upload();
        g2g_solve_stats s{};
        const double t0 = now_ms();
        check(g2g_batch_solve(batch_->h, max_iter, tol, G2G_SCHED_CONTINUOUS, &s));
        int32_t st = 0, it = 0;
        double r = 0.0;
        check(g2g_batch_get_results(batch_->h, &st, &it, &r));
        timings_.t_mismatch = {s.ms_mismatch, s.ms_mismatch};
        timings_.t_refactorize = {s.ms_factor, s.ms_factor};
        timings_.t_solve = {s.ms_backward, s.ms_backward};
        timings_.t_total_ms = now_ms() - t0;
        timings_.nb_iter = it;
        timings_.n_refactorize = s.n_factor_launches;
        timings_.converged = st == 0;
        timings_.residual = r;
        keep_v();
    }
    // the solved V into the session's device vector (what v_dlpack aliases)
    void keep_v()
    {
        ensure(dV_, static_cast<size_t>(grid_->n()) * sizeof(cplx));
        check(g2g_batch_get_v_rows(batch_->h, reinterpret_cast<double*>(dV_->p), grid_->n(), 1, 0, 1));
    }

    std::unique_ptr<Grid> grid_;
    std::unique_ptr<Batch> batch_;
    std::vector<cplx> S_, V_;
    std::vector<std::pair<int, double>> vm_set_;  // (bus, |V| set-point) without a |V| unknown
    DevBufPtr dV_;                                 // the solved V on the device
    int max_iter_;
    double tol_;
    int precision_;
    AcPfTimings timings_;
};

// branch data and outage edits (contingency_analysis_helper.cpp:23-83,131-163)
struct Branches {
    std::vector<int32_t> f, t;
    std::vector<cplx> yff, yft, ytf, ytt;
    std::vector<double> base_or, base_ex;  // A per unit current (sn_mva 1e6 / (sqrt(3) vn_kv 1e3))
    std::vector<double> vn_kv_bus;         // [n] nominal kV
    std::vector<char> bridge;
    bool set = false;

    int m() const { return static_cast<int>(f.size()); }

    // yff, yft, ytf, ytt per branch
    std::vector<cplx> y4() const
    {
        std::vector<cplx> y(4 * f.size());
        for (size_t l = 0; l < f.size(); ++l) {
            y[4 * l] = yff[l];
            y[4 * l + 1] = yft[l];
            y[4 * l + 2] = ytf[l];
            y[4 * l + 3] = ytt[l];
        }
        return y;
    }

    void assign(const IArr& bf, const IArr& bt, const CArr& ff, const CArr& ft, const CArr& tf, const CArr& tt,
                const DArr& vn_kv, double sn_mva, int n)
    {
        if (vn_kv.size() != n) throw py::value_error("bus_vn_kv must have n_bus entries");
        if (!(sn_mva > 0)) throw py::value_error("sn_mva must be positive");
        const auto m = bf.size();
        if (bt.size() != m || ff.size() != m || ft.size() != m || tf.size() != m || tt.size() != m)
            throw py::value_error("branch arrays differ in length");
        f = ivec(bf);
        t = ivec(bt);
        for (py::ssize_t l = 0; l < m; ++l)
            if (f[l] < 0 || f[l] >= n || t[l] < 0 || t[l] >= n) throw py::value_error("branch bus out of range");
        yff.assign(ff.data(), ff.data() + m);
        yft.assign(ft.data(), ft.data() + m);
        ytf.assign(tf.data(), tf.data() + m);
        ytt.assign(tt.data(), tt.data() + m);
        base_or.resize(m);
        base_ex.resize(m);
        const double* vn = vn_kv.data();
        vn_kv_bus.assign(vn, vn + n);
        for (py::ssize_t l = 0; l < m; ++l) {
            base_or[l] = sn_mva * 1e6 / (std::sqrt(3.0) * vn[f[l]] * 1e3);
            base_ex[l] = sn_mva * 1e6 / (std::sqrt(3.0) * vn[t[l]] * 1e3);
        }
        bridge = find_bridges(n);
        set = true;
    }

    // Tarjan low-link over the multigraph (a branch parallel to another is no bridge)
    std::vector<char> find_bridges(int n) const
    {
        std::vector<std::vector<std::pair<int, int>>> adj(n);
        for (int l = 0; l < m(); ++l) {
            if (f[l] == t[l]) continue;
            adj[f[l]].push_back({t[l], l});
            adj[t[l]].push_back({f[l], l});
        }
        std::vector<int> disc(n, -1), low(n, 0);
        std::vector<char> br(m(), 0);
        int timer = 0;
        struct Frame { int v, pe; size_t k; };
        for (int s = 0; s < n; ++s) {
            if (disc[s] >= 0) continue;
            std::vector<Frame> st{{s, -1, 0}};
            disc[s] = low[s] = timer++;
            while (!st.empty()) {
                Frame& fr = st.back();
                if (fr.k < adj[fr.v].size()) {
                    auto [w, l] = adj[fr.v][fr.k++];
                    if (l == fr.pe) continue;
                    if (disc[w] < 0) {
                        disc[w] = low[w] = timer++;
                        st.push_back({w, l, 0});
                    } else {
                        low[fr.v] = std::min(low[fr.v], disc[w]);
                    }
                } else {
                    const Frame done = fr;
                    st.pop_back();
                    if (!st.empty()) {
                        Frame& p = st.back();
                        low[p.v] = std::min(low[p.v], low[done.v]);
                        if (low[done.v] > disc[p.v]) br[done.pe] = 1;
                    }
                }
            }
        }
        return br;
    }

    // does removing `out` disconnect the graph? (union-find over the rest)
    bool islands(const std::vector<int>& out, int n) const
    {
        if (out.empty()) return false;
        if (out.size() == 1) return bridge[out[0]] != 0;
        std::vector<char> gone(m(), 0);
        for (int l : out) gone[l] = 1;
        std::vector<int> par(n);
        std::iota(par.begin(), par.end(), 0);
        auto find = [&](int x) {
            while (par[x] != x) x = par[x] = par[par[x]];
            return x;
        };
        int comps = n;
        for (int l = 0; l < m(); ++l) {
            if (gone[l]) continue;
            const int a = find(f[l]), b = find(t[l]);
            if (a != b) {
                par[a] = b;
                --comps;
            }
        }
        return comps > 1;
    }

    // the buses outside the largest connected component once `out` is tripped
    // (compute_component_masks, contingency_analysis_helper.cpp:657-692: the
    // largest by bus count, a tie to the component met first in bus order)
    std::vector<int> off_main(const std::vector<int>& out, int n) const
    {
        std::vector<char> gone(m(), 0);
        for (int l : out) gone[l] = 1;
        std::vector<int> par(n);
        std::iota(par.begin(), par.end(), 0);
        auto find = [&](int x) {
            while (par[x] != x) x = par[x] = par[par[x]];
            return x;
        };
        for (int l = 0; l < m(); ++l)
            if (!gone[l]) {
                const int a = find(f[l]), b = find(t[l]);
                if (a != b) par[a] = b;
            }
        std::vector<int> label(n, -1), count;
        std::vector<int> comp(n);
        for (int i = 0; i < n; ++i) {
            const int r = find(i);
            if (label[r] < 0) {
                label[r] = static_cast<int>(count.size());
                count.push_back(0);
            }
            comp[i] = label[r];
            ++count[comp[i]];
        }
        const int main = static_cast<int>(std::max_element(count.begin(), count.end()) - count.begin());
        std::vector<int> off;
        for (int i = 0; i < n; ++i)
            if (comp[i] != main) off.push_back(i);
        return off;
    }

    // value edits of one outage, merged by CSR position
    void edits(const Grid& g, const std::vector<int>& out, std::vector<int32_t>& ent, std::vector<cplx>& del) const
    {
        std::vector<std::pair<int64_t, cplx>> e;
        auto put = [&](int r, int c, cplx v) {
            const int64_t p = g.pos(r, c);
            if (p < 0) throw py::value_error("a tripped branch's entry is not in Ybus's pattern");
            e.push_back({p, v});
        };
        for (int l : out) {
            if (l < 0 || l >= m()) throw py::value_error("branch id out of range");
            put(f[l], f[l], yff[l]);
            put(t[l], t[l], ytt[l]);
            if (f[l] != t[l]) {
                put(f[l], t[l], yft[l]);
                put(t[l], f[l], ytf[l]);
            }
        }
        std::sort(e.begin(), e.end(), [](auto& a, auto& b) { return a.first < b.first; });
        for (size_t k = 0; k < e.size(); ++k) {
            if (!ent.empty() && k > 0 && e[k].first == e[k - 1].first) {
                del.back() += e[k].second;
                continue;
            }
            ent.push_back(static_cast<int32_t>(e[k].first));
            del.push_back(e[k].second);
        }
    }
};

// ---- batch sessions (Injection sweep / contingency analysis / scenario sweep)
class BatchSession {
public:
    BatchSession(py::object Ybus, CArr Vinit, CArr Sbus, IArr slack_ids, IArr pv, IArr pq, int batch_size,
                 int nb_iter, int max_iter_base, double tol_base, int device, bool presolved_v, StepControl step,
                 Precision precision, py::object extensions = py::none())
        : batch_size_(batch_size), nb_iter_(nb_iter), precision_(static_cast<int>(precision)), step_(step)
    {
        if (batch_size <= 0) throw py::value_error("batch_size must be positive");
        if (nb_iter < 0) throw py::value_error("nb_iter must be >= 0");
        set_device(device);
        const double t0 = now_ms();
        grid_ = std::make_unique<Grid>(Ybus, slack_ids, pv, pq, extensions);
        timings_.t_analysis_ms = grid_->t_analyze_ms;
        const int n = grid_->n();
        if (Vinit.size() != n || Sbus.size() != n) throw py::value_error("Vinit / Sbus must have n_bus entries");
        S_base_.assign(Sbus.data(), Sbus.data() + n);
        V_base_.assign(Vinit.data(), Vinit.data() + n);
        // the base case, solved once (or checked, when presolved)
        {
            py::gil_scoped_release nogil;
            Batch b(*grid_, 1, G2G_FP64, 0);
            step_.apply(b.h);
            check(g2g_batch_set_sbus(b.h, reinterpret_cast<const double*>(S_base_.data()), n, 1, 0, 0));
            check(g2g_batch_set_v(b.h, reinterpret_cast<const double*>(V_base_.data()), n, 1, 0, 0));
            int32_t st = 0, it = 0;
            double r = 0.0;
            if (presolved_v) {
                check(g2g_batch_eval(b.h, nullptr, nullptr, &r));
                if (!(r < tol_base))
                    throw std::runtime_error("presolved_v: ||F(Vinit)||inf = " + std::to_string(r) +
                                             " is not below tol_base");
            } else {
                check(g2g_batch_solve(b.h, max_iter_base, tol_base, G2G_SCHED_CONTINUOUS, nullptr));
                check(g2g_batch_get_results(b.h, &st, &it, &r));
                if (st != 0)
                    throw std::runtime_error("the base case did not converge (residual " + std::to_string(r) + ")");
                check(g2g_batch_get_v(b.h, reinterpret_cast<double*>(V_base_.data()), n, 0));
            }
        }
        timings_.t_base_case_ms = now_ms() - t0;
    }
    virtual ~BatchSession() = default;

    // settings (mutable, next run; CLAUDE.md:130-138)
    int batch_size = 0;
    int get_batch_size() const { return batch_size_; }
    void set_batch_size(int v)
    {
        if (v <= 0) throw py::value_error("batch_size must be positive");
        batch_size_ = v;
    }
    int get_nb_iter() const { return nb_iter_; }
    void set_nb_iter(int v)
    {
        if (v < 0) throw py::value_error("nb_iter must be >= 0");
        nb_iter_ = v;
    }
    BatchScheduling get_scheduling() const { return sched_; }
    void set_scheduling(BatchScheduling s) { sched_ = s; }
    double get_tol() const { return tol_; }
    void set_tol(double t)
    {
        if (!(t > 0)) throw py::value_error("tol must be positive");
        tol_ = t;
    }
    Precision get_precision() const { return static_cast<Precision>(precision_); }
    void set_precision(Precision p) { precision_ = static_cast<int>(p); }
    void set_warm_start(uintptr_t mlp, int k_warm, double growth)
    {
        if (k_warm < 0 || growth < 0) throw py::value_error("k_warm and growth must be >= 0");
        mlp_ = mlp;
        k_warm_ = k_warm;
        growth_ = growth;
    }
    void clear_warm_start() { mlp_ = 0; }
    bool handle_disconnected_grid = false;

    // compute_limit_violations (contingency_analysis/__init__.py:589-735)
    bool compute_limit_violations = false;
    double violation_tol = 1e-6;
    double get_violation_rel_tol() const { return violation_rel_tol_; }
    void set_violation_rel_tol(double v)
    {
        if (!(v >= 0.0 && v < 1.0)) throw py::value_error("violation_rel_tol must be a real number in [0, 1[");
        violation_rel_tol_ = v;
    }
    int get_violation_capacity() const { return violation_capacity_; }
    void set_violation_capacity(int k)
    {
        if (k < 1 || k > G2G_MAX_VIOLATION_CAPACITY)
            throw py::value_error("violation_capacity must be in [1, " + std::to_string(G2G_MAX_VIOLATION_CAPACITY) +
                                  "], got " + std::to_string(k));
        violation_capacity_ = k;
    }
    void set_limits(DArr vmin_kv, DArr vmax_kv, DArr lim1_ka, DArr lim2_ka, int n_lines)
    {
        if (!br_.set) throw std::runtime_error("call set_branch_data() first");
        const int n = n_bus(), m = br_.m();
        if (vmin_kv.size() != n || vmax_kv.size() != n) throw py::value_error("bus limits must have n_bus entries");
        if (lim1_ka.size() != m || lim2_ka.size() != m)
            throw py::value_error("branch limits must have n_branches entries");
        if (n_lines < 0 || n_lines > m) throw py::value_error("n_lines must be in [0, n_branches]");
        vmin_.assign(vmin_kv.data(), vmin_kv.data() + n);
        vmax_.assign(vmax_kv.data(), vmax_kv.data() + n);
        lim1_.assign(lim1_ka.data(), lim1_ka.data() + m);
        lim2_.assign(lim2_ka.data(), lim2_ka.data() + m);
        vn_kv_ = br_.vn_kv_bus;
        n_lines_ = n_lines;
        limits_set_ = true;
    }
    py::array_t<int32_t> viol_i(const std::vector<int32_t> ViolResults::*f) const
    {
        need_violations();
        return to_array(res_.viol.*f);
    }
    py::array_t<double> viol_d(const std::vector<double> ViolResults::*f) const
    {
        need_violations();
        return to_array(res_.viol.*f);
    }

    void run()
    {
        const int rows = n_rows();
        if (rows <= 0) throw std::runtime_error("nothing to run: set the rows first");
        RowJob job;
        prepare(job);
        job.max_iter = nb_iter_;
        job.tol = tol_;
        job.sched = sched_ == BatchScheduling::Chunked ? G2G_SCHED_CHUNKED : G2G_SCHED_CONTINUOUS;
        job.mlp = mlp_;
        job.k_warm = k_warm_;
        job.growth = growth_;
        job.step = step_;
        BatchTimings t = timings_;
        t.t_mismatch = t.t_refactorize = t.t_solve = t.t_warm_start = TimingEntry{};
        t.n_chunks = 0;
        t.rows_factored = 0;
        t.n_requeued = 0;
        t.t_copy_V_to_host_ms = 0.0;
        t.t_violations_ms = 0.0;
        ViolJob vj;
        std::vector<cplx> y4;
        if (compute_limit_violations) {
            if (!limits_set_) throw std::runtime_error("compute_limit_violations: call set_limits() first");
            y4 = br_.y4();
            vj.table = g2g_limit_table{br_.m(), br_.f.data(), br_.t.data(), reinterpret_cast<const double*>(y4.data()),
                                       br_.base_or.data(), br_.base_ex.data(), lim1_.data(), lim2_.data(), n_lines_,
                                       vn_kv_.data(), vmin_.data(), vmax_.data()};
            vj.trips = &trips_;
            vj.tol = violation_tol;
            vj.rel_tol = violation_rel_tol_;
            vj.K = violation_capacity_;
            job.viol = &vj;
        }
        const double t0 = now_ms();
        const int n = n_bus();
        // the rows on the device: injections once per change of the session's
        // rows (shared rows and start voltages are small and go every run)
        if (job.S_stride == 0) {
            upload(dS1_, job.S, static_cast<size_t>(n) * sizeof(cplx));
            job.dS = dS1_->c();
        } else {
            if (s_dirty_ || !dS_ || dS_->bytes < static_cast<size_t>(rows) * n * sizeof(cplx)) {
                upload(dS_, job.S, static_cast<size_t>(rows) * job.S_stride * sizeof(cplx));
                s_dirty_ = false;
            }
            job.dS = dS_->c();
        }
        upload(dV0_, job.V0, static_cast<size_t>(job.V_stride ? rows : 1) * n * sizeof(cplx));
        job.dV0 = dV0_->c();
        ensure(dVres_, static_cast<size_t>(rows) * n * sizeof(cplx));
        job.dV = dVres_->c();
        bool any_skip = false;
        for (char c : job.skip) any_skip |= c != 0;
        if (any_skip)  // rows not simulated read NaN (all-ones bytes)
            hipchk(hipMemset(dVres_->p, 0xFF, static_cast<size_t>(rows) * n * sizeof(cplx)), "hipMemset");
        res_rows_ = rows;
        run_rows(*grid_, batch_, precision_, batch_size_, rows, job, res_, t);
        t.t_run_ms = now_ms() - t0;
        timings_ = t;
        disc_.assign(static_cast<size_t>(rows), 0);
        for (size_t r = 0; r < job.skip.size() && r < disc_.size(); ++r) disc_[r] = job.skip[r] ? 1 : 0;
        ran_ = true;
        flows_ = false;
    }

    // the result voltages, downloaded through pinned staging
    CArr get_V_results()
    {
        need_run();
        const size_t cnt = static_cast<size_t>(res_rows_) * n_bus();
        CArr a(static_cast<py::ssize_t>(cnt));
        if (cnt) {
            void* st = pin_.get(cnt * sizeof(cplx));
            hipchk(hipMemcpy(st, dVres_->p, cnt * sizeof(cplx), hipMemcpyDeviceToHost), "hipMemcpy D2H");
            std::memcpy(a.mutable_data(), st, cnt * sizeof(cplx));
        }
        return a;
    }
    // a DLPack (n_rows, n_bus) complex128 tensor aliasing the session's device
    // results: the next run() overwrites it (clone to keep it)
    py::object v_results_dlpack()
    {
        need_run();
        return alias_capsule(dVres_, res_rows_, n_bus(), 2);
    }
    // extension state of every row (rows x n_xs: sa, then each controller's Q; NaN if not simulated)
    py::array_t<double> get_ext_state() const
    {
        need_run();
        py::array_t<double> a({static_cast<py::ssize_t>(res_rows_), static_cast<py::ssize_t>(grid_->n_xs)});
        std::copy(res_.xs.begin(), res_.xs.end(), a.mutable_data());
        return a;
    }
    py::array_t<double> get_residuals() const { need_run(); return to_array(res_.res); }
    py::array_t<int32_t> get_row_iterations() const { need_run(); return to_array(res_.iters); }
    py::array_t<int32_t> get_row_status() const { need_run(); return to_array(res_.status); }
    // per row, 1 = its trips split the grid and it was not simulated (empty
    // before run(); scenario_sweep_session.hpp:867-870)
    py::array_t<int32_t> get_disconnected() const { return to_array(disc_); }
    const BatchTimings& get_timings() const { return timings_; }
    int n_bus() const { return grid_->n(); }
    int dim_J() const { return grid_->info.d; }
    // branch data (python_bindings.cpp:1819-1832)
    void set_branch_data(IArr bf, IArr bt, CArr yff, CArr yft, CArr ytf, CArr ytt, DArr bus_vn_kv, double sn_mva)
    {
        br_.assign(bf, bt, yff, yft, ytf, ytt, bus_vn_kv, sn_mva, n_bus());
    }
    int n_branches() const { return br_.m(); }

    // branch currents of every simulated row at its result V, on the device
    // (python_bindings.cpp:1869-1872, 1915-1922)
    void compute_flows()
    {
        need_run();
        if (!br_.set) throw std::runtime_error("call set_branch_data() first");
        const int n = n_bus(), m = br_.m(), rows = n_rows();
        or_.assign(static_cast<size_t>(rows) * m, std::nan(""));
        ex_.assign(static_cast<size_t>(rows) * m, std::nan(""));
        const std::vector<cplx> y4 = br_.y4();
        std::vector<int> live;
        for (int r = 0; r < rows; ++r)
            if (res_.status[r] != static_cast<int32_t>(RowStatus::NotSimulated)) live.push_back(r);
        if (live.empty()) return;
        const int cap = std::min<int>(std::max(1, batch_size_), static_cast<int>(live.size()));
        if (!batch_ || batch_->rows < cap) {
            batch_.reset();
            batch_ = std::make_unique<Batch>(*grid_, cap, precision_, 0);
        }
        const int B = batch_->rows;
        std::vector<double> o(static_cast<size_t>(B) * m), e(static_cast<size_t>(B) * m);
        std::vector<int32_t> tp, tb;
        for (size_t c0 = 0; c0 < live.size(); c0 += B) {
            const int k = static_cast<int>(std::min<size_t>(B, live.size() - c0));
            tp.assign(1, 0);
            tb.clear();
            for (int q = 0; q < k; ++q) {
                const int r = live[c0 + q];
                if (static_cast<size_t>(r) < trips_.size())
                    for (int l : trips_[r]) tb.push_back(l);
                tp.push_back(static_cast<int32_t>(tb.size()));
            }
            check(g2g_batch_set_active_rows(batch_->h, k));
            for (const auto& rn : row_runs(live, c0, k))  // the rows' results, device to device
                check(g2g_batch_set_v(batch_->h, reinterpret_cast<const double*>(dVres_->c() + static_cast<size_t>(rn[0]) * n),
                                      n, rn[1], rn[2], 1));
            check(g2g_batch_branch_flows(batch_->h, m, br_.f.data(), br_.t.data(),
                                         reinterpret_cast<const double*>(y4.data()), br_.base_or.data(),
                                         br_.base_ex.data(), tp.data(), tb.empty() ? nullptr : tb.data(), o.data(),
                                         e.data()));
            for (int q = 0; q < k; ++q) {
                const int r = live[c0 + q];
                std::copy_n(o.begin() + static_cast<size_t>(q) * m, m, or_.begin() + static_cast<size_t>(r) * m);
                std::copy_n(e.begin() + static_cast<size_t>(q) * m, m, ex_.begin() + static_cast<size_t>(r) * m);
            }
        }
        flows_ = true;
    }
    py::array_t<double> get_or_amps() const { return flows(or_); }
    py::array_t<double> get_ex_amps() const { return flows(ex_); }

    CArr v_base() const
    {
        CArr a(grid_->n());
        std::copy(V_base_.begin(), V_base_.end(), a.mutable_data());
        return a;
    }
    virtual int n_rows() const = 0;

protected:
    virtual void prepare(RowJob& job) = 0;
    void need_run() const
    {
        if (!ran_) throw std::runtime_error("call run() first");
    }
    py::array_t<double> flows(const std::vector<double>& v) const
    {
        if (!flows_) throw std::runtime_error("call compute_flows() first");
        py::array_t<double> a({static_cast<py::ssize_t>(n_rows()), static_cast<py::ssize_t>(br_.m())});
        std::copy(v.begin(), v.end(), a.mutable_data());
        return a;
    }
    Branches br_;
    std::vector<std::vector<int>> trips_;  // tripped branches per row
    std::vector<double> or_, ex_;
    bool flows_ = false;

    std::unique_ptr<Grid> grid_;
    std::unique_ptr<Batch> batch_;
    std::vector<cplx> S_base_, V_base_;
    int batch_size_;
    int nb_iter_;
    int precision_;
    StepControl step_;
    BatchScheduling sched_ = BatchScheduling::Chunked;
    double tol_ = 1e-8;
    uintptr_t mlp_ = 0;
    int k_warm_ = 4;
    double growth_ = 2.0;
    RowResults res_;
    BatchTimings timings_;
    bool ran_ = false;
    std::vector<int32_t> disc_;
    // device rows: injections (and their change flag), a shared injection
    // row, start voltages, results (aliased by v_results_dlpack), staging
    DevBufPtr dS_, dS1_, dV0_, dVres_;
    bool s_dirty_ = true;
    int res_rows_ = 0;
    Pinned pin_;
    // limits of compute_limit_violations
    std::vector<double> vmin_, vmax_, lim1_, lim2_, vn_kv_;
    int n_lines_ = 0;
    bool limits_set_ = false;
    double violation_rel_tol_ = 1e-9;
    int violation_capacity_ = 16;
    void need_violations() const
    {
        need_run();
        if (res_.viol.count.empty() && n_rows() > 0)
            throw std::runtime_error("no limit violations: run() with compute_limit_violations = True");
    }

    // Trips per row: the value edits, and for rows they split, the buses off
    // the largest component (masked under handle_disconnected_grid unless
    // the slack is among them: then the row is skipped either way).
    struct Topology {
        std::vector<int32_t> eptr{0}, eent;
        std::vector<cplx> edel;
        std::vector<char> split, strands_slack;
        std::vector<int32_t> mptr{0}, mbus;

        void build(const Grid& g, const Branches& br, const std::vector<std::vector<int>>& ids)
        {
            const int n = g.n();
            eptr.assign(1, 0);
            eent.clear();
            edel.clear();
            mptr.assign(1, 0);
            mbus.clear();
            split.assign(ids.size(), 0);
            strands_slack.assign(ids.size(), 0);
            for (size_t r = 0; r < ids.size(); ++r) {
                br.edits(g, ids[r], eent, edel);
                eptr.push_back(static_cast<int32_t>(eent.size()));
                if (br.islands(ids[r], n)) {
                    split[r] = 1;
                    const std::vector<int> off = br.off_main(ids[r], n);
                    if (std::find(off.begin(), off.end(), g.slack) != off.end()) strands_slack[r] = 1;
                    else mbus.insert(mbus.end(), off.begin(), off.end());
                }
                mptr.push_back(static_cast<int32_t>(mbus.size()));
            }
        }
        void apply(RowJob& job, bool handle) const
        {
            job.edit_ptr = &eptr;
            job.edit_ent = &eent;
            job.edit_delta = &edel;
            job.skip.assign(split.size(), 0);
            for (size_t r = 0; r < split.size(); ++r) job.skip[r] = split[r] && (!handle || strands_slack[r]);
            if (handle && !mbus.empty()) {
                job.mask_ptr = &mptr;
                job.mask_bus = &mbus;
            }
        }
    };
};

// InjectionSweepSession (python_bindings.cpp:2167-2423)
class InjectionSweepSession : public BatchSession {
public:
    using BatchSession::BatchSession;
    void set_injections(DArr p_mw, DArr q_mvar, double sn_mva)
    {
        if (p_mw.ndim() != 2 || q_mvar.ndim() != 2 || p_mw.shape(0) != q_mvar.shape(0) ||
            p_mw.shape(1) != q_mvar.shape(1))
            throw py::value_error("p_mw and q_mvar must be (n_scenarios, n_bus) arrays of one shape");
        if (p_mw.shape(1) != n_bus()) throw py::value_error("injection rows must have n_bus entries");
        if (!(sn_mva > 0)) throw py::value_error("sn_mva must be positive");
        const py::ssize_t rows = p_mw.shape(0), n = p_mw.shape(1);
        S_.resize(static_cast<size_t>(rows * n));
        for (py::ssize_t i = 0; i < rows * n; ++i) S_[i] = cplx(p_mw.data()[i], q_mvar.data()[i]) / sn_mva;
        rows_ = static_cast<int>(rows);
        s_dirty_ = true;  // uploaded once, at the next run
        ran_ = false;
    }
    void set_sbus(CArr S)
    {
        if (S.ndim() != 2 || S.shape(1) != n_bus()) throw py::value_error("Sbus must be (n_scenarios, n_bus)");
        S_.assign(S.data(), S.data() + S.size());
        rows_ = static_cast<int>(S.shape(0));
        s_dirty_ = true;
        ran_ = false;
    }
    // per-row start voltages (None: the base case), as ScenarioSweep's set_v_init_from
    void set_v_init_from(py::object V)
    {
        if (V.is_none()) {
            V0_.clear();
            return;
        }
        CArr a = V.cast<CArr>();
        if (a.size() % n_bus()) throw py::value_error("V must hold whole rows of n_bus entries");
        V0_.assign(a.data(), a.data() + a.size());
    }
    int n_rows() const override { return rows_; }
    int n_scenarios() const { return rows_; }
    // per-row generator voltage set-points (python_bindings.cpp:2232-2243): gen_v
    // [n_scenarios][n_gen] in per unit, gen_bus[g] the bus generator g holds; at
    // a Vm-fixed bus (PV or slack) the row's start |V| becomes gen_v (angle
    // kept); other buses and NaN entries are ignored
    void set_gen_v(DArr gen_v, IArr gen_bus)
    {
        if (gen_v.ndim() != 2 || gen_v.shape(1) != gen_bus.size())
            throw py::value_error("gen_v must be (n_scenarios, n_gen) with n_gen = len(gen_bus)");
        for (py::ssize_t g = 0; g < gen_bus.size(); ++g)
            if (gen_bus.data()[g] < -1 || gen_bus.data()[g] >= n_bus())
                throw py::value_error("gen_bus entry out of range");
        gv_.assign(gen_v.data(), gen_v.data() + gen_v.size());
        gv_rows_ = static_cast<int>(gen_v.shape(0));
        gb_.assign(gen_bus.data(), gen_bus.data() + gen_bus.size());
    }
    void clear_gen_v()
    {
        gv_.clear();
        gb_.clear();
        gv_rows_ = 0;
    }
    py::array_t<bool> is_vm_fixed_bus() const
    {
        py::array_t<bool> a(n_bus());
        bool* f = a.mutable_data();
        std::fill(f, f + n_bus(), false);
        for (int b : grid_->pv) f[b] = true;
        for (int b : grid_->fixed) f[b] = true;
        return a;
    }

protected:
    void prepare(RowJob& job) override
    {
        const int n = n_bus();
        job.S = S_.data();
        job.S_stride = n;
        start(job, n);
    }
    void start(RowJob& job, int n)
    {
        start_v(job, n);
        if (gb_.empty()) return;
        if (gv_rows_ != rows_) throw py::value_error("set_gen_v: one row per scenario");
        // the start voltages with the set-points applied, one row per scenario
        std::vector<char> fixed(static_cast<size_t>(n), 0);
        for (int b : grid_->pv) fixed[b] = 1;
        for (int b : grid_->fixed) fixed[b] = 1;
        Vg_.resize(static_cast<size_t>(rows_) * n);
        const int ng = static_cast<int>(gb_.size());
        for (int r = 0; r < rows_; ++r) {
            cplx* v = Vg_.data() + static_cast<size_t>(r) * n;
            std::copy_n(job.V0 + r * job.V_stride, n, v);
            for (int g = 0; g < ng; ++g) {
                const int b = gb_[g];
                const double vm = gv_[static_cast<size_t>(r) * ng + g];
                if (b < 0 || !fixed[b] || std::isnan(vm)) continue;
                const double a = std::abs(v[b]) > 0.0 ? std::arg(v[b]) : 0.0;
                v[b] = std::polar(vm, a);
            }
        }
        job.V0 = Vg_.data();
        job.V_stride = n;
    }
    void start_v(RowJob& job, int n)
    {
        if (V0_.empty()) {
            job.V0 = V_base_.data();
            job.V_stride = 0;
        } else if (V0_.size() == static_cast<size_t>(n)) {
            job.V0 = V0_.data();
            job.V_stride = 0;
        } else {
            if (V0_.size() != static_cast<size_t>(rows_) * n)
                throw py::value_error("set_v_init_from: one row or one per scenario");
            job.V0 = V0_.data();
            job.V_stride = n;
        }
    }
    std::vector<cplx> S_, V0_, Vg_;
    std::vector<double> gv_;
    std::vector<int32_t> gb_;
    int gv_rows_ = 0;
    int rows_ = 0;
};

// ContingencyAnalysisSession (python_bindings.cpp:1745-2166)
class ContingencyAnalysisSession : public BatchSession {
public:
    using BatchSession::BatchSession;
    void build_contingencies(const std::vector<std::vector<int>>& ids)
    {
        if (!br_.set) throw std::runtime_error("call set_branch_data() first");
        const double t0 = now_ms();
        topo_.build(*grid_, br_, ids);
        ctg_ = ids;
        trips_ = ids;
        timings_.t_preprocess_ms = now_ms() - t0;
        ran_ = false;
    }
    int n_rows() const override { return static_cast<int>(ctg_.size()); }
    int n_contingencies() const { return n_rows(); }

protected:
    void prepare(RowJob& job) override
    {
        job.S = S_base_.data();
        job.S_stride = 0;
        job.V0 = V_base_.data();
        job.V_stride = 0;
        topo_.apply(job, handle_disconnected_grid);
    }
    std::vector<std::vector<int>> ctg_;
    Topology topo_;
};

// ScenarioSweepSession (python_bindings.cpp:2424-3207): per-row injections and
// per-row branch trips together, optional per-row start voltages
class ScenarioSweepSession : public InjectionSweepSession {
public:
    using InjectionSweepSession::InjectionSweepSession;
    void set_topology(const std::vector<std::vector<int>>& ids)
    {
        if (!br_.set) throw std::runtime_error("call set_branch_data() first");
        tp_.build(*grid_, br_, ids);
        topo_ = true;
        trips_ = ids;
        ran_ = false;
    }
    void clear_topology()
    {
        topo_ = false;
        tp_ = Topology{};
        trips_.clear();
    }
    void clear_v_init() { V0_.clear(); }
    bool has_v_init() const { return !V0_.empty(); }

protected:
    void prepare(RowJob& job) override
    {
        InjectionSweepSession::prepare(job);
        if (topo_) {
            if (tp_.eptr.size() != static_cast<size_t>(rows_) + 1)
                throw py::value_error("set_topology: one entry per scenario");
            tp_.apply(job, handle_disconnected_grid);
        }
    }
    bool topo_ = false;
    Topology tp_;
};

// ---- raw batched LU (raw_cudss_solve.hpp:31-121) --------------------------
struct RawPlan {
    g2g_plan_t h = nullptr;
    g2g_plan_info info{};
    RawPlan(int dim, const std::vector<int>& indptr, const std::vector<int>& indices)
    {
        if (dim <= 0) throw py::value_error("dim must be positive");
        if (static_cast<int>(indptr.size()) != dim + 1) throw py::value_error("indptr must have dim + 1 entries");
        if (static_cast<int>(indices.size()) != indptr[dim]) throw py::value_error("indices must have nnz entries");
        check(g2g_lu_plan_create(dim, indptr.data(), indices.data(), nullptr, &h));
        check(g2g_plan_get_info(h, &info));
    }
    ~RawPlan() { g2g_plan_destroy(h); }
};

// max over rows of ||A x - b||inf / ||b||inf, and the rows with a non-finite x
std::pair<double, int> raw_residual(int dim, const std::vector<int>& p, const std::vector<int>& ix,
                                    const double* vals, const double* b, const double* x, int rows, int nnz)
{
    double worst = 0.0;
    int bad = 0;
    for (int r = 0; r < rows; ++r) {
        const double* v = vals + static_cast<size_t>(r) * nnz;
        const double* br = b + static_cast<size_t>(r) * dim;
        const double* xr = x + static_cast<size_t>(r) * dim;
        bool finite = true;
        double num = 0.0, den = 0.0;
        for (int i = 0; i < dim; ++i) {
            if (!std::isfinite(xr[i])) finite = false;
            double ax = 0.0;
            for (int q = p[i]; q < p[i + 1]; ++q) ax += v[q] * xr[ix[q]];
            num = std::max(num, std::fabs(ax - br[i]));
            den = std::max(den, std::fabs(br[i]));
        }
        if (!finite) {
            ++bad;
            continue;
        }
        worst = std::max(worst, den > 0.0 ? num / den : num);
    }
    return {worst, bad};
}

std::vector<double> solve_raw(int dim, const std::vector<int>& indptr, const std::vector<int>& indices,
                              const std::vector<double>& data, const std::vector<double>& rhs, int device,
                              ReorderingAlg, MatchingAlg, PivotEpsilonAlg, int ir_n_steps)
{
    set_device(device);
    if (static_cast<int>(rhs.size()) != dim) throw py::value_error("rhs must have dim entries");
    if (ir_n_steps < 0) throw py::value_error("ir_n_steps must be >= 0");
    RawPlan P(dim, indptr, indices);
    if (data.size() != indices.size()) throw py::value_error("data must have nnz entries");
    std::vector<double> x(static_cast<size_t>(dim));
    check(g2g_lu_solve_batch(P.h, 1, G2G_FP64, data.data(), rhs.data(), x.data(), 1e-12, ir_n_steps, nullptr,
                             nullptr));
    return x;
}

struct RawBench {
    int dim = 0, nnz = 0, batch_size = 0;
    double context_init_ms = 0.0, analysis_ms = 0.0, factorize_ms = 0.0, refactorize_ms = 0.0, solve_ms = 0.0;
    int n_refactorize = 0;
    long long lu_nnz = -1, mem_device_permanent = -1, mem_device_peak = -1, mem_host_permanent = -1,
              mem_host_peak = -1;
    double max_rel_residual = 0.0;
    int n_nonfinite_slots = 0, n_groups = 1;
    double group_cycle_ms = 0.0;
    int n_handles = 1;
    std::string handle_launch = "serial";
    double handle_setup_ms = 0.0, handle_round_ms = 0.0;
    int n_active = 0;
    double active_max_rel_residual = -1.0;
    long long n_inactive_touched = -1;
    double inactive_max_rel_residual = -1.0;
};

RawBench bench_raw(int dim, const std::vector<int>& indptr, const std::vector<int>& indices,
                   const std::vector<double>& data, int batch_size, int n_refactorize, int device, ReorderingAlg,
                   MatchingAlg, PivotEpsilonAlg, int ir_n_steps, int n_groups, int n_handles,
                   const std::string& handle_launch, LinearSolver, int n_active)
{
    set_device(device);
    if (batch_size <= 0) throw py::value_error("batch_size must be positive");
    if (n_refactorize < 0 || ir_n_steps < 0 || n_active < 0 || n_active > batch_size)
        throw py::value_error("n_refactorize, ir_n_steps must be >= 0 and n_active in [0, batch_size]");
    if (n_groups != 1 || n_handles != 1)
        throw py::value_error("n_groups / n_handles > 1 cycle cuDSS handles; this build's batched LU has none");
    RawBench R;
    R.dim = dim;
    R.nnz = static_cast<int>(indices.size());
    R.batch_size = batch_size;
    R.n_refactorize = n_refactorize;
    R.handle_launch = handle_launch;
    R.n_active = n_active;
    double t0 = now_ms();
    RawPlan P(dim, indptr, indices);
    R.analysis_ms = now_ms() - t0;
    R.lu_nnz = P.info.nLU;
    if (data.size() != indices.size()) throw py::value_error("data must have nnz entries");
    const int nnz = R.nnz;
    // every slot holds the same values; rhs = A 1, so x = 1
    std::vector<double> vals(static_cast<size_t>(batch_size) * nnz), b(static_cast<size_t>(batch_size) * dim, 0.0),
        x(b.size());
    for (int r = 0; r < batch_size; ++r) std::copy(data.begin(), data.end(), vals.begin() + static_cast<size_t>(r) * nnz);
    for (int i = 0; i < dim; ++i) {
        double s = 0.0;
        for (int q = indptr[i]; q < indptr[i + 1]; ++q) s += data[q];
        for (int r = 0; r < batch_size; ++r) b[static_cast<size_t>(r) * dim + i] = s;
    }
    auto solve = [&](int rows) {
        check(g2g_lu_solve_batch(P.h, rows, G2G_FP64, vals.data(), b.data(), x.data(), 1e-12, ir_n_steps, nullptr,
                                 nullptr));
    };
    {
        py::gil_scoped_release nogil;
        t0 = now_ms();
        solve(batch_size);
        R.factorize_ms = now_ms() - t0;
        double tot = 0.0;
        for (int k = 0; k < n_refactorize; ++k) {
            t0 = now_ms();
            solve(batch_size);
            tot += now_ms() - t0;
        }
        // one call factors and solves: its time is the refactorization's
        R.refactorize_ms = n_refactorize ? tot / n_refactorize : 0.0;
    }
    auto rr = raw_residual(dim, indptr, indices, vals.data(), b.data(), x.data(), batch_size, nnz);
    R.max_rel_residual = rr.first;
    R.n_nonfinite_slots = rr.second;
    if (n_active > 0) {
        // the prefix alone, with its values tripled; the other rows are not touched
        for (size_t q = 0; q < static_cast<size_t>(n_active) * nnz; ++q) vals[q] *= 3.0;
        for (size_t q = 0; q < static_cast<size_t>(n_active) * dim; ++q) b[q] *= 3.0;
        std::vector<double> keep(x.begin() + static_cast<size_t>(n_active) * dim, x.end());
        solve(n_active);
        R.active_max_rel_residual =
            raw_residual(dim, indptr, indices, vals.data(), b.data(), x.data(), n_active, nnz).first;
        long long touched = 0;
        for (size_t q = 0; q < keep.size(); ++q) touched += keep[q] != x[static_cast<size_t>(n_active) * dim + q];
        R.n_inactive_touched = touched;
    }
    return R;
}

}  // namespace

PYBIND11_MODULE(_gpusim2grid, m)
{
    m.doc() = "MI355X batched Newton-Raphson power flow (HIP kernels behind the g2g C ABI)";

    py::enum_<RowStatus>(m, "RowStatus")
        .value("Converged", RowStatus::Converged)
        .value("MaxIter", RowStatus::MaxIter)
        .value("Diverged", RowStatus::Diverged)
        .value("NotSimulated", RowStatus::NotSimulated);
    py::enum_<BatchScheduling>(m, "BatchScheduling")
        .value("Chunked", BatchScheduling::Chunked)
        .value("Continuous", BatchScheduling::Continuous);
    py::enum_<LinearSolver>(m, "LinearSolver")
        .value("Cudss", LinearSolver::Cudss, "accepted for compatibility: the batched LU of this build runs")
        .value("OwnLu", LinearSolver::OwnLu);
    py::enum_<ReorderingAlg>(m, "ReorderingAlg")
        .value("Default", ReorderingAlg::Default)
        .value("BtfColamd", ReorderingAlg::BtfColamd)
        .value("Colamd", ReorderingAlg::Colamd)
        .value("Amd", ReorderingAlg::Amd)
        .value("NestedDissection", ReorderingAlg::NestedDissection)
        .value("None_", ReorderingAlg::None);
    py::enum_<MatchingAlg>(m, "MatchingAlg")
        .value("None_", MatchingAlg::None)
        .value("MaxDiagCount", MatchingAlg::MaxDiagCount)
        .value("MaxMinDiag", MatchingAlg::MaxMinDiag)
        .value("MaxMinDiagAlt", MatchingAlg::MaxMinDiagAlt)
        .value("MaxDiagSum", MatchingAlg::MaxDiagSum)
        .value("MaxDiagProduct", MatchingAlg::MaxDiagProduct)
        .value("Auto", MatchingAlg::Auto);
    py::enum_<PivotEpsilonAlg>(m, "PivotEpsilonAlg")
        .value("Default", PivotEpsilonAlg::Default)
        .value("Scaled", PivotEpsilonAlg::Scaled)
        .value("Static", PivotEpsilonAlg::Static);
    py::enum_<ContingencySolverType>(m, "ContingencySolverType")
        .value("DirectRefactorEvery", ContingencySolverType::DirectRefactorEvery)
        .value("DirectIter0Only", ContingencySolverType::DirectIter0Only)
        .value("DirectRefactorEveryN", ContingencySolverType::DirectRefactorEveryN);
    py::enum_<Precision>(m, "Precision").value("Fp64", Precision::Fp64).value("Mixed", Precision::Mixed);

    py::class_<TimingEntry>(m, "TimingEntry")
        .def_readonly("gpu_ms", &TimingEntry::gpu_ms)
        .def_readonly("wall_ms", &TimingEntry::wall_ms);
    py::class_<AcPfTimings>(m, "AcPfTimings")
        .def_readonly("t_init_ms", &AcPfTimings::t_init_ms)
        .def_readonly("t_analyze_ms", &AcPfTimings::t_analyze_ms)
        .def_readonly("t_mismatch", &AcPfTimings::t_mismatch)
        .def_readonly("t_refactorize", &AcPfTimings::t_refactorize)
        .def_readonly("t_solve", &AcPfTimings::t_solve)
        .def_readonly("t_total_ms", &AcPfTimings::t_total_ms)
        .def_readonly("t_copy_v_to_host_ms", &AcPfTimings::t_copy_v_to_host_ms)
        .def_readonly("nb_iter", &AcPfTimings::nb_iter)
        .def_readonly("n_refactorize", &AcPfTimings::n_refactorize)
        .def_readonly("converged", &AcPfTimings::converged)
        .def_readonly("residual", &AcPfTimings::residual);
    py::class_<BatchTimings>(m, "BatchTimings")
        .def_readonly("t_base_case_ms", &BatchTimings::t_base_case_ms)
        .def_readonly("t_analysis_ms", &BatchTimings::t_analysis_ms)
        .def_readonly("t_alloc_ms", &BatchTimings::t_alloc_ms)
        .def_readonly("t_preprocess_ms", &BatchTimings::t_preprocess_ms)
        .def_readonly("t_mismatch", &BatchTimings::t_mismatch)
        .def_readonly("t_refactorize", &BatchTimings::t_refactorize)
        .def_readonly("t_solve", &BatchTimings::t_solve)
        .def_readonly("t_warm_start", &BatchTimings::t_warm_start)
        .def_readonly("t_run_ms", &BatchTimings::t_run_ms)
        .def_readonly("t_copy_V_to_host_ms", &BatchTimings::t_copy_V_to_host_ms)
        .def_readonly("t_violations_ms", &BatchTimings::t_violations_ms)
        .def_readonly("n_chunks", &BatchTimings::n_chunks)
        .def_readonly("n_rows", &BatchTimings::n_rows)
        .def_readonly("n_converged", &BatchTimings::n_converged)
        .def_readonly("n_requeued", &BatchTimings::n_requeued)
        .def_readonly("rows_factored", &BatchTimings::rows_factored)
        .def_property_readonly("t_total_ms", &BatchTimings::total_ms);
    m.attr("ContingencyTimings") = m.attr("BatchTimings");

    py::class_<AcPfNrSession, std::shared_ptr<AcPfNrSession>>(m, "AcPfNrSession")
        .def(py::init([](py::object Ybus, CArr Vinit, CArr Sbus, IArr slack_ids, DArr slack_weights, IArr pv,
                         IArr pq, int max_iter, double tol, int device, bool presolved_v, ReorderingAlg,
                         MatchingAlg, PivotEpsilonAlg, int ir_n_steps, bool, bool scaling, double max_dva,
                         double max_dvm, Precision precision, py::object extensions) {
                 return std::make_shared<AcPfNrSession>(Ybus, Vinit, Sbus, slack_ids, slack_weights, pv, pq,
                                                        max_iter, tol, device, presolved_v,
                                                        StepControl(ir_n_steps, scaling, max_dva, max_dvm), precision,
                                                        extensions);
             }),
             py::arg("Ybus"), py::arg("Vinit"), py::arg("Sbus"), py::arg("slack_ids"), py::arg("slack_weights"),
             py::arg("pv"), py::arg("pq"), py::arg("max_iter"), py::arg("tol"), py::arg("device") = -1,
             py::arg("presolved_v") = false, py::arg("reordering_alg") = ReorderingAlg::Default,
             py::arg("matching_alg") = MatchingAlg::None, py::arg("pivot_epsilon_alg") = PivotEpsilonAlg::Default,
             py::arg("ir_n_steps") = 0, py::arg("debug_base_case") = false,
             py::arg("scaling_max_voltage_change") = false, py::arg("max_dVa") = 0.5, py::arg("max_dVm") = 0.1,
             py::arg("precision") = Precision::Fp64, py::arg("extensions") = py::none())
        .def_property_readonly("timings", &AcPfNrSession::timings)
        .def("get_ext_state", &AcPfNrSession::get_ext_state)
        .def("resolve", &AcPfNrSession::resolve, py::arg("Vinit"), py::arg("Sbus") = py::none(),
             py::arg("max_iter") = -1, py::arg("tol") = -1.0)
        .def("solve_at_setpoints", &AcPfNrSession::solve_at_setpoints, py::arg("Vinit"), py::arg("max_iter"),
             py::arg("tol"))
        .def("get_v", &AcPfNrSession::get_v)
        .def("v_dlpack", &AcPfNrSession::v_dlpack)
        .def("get_F", &AcPfNrSession::get_F)
        .def("get_J", &AcPfNrSession::get_J)
        .def("solve_JT", &AcPfNrSession::solve_JT, py::arg("rhs"))
        .def_property_readonly("n_pvpq", &AcPfNrSession::n_pvpq)
        .def_property_readonly("n_pq", &AcPfNrSession::n_pq)
        .def_property_readonly("dim_J", &AcPfNrSession::dim_J)
        .def_property_readonly("pvpq", &AcPfNrSession::pvpq)
        .def_property_readonly("pq", &AcPfNrSession::pq)
        .def_property_readonly("p_row_of_bus", &AcPfNrSession::p_row_of_bus)
        .def_property_readonly("q_row_of_bus", &AcPfNrSession::q_row_of_bus)
        .def_property_readonly("theta_col_of_bus", &AcPfNrSession::p_row_of_bus)
        .def_property_readonly("vm_col_of_bus", &AcPfNrSession::q_row_of_bus);

    m.def(
        "acpf_nr_gpu",
        [](py::array_t<cplx> out, py::object Ybus, CArr V, CArr Sbus, IArr slack_ids, DArr slack_weights, IArr pv,
           IArr pq, int max_iter, double tol, int nb_solve, int device) {
            if (nb_solve < 1) throw py::value_error("nb_solve must be >= 1");
            AcPfNrSession s(Ybus, V, Sbus, slack_ids, slack_weights, pv, pq, max_iter, tol, device, false, StepControl(),
                            Precision::Fp64);
            for (int k = 1; k < nb_solve; ++k) s.resolve(V, py::none(), -1, -1.0);
            CArr v = s.get_v();
            if (out.ndim() != 1 || out.shape(0) != v.shape(0))
                throw py::value_error("out must be a complex128 array of n_bus entries");
            std::copy(v.data(), v.data() + v.size(), out.mutable_data());
            return s.timings();
        },
        py::arg("out"), py::arg("ac_Ybus"), py::arg("V"), py::arg("Sbus"), py::arg("slack_ids"),
        py::arg("slack_weights"), py::arg("pv"), py::arg("pq"), py::arg("max_iter"), py::arg("tol"),
        py::arg("nb_solve"), py::arg("device") = -1);

    auto bind_batch = [](auto& c) {
        using S = typename std::remove_reference_t<decltype(c)>::type;
        c.def("run", &S::run, py::call_guard<py::gil_scoped_release>())
            .def("get_V_results", &S::get_V_results)
            .def("get_residuals", &S::get_residuals)
            .def("get_row_iterations", &S::get_row_iterations)
            .def("get_row_status", &S::get_row_status)
            .def("get_timings", &S::get_timings)
            .def_property_readonly("timings", &S::get_timings)
            .def_property_readonly("n_bus", &S::n_bus)
            .def_property_readonly("dim_J", &S::dim_J)
            .def("get_v_base", &S::v_base)
            .def_property("batch_size", &S::get_batch_size, &S::set_batch_size)
            .def_property("nb_iter", &S::get_nb_iter, &S::set_nb_iter)
            .def_property("scheduling", &S::get_scheduling, &S::set_scheduling)
            .def_property("tol", &S::get_tol, &S::set_tol)
            .def_property("precision", &S::get_precision, &S::set_precision)
            .def("set_warm_start", &S::set_warm_start, py::arg("mlp_handle"), py::arg("k_warm") = 4,
                 py::arg("growth") = 2.0)
            .def("set_branch_data", &S::set_branch_data, py::arg("branch_from"), py::arg("branch_to"),
                 py::arg("yff_eff"), py::arg("yft_eff"), py::arg("ytf_eff"), py::arg("ytt_eff"), py::arg("bus_vn_kv"),
                 py::arg("sn_mva"))
            .def_property_readonly("n_branches", &S::n_branches)
            .def("compute_flows", &S::compute_flows, py::call_guard<py::gil_scoped_release>())
            .def("get_or_amps", &S::get_or_amps)
            .def("get_ex_amps", &S::get_ex_amps)
            .def("clear_warm_start", &S::clear_warm_start)
            .def_readwrite("handle_disconnected_grid", &S::handle_disconnected_grid)
            .def("v_results_dlpack", &S::v_results_dlpack)
            .def("get_ext_state", &S::get_ext_state)
            .def_readwrite("compute_limit_violations", &S::compute_limit_violations)
            .def_readwrite("violation_tol", &S::violation_tol)
            .def_property("violation_rel_tol", &S::get_violation_rel_tol, &S::set_violation_rel_tol)
            .def_property("violation_capacity", &S::get_violation_capacity, &S::set_violation_capacity)
            .def("set_limits", &S::set_limits, py::arg("bus_vmin_kv"), py::arg("bus_vmax_kv"),
                 py::arg("branch_limit_a1_ka"), py::arg("branch_limit_a2_ka"), py::arg("n_lines"))
            .def("get_violation_count", [](const S& s) { return s.viol_i(&ViolResults::count); })
            .def("get_violation_truncated", [](const S& s) { return s.viol_i(&ViolResults::truncated); })
            .def("get_violation_count_low_voltage", [](const S& s) { return s.viol_i(&ViolResults::n_low); })
            .def("get_violation_count_high_voltage", [](const S& s) { return s.viol_i(&ViolResults::n_high); })
            .def("get_violation_count_current", [](const S& s) { return s.viol_i(&ViolResults::n_current); })
            .def("get_violation_element_type", [](const S& s) { return s.viol_i(&ViolResults::etype); })
            .def("get_violation_element_id", [](const S& s) { return s.viol_i(&ViolResults::eid); })
            .def("get_violation_side", [](const S& s) { return s.viol_i(&ViolResults::side); })
            .def("get_violation_type", [](const S& s) { return s.viol_i(&ViolResults::type); })
            .def("get_violation_value", [](const S& s) { return s.viol_d(&ViolResults::value); })
            .def("get_violation_limit", [](const S& s) { return s.viol_d(&ViolResults::limit); })
            .def("get_disconnected", &S::get_disconnected);
    };
    auto batch_init = [](auto* tag) {
        using S = std::remove_pointer_t<decltype(tag)>;
        return py::init([](py::object Ybus, CArr Vinit, CArr Sbus, IArr slack_ids, DArr, IArr pv, IArr pq,
                           int batch_size, int nb_iter, int max_iter_base, double tol_base, int device,
                           bool presolved_v, ReorderingAlg, MatchingAlg, PivotEpsilonAlg, int ir_n_steps, bool,
                           bool scaling, double max_dva, double max_dvm, Precision precision, py::object extensions) {
            return std::make_shared<S>(Ybus, Vinit, Sbus, slack_ids, pv, pq, batch_size, nb_iter, max_iter_base,
                                       tol_base, device, presolved_v, StepControl(ir_n_steps, scaling, max_dva, max_dvm),
                                       precision, extensions);
        });
    };
#define G2G_BATCH_ARGS                                                                                          \
    py::arg("Ybus"), py::arg("Vinit"), py::arg("Sbus"), py::arg("slack_ids"), py::arg("slack_weights"),         \
        py::arg("pv"), py::arg("pq"), py::arg("batch_size"), py::arg("nb_iter"), py::arg("max_iter_base") = 10, \
        py::arg("tol_base") = 1e-6, py::arg("device") = -1, py::arg("presolved_v") = false,                     \
        py::arg("reordering_alg") = ReorderingAlg::Default, py::arg("matching_alg") = MatchingAlg::None,        \
        py::arg("pivot_epsilon_alg") = PivotEpsilonAlg::Default, py::arg("ir_n_steps") = 0,                     \
        py::arg("debug_base_case") = false, py::arg("scaling_max_voltage_change") = false,                      \
        py::arg("max_dVa") = 0.5, py::arg("max_dVm") = 0.1, py::arg("precision") = Precision::Fp64,             \
        py::arg("extensions") = py::none()

    py::class_<InjectionSweepSession, std::shared_ptr<InjectionSweepSession>> inj(m, "InjectionSweepSession");
    inj.def(batch_init(static_cast<InjectionSweepSession*>(nullptr)), G2G_BATCH_ARGS)
        .def("set_injections", &InjectionSweepSession::set_injections, py::arg("p_mw"), py::arg("q_mvar"),
             py::arg("sn_mva"))
        .def("set_sbus", &InjectionSweepSession::set_sbus, py::arg("Sbus"))
        .def("set_v_init_from", &InjectionSweepSession::set_v_init_from, py::arg("V"))
        .def("set_gen_v", &InjectionSweepSession::set_gen_v, py::arg("gen_v"), py::arg("gen_bus"))
        .def("clear_gen_v", &InjectionSweepSession::clear_gen_v)
        .def_property_readonly("is_vm_fixed_bus", &InjectionSweepSession::is_vm_fixed_bus)
        .def_property_readonly("n_scenarios", &InjectionSweepSession::n_scenarios);
    bind_batch(inj);

    py::class_<ContingencyAnalysisSession, std::shared_ptr<ContingencyAnalysisSession>> ca(
        m, "ContingencyAnalysisSession");
    ca.def(batch_init(static_cast<ContingencyAnalysisSession*>(nullptr)), G2G_BATCH_ARGS)
        .def("build_contingencies", &ContingencyAnalysisSession::build_contingencies,
             py::arg("branch_ids_per_ctg"))
        .def_property_readonly("n_contingencies", &ContingencyAnalysisSession::n_contingencies);
    bind_batch(ca);

    py::class_<ScenarioSweepSession, std::shared_ptr<ScenarioSweepSession>> ss(m, "ScenarioSweepSession");
    ss.def(batch_init(static_cast<ScenarioSweepSession*>(nullptr)), G2G_BATCH_ARGS)
        .def("set_injections", &ScenarioSweepSession::set_injections, py::arg("p_mw"), py::arg("q_mvar"),
             py::arg("sn_mva"))
        .def("set_sbus", &ScenarioSweepSession::set_sbus, py::arg("Sbus"))
        .def("set_topology", &ScenarioSweepSession::set_topology, py::arg("branch_ids_per_row"))
        .def("clear_topology", &ScenarioSweepSession::clear_topology)
        .def("set_v_init_from", &ScenarioSweepSession::set_v_init_from, py::arg("V"))
        .def("clear_v_init", &ScenarioSweepSession::clear_v_init)
        .def("set_gen_v", &ScenarioSweepSession::set_gen_v, py::arg("gen_v"), py::arg("gen_bus"))
        .def("clear_gen_v", &ScenarioSweepSession::clear_gen_v)
        .def_property_readonly("is_vm_fixed_bus", &ScenarioSweepSession::is_vm_fixed_bus)
        .def_property_readonly("has_v_init", &ScenarioSweepSession::has_v_init)
        .def_property_readonly("n_scenarios", &ScenarioSweepSession::n_scenarios);
    bind_batch(ss);
#undef G2G_BATCH_ARGS

    m.def(
        "acpf_nr_gpu_injection",
        [](py::object Ybus, CArr Vinit, CArr Sbus, IArr slack_ids, DArr slack_weights, IArr pv, IArr pq, DArr p_mw,
           DArr q_mvar, double sn_mva, int batch_size, int nb_iter, int max_iter_base, double tol_base,
           ContingencySolverType, int) {
            auto s = std::make_shared<InjectionSweepSession>(Ybus, Vinit, Sbus, slack_ids, pv, pq, batch_size,
                                                             nb_iter, max_iter_base, tol_base, -1, false, StepControl(),
                                                             Precision::Fp64);
            s->set_injections(p_mw, q_mvar, sn_mva);
            {
                py::gil_scoped_release nogil;
                s->run();
            }
            return py::make_tuple(s->get_V_results(), s->get_residuals(), s->get_timings());
        },
        py::arg("Ybus"), py::arg("Vinit"), py::arg("Sbus"), py::arg("slack_ids"), py::arg("slack_weights"),
        py::arg("pv"), py::arg("pq"), py::arg("p_mw"), py::arg("q_mvar"), py::arg("sn_mva"), py::arg("batch_size"),
        py::arg("nb_iter"), py::arg("max_iter_base") = 10, py::arg("tol_base") = 1e-6,
        py::arg("strategy_type") = ContingencySolverType::DirectRefactorEvery, py::arg("refactor_period") = 1);

    m.def(
        "warmup",
        [](int device) {
            set_device(device);
            // one tiny solve loads the kernels' code objects
            py::object sp = py::module_::import("scipy.sparse");
            std::vector<cplx> y{cplx(1, -10), cplx(-1, 10), cplx(-1, 10), cplx(1, -10)};
            py::array_t<cplx> data(4);
            std::copy(y.begin(), y.end(), data.mutable_data());
            py::array_t<int32_t> ind(std::vector<py::ssize_t>{4});
            py::array_t<int32_t> ptr(std::vector<py::ssize_t>{3});
            const int32_t ii[4] = {0, 1, 0, 1}, pp[3] = {0, 2, 4};
            std::copy(ii, ii + 4, ind.mutable_data());
            std::copy(pp, pp + 3, ptr.mutable_data());
            py::object Y = sp.attr("csr_matrix")(py::make_tuple(data, ind, ptr), py::make_tuple(2, 2));
            CArr V(2), S(2);
            V.mutable_data()[0] = V.mutable_data()[1] = cplx(1, 0);
            S.mutable_data()[0] = cplx(0, 0);
            S.mutable_data()[1] = cplx(-0.1, -0.05);
            IArr sl(1), pv(0), pq(1);
            sl.mutable_data()[0] = 0;
            pq.mutable_data()[0] = 1;
            DArr w(1);
            w.mutable_data()[0] = 1.0;
            AcPfNrSession s(Y, V, S, sl, w, pv, pq, 10, 1e-8, device, false, StepControl(), Precision::Fp64);
            AcPfNrSession s2(Y, V, S, sl, w, pv, pq, 10, 1e-8, device, false, StepControl(), Precision::Mixed);
        },
        py::arg("device") = -1);
    m.def("solve_cudss_raw", &solve_raw, py::call_guard<py::gil_scoped_release>(), py::arg("dim"),
          py::arg("indptr"), py::arg("indices"), py::arg("data"), py::arg("rhs"), py::arg("device") = -1,
          py::arg("reordering_alg") = ReorderingAlg::Default, py::arg("matching_alg") = MatchingAlg::None,
          py::arg("pivot_epsilon_alg") = PivotEpsilonAlg::Default, py::arg("ir_n_steps") = 0,
          "Solve J dx = rhs for one CSR matrix through the batched LU (plan, factor, solve).");
    py::class_<RawBench>(m, "CudssBatchBenchResult")
        .def_readonly("dim", &RawBench::dim)
        .def_readonly("nnz", &RawBench::nnz)
        .def_readonly("batch_size", &RawBench::batch_size)
        .def_readonly("context_init_ms", &RawBench::context_init_ms)
        .def_readonly("analysis_ms", &RawBench::analysis_ms)
        .def_readonly("factorize_ms", &RawBench::factorize_ms)
        .def_readonly("refactorize_ms", &RawBench::refactorize_ms)
        .def_readonly("solve_ms", &RawBench::solve_ms)
        .def_readonly("n_refactorize", &RawBench::n_refactorize)
        .def_readonly("lu_nnz", &RawBench::lu_nnz)
        .def_readonly("mem_device_permanent", &RawBench::mem_device_permanent)
        .def_readonly("mem_device_peak", &RawBench::mem_device_peak)
        .def_readonly("mem_host_permanent", &RawBench::mem_host_permanent)
        .def_readonly("mem_host_peak", &RawBench::mem_host_peak)
        .def_readonly("max_rel_residual", &RawBench::max_rel_residual)
        .def_readonly("n_nonfinite_slots", &RawBench::n_nonfinite_slots)
        .def_readonly("n_groups", &RawBench::n_groups)
        .def_readonly("group_cycle_ms", &RawBench::group_cycle_ms)
        .def_readonly("n_handles", &RawBench::n_handles)
        .def_readonly("handle_launch", &RawBench::handle_launch)
        .def_readonly("handle_setup_ms", &RawBench::handle_setup_ms)
        .def_readonly("handle_round_ms", &RawBench::handle_round_ms)
        .def_readonly("n_active", &RawBench::n_active)
        .def_readonly("active_max_rel_residual", &RawBench::active_max_rel_residual)
        .def_readonly("n_inactive_touched", &RawBench::n_inactive_touched)
        .def_readonly("inactive_max_rel_residual", &RawBench::inactive_max_rel_residual);
    m.def("benchmark_cudss_batch_raw", &bench_raw, py::arg("dim"), py::arg("indptr"), py::arg("indices"),
          py::arg("data"), py::arg("batch_size"), py::arg("n_refactorize") = 4, py::arg("device") = -1,
          py::arg("reordering_alg") = ReorderingAlg::Default, py::arg("matching_alg") = MatchingAlg::None,
          py::arg("pivot_epsilon_alg") = PivotEpsilonAlg::Default, py::arg("ir_n_steps") = 0,
          py::arg("n_groups") = 1, py::arg("n_handles") = 1, py::arg("handle_launch") = "serial",
          py::arg("linear_solver") = LinearSolver::Cudss, py::arg("n_active") = 0,
          "Time the batched LU on one CSR matrix replicated batch_size times (rhs = A 1, so x = 1): "
          "plan (analysis_ms), first call (factorize_ms), then n_refactorize calls (refactorize_ms: "
          "a call factors and solves; solve_ms stays 0).");
    m.attr("is_fp32") = false;
    m.attr("have_ls2g_bridge") = false;
    m.attr("abi_version") = g2g_version();
}
