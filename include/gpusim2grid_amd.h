/*
 * gpusim2grid_amd.h -- C ABI of the MI355X batched Newton-Raphson power flow.
 *
 * Plain C types only (int32 indices, float64 values, complex numbers as
 * interleaved (re, im) float64 pairs, opaque handles, a HIP stream passed as
 * void*).  Every function returns G2G_OK or an error code; g2g_last_error()
 * gives the message of the last failure on the calling thread.  Per-row
 * numerical failures are statuses, never errors (gpusim2grid CLAUDE.md:172).
 *
 * The reference binds its solvers through pybind11 (src/_cpp/python_bindings.cpp)
 * onto C++ sessions; the seams these entry points replace are cited per
 * function.  Handles are not thread-safe: one stream per batch, like the
 * reference's sessions (src/_cpp/acpf_nr_state.cuh:29).
 */
#ifndef GPUSIM2GRID_AMD_H
#define GPUSIM2GRID_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define G2G_OK 0
#define G2G_ERR_INVALID 1 /* bad argument (Python: ValueError)            */
#define G2G_ERR_RUNTIME 2 /* unsupported case or internal failure         */
#define G2G_ERR_HIP 3     /* HIP runtime error                              */

/* precision of the Jacobian, the factorization and the triangular solves;
 * the mismatch, the convergence test and the V state are always FP64 */
#define G2G_FP64 0
#define G2G_MIXED 1

/* schedules (src/_cpp/contingency/row_status.hpp:28-31) */
#define G2G_SCHED_CHUNKED 0    /* every row runs exactly max_iter updates  */
#define G2G_SCHED_CONTINUOUS 1 /* a row stops once ||F||inf < tol           */

/* row status (src/_cpp/contingency/row_status.hpp:33-43) */
#define G2G_CONVERGED 0
#define G2G_MAXITER 1
#define G2G_DIVERGED 2
#define G2G_NOTSIMULATED 3

/* arrays g2g_plan_get_array can copy out (test and integration introspection) */
#define G2G_ARR_JP 0        /* reference J CSR indptr [d+1]                   */
#define G2G_ARR_JI 1        /* reference J CSR indices [nJ]                   */
#define G2G_ARR_MAPS 2      /* J position of each Ybus entry, blocks 11,12,21,22 [4 nY] */
#define G2G_ARR_PVPQ 3      /* sorted pvpq [npvpq]                            */
#define G2G_ARR_PERM 4      /* pivot -> reference unknown [d]                 */
#define G2G_ARR_COLP 5      /* L column structure indptr [d+1] (pivot labels) */
#define G2G_ARR_COLI 6      /* L column structure indices                     */
#define G2G_ARR_PROG_F 7    /* factorization program                          */
#define G2G_ARR_PROG_B 8    /* backward program                               */
#define G2G_ARR_APOS 9      /* A-stream position of each input CSR entry [nnz] */
#define G2G_ARR_FPOS 10     /* A-stream position of each rhs entry [d]        */
#define G2G_ARR_UOFF 11     /* U-stream record start per pivot [d+1]          */
#define G2G_ARR_SEGS 12     /* per walker: prog_f off/len, prog_b off/len, A off/len, U off/len, pivots [9 walkers] */
#define G2G_ARR_BSEGS 13    /* the same per backward-substitution walker */
#define G2G_ARR_PROG_FS 14  /* split walk: factorization programs of the sets and the top part */
#define G2G_ARR_PROG_BS 15  /* split walk: backward programs */
#define G2G_ARR_SPLIT_SEGS 16  /* split walk: factorization segments (9 ints per set, then the top) */
#define G2G_ARR_SPLIT_BSEGS 17 /* split walk: backward segments */
#define G2G_ARR_SPLIT_APOS 18  /* split walk: A-stream position of each input CSR entry */
#define G2G_ARR_SPLIT_FPOS 19  /* split walk: A-stream position of each rhs entry */

typedef struct g2g_plan_s *g2g_plan_t;
typedef struct g2g_batch_s *g2g_batch_t;
typedef struct g2g_mlp_s *g2g_mlp_t;

typedef struct {
    int32_t n, nY, d, nJ, npvpq, npq, slack; /* grid (0 / -1 for a raw matrix plan)   */
    int32_t nnz;                             /* input pattern entries (nJ for grids)  */
    int32_t nLU, n_updates, max_c, height;   /* symbolic factorization                */
    int32_t slots_f, slots_b, a_len, u_len;  /* program resources                     */
    int32_t n_steps;                         /* factorization walk steps (pivots and 2-column supernodes) */
    int32_t walkers, n_top;                  /* walker wavefronts per tile; pivots of the top part */
    int32_t bwalkers;                        /* walker wavefronts per tile of the backward substitution */
    int32_t u_stride_f64, u_stride_f32;      /* U positions a tile holds in a batch of this precision */
    int32_t n_ext;                           /* extension rows (= columns) after [P(pvpq) | Q(pq)] */
    int32_t n_xs;                            /* extension state values per row (sa, then each controller's Q) */
    int32_t sa_col;                          /* J column of sa (= the slack's P row), -1 without distributed slack */
    int32_t n_split;                         /* split walk: subtree sets per tile (0: none) */
    int32_t slots_fs, slots_bs, pend_len;    /* split walk: slots and pending contributions per tile */
    int32_t split_a_len, split_u_len;        /* split walk: its A- and U-stream lengths */
    int32_t a_stride_f64, a_stride_f32;      /* A positions a tile holds in a batch of this precision */
} g2g_plan_info;

typedef struct {
    int32_t iterations;     /* Newton iterations launched                        */
    int32_t n_converged, n_maxiter, n_diverged;
    double ms_total;        /* GPU time of the whole solve (hipEvents)           */
    double ms_mismatch;     /* mismatch + Jacobian fill                          */
    double ms_factor;       /* factorization + forward substitution              */
    double ms_backward;     /* backward substitution + V update                  */
    int64_t bytes_factor;   /* A- and U-stream bytes of one factorization launch  */
    int64_t rows_factored;  /* rows summed over the factorization launches       */
    int32_t n_factor_launches;
    int32_t n_requeued;     /* warm solves: rows restarted from the caller's V   */
    double ms_warm_start;   /* warm solves: MLP inference                        */
    /* the main-layout factorization launches alone (whole-tile walks; the
     * split walk's tail launches excluded): time, rows, launches */
    double ms_factor_main;
    int64_t rows_factored_main;
    int32_t n_factor_main;
} g2g_solve_stats;

const char *g2g_last_error(void);
int g2g_version(void);
/* Makes `device` the calling thread's current device (-1: leave it), as the
 * reference's `device` arguments do (acpf_nr.cu:23-39). */
int g2g_set_device(int32_t device);

/* Builds every row-independent table of a grid: the reference's [pvpq | pq]
 * Jacobian pattern and scatter maps (replaces build_J_structure,
 * src/_cpp/acpf_nr.cu:145-236, and the host part of AcPfNrState's
 * constructor, acpf_nr.cu:567-806), the fill-reducing pivot order and the
 * factorization / substitution programs (replaces cuDSS ANALYSIS,
 * src/_cpp/contingency/strategies/cudss_batch_solver.cuh:339-440, and the own
 * LU's build_lu_plan, src/_cpp/own_lu/lu_plan.cpp:88-285).  Ybus is CSR with
 * sorted unique columns; Yx holds 2*nY doubles; pv/pq/slack are bus indices.
 * Every bus in neither pv nor pq is fixed (no unknowns: its V stays as the
 * row's start V), as the reference's array mode treats them
 * (acpf_nr.cu:644-647,1732-1733); `slack` is one of them. */
int g2g_plan_create(int32_t n, const int32_t *Yp, const int32_t *Yi, const double *Yx, int32_t npv,
                    const int32_t *pv, int32_t npq, const int32_t *pq, int32_t slack, g2g_plan_t *out);

/* Extension rows of the Newton system (the reference's MultiSlack, HVDC
 * angle-droop and VoltageControl features of its augmented ledger,
 * ledger_data.hpp:61-155, acpf_nr_kernels.cu:793-1068, nr_iter_step.cuh:302-413).
 * Each part is off when its arrays are NULL / its count is 0.
 *
 * Layout: rows [P of sorted pvpq | Q of pq | extension rows] and columns
 * [theta of pvpq | |V| of pq | extension columns]; the extension block is,
 * in this order: the distributed slack (the slack's P row, column sa), then
 * per voltage-control group its voltage row (column Q of its first
 * controller) and one sharing row per further controller (column Q of that
 * controller).  Per row the solve carries the state sa (starting at
 * sum Re(Sbus)) and each controller's Q (starting at 0), all FP64.
 *
 * Distributed slack: F_P(i) = -(P_i(V) - Re S_i) - sa w_i at every bus with a
 * P row, J(P_i, sa) = w_i.  Participants (w_i != 0) are pv / pq buses or the
 * slack, which gets a P row.
 * HVDC line e (bus1 -> bus2): raw = p0 + k (theta1 - theta2); linear regime
 * (status 0): the sending end (raw >= 0: bus1) draws |raw| and the other end
 * receives (1 - lf_r)(L - r L^2) with L = (1 - lf_s)|raw|; status +1 / -1
 * fixes the sent power at pmax12 / pmax21, 2 trips the line.  The flows
 * leaving each end enter its P residual; in the linear regime J gets the
 * exact slopes at (P1, theta1), (P1, theta2), (P2, theta1), (P2, theta2).
 * Voltage control group g: |V(reg)| + sum_c s_c Q_c = v_set (voltage row),
 * w_1 Q_c - w_c Q_1 = 0 (sharing rows), and each controller's Q enters the Q
 * residual of its bus (J(Q_bus, Q_c) = -1).  Regulated and controller buses
 * are pq buses; weights are positive. */
typedef struct g2g_extensions {
    const double *slack_weights;                     /* [n] or NULL */
    int32_t n_hvdc;
    const int32_t *hvdc_bus1, *hvdc_bus2, *hvdc_status;
    const double *hvdc_p0, *hvdc_k, *hvdc_lf1, *hvdc_lf2, *hvdc_r, *hvdc_pmax12, *hvdc_pmax21;
    int32_t n_vc_groups;
    const int32_t *vc_reg_bus;                       /* [groups] */
    const double *vc_v_set;                          /* [groups] */
    const int32_t *vc_ptr;                           /* [groups + 1] controllers of each group */
    const int32_t *vc_bus;                           /* [controllers] */
    const double *vc_slope, *vc_weight;              /* [controllers] */
} g2g_extensions;

/* g2g_plan_create with extension rows (ext nullable: the bare system).
 * Batches of such a plan run the plain solve (g2g_batch_solve), evaluation
 * and results; step scaling, warm starts, streamed rows, bus masks and the
 * adjoint reject them. */
int g2g_plan_create_ext(int32_t n, const int32_t *Yp, const int32_t *Yi, const double *Yx, int32_t npv,
                        const int32_t *pv, int32_t npq, const int32_t *pq, int32_t slack, const g2g_extensions *ext,
                        g2g_plan_t *out);

/* Plan of a bare square matrix pattern for the raw batched LU (replaces
 * CudssBatchSolver::initialize, cudss_batch_solver.cuh:273-440, behind
 * benchmark_cudss_batch_raw, src/_cpp/raw_cudss_solve.hpp:105-121).  `order`
 * (nullable) fixes the elimination order before the postorder relabelling. */
int g2g_lu_plan_create(int32_t d, const int32_t *Ap, const int32_t *Ai, const int32_t *order, g2g_plan_t *out);

void g2g_plan_destroy(g2g_plan_t plan);
int g2g_plan_get_info(g2g_plan_t plan, g2g_plan_info *info);
int g2g_plan_get_array(g2g_plan_t plan, int32_t which, int32_t *out, int64_t capacity, int64_t *length);

/* Device state of `rows` power flows sharing a plan (replaces BatchPfDriver's
 * allocations, src/_cpp/contingency/batch_pf_driver.cu:122-330).
 * per_row_ybus = 1 keeps one Ybus value vector per row (outage batches).
 * stream: hipStream_t (NULL = default stream). */
int g2g_batch_create(g2g_plan_t plan, int32_t rows, int32_t precision, int32_t per_row_ybus, void *stream,
                     g2g_batch_t *out);
void g2g_batch_destroy(g2g_batch_t batch);

/* Start voltages: `rows` complex rows of n values, row stride ld (complex
 * numbers; ld = 0 broadcasts one row), written to rows [row0, row0 + rows).
 * on_device = 1 when V points to device memory. */
int g2g_batch_set_v(g2g_batch_t batch, const double *V, int64_t ld, int32_t rows, int32_t row0, int32_t on_device);
/* Per-unit injections Sbus, same conventions (the reference's per-row Sbus,
 * src/_cpp/injection_sweep_session.cu:235-247; ld = 0: shared, as a
 * contingency batch's Sbus with stride 0, acpf_nr_kernels.cuh:30-41). */
int g2g_batch_set_sbus(g2g_batch_t batch, const double *S, int64_t ld, int32_t rows, int32_t row0,
                       int32_t on_device);
/* Outage value edits: row r subtracts delta[q] (complex) from Ybus entry
 * entry[q] for q in [row_ptr[r], row_ptr[r+1]) (replaces the flat patches and
 * apply_contingencies_kernel, contingency_analysis_helper.cpp:163 and
 * acpf_nr_kernels.cu:328-348).  Needs per_row_ybus; resets every row to the
 * plan's base values first.  A row's repeated entries are summed into one
 * edit.  Host arrays. */
int g2g_batch_set_ybus_edits(g2g_batch_t batch, const int32_t *row_ptr, const int32_t *entry, const double *delta,
                             int32_t rows);
/* Islanded buses (handle_disconnected_grid; replaces build_mask_entries,
 * apply_bus_mask_kernel and mask_V_nan_kernel, contingency_analysis_helper.hpp:
 * 492-505 and acpf_nr_kernels.cu:219-307): row r masks buses[q] for q in
 * [row_ptr[r], row_ptr[r+1]).  A masked bus's P/Q equations become identity
 * rows of J with zero mismatch (its voltage stays where it started and it is
 * left out of the residual) and its result voltage is NaN.  The slack cannot
 * be masked.  Host arrays; row_ptr = NULL clears the mask.  Not with streamed
 * rows. */
int g2g_batch_set_bus_mask(g2g_batch_t batch, const int32_t *row_ptr, const int32_t *buses, int32_t rows);
/* Rows [0, active_rows) take part in the next solve; the others are left
 * untouched and reported NotSimulated. */
int g2g_batch_set_active_rows(g2g_batch_t batch, int32_t active_rows);

/* Newton-Raphson on every row (replaces run_nr_loop, src/_cpp/contingency/driver.cuh:136-259,
 * with the chunked status rule batch_pf_driver.cu:1361-1377 and the
 * continuous one :2895-2913).  A row converges when ||F||inf < tol (per
 * unit); its iteration count is the number of Newton updates applied. */
int g2g_batch_solve(g2g_batch_t batch, int32_t max_iter, double tol, int32_t scheduling, g2g_solve_stats *stats);

/* Step control of the next solves (AcPfNrSession's ir_n_steps and
 * scaling_max_voltage_change / max_dVa / max_dVm, python_bindings.cpp:1369-1560).
 * ir_steps: rounds of iterative refinement of every Newton step, in the
 * factorization's precision (BatchedLu::refine, own_lu/batched_lu.cu:391-411).
 * scaling: each row's step is multiplied by alpha = min(1, max_dva / max|dVa|,
 * max_dvm / max|dVm|) before the update (reduce_step_norms_kernel /
 * apply_step_scale_kernel, acpf_nr_kernels.cu:677-723; driver.cuh:225-235). */
int g2g_batch_set_step_control(g2g_batch_t batch, int32_t ir_steps, int32_t scaling, double max_dva, double max_dvm);

/* Results.  V: rows x n complex, row stride ld (complex numbers). */
/* Learned warm start with fallback (no counterpart in the reference, whose
 * rows start from the base case or a caller's V, python_bindings.cpp:2722-2736).
 * The MLP (g2g_mlp_*, below) maps each active row's Sbus (and its outage
 * mask: host, active_rows x n_mask bytes, 0/1; n_mask = 0 for none) to start
 * voltages; continuous Newton runs at most k_warm updates from there; a row
 * whose ||F||inf exceeds growth x its previous value (growth > 0) or that is
 * not converged after k_warm updates restarts from the batch's V as set by
 * g2g_batch_set_v, with max_iter updates.  Statuses follow the second pass;
 * iteration counts add both passes. */
int g2g_batch_solve_warm(g2g_batch_t batch, g2g_mlp_t mlp, const uint8_t *mask, int32_t n_mask, int32_t k_warm,
                         double growth, int32_t max_iter, double tol, g2g_solve_stats *stats);

/* The residual of the refinement rounds of Newton steps (ir_n_steps of
 * g2g_batch_set_step_control): 0, the default, in the solve's precision from
 * the A stream's J values (BatchedLu::refine, own_lu/batched_lu.cu:391-411);
 * 1, in FP64 from V: r = F(V) - J(V) dx with J dx formed matrix-free, the
 * correction solved on the FP32 factors in mixed precision.  Not available
 * with extension rows. */
int g2g_batch_set_refinement_residual(g2g_batch_t batch, int32_t fp64);

/* A resident outage mask for the batch's warm starts: rows x n_mask bytes
 * (0/1, one row per batch row), uploaded once; g2g_batch_solve_warm and
 * g2g_batch_warm_start called with no mask (NULL, 0) use it.  NULL / 0 drops
 * it.  No counterpart in the reference (its starts are not learned). */
int g2g_batch_set_warm_mask(g2g_batch_t batch, const uint8_t *mask, int32_t n_mask);

/* Warm solves' early rejection: a row on its learned start whose ||F||inf,
 * from its second update on, exceeds q x the previous one squared (it is
 * not converging quadratically) restarts at once, as a failed row does
 * (0, the default: off; applies when growth > 0). */
int g2g_batch_set_warm_contraction(g2g_batch_t batch, double q);

/* The MLP's start voltages for the active rows, written as the batch's V
 * (no counterpart in the reference; same mask convention as above). */
int g2g_batch_warm_start(g2g_batch_t batch, g2g_mlp_t mlp, const uint8_t *mask, int32_t n_mask);

/* Branch terminal currents of the active rows at the batch's V (replaces
 * compute_branch_flows_kernel, acpf_nr_kernels.cu:122-182): for branch l,
 * I_or = yff V_from + yft V_to, I_ex = ytf V_from + ytt V_to, and
 * amps = |I| * base (base_or / base_ex per branch, e.g. sn_mva 1e6 /
 * (sqrt(3) vn_kv 1e3)).  y4 holds 4 m complex values per branch in the order
 * yff, yft, ytf, ytt; a bus of -1 marks a missing terminal (0 A).  Row r's
 * tripped branches trip_branch[trip_ptr[r] .. trip_ptr[r+1]) report 0
 * (trip_ptr / trip_branch nullable).  Outputs rows x m, host arrays. */
int g2g_batch_branch_flows(g2g_batch_t batch, int32_t m, const int32_t *from, const int32_t *to, const double *y4,
                           const double *base_or, const double *base_ex, const int32_t *trip_ptr,
                           const int32_t *trip_branch, double *or_amps, double *ex_amps);

/* Limit checks of the active rows at the batch's V (compute_limit_violations;
 * replaces check_limit_violations_kernel, contingency/violation_kernels.cuh:
 * 80-160, with the record layout of ContingencyAnalysisSession's
 * get_violation_*, contingency_analysis/__init__.py:616-673).
 *
 * A row whose residual is NaN or above tol gets one record {GRID, -1, 0,
 * DIVERGENCE, residual, tol}.  Otherwise, for each branch l not tripped in the
 * row: side 1 is a CURRENT violation when |I_or| * base_or * 1e-3 (kA) >
 * limit1_ka * (1 + rel_tol), side 2 likewise with I_ex / base_ex / limit2_ka;
 * for each bus b with vm = |V_b| * vn_kv: LOW_VOLTAGE when vm < vmin_kv *
 * (1 - rel_tol), else HIGH_VOLTAGE when vm > vmax_kv * (1 + rel_tol).  NaN
 * limits are unset; NaN voltages (islanded buses) never violate.  Each type
 * keeps its `capacity` most severe records (severity |value / limit - 1|, ties
 * to the element met first: branch sides, then buses), most severe first, in
 * the order CURRENT, LOW_VOLTAGE, HIGH_VOLTAGE; element codes BUS 0, LINE 1,
 * TRAFO 2, GRID 3 (branches l >= n_lines are transformers, numbered
 * l - n_lines); type codes LOW_VOLTAGE 0, HIGH_VOLTAGE 1, CURRENT 2,
 * DIVERGENCE 4.  A row reported NotSimulated gets count -1 and type counts -1.
 * Outputs are host arrays: per row count, truncated (a type had more than
 * capacity), n_low / n_high / n_current (true counts); records [rows][3 capacity]. */
#define G2G_MAX_VIOLATION_CAPACITY 1024
typedef struct g2g_limit_table {
    int32_t m;               /* branches */
    const int32_t *from;     /* [m] bus, -1 = missing terminal */
    const int32_t *to;
    const double *y4;        /* [m][4] complex yff, yft, ytf, ytt */
    const double *base_or;   /* [m] A per unit current, origin side */
    const double *base_ex;   /* [m] extremity side */
    const double *limit1_ka; /* [m] NaN = none */
    const double *limit2_ka;
    int32_t n_lines;         /* branches [0, n_lines) are lines, the rest transformers */
    const double *vn_kv;     /* [n] */
    const double *vmin_kv;   /* [n] NaN = none */
    const double *vmax_kv;
} g2g_limit_table;
typedef struct g2g_violations {
    int32_t *count, *truncated, *n_low, *n_high, *n_current;       /* [rows] */
    int32_t *element_type, *element_id, *side, *type;             /* [rows][3 capacity] */
    double *value, *limit;
} g2g_violations;
int g2g_batch_limit_violations(g2g_batch_t batch, const g2g_limit_table *table, const int32_t *trip_ptr,
                               const int32_t *trip_branch, double tol, double rel_tol, int32_t capacity,
                               g2g_violations *out);

int g2g_batch_get_v(g2g_batch_t batch, double *V, int64_t ld, int32_t on_device);
/* Rows [row0, row0 + rows) of the batch's V into V (row stride ld complex;
 * on_device = 1: V is device memory, no host round trip). */
int g2g_batch_get_v_rows(g2g_batch_t batch, double *V, int64_t ld, int32_t rows, int32_t row0, int32_t on_device);
/* Extension state of every row (rows x n_xs host arrays: sa, then each
 * controller's Q).  A solve starts it afresh; the setter serves evaluation. */
int g2g_batch_get_ext_state(g2g_batch_t batch, double *out);
int g2g_batch_set_ext_state(g2g_batch_t batch, const double *in);
int g2g_batch_get_results(g2g_batch_t batch, int32_t *status, int32_t *iters, double *residual);

/* Evaluates, at the current V, F (reference row layout, rows x d) and, when
 * J != NULL, the Jacobian values in the reference CSR order (rows x nJ), on
 * the device kernels; residual (nullable) gets ||F||inf.  Host arrays. */
int g2g_batch_eval(g2g_batch_t batch, double *F, double *J, double *residual);

/* Batched adjoint: J^T x = rhs for every active row, J at the batch's V
 * (BatchAdjoint / solve_JT_batch, contingency/batch_pf_driver.cuh:177-217,807-830;
 * AcPfNrSession.solve_JT_dlpack, python_bindings.cpp:1525).  rhs and x are
 * host arrays rows x d in the reference's J row / column layout.  The J^T
 * values fill the same factorization programs (the pattern is structurally
 * symmetric); the solve's statuses are kept. */
int g2g_batch_solve_adjoint(g2g_batch_t batch, const double *rhs, double *x);

/* Raw batched LU solve of `rows` matrices on a g2g_lu_plan_create pattern:
 * vals rows x nnz (CSR order), rhs rows x d, x rows x d (host arrays);
 * pivots below pivot_eps * max|A_row| are clamped (own_lu/batched_lu.cu:45-53)
 * and counted per row in `clamped` (nullable).  ir_steps rounds of iterative
 * refinement x += (LU)^-1 (rhs - A x) follow, in the solve's precision
 * (BatchedLu::refine, own_lu/batched_lu.cu:391-411; solve_cudss_raw's
 * ir_n_steps, python_bindings.cpp:3827). */
int g2g_lu_solve_batch(g2g_plan_t plan, int32_t rows, int32_t precision, const double *vals, const double *rhs,
                       double *x, double pivot_eps, int32_t ir_steps, int32_t *clamped, void *stream);

/* g2g_lu_solve_batch on device arrays, with no allocation: no counterpart in
 * the reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * d_vals (rows x nnz, CSR order), d_rhs (rows x d), d_x (rows x d) and
 * d_clamped (rows, nullable) are device arrays; results match
 * g2g_lu_solve_batch's bit for bit.  Every intermediate lives in a file-scope
 * device array of 64 MiB per GPU (held until the process ends): a call whose
 * rows need more is rejected (G2G_ERR_INVALID, the message names the row
 * limit of this plan and precision).  The plan's device tables for this
 * precision must already be resident (one g2g_lu_solve_batch call with the
 * plan and precision makes them so).  Calls that overlap on one GPU share the
 * array and race; the call returns once its work on `stream` is done. */
int g2g_lu_solve_batch_noalloc(g2g_plan_t plan, int32_t rows, int32_t precision, const double *d_vals,
                               const double *d_rhs, double *d_x, double pivot_eps, int32_t ir_steps,
                               int32_t *d_clamped, void *stream);

/* ---- Streamed rows (continuous batching with slot refill; replaces the
 * reference's _solve_continuous refill, contingency/batch_pf_driver.cu:2720-3100,
 * and slot_schedule.cuh:82-156).  The batch's rows become slots: a caller
 * fills free slots with new rows (Sbus, start V and, for per-row Ybus
 * batches, the row's outage edits), runs rounds of k Newton iterations on
 * every active slot, reads the slots' statuses with g2g_batch_get_results
 * (G2G_NOTSIMULATED = free, -1 = still iterating) and takes finished slots'
 * V out, which frees them.  Each row counts its own updates against max_iter;
 * a row whose last update of a round converged it is reported by the next
 * round.  Rows are host or device arrays (on_device), row-major n complex
 * with row strides lds / ldv in complex numbers (0 broadcasts one row). */
int g2g_batch_stream_begin(g2g_batch_t batch);
/* Streamed rounds as HIP graphs (default on): a round's launches (k
 * iterations of mismatch, factorization, substitution and update, and the
 * iteration clock's step, which lives on the device) are captured on the
 * first round of a shape (k, max_iter, tol, refinement, step scaling) and
 * replayed after; 0 launches them one by one.  No counterpart in the
 * reference, whose continuous rounds launch kernels each time
 * (batch_pf_driver.cu:2720-3100).  The info call reports the setting and the
 * number of replays so far. */
int g2g_batch_set_stream_graph(g2g_batch_t batch, int32_t on);
int g2g_batch_stream_graph_info(g2g_batch_t batch, int32_t *on, int64_t *replays);

/* The learned start voltages of n freshly filled slots, predicted on the
 * device from the slots' Sbus (and a host 0/1 outage mask [n][n_mask], or
 * NULL) and written as their V; no counterpart in the reference. */
int g2g_batch_stream_warm(g2g_batch_t batch, g2g_mlp_t mlp, int32_t n, const int32_t *slots, const uint8_t *mask,
                          int32_t n_mask);
int g2g_batch_stream_fill(g2g_batch_t batch, int32_t n, const int32_t *slots, const double *S, int64_t lds,
                          const double *V0, int64_t ldv, int32_t on_device, const int32_t *edit_ptr,
                          const int32_t *edit_entry, const double *edit_delta);
int g2g_batch_stream_round(g2g_batch_t batch, int32_t k, int32_t max_iter, double tol);
int g2g_batch_stream_take(g2g_batch_t batch, int32_t n, const int32_t *slots, double *V, int64_t ldv,
                          int32_t on_device);

/* Device pointer of the batch's tiled Sbus array [ceil(rows/64)][n][64]
 * complex (bench and zero-copy callers fill it directly). */
int g2g_batch_device_arrays(g2g_batch_t batch, void **V_tiles, void **S_tiles);

/* ---- Learned warm start (no counterpart in the reference, whose ML models are
 * random-init timing stand-ins,
 * 20260831_neurips_paper/scripts/ml_injection.py).  A 3-layer GELU MLP,
 * in -> H -> H -> out, bf16 MFMA products with fp32 accumulation, fp32
 * master weights and Adam state on the device.  Matrices are row-major
 * float32 (inputs M x n_in, outputs M x n_out); on_device = 1 when they point
 * to device memory. */
int g2g_mlp_create(int32_t n_in, int32_t n_hidden, int32_t n_out, uint64_t seed, void *stream, g2g_mlp_t *out);
void g2g_mlp_destroy(g2g_mlp_t mlp);
int64_t g2g_mlp_num_params(g2g_mlp_t mlp);
/* offsets of W1 [n_in x H], b1, W2 [H x H], b2, W3 [H x n_out], b3 in the
 * flat parameter vector (6 values) */
int g2g_mlp_param_offsets(g2g_mlp_t mlp, int64_t *offsets);
int g2g_mlp_get_params(g2g_mlp_t mlp, float *out, int64_t capacity);
int g2g_mlp_set_params(g2g_mlp_t mlp, const float *in, int64_t n);
/* feature and output normalisation: x_feat = (x - mu) / sigma, y = mu + sigma * out */
int g2g_mlp_set_normalization(g2g_mlp_t mlp, const float *in_mu, const float *in_sigma, const float *out_mu,
                              const float *out_sigma);
/* bus types for warm starts: kind 0 PV (|V| = vfix, angle predicted), 1 PQ
 * (both predicted), 2 fixed (every bus outside pv and pq: keeps the row's
 * start V) */
int g2g_mlp_set_buses(g2g_mlp_t mlp, const int32_t *kind, const double *vfix, int32_t n);
int g2g_mlp_forward(g2g_mlp_t mlp, const float *X, int32_t M, float *Y, int32_t on_device);
/* mean-squared-error loss over all M x n_out outputs and its gradient (flat,
 * num_params floats; loss and grads nullable) */
int g2g_mlp_loss_grad(g2g_mlp_t mlp, const float *X, const float *T, int32_t M, float *loss, float *grads,
                      int32_t on_device);
/* one Adam step (beta 0.9 / 0.999, eps 1e-8) on a minibatch */
int g2g_mlp_train_step(g2g_mlp_t mlp, const float *X, const float *T, int32_t M, float lr, float *loss,
                       int32_t on_device);
/* Training steps scale the gradient by min(1, max_norm / ||g||2) before the
 * Adam update (0: off, the default). */
int g2g_mlp_set_grad_clip(g2g_mlp_t mlp, float max_norm);
/* C = op(A) op(B) with bf16 operands (raw bit patterns) and fp32 result, on the
 * MLP's MFMA kernels (test entry; host arrays).  op(A) = A, op(B) = B^T with
 * K, lda and ldb multiples of 8 runs the forward pass's K-contiguous kernel. */
int g2g_gemm_bf16(int32_t trans_a, int32_t trans_b, int32_t M, int32_t N, int32_t K, const uint16_t *A, int64_t lda,
                  const uint16_t *B, int64_t ldb, float *C, int64_t ldc);

#ifdef __cplusplus
}
#endif

#endif /* GPUSIM2GRID_AMD_H */
