// Kernel argument bundles and constants (device side).
#pragma once

#include <hip/hip_runtime.h>

#include "program_format.hpp"

namespace g2g {

// Each row's Ybus value edits (outage batches): row r's Ybus is Y minus its
// deltas, the tripped branches' terms.  Tiled like the rows: n_ed[row],
// ent / val [T][cap][64].  flag[tile][bus] is set when some row of the tile
// edits an entry of Ybus row `bus`: only those buses look the edits up.
struct RowEdits {
    const int* n_ed;
    const int* ent;
    const double2* val;
    const int* flag;
    int cap;
};

template <typename Real>
struct MismatchArgs {
    const int* k1_bus;      // kK1BusWords ints per bus (mismatch program)
    const int* k1_ent;      // 6 ints per Ybus entry
    int nbus;
    const double2* Y;       // shared Ybus values [nY]
    RowEdits ed;            // outage batches: each row's value edits on Y
    const double2* V;       // [T][n][64]
    const double2* S;       // [T][n][64]
    Real* A;                // A stream [T][a_stride][64]
    int n, nY, a_stride;
    double* res;            // [rows_pad] ||F||inf
    double* jmax;           // [rows_pad] max |J| (pivot clamp scale)
    int* status;            // [rows_pad]
    int* iters;             // [rows_pad]
    int* tile_active;       // [T]
    int* n_active;          // rows, then tiles, still iterating (accumulated)
    int rows;
    int iter, max_iter;
    double tol;
    int freeze;             // 1: a converged row stops (continuous); 0: fixed count (chunked)
    double growth;          // > 0: a row whose ||F||inf exceeds growth x its previous one is rejected
    double contraction;     // > 0: a row on its learned start is rejected from its second update on
                            // when ||F||inf exceeds contraction x its previous value squared
    const int* it0;         // nullable: per row, the iteration its solve started at (streamed rows)
    const unsigned char* mask;  // nullable: [T][n][64], 1 = bus islanded in this row (its equations
                                // become identity rows: F = 0, J row = unit diagonal)
    const int* budget;          // nullable: per row, its update budget (else max_iter)
    const unsigned char* warm;  // nullable: per row, 1 = still on its learned start (a failure
                                // marks it for a restart instead of ending it)
    int splits;                 // workgroups per tile (bus ranges); <= 1: one, which also
                                // applies the status rule
    double* part;               // nullable: per split, ||F||inf then max |J| partials
                                // [2][splits][part_stride]; set, the status rule is left
                                // to k_mis_status (several splits, or extension rows)
    int part_stride;            // the launch's tiles x 64
    const double2* sx;          // nullable: extension buses' residual terms [T][n_xb][64]
    int n_xb;
    int tiles;                  // tiles of the launch
    int xcd_group;              // 1: a tile's split workgroups grouped on one XCD
    const int* only;            // nullable: evaluate only the tiles with only[tile] > 0
    const int* clock;           // nullable: the iteration is *clock + iter (a graph-replayed round)
};

// Extension rows (distributed slack, HVDC droop, voltage control): one lane
// per row, one wavefront per tile; the tables are the plan's ExtTables.
template <typename Real>
struct ExtArgs {
    const double2* V;           // [T][n][64]
    const double2* S;           // [T][n][64]
    double* xs;                 // [T][n_xs][64] state: sa, then Q of each controller
    double2* sx;                // [T][n_xb][64] residual terms (dF_P, dF_Q) per extension bus
    Real* A;                    // A stream [T][a_stride][64]
    const Real* X;              // solved step [T][d][64]
    const int* tile_active;
    const int* status;
    int rows, n, d, a_stride, n_xs, n_xb, sa, q0;
    const int* inj_ptr;         // [n_xb + 1]
    const int* inj;             // {kind, index} per record
    const double* ds_w;
    int n_jrec;
    const int* jrec;            // {A position, kind, index, add}
    const double* jconst;
    int n_frec;
    const int* frec;            // {A position of F, kind, index}
    const int* xcol;            // [n_xs] J column of each state value
    const int* h_bus1;
    const int* h_bus2;
    const int* h_status;
    const double* h_par;        // per line: p0, k, lf1, lf2, r, pmax12, pmax21
    const int* vc_reg;
    const int* vc_ptr;
    const double* vc_vset;
    const double* vc_slope;
    const double* vc_w;
    const int* sh_c1;
    const int* sh_ck;
    double* part;               // mismatch partials (see MismatchArgs)
    int part_stride, splits;
};

// limit-violation check of solved rows (one wavefront per row)
struct ViolationArgs {
    const double2* V;           // [T][n][64]
    int n, rows;
    const int* status;          // [rows_pad]
    const double* res;          // [rows_pad]
    double tol;                 // DIVERGENCE above this residual
    double up, dn;              // 1 + rel_tol, 1 - rel_tol
    int m, n_lines;
    const int* from;            // [m] (-1: missing terminal)
    const int* to;
    const double2* y4;          // [m][4] yff, yft, ytf, ytt
    const double* base_or;      // [m] A per unit current
    const double* base_ex;
    const double* lim1;         // [m] kA, NaN = none
    const double* lim2;
    const double* vn_kv;        // [n]
    const double* vmin;         // [n] kV, NaN = none
    const double* vmax;
    const int* trip_ptr;        // nullable [rows + 1]
    const int* trip;            // tripped branches per row
    int K;                      // records kept per type
    int* count;                 // [rows] records kept (-1: not simulated)
    int* truncated;             // [rows]
    int* n_type;                // [rows][3] true counts: low, high, current
    int* etype;                 // [rows][3K]
    int* eid;
    int* side;
    int* vtype;
    double* value;
    double* limit;
};

template <typename Real>
struct FactorArgs {
    const Real* A;          // A stream [T][a_stride][64]
    Real* U;                // U stream [T][u_stride][64]
    const int* prog;        // padded to prog_stride words
    const int* tile_active;
    const double* jmax;     // nullable: scale 1
    int* clamped;           // nullable: clamped pivots per row
    int* error;             // set when a wavefront gave up waiting on its stream
    int d, a_stride, u_stride, prog_stride, slots, rows;
    int max_c;              // longest column (sizes the long-column scratch rows)
    double pivot_eps;
    const int* seg;         // per walker: prog_f off/len, prog_b off/len, A off/len, U off/len, pivots
    // split walk: the launch covers programs seg0 .. seg0 + gridDim.x / T - 1 (program-major,
    // T tiles each); pending update targets go through pend [T][pend_stride][64]
    int T, seg0;
    Real* pend;
    int pend_stride;
    Real* Lw;               // kFactorStoreL: the multipliers' stream (the A stream's layout)
};

template <typename Real>
struct BackwardArgs {
    const Real* U;
    const int* prog;        // padded to prog_stride words
    const int* tile_active;
    Real* X;                // solution [T][d][64] in input labels
    int* error;             // set when a wavefront gave up waiting on its stream
    int d, u_len, u_stride, prog_stride, rows;
    int max_c;              // longest column (a longer record header means a corrupt walk)
    const int* seg;         // per walker segments (as FactorArgs::seg)
    int n_top;              // pivots walker 0 walks before the others start
    int T, seg0;            // split walk (as FactorArgs)
};

template <typename Real>
struct UpdateArgs {
    const Real* X;          // solved step [T][d][64] (reference unknown labels)
    const int* upd;         // per bus {bus, theta unknown, |V| unknown or -1}
    int nbus;
    const int* tile_active;
    const int* status;
    double2* V;
    int d, n, rows;
    double max_dva, max_dvm;  // step scaling bounds; max_dva <= 0: off
    int splits;               // workgroups per tile (bus ranges); > 1 only with step scaling off
    const Real* Xadd;         // nullable: the step is X + Xadd (added in Real, the refinement's last correction)
};

// Refinement with an FP64 residual: r = F(V) - J(V) dx, both from V in FP64
// (J dx matrix-free: dS_i = dV_i conj(I_i) + V_i conj(sum_j Y_ij dV_j) with
// dV_j = V_j (j dtheta_j + dVm_j / |V_j|)), written over the A stream's rhs.
template <typename Real>
struct IrArgs {
    const int* k1_bus;          // the mismatch's bus and entry tables
    const int* k1_ent;
    int nbus;
    const double2* Y;           // shared Ybus values
    RowEdits ed;                // outage batches: each row's edits
    const double2* V;           // [T][n][64]
    const double2* S;           // [T][n][64]
    const unsigned char* mask;  // nullable: islanded buses [T][n][64] (identity rows)
    const int* colmap;          // [n][2]: each bus's theta and |V| unknown (-1: none)
    Real* A;                    // rhs positions overwritten with r
    const Real* X;              // the current step [T][d][64]
    Real* Xs;                   // its copy
    const int* tile_active;
    int n, nY, d, a_stride;
    int splits;                 // workgroups per tile (split-major), each over every splits-th bus group
};

template <typename Real>
struct RefineArgs {
    Real* A;                // A stream [T][a_stride][64]: J values read, rhs overwritten
    const Real* X;          // current solution [T][d][64] (unknown labels)
    Real* Bs;               // saved rhs [T][d][64]
    Real* Xs;               // copy of X [T][d][64]
    const int* Jp;          // CSR of the matrix in unknown labels
    const int* Ji;
    const int* apos;        // A-stream position of each CSR entry
    const int* fpos;        // A-stream position of each rhs entry
    const int* tile_active;
    int d, a_stride, first;
};

}  // namespace g2g
