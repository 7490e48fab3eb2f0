// Host-side symbolic layer of the batched Newton-Raphson power flow.
//
// Everything here runs once per grid pattern on the CPU and produces the
// shared (row-independent) tables the device kernels stream through:
//
//   * the reference's Jacobian layout (rows/cols [P of sorted pvpq | Q of pq],
//     gpusim2grid src/_cpp/acpf_nr.cu:145-236) and its scatter maps;
//   * a fill-reducing pivot order (bus-level minimum degree, each bus's
//     unknowns kept adjacent), relabelled by a postorder of the elimination
//     tree so that partial updates live only briefly;
//   * the factorization program: per pivot, where each value of its row and
//     column comes from (the value stream and/or an LDS update slot) and which
//     slots its rank-1 update feeds (replaces the reference's left-looking
//     items and contribution lists, own_lu/lu_plan.cpp:88-285);
//   * the backward-substitution program, including the fused polar V update;
//   * the mismatch/Jacobian program (buses in reverse Cuthill-McKee order).
//
// Stream layout ("tiled interleaved"): a per-row array of L values is stored
// as [row / 64][position][row % 64], so one wavefront (64 rows) reads or
// writes 64 consecutive elements per position.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "program_format.hpp"

namespace g2g {

struct SymbolicLU {
    int d = 0;
    int nA = 0;         // pattern entries of the (symmetrised) input present in A
    int nLU = 0;        // L + U + diagonal entries
    int n_updates = 0;  // sum over pivots of c_k^2 (rank-1 update terms)
    int n_steps = 0;    // walk steps of the factorization (pivots, 2-column supernodes)
    std::vector<int> step_w;  // per pivot: 1 single, 2 first of a 2-column supernode, 0 its second
    int max_c = 0;      // longest off-diagonal column
    int height = 0;     // elimination tree height
    std::vector<int> perm;     // pivot k -> input index
    std::vector<int> iperm;    // input index -> pivot k
    std::vector<int> colp, coli;  // struct of L(:,k) (pivot labels > k, sorted), CSR over k

    // factorization + forward substitution program (one record per pivot):
    //   c, bref(b(k) slot or none), dref, lref[c], uref[c], upd[c*c], bupd[c]
    std::vector<int> prog_f;
    std::vector<int> prog_f_off;  // [d+1]
    int slots_f = 0;
    // A stream: per pivot [F(k)] [diag] [L present...] [U present...]
    int a_len = 0;                 // A-stream length (d + nA)
    std::vector<int> a_pos_of_entry;  // input CSR position -> A-stream position
    std::vector<int> f_pos;           // input index (equation) -> A-stream position of its rhs
    // U stream: per pivot [y(k)] [U(k,k)] [U(k,j)...]
    int u_len = 0;
    std::vector<int> u_off;        // [d+1] start of pivot k's U record
    // walkers (program_format.hpp): per walker, its program, A-stream and
    // U-stream segments (offsets into prog_f / prog_b in words, into a tile's
    // streams in positions) and the pivots it walks; walker 0 walks the top
    // part's n_top pivots after its first phase (factorization) or before it
    // (backward substitution)
    struct Seg {
        int prog_f_off = 0, prog_f_len = 0, prog_b_off = 0, prog_b_len = 0;
        int a_off = 0, a_len = 0, u_off = 0, u_len = 0, n_piv = 0;
    };
    int walkers = 1;
    int n_top = 0;
    std::vector<Seg> seg;
    // the backward substitution's walkers (a multiple of the factorization's:
    // the forest is cut into bwalkers subtree sets, factorization walker w
    // walks sets w * bwalkers / walkers ... in turn): their program and
    // U-stream segments; walker 0's covers set 0 and then the top part
    int bwalkers = 1;
    std::vector<Seg> bseg;
    // backward substitution program (one record per pivot, pivots d-1 .. 0):
    //   c, xref[c] (slot of x(j) for j in struct_k), out (slot for x(k) or none),
    //   op, and for op != 0 two further ints (see below)
    std::vector<int> prog_b;
    int slots_b = 0;
    // Split walk (tail launches): every walker set cut into fine sets of
    // whole subtrees and a middle part above them; each fine set (level 0),
    // middle part (level 1) and the top part (level 2) is a program walked
    // by a workgroup of its own, a level per launch.  Factorization: a
    // program ends with dump records writing the update targets it leaves
    // pending to a per-tile buffer (pend_len positions); a higher program
    // gathers each pending value right before the step that consumes it.
    // Backward: the top part first, then the middles, then the fine sets;
    // a program starts with preload records reading the higher programs' x
    // it needs.  split_seg / split_bseg hold one segment per program (fine
    // sets, middles, top), in a stream layout of the split walk's own.
    int n_split = 0;
    int split_fine = 0, split_mid = 0;  // programs of levels 0 and 1 (the top is the last)
    std::vector<Seg> split_seg, split_bseg;
    std::vector<int> prog_fs, prog_bs;
    int slots_fs = 0, slots_bs = 0;
    int pend_len = 0;
    // the split walk's own stream layout (its programs one after the other)
    int split_a_len = 0, split_u_len = 0;
    std::vector<int> split_apos;  // input CSR position -> A-stream position
    std::vector<int> split_fpos;  // equation -> A-stream position of its right-hand side
};

// Builds the symbolic factorization of a square CSR pattern (structurally
// symmetrised).  `order` is the pivot order (a permutation of 0..d-1) before
// the postorder relabelling; empty = minimum degree on the pattern itself.
// `walkers` (1 or 2) splits each tile's walk over that many walker wavefronts.
// `bwalkers` (1, 2 or 4, a multiple of `walkers`; <= 0: as `walkers`) does the
// same for the backward substitution.
SymbolicLU build_symbolic(int d, const int* Ap, const int* Ai, std::vector<int> order = {}, int walkers = 1,
                          int bwalkers = 0);

// Minimum degree order of an undirected graph given by symmetric adjacency
// lists (CSR, no self loops) with vertex weights; returns the elimination order.
std::vector<int> min_degree_order(int n, const std::vector<int>& adjp, const std::vector<int>& adji,
                                  const std::vector<int>& weight);

// Reverse Cuthill-McKee order of an undirected graph (CSR adjacency).
std::vector<int> rcm_order(int n, const std::vector<int>& adjp, const std::vector<int>& adji);

// Backward-program records hold [c, xref[c], out, op, bus, partner slot]; a
// nonzero op (program_format.hpp) completes the bus: with the partner slot's
// x it applies V <- V ((|V| + dVm) / |V|) exp(j dVa).

// Extension rows of the Newton system (host copy of g2g_extensions; the
// reference's MultiSlack / HVDC droop / VoltageControl features,
// acpf_nr_kernels.cu:793-1068, ledger_data.hpp:61-155).  Empty vectors = off.
struct ExtSpec {
    std::vector<double> slack_w;  // [n] distributed-slack weights (empty: off)
    // HVDC lines in angle droop: p = p0 + k (theta1 - theta2), losses and a frozen regime
    std::vector<int> h_bus1, h_bus2, h_status;  // status 0 linear, +1 / -1 saturated, 2 tripped
    std::vector<double> h_p0, h_k, h_lf1, h_lf2, h_r, h_pmax12, h_pmax21;
    // voltage control: group g regulates |V(vc_reg[g])| at vc_vset[g] with
    // controllers [vc_ptr[g], vc_ptr[g+1]) of vc_bus (their reactive output is
    // an unknown), SVC slopes and sharing weights
    std::vector<int> vc_reg, vc_ptr, vc_bus;
    std::vector<double> vc_vset, vc_slope, vc_w;
    bool has_ds() const { return !slack_w.empty(); }
    int n_hvdc() const { return static_cast<int>(h_bus1.size()); }
    int n_grp() const { return static_cast<int>(vc_reg.size()); }
    int n_ctrl() const { return static_cast<int>(vc_bus.size()); }
    bool empty() const { return !has_ds() && n_hvdc() == 0 && n_grp() == 0; }
};

// Device-side description of the extension rows of a plan (built with the
// plan; see program_format.hpp for the record kinds).
struct ExtTables {
    int n_xs = 0;                 // per-row state values: [sa (distributed slack)] [Q_c per controller]
    int sa = -1;                  // state index of sa, -1: no distributed slack
    int q0 = 0;                   // state index of the first controller's Q
    int sa_col = -1;              // J column of sa (also the slack's P row)
    std::vector<int> xb_bus;      // buses whose mismatch takes an extension injection
    std::vector<int> inj_ptr;     // [n_xb + 1] records of ext bus xb
    std::vector<int> inj;         // records {kind, index} (kExtInj*)
    std::vector<double> ds_w;     // weight per distributed-slack record
    std::vector<int> jrec;        // J records {A position, kind, index, add}
    std::vector<double> jconst;   // values of the kExtJConst records
    std::vector<int> frec;        // custom-row records {A position of F, kind, group / share index}
    std::vector<int> xcol;        // per state value, its J column (the step's unknown)
    // HVDC lines [n_hvdc] and VC groups / shares (copied from the spec)
    std::vector<int> h_bus1, h_bus2, h_status;
    std::vector<double> h_par;    // per line: p0, k, lf1, lf2, r, pmax12, pmax21
    std::vector<int> vc_reg, vc_ptr;
    std::vector<double> vc_vset, vc_slope, vc_w;
    std::vector<int> sh_c1, sh_ck;  // per share row: its group's first controller and controller k
    bool on() const { return n_xs > 0 || !h_bus1.empty(); }
};

struct NrPlan {
    int n = 0, nY = 0, npv = 0, npq = 0, npvpq = 0, d = 0, nJ = 0, slack = -1;
    int n_ext = 0;                       // extension rows (= columns) after [P(pvpq) | Q(pq)]
    ExtSpec ext;
    ExtTables xt;
    std::vector<int> col_th;             // per bus: J column of its angle, -1 if none
    std::vector<int> Yp, Yi;
    std::vector<double> Yx;  // complex, interleaved
    std::vector<int> pvpq, pq;           // reference layout
    std::vector<int> row_p, row_q;       // per bus: reference J row of P / Q, -1 if none
    std::vector<int> Jp, Ji;             // reference J CSR
    std::vector<int> map[4];             // per Ybus entry: J CSR position (11,12,21,22) or -1
    SymbolicLU lu;
    // mismatch/Jacobian program: buses with a P equation in RCM order;
    // k1_bus[i] = {bus, entry begin, entry end, fpos_p, fpos_q, diag entry,
    // extension bus index or -1, first Ybus entry of the bus's row}
    std::vector<int> k1_bus;             // kK1BusWords ints per bus
    // per Ybus entry (in k1 order): {col j, ybus k, a11, a12, a21, a22}
    std::vector<int> k1_ent;             // 6 ints per entry
    int k1_nbus = 0;
    // the same tables in the split walk's stream layout (empty: no split walk)
    std::vector<int> k1_bus_s, k1_ent_s;
    bool split_ok() const { return !k1_bus_s.empty(); }
};

// walkers <= 0: default_walkers(d)
constexpr int kWalkersMinD = 2000;
// two walkers when d >= kWalkersMinD, else one; G2G_WALKERS=1|2 in the
// environment overrides
int default_walkers(int d);
// backward walkers: four when d >= kWalkersMinD, else one; G2G_BWALKERS=1|2|4 overrides
int default_bwalkers(int d);
NrPlan build_nr_plan(int n, const int* Yp, const int* Yi, const double* Yx, int npv, const int* pv,
                     int npq, const int* pq, int slack, int walkers = 0, int bwalkers = 0,
                     const ExtSpec* ext = nullptr);

}  // namespace g2g
