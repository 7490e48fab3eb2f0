// Device kernels of the batched Newton-Raphson power flow (gfx950, wave64).
//
// Every kernel maps one lane to one batch row: a wavefront owns a tile of 64
// rows and walks a row-independent program whose control flow and indices
// are wave-uniform (scalar loads), while the per-row values are stored
// "tiled interleaved" -- [tile][position][lane] -- so that each per-row
// access of the wave is one contiguous 64-element segment.
//
//   k_mismatch      F = -[Re/Im(V conj(Ybus V) - Sbus)], ||F||inf, per-row
//                   status, and (optionally) the Jacobian values dS/dVa,
//                   dS/dVm scattered straight into the factorization's input
//                   stream (gpusim2grid acpf_nr_kernels.cu:493-654 compute the
//                   same quantities in three kernels plus a cuSPARSE SpMV).
//   k_factor        static-pivot LU of every row's J, right-looking in the
//                   elimination-tree postorder, rank-1 updates accumulated in
//                   LDS slots, forward substitution fused (the reference's
//                   own_lu/batched_lu.cu:56-84 left-looking items and
//                   :86-98 forward solve).
//   k_backward      U x = y in reverse postorder with x(j) kept in LDS slots
//                   (own_lu/batched_lu.cu:100-109's backward solve).
//   k_update        polar V update of every bus (acpf_nr_kernels.cu:728-788).
//
// k_factor and k_backward read their program and value streams through LDS
// rings filled ahead of use (streams.hpp): a wavefront's walk is sequential,
// so its HBM traffic is a pure prefetched stream.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_common.hpp"
#include "streams.hpp"

namespace g2g {

// ---------------------------------------------------------------------------
// Layout helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ size_t tidx(size_t tile, size_t len, size_t pos, int lane)
{
    return (tile * len + pos) * kLanes + static_cast<size_t>(lane);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}

// a * conj(b)
__device__ __forceinline__ double2 cmulc(double2 a, double2 b)
{
    return make_double2(fma(a.x, b.x, a.y * b.y), fma(a.y, b.x, -a.x * b.y));
}

// a store of streamed output (the mismatch's A stream, the factorization's U
// stream) that no later access of the kernel reads: non-temporal, so it does
// not evict what the kernel reuses from the L2 (the neighbours' V, the
// program)
template <typename T>
__device__ __forceinline__ void stream_store(T* p, T v)
{
    __builtin_nontemporal_store(v, p);
}

__device__ __forceinline__ double nan_max(double m, double v)
{
    // NaN-sticky max of |v| (acpf_nr_kernels.cu:1101-1114)
    const double a = fabs(v);
    return (isnan(m) || isnan(a)) ? __builtin_nan("") : (a > m ? a : m);
}

// ---------------------------------------------------------------------------
// k_mismatch: one workgroup per tile, kMisWaves waves split the buses
// ---------------------------------------------------------------------------

constexpr int kMisBatch = 4;  // Ybus entries of a bus gathered together
constexpr int kMisBuses = 1;  // buses a wave works on at once

// Ybus entry k of this lane's row: the shared value minus the row's edits of
// entry k (an edited bus's entries only)
__device__ __forceinline__ double2 edited_y(const RowEdits& ed, size_t tile, int lane, int ned, int k, double2 y)
{
    for (int t = 0; t < ned; ++t)
        if (ed.ent[tidx(tile, ed.cap, t, lane)] == k) {
            const double2 d = ed.val[tidx(tile, ed.cap, t, lane)];
            y.x -= d.x;
            y.y -= d.y;
        }
    return y;
}

// The running sums of one bus of k_mismatch.
struct MisBus {
    int i, e0, e1, fp, fq, xb;
    // the lane row's edits of this bus's Ybus row: up to two kept in
    // registers (entry ke, delta kv; -1: none), or, with more, ned > 0 and
    // every entry scans the row's list
    int ke0, ke1, ned;
    double2 kv0, kv1;
    bool ok, masked;
    double2 vi, sb, I;
    double dr, di;
    int dent;
};

// Adds Ybus entries [eb, eb + kMisBatch) of bus b (y, vj already gathered;
// j < 0 marks a padding entry): I += Y_ij V_j, and the off-diagonal J values
// into the A stream.
template <typename Real, bool WithJ>
__device__ __forceinline__ void mis_entries(const MismatchArgs<Real>& a, size_t tile, int lane, MisBus& b, int eb,
                                            const int (&j)[kMisBatch], const double2 (&y)[kMisBatch],
                                            const double2 (&vj)[kMisBatch], double& jmax)
{
#pragma unroll
    for (int q = 0; q < kMisBatch; ++q) {
        if (j[q] < 0) continue;
        if (j[q] == b.i) {
            b.dr = y[q].x;
            b.di = y[q].y;
            b.dent = eb + q;
            b.I.x = fma(y[q].x, b.vi.x, fma(-y[q].y, b.vi.y, b.I.x));
            b.I.y = fma(y[q].x, b.vi.y, fma(y[q].y, b.vi.x, b.I.y));
            continue;
        }
        const double2 t = cmul(y[q], vj[q]);
        b.I.x += t.x;
        b.I.y += t.y;
        if (WithJ) {
            const int* en = a.k1_ent + 6 * (eb + q);
            const double2 w = cmulc(b.vi, t);  // V_i conj(Y_ij V_j)
            const double inv = 1.0 / sqrt(fma(vj[q].x, vj[q].x, vj[q].y * vj[q].y));
            const int a11 = en[2], a12 = en[3], a21 = en[4], a22 = en[5];
            double j11 = w.y, j21 = -w.x, j12 = w.x * inv, j22 = w.y * inv;
            if (b.masked) j11 = j21 = j12 = j22 = 0.0;
            if (a11 >= 0) stream_store(&a.A[tidx(tile, a.a_stride, a11, lane)], static_cast<Real>(j11));
            if (a12 >= 0) stream_store(&a.A[tidx(tile, a.a_stride, a12, lane)], static_cast<Real>(j12));
            if (a21 >= 0) stream_store(&a.A[tidx(tile, a.a_stride, a21, lane)], static_cast<Real>(j21));
            if (a22 >= 0) stream_store(&a.A[tidx(tile, a.a_stride, a22, lane)], static_cast<Real>(j22));
            jmax = fmax(jmax, fmax(fmax(fabs(j11), fabs(j12)), fmax(fabs(j21), fabs(j22))));
        }
    }
}

// Gathers entries [eb, eb + kMisBatch) of bus b (clamped to its last entry).
template <typename Real, bool PerRowY>
__device__ __forceinline__ void mis_gather(const MismatchArgs<Real>& a, size_t tile, int lane, const MisBus& b,
                                           int eb, int (&j)[kMisBatch], double2 (&y)[kMisBatch],
                                           double2 (&vj)[kMisBatch])
{
#pragma unroll
    for (int q = 0; q < kMisBatch; ++q) {
        const int e = eb + q < b.e1 ? eb + q : b.e1 - 1;
        const int* en = a.k1_ent + 6 * e;
        j[q] = eb + q < b.e1 ? en[0] : -1;
        const int yk = en[1];
        y[q] = a.Y[yk];
        if (PerRowY) {
            if (b.ned > 0) {
                y[q] = edited_y(a.ed, tile, lane, b.ned, yk, y[q]);
            } else {
                if (b.ke0 == yk) y[q] = make_double2(y[q].x - b.kv0.x, y[q].y - b.kv0.y);
                if (b.ke1 == yk) y[q] = make_double2(y[q].x - b.kv1.x, y[q].y - b.kv1.y);
            }
        }
        vj[q] = (j[q] >= 0 && j[q] != b.i) ? a.V[tidx(tile, a.n, j[q], lane)] : b.vi;
    }
}

// The status rule of one row from its ||F||inf (r) and max |J| (jm), run by
// one wavefront per tile; ends with the tile's count of rows still iterating.
template <typename Real, bool WithJ>
__device__ __forceinline__ void mis_status(const MismatchArgs<Real>& a, size_t tile, int lane, double r, double jm)
{
    const int row = static_cast<int>(tile) * kLanes + lane;
    int still = 0;
    if (a.status[row] == kActive && row < a.rows) {
        const double prev = a.res[row];
        a.res[row] = r;
        if (WithJ) a.jmax[row] = jm;
        const int iter = a.clock ? *a.clock + a.iter : a.iter;
        const int it = a.it0 ? iter - a.it0[row] : iter;  // updates this row has had
        const int budget = a.budget ? a.budget[row] : a.max_iter;
        // a row on its learned start that fails (non-finite, growing, or out
        // of its budget) is marked for a restart rather than ended
        const bool wrow = a.warm && a.warm[row];
        const bool growth_on = a.growth > 0.0 && (!a.warm || wrow);
        int st = kActive;
        if (!isfinite(r)) st = wrow ? kRejected : kDiverged;
        else if (a.freeze && r < a.tol) st = kConverged;
        else if (growth_on && it > 0 && r > a.growth * prev) st = kRejected;
        // not in Newton's quadratic regime by now: on its learned start, the
        // row would take too many updates
        else if (wrow && a.contraction > 0.0 && it >= 2 && r >= a.tol && r > a.contraction * prev * prev)
            st = kRejected;
        else if (it >= budget) st = r < a.tol ? kConverged : (wrow ? kRejected : kMaxIter);
        if (st != kActive) {
            a.status[row] = st;
            if (!(wrow && st == kRejected)) a.iters[row] = it;
        }
        still = st == kActive || (wrow && st == kRejected);
    }
    // wave-level count of rows still iterating
    const unsigned long long m = __ballot(still);
    if (lane == 0) {
        const int cnt = __popcll(m);
        a.tile_active[tile] = cnt;
        if (cnt) {
            atomicAdd(a.n_active, cnt);
            atomicAdd(a.n_active + 1, 1);
        }
    }
}

// NaN-sticky maximum of residual partials (a NaN anywhere makes the row's
// residual NaN, so the status rule sees the non-finite step)
__device__ __forceinline__ double res_max(double r, double v)
{
    return (isnan(r) || isnan(v)) ? __builtin_nan("") : fmax(r, v);
}

template <typename Real, bool PerRowY, bool WithJ>
__global__ __launch_bounds__(kLanes* kMismatchWaves) void k_mismatch(MismatchArgs<Real> a)
{
    const int lane = threadIdx.x & (kLanes - 1);
    // wave-uniform, so the bus and entry tables are read with scalar loads
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x / kLanes));
    // splits workgroups per tile; split s takes every splits-th group of
    // kMismatchWaves buses.  The launch's workgroups go to the XCDs round
    // robin (workgroup b to XCD b % 8); with xcd_group, the splits of tile t
    // are workgroups ((t / 8) splits + s) 8 + t % 8, so they run on one XCD
    // at about the same time and the neighbour V they gather is shared in its
    // L2.  Else split-major: split s of tile t is workgroup s T + t.
    const int splits = a.splits > 1 ? a.splits : 1;
    int split;
    size_t tile;
    if (a.xcd_group) {
        const int b = static_cast<int>(blockIdx.x);
        const int q = b / kXcds;
        split = q % splits;
        tile = static_cast<size_t>(q / splits) * kXcds + b % kXcds;
        if (tile >= static_cast<size_t>(a.tiles)) return;  // the grid is padded to whole XCD groups
    } else {
        const int T = static_cast<int>(gridDim.x) / splits;
        split = static_cast<int>(blockIdx.x) / T;
        tile = blockIdx.x - static_cast<size_t>(split) * T;
    }
    const int stride = kMismatchWaves * splits;
    __shared__ double s_res[kMismatchWaves][kLanes];
    __shared__ double s_jmax[kMismatchWaves][kLanes];

    // a pass over some tiles only (the rows restarted at this iteration):
    // the others keep what the iteration's evaluation wrote
    if (a.only && a.only[tile] <= 0) return;
    double rmax = 0.0, jmax = 0.0;
    if (a.tile_active[tile] > 0) {
        // The waves take the buses round-robin in the program's (RCM) order,
        // so they all work in one neighbourhood of the grid at a time and the
        // neighbours' V they gather stay in cache.  A wave works on
        // kMisBuses buses at once: every load of their first kMisBatch
        // entries is issued before any is used, which keeps that many
        // gathers in flight per wave instead of one bus's dependent chain.
        for (int bb0 = split * kMismatchWaves + wave; bb0 < a.nbus; bb0 += stride * kMisBuses) {
            MisBus b[kMisBuses];
#pragma unroll
            for (int g = 0; g < kMisBuses; ++g) {
                const int bb = bb0 + g * stride;
                b[g].ok = bb < a.nbus;
                const int* bus = a.k1_bus + kK1BusWords * (b[g].ok ? bb : a.nbus - 1);
                b[g].i = bus[0];
                b[g].e0 = bus[1];
                b[g].e1 = bus[2];
                b[g].fp = bus[3];
                b[g].fq = bus[4];
                b[g].xb = a.sx ? bus[6] : -1;
                b[g].vi = a.V[tidx(tile, a.n, b[g].i, lane)];
                b[g].sb = a.S[tidx(tile, a.n, b[g].i, lane)];
                // an islanded bus (apply_bus_mask_kernel, acpf_nr_kernels.cu:219-244):
                // identity rows in J, zero mismatch, out of the residual
                b[g].masked = a.mask && a.mask[tidx(tile, a.n, b[g].i, lane)];
                b[g].ke0 = b[g].ke1 = -1;
                b[g].ned = 0;
                if (PerRowY && a.ed.flag[tile * a.n + b[g].i]) {
                    // the lane row's edits within this bus's Ybus row, scanned once
                    const int yp0 = bus[7], yp1 = yp0 + (b[g].e1 - b[g].e0);
                    const int cnt = a.ed.n_ed[tile * kLanes + lane];
                    int nk = 0;
                    for (int t = 0; t < cnt; ++t) {
                        const int en = a.ed.ent[tidx(tile, a.ed.cap, t, lane)];
                        if (en < yp0 || en >= yp1) continue;
                        if (nk == 0) {
                            b[g].ke0 = en;
                            b[g].kv0 = a.ed.val[tidx(tile, a.ed.cap, t, lane)];
                        } else if (nk == 1) {
                            b[g].ke1 = en;
                            b[g].kv1 = a.ed.val[tidx(tile, a.ed.cap, t, lane)];
                        }
                        ++nk;
                    }
                    if (nk > 2) b[g].ned = cnt;
                }
                b[g].I = make_double2(0.0, 0.0);
                b[g].dr = b[g].di = 0.0;
                b[g].dent = -1;
            }
            int j[kMisBuses][kMisBatch];
            double2 y[kMisBuses][kMisBatch], vj[kMisBuses][kMisBatch];
#pragma unroll
            for (int g = 0; g < kMisBuses; ++g) mis_gather<Real, PerRowY>(a, tile, lane, b[g], b[g].e0, j[g], y[g], vj[g]);
#pragma unroll
            for (int g = 0; g < kMisBuses; ++g) {
                if (!b[g].ok) continue;
                mis_entries<Real, WithJ>(a, tile, lane, b[g], b[g].e0, j[g], y[g], vj[g], jmax);
                // the rest of a bus with more than kMisBatch entries
                for (int eb = b[g].e0 + kMisBatch; eb < b[g].e1; eb += kMisBatch) {
                    int jj[kMisBatch];
                    double2 yy[kMisBatch], vv[kMisBatch];
                    mis_gather<Real, PerRowY>(a, tile, lane, b[g], eb, jj, yy, vv);
                    mis_entries<Real, WithJ>(a, tile, lane, b[g], eb, jj, yy, vv, jmax);
                }
                const MisBus& c = b[g];
                const double2 sc = cmulc(c.vi, c.I);  // V conj(I)
                // an extension bus's residual terms (distributed slack share,
                // HVDC flows, controller reactive output)
                const double2 ex = c.xb >= 0 ? a.sx[tidx(tile, a.n_xb, c.xb, lane)] : make_double2(0.0, 0.0);
                const double fpv = c.masked ? 0.0 : -(sc.x - c.sb.x) + ex.x;
                rmax = nan_max(rmax, fpv);
                stream_store(&a.A[tidx(tile, a.a_stride, c.fp, lane)], static_cast<Real>(fpv));
                if (c.fq >= 0) {
                    const double fqv = c.masked ? 0.0 : -(sc.y - c.sb.y) + ex.y;
                    rmax = nan_max(rmax, fqv);
                    stream_store(&a.A[tidx(tile, a.a_stride, c.fq, lane)], static_cast<Real>(fqv));
                }
                if (WithJ) {
                    const int* en = a.k1_ent + 6 * c.dent;
                    const double2 t = make_double2(fma(c.dr, c.vi.x, -c.di * c.vi.y), fma(c.dr, c.vi.y, c.di * c.vi.x));
                    // dS/dVa = j V_i conj(I - t); dS/dVm = V_i conj(t + I) / |V_i|
                    const double2 z = cmulc(c.vi, make_double2(c.I.x - t.x, c.I.y - t.y));
                    const double2 u = cmulc(c.vi, make_double2(c.I.x + t.x, c.I.y + t.y));
                    const double inv = 1.0 / sqrt(fma(c.vi.x, c.vi.x, c.vi.y * c.vi.y));
                    double j11 = -z.y, j21 = z.x, j12 = u.x * inv, j22 = u.y * inv;
                    if (c.masked) {
                        j11 = j22 = 1.0;
                        j12 = j21 = 0.0;
                    }
                    const int a11 = en[2], a12 = en[3], a21 = en[4], a22 = en[5];
                    if (a11 >= 0) stream_store(&a.A[tidx(tile, a.a_stride, a11, lane)], static_cast<Real>(j11));
                    if (a12 >= 0) stream_store(&a.A[tidx(tile, a.a_stride, a12, lane)], static_cast<Real>(j12));
                    if (a21 >= 0) stream_store(&a.A[tidx(tile, a.a_stride, a21, lane)], static_cast<Real>(j21));
                    if (a22 >= 0) stream_store(&a.A[tidx(tile, a.a_stride, a22, lane)], static_cast<Real>(j22));
                    jmax = fmax(jmax, fmax(fmax(fabs(j11), fabs(j12)), fmax(fabs(j21), fabs(j22))));
                }
            }
        }
    }
    s_res[wave][lane] = rmax;
    s_jmax[wave][lane] = jmax;
    __syncthreads();
    if (wave != 0) return;
    if (a.tile_active[tile] <= 0) return;
    double r = 0.0, jm = 0.0;
    for (int w = 0; w < kMismatchWaves; ++w) {
        r = res_max(r, s_res[w][lane]);
        jm = fmax(jm, s_jmax[w][lane]);
    }
    if (a.part) {
        // k_mis_status combines the splits (and the extension rows) and
        // applies the status rule
        const size_t q = static_cast<size_t>(split) * a.part_stride + tile * kLanes + lane;
        a.part[q] = r;
        a.part[q + static_cast<size_t>(splits) * a.part_stride] = jm;
        return;
    }
    mis_status<Real, WithJ>(a, tile, lane, r, jm);
}

// k_mis_status: one wavefront per tile combines the splits' partials of a
// split k_mismatch and applies the status rule
template <typename Real, bool WithJ>
__global__ __launch_bounds__(kLanes) void k_mis_status(MismatchArgs<Real> a)
{
    const int lane = threadIdx.x;
    const size_t tile = blockIdx.x;
    if (a.tile_active[tile] <= 0 || (a.only && a.only[tile] <= 0)) return;
    const size_t row = tile * kLanes + lane;
    double r = 0.0, jm = 0.0;
    for (int s = 0; s < a.splits; ++s) {
        r = res_max(r, a.part[static_cast<size_t>(s) * a.part_stride + row]);
        jm = fmax(jm, a.part[static_cast<size_t>(a.splits + s) * a.part_stride + row]);
    }
    mis_status<Real, WithJ>(a, tile, lane, r, jm);
}

// ---------------------------------------------------------------------------
// Extension rows (the reference's MultiSlack, HVDC droop and VoltageControl
// features, acpf_nr_kernels.cu:793-1068, wired as nr_iter_step.cuh:302-413).
// One lane per row, one wavefront per tile.
//   k_ext_pre     before the mismatch: each extension bus's residual terms
//                 dF_P = -sa w - (HVDC flows leaving the bus), dF_Q = +Q(c)
//   k_ext_post    after it: the voltage-control rows' F, the extension J
//                 entries (weights, HVDC slopes, +-1, SVC slopes, sharing
//                 weights), folded into the row's residual and |J| partials
//   k_ext_update  after the V update: sa and Q(c) take their step
//   k_ext_init    at the start: sa = sum Re(Sbus), Q(c) = 0
// ---------------------------------------------------------------------------

// power the non-controlling end receives for |p| sent by the controlling end
__device__ __forceinline__ double hvdc_received(double p_abs, double lf_send, double lf_recv, double r)
{
    const double line = (1.0 - lf_send) * p_abs;
    return (1.0 - lf_recv) * (line - r * line * line);
}

__device__ __forceinline__ double bus_angle(double2 v) { return atan2(v.y, v.x); }

// active power leaving each end bus into line e (p1 at bus1, p2 at bus2)
template <typename Real>
__device__ __forceinline__ void hvdc_flows(const ExtArgs<Real>& a, size_t tile, int lane, int e, double& p1,
                                           double& p2)
{
    const double* q = a.h_par + 7 * e;  // p0, k, lf1, lf2, r, pmax12, pmax21
    const int st = a.h_status[e];
    if (st == kHvdcTripped) {
        p1 = p2 = 0.0;
    } else if (st > 0) {
        p1 = q[5];
        p2 = -hvdc_received(q[5], q[2], q[3], q[4]);
    } else if (st < 0) {
        p1 = -hvdc_received(q[6], q[3], q[2], q[4]);
        p2 = q[6];
    } else {
        const double t1 = bus_angle(a.V[tidx(tile, a.n, a.h_bus1[e], lane)]);
        const double t2 = bus_angle(a.V[tidx(tile, a.n, a.h_bus2[e], lane)]);
        const double raw = q[0] + q[1] * (t1 - t2);
        if (raw >= 0.0) {
            p1 = raw;
            p2 = -hvdc_received(raw, q[2], q[3], q[4]);
        } else {
            p1 = -hvdc_received(-raw, q[3], q[2], q[4]);
            p2 = -raw;
        }
    }
}

// dp1/dtheta1 and dp2/dtheta1 of a line in its linear regime (d/dtheta2 is the negative)
template <typename Real>
__device__ __forceinline__ void hvdc_slopes(const ExtArgs<Real>& a, size_t tile, int lane, int e, double& s1,
                                            double& s2)
{
    const double* q = a.h_par + 7 * e;
    const double t1 = bus_angle(a.V[tidx(tile, a.n, a.h_bus1[e], lane)]);
    const double t2 = bus_angle(a.V[tidx(tile, a.n, a.h_bus2[e], lane)]);
    const double raw = q[0] + q[1] * (t1 - t2);
    const bool from1 = raw >= 0.0;
    const double line = (1.0 - (from1 ? q[2] : q[3])) * fabs(raw);
    const double recv_slope = (1.0 - q[2]) * (1.0 - q[3]) * (1.0 - 2.0 * q[4] * line);
    s1 = from1 ? q[1] : q[1] * recv_slope;
    s2 = from1 ? -q[1] * recv_slope : -q[1];
}

template <typename Real>
__global__ __launch_bounds__(kLanes) void k_ext_pre(ExtArgs<Real> a)
{
    const int lane = threadIdx.x;
    const size_t tile = blockIdx.x;
    if (a.tile_active[tile] <= 0) return;
    const double sa = a.sa >= 0 ? a.xs[tidx(tile, a.n_xs, a.sa, lane)] : 0.0;
    for (int xb = 0; xb < a.n_xb; ++xb) {
        double2 acc = make_double2(0.0, 0.0);
        for (int r = a.inj_ptr[xb]; r < a.inj_ptr[xb + 1]; ++r) {
            const int kind = a.inj[2 * r], idx = a.inj[2 * r + 1];
            if (kind == kExtInjDs) {
                acc.x -= sa * a.ds_w[idx];
            } else if (kind == kExtInjVc) {
                acc.y += a.xs[tidx(tile, a.n_xs, a.q0 + idx, lane)];
            } else {
                double p1, p2;
                hvdc_flows(a, tile, lane, idx, p1, p2);
                acc.x -= kind == kExtInjHvdc1 ? p1 : p2;
            }
        }
        a.sx[tidx(tile, a.n_xb, xb, lane)] = acc;
    }
}

template <typename Real>
__global__ __launch_bounds__(kLanes) void k_ext_post(ExtArgs<Real> a)
{
    const int lane = threadIdx.x;
    const size_t tile = blockIdx.x;
    if (a.tile_active[tile] <= 0) return;
    auto Q = [&](int c) { return a.xs[tidx(tile, a.n_xs, a.q0 + c, lane)]; };
    double rmax = 0.0, jm = 0.0;
    for (int f = 0; f < a.n_frec; ++f) {
        const int* fr = a.frec + 3 * f;
        double F;
        if (fr[1] == kExtFVc) {
            const int g = fr[2];
            const double2 v = a.V[tidx(tile, a.n, a.vc_reg[g], lane)];
            double t = sqrt(fma(v.x, v.x, v.y * v.y)) - a.vc_vset[g];
            for (int c = a.vc_ptr[g]; c < a.vc_ptr[g + 1]; ++c) t = fma(a.vc_slope[c], Q(c), t);
            F = -t;
        } else {
            const int c1 = a.sh_c1[fr[2]], ck = a.sh_ck[fr[2]];
            F = -(a.vc_w[c1] * Q(ck) - a.vc_w[ck] * Q(c1));
        }
        rmax = nan_max(rmax, F);
        a.A[tidx(tile, a.a_stride, fr[0], lane)] = static_cast<Real>(F);
    }
    for (int j = 0; j < a.n_jrec; ++j) {
        const int* jr = a.jrec + 4 * j;
        double val;
        if (jr[1] == kExtJConst) {
            val = a.jconst[jr[2]];
        } else {
            double s1, s2;
            hvdc_slopes(a, tile, lane, jr[2] >> 2, s1, s2);
            const int q = jr[2] & 3;  // 0: d1/d1, 1: d1/d2, 2: d2/d1, 3: d2/d2
            val = q == 0 ? s1 : q == 1 ? -s1 : q == 2 ? s2 : -s2;
        }
        Real* p = &a.A[tidx(tile, a.a_stride, jr[0], lane)];
        const Real v = jr[3] ? *p + static_cast<Real>(val) : static_cast<Real>(val);
        *p = v;
        jm = fmax(jm, fabs(static_cast<double>(v)));
    }
    const size_t row = tile * kLanes + lane;
    const size_t jq = static_cast<size_t>(a.splits) * a.part_stride + row;
    a.part[row] = res_max(a.part[row], rmax);
    a.part[jq] = fmax(a.part[jq], jm);
}

template <typename Real>
__global__ __launch_bounds__(kLanes) void k_ext_update(ExtArgs<Real> a)
{
    const int lane = threadIdx.x;
    const size_t tile = blockIdx.x;
    if (a.tile_active[tile] <= 0) return;
    const int row = static_cast<int>(tile) * kLanes + lane;
    if (row >= a.rows || a.status[row] != kActive) return;
    for (int q = 0; q < a.n_xs; ++q)
        a.xs[tidx(tile, a.n_xs, q, lane)] += static_cast<double>(a.X[tidx(tile, a.d, a.xcol[q], lane)]);
}

template <typename Real>
__global__ __launch_bounds__(kLanes) void k_ext_init(ExtArgs<Real> a)
{
    const int lane = threadIdx.x;
    const size_t tile = blockIdx.x;
    if (a.sa >= 0) {
        double s = 0.0;
        for (int i = 0; i < a.n; ++i) s += a.S[tidx(tile, a.n, i, lane)].x;
        a.xs[tidx(tile, a.n_xs, a.sa, lane)] = s;
    }
    for (int q = a.q0; q < a.n_xs; ++q) a.xs[tidx(tile, a.n_xs, q, lane)] = 0.0;
}

// ---------------------------------------------------------------------------
// k_factor: right-looking static-pivot LU + forward substitution
//
// Workgroup of three wavefronts over one tile of 64 rows: wave 0 walks the
// program (LDS only, plus the U-stream stores), wave 1 streams the tile's A
// values into an LDS ring, wave 2 streams the program words (streams.hpp).
// ---------------------------------------------------------------------------

template <typename Real>
struct StreamShape;
template <>
struct StreamShape<double> {
    static constexpr int kCh = kChunkF64, kNch = kRingSlotsF64;
};
template <>
struct StreamShape<float> {
    static constexpr int kCh = kChunkF32, kNch = kRingSlotsF32;
};
using ProgRing = Ring<int, kProgChunk, kProgSlots, false, false, kProgMirror>;
using ProgReader = ProgRing::Reader;

// The first NU of a record's NW program words from position p (4-aligned)
// as broadcast reads: one base address and immediate offsets (a record
// running past the ring's end reads on into its mirror).  Only the used
// words are read: a padding word loaded into a register is dead at once, and
// the register allocator may give it to a later load or address, which then
// waits for this read and every LDS read issued before it.
template <int NW, int NU = NW>
__device__ __forceinline__ void read_words(const ProgReader& P, int p, v4u (&w)[NW / 4])
{
    static_assert(NW <= kProgMirror, "a record read at once fits the ring's mirror");
    static_assert(NU <= NW && NW - NU < 4, "the unused words are the record's padding");
    const v4u* base = reinterpret_cast<const v4u*>(P.ring) + ((p >> 2) & (ProgRing::kRing / 4 - 1));
    constexpr int NF = NU / 4, R = NU % 4;
#pragma unroll
    for (int i = 0; i < NF; ++i) w[i] = base[i];
    if constexpr (R > 0) {
        const unsigned* t = reinterpret_cast<const unsigned*>(base + NF);
        if constexpr (R >= 2) {
            const v2u t2 = *reinterpret_cast<const v2u*>(t);
            w[NF][0] = t2[0];
            w[NF][1] = t2[1];
        } else {
            w[NF][0] = t[0];
        }
        if constexpr (R == 3) w[NF][2] = t[2];
    }
}

__device__ __forceinline__ int prog_word(ProgReader& P, int p)
{
    P.window(p, p);
    return __builtin_amdgcn_readfirstlane(P.at(p));
}

// word q of a record section held in registers
template <int N>
__device__ __forceinline__ unsigned word(const v4u (&w)[N], int q)
{
    return w[q / 4][q % 4];
}

// this lane's element of the slot row at byte offset `ref` (in FP32 rows)
template <typename Real>
__device__ __forceinline__ Real* slot_at(unsigned char* lane_base, unsigned ref)
{
    return reinterpret_cast<Real*>(lane_base + ref * (sizeof(Real) / 4));
}

// N consecutive stream positions from p (already inside the reader's window):
// one base address and immediate offsets unless the run wraps the ring
template <int N, typename RD, typename T>
__device__ __forceinline__ void load_values(const RD& R, int p, T (&v)[N])
{
    const int b = p & (RD::kRingLen - 1);
    if (b + N > RD::kRingLen) {
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] = R.at(p + i);
    } else {
        const T* q = R.ring + static_cast<size_t>(b) * kLanes;
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] = q[i * kLanes];
    }
}

// 1 / pivot: the hardware reciprocal in FP32 (1 ulp; the FP32 factors feed
// a Newton step the FP64 mismatch corrects), the correctly rounded quotient
// in FP64
template <typename Real>
__device__ __forceinline__ Real pivot_inv(Real d)
{
    if constexpr (sizeof(Real) == 4) return __builtin_amdgcn_rcpf(d);
    else return Real(1) / d;
}

template <typename Real>
struct FactorState {
    unsigned char* lb;  // this lane's slot base
    Real* U;            // U stream of this lane
    Real* Aw;           // kFactorStoreL: this walker's segment of the multipliers' stream, this lane
    Real eps;
    int clamped;
    int p, ap, up;
};

// Walk modes of k_factor.  kFactorStoreL also writes each pivot's final L
// multipliers into a second stream of the A stream's layout (Lw), at the
// positions of the L values they came from.  kForwardOnly then walks the
// same program over that stream, with a right-hand side in its F positions,
// as a forward substitution alone: y(k) = F(k) + b(k), b(S) += L(S,k) y(k),
// y into the U stream.  The operations and their order are the
// factorization's own, so y comes out bitwise the same as from a second full
// factorization with that right-hand side; the Schur updates, the pivots and
// the U rows are skipped.  (kFactorFull: the factorization alone.)

// One pivot with C off-diagonal entries.  The whole record is read at once;
// then every slot the pivot reads -- its sources and, for short columns, the
// old values of all its update targets -- is read in one batch (a pivot's
// sources and targets never share a slot), so the pivot pays about two LDS
// round trips.  Sources are cleared after the read; targets accumulate.
// Longer columns read their update targets row by row, one row ahead of the
// writes.  The U record stores y(k), 1 / U(k,k) and the U row.
// Longest columns read in one batch (pivot, 2-column supernode), per value
// size: the FP32 walk has the registers for more.
template <typename Real>
constexpr int kAllReadC = sizeof(Real) == 4 ? 5 : 5;
template <typename Real>
constexpr int kAllRead2C = sizeof(Real) == 4 ? 4 : 3;

template <int C, int MODE, typename Real, typename AR>
__device__ __forceinline__ int factor_pivot(AR& A, ProgReader& P, FactorState<Real>& st)
{
    constexpr int NS = frec_src_words(C), NW = frec_words(C), NV = 2 + 2 * C;
    constexpr bool kAll = C <= kAllReadC<Real>;
    P.window(st.p, st.p + NW - 1);
    A.window(st.ap, st.ap + NV - 1);
    v4u w[NW / 4];
    read_words<NW, NS + (C + 1) * C>(P, st.p, w);
    Real av[NV];
    load_values<NV>(A, st.ap, av);
    auto at = [&](int q) -> Real* { return slot_at<Real>(st.lb, word(w, q)); };
    if constexpr (MODE == kForwardOnly) {
        // sources [diag, b(k), ...]; A stream [F, diag, -L / U(k,k) (stored), ...];
        // the refs it needs read one word each
        const unsigned* wr = reinterpret_cast<const unsigned*>(P.ring) + (st.p & (ProgRing::kRing - 1));
        unsigned rt[C > 0 ? C : 1];
#pragma unroll
        for (int a = 0; a < C; ++a) rt[a] = wr[NS + C * C + a];
        const unsigned rb = wr[2], head = wr[0];
        __builtin_amdgcn_sched_barrier(0);  // every ref's read issued before any slot read
        auto at = [&](unsigned ref) -> Real* { return slot_at<Real>(st.lb, ref); };
        Real tb[C > 0 ? C : 1];
#pragma unroll
        for (int a = 0; a < C; ++a) tb[a] = *at(rt[a]);
        const Real bk = *at(rb);
        *at(rb) = Real(0);
        const Real fk = av[0] + bk;
        stream_store(&st.U[static_cast<size_t>(st.up) * kLanes], fk);
#pragma unroll
        for (int a = 0; a < C; ++a) *at(rt[a]) = fma(av[2 + a], fk, tb[a]);
        st.up += 2 + C;
        st.ap += NV;
        st.p += NW;
        return static_cast<int>(static_cast<unsigned>(__builtin_amdgcn_readfirstlane(head)) >> 16);
    }
    const int ap0 = st.ap;
    Real sv[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) sv[q] = *at(1 + q);
    Real tv[kAll && C > 0 ? C * C : 1], tb[C > 0 ? C : 1];
    if constexpr (kAll) {
#pragma unroll
        for (int i = 0; i < C * C; ++i) tv[i] = *at(NS + i);
    }
#pragma unroll
    for (int a = 0; a < C; ++a) tb[a] = *at(NS + C * C + a);
#pragma unroll
    for (int q = 0; q < NV; ++q) *at(1 + q) = Real(0);
    Real dv = av[1] + sv[0];
    const Real fk = av[0] + sv[1];
    Real lv[C > 0 ? C : 1], uv[C > 0 ? C : 1];
#pragma unroll
    for (int q = 0; q < C; ++q) lv[q] = av[2 + q] + sv[2 + q];
#pragma unroll
    for (int q = 0; q < C; ++q) uv[q] = av[2 + C + q] + sv[2 + C + q];
    st.ap += NV;
    if (fabs(dv) < st.eps) {  // static-pivot clamp (own_lu/batched_lu.cu:45-53); NaN passes
        dv = dv < Real(0) ? -st.eps : st.eps;
        ++st.clamped;
    }
    const Real inv = pivot_inv(dv);
#pragma unroll
    for (int q = 0; q < C; ++q) lv[q] *= -inv;  // negated multipliers
    if constexpr (MODE == kFactorStoreL) {
#pragma unroll
        for (int q = 0; q < C; ++q) stream_store(&st.Aw[static_cast<size_t>(ap0 + 2 + q) * kLanes], lv[q]);
    }
    Real* u = st.U + static_cast<size_t>(st.up) * kLanes;
    stream_store(&u[0], fk);
    stream_store(&u[kLanes], inv);
#pragma unroll
    for (int q = 0; q < C; ++q) stream_store(&u[(2 + q) * kLanes], uv[q]);
    st.up += 2 + C;
    if constexpr (kAll) {
#pragma unroll
        for (int a = 0; a < C; ++a)
#pragma unroll
            for (int b = 0; b < C; ++b) *at(NS + a * C + b) = fma(lv[a], uv[b], tv[a * C + b]);
    } else {
        Real t0[C], t1[C];
#pragma unroll
        for (int q = 0; q < C; ++q) t0[q] = *at(NS + q);
#pragma unroll
        for (int a = 0; a < C; ++a) {
            Real(&cur)[C] = (a & 1) ? t1 : t0;
            Real(&nxt)[C] = (a & 1) ? t0 : t1;
            if (a + 1 < C) {
#pragma unroll
                for (int q = 0; q < C; ++q) nxt[q] = *at(NS + (a + 1) * C + q);
            }
#pragma unroll
            for (int b = 0; b < C; ++b) *at(NS + a * C + b) = fma(lv[a], uv[b], cur[b]);
        }
    }
#pragma unroll
    for (int a = 0; a < C; ++a) *at(NS + C * C + a) = fma(lv[a], fk, tb[a]);
    st.p += NW;
    return static_cast<int>(static_cast<unsigned>(__builtin_amdgcn_readfirstlane(word(w, 0))) >> 16);  // next record's c
}

// 2-column supernode (pivots k, k+1 with struct(k) = {k+1} + S, |S| = C): all
// sources of both pivots are read at once; pivot k is applied to pivot k+1's
// row, column and diagonal in registers; S x S and b(S) receive one rank-2
// update.  The U stream gets the two pivots' records, as separate pivots
// would write them.  The A stream holds pivot k's then pivot k+1's values:
//   [F(k), a(k,k), a(k+1,k), L(S,k)[C], a(k,k+1), U(k,S)[C],
//    F(k+1), a(k+1,k+1), L(S,k+1)[C], U(k+1,S)[C]]
template <int C, int MODE, typename Real, typename AR>
__device__ __forceinline__ int factor_pivot2(AR& A, ProgReader& P, FactorState<Real>& st)
{
    constexpr int NS = frec2_src_words(C), NW = frec2_words(C), NV = 6 + 4 * C;
    constexpr bool kAll = C <= kAllRead2C<Real>;
    constexpr int K1 = 4 + 2 * C;  // pivot k+1's first A-stream value
    P.window(st.p, st.p + NW - 1);
    A.window(st.ap, st.ap + NV - 1);
    v4u w[NW / 4];
    read_words<NW, NS + (C + 1) * C>(P, st.p, w);
    Real av[NV];
    load_values<NV>(A, st.ap, av);
    auto at = [&](int q) -> Real* { return slot_at<Real>(st.lb, word(w, q)); };
    if constexpr (MODE == kForwardOnly) {
        // stored: l at 2, L(S,k) / U(k,k) at 3 .., L(S,k+1) / U(k+1,k+1) at K1 + 2 ..;
        // the refs it needs read one word each (a wide read whose other words
        // are dead lets the compiler reuse their registers and wait in between)
        const unsigned* wr = reinterpret_cast<const unsigned*>(P.ring) + (st.p & (ProgRing::kRing - 1));
        unsigned rt[C > 0 ? C : 1];
#pragma unroll
        for (int a = 0; a < C; ++a) rt[a] = wr[NS + C * C + a];
        const unsigned r0 = wr[2], r1 = wr[2 + K1], head = wr[0];
        __builtin_amdgcn_sched_barrier(0);  // every ref's read issued before any slot read
        auto at = [&](unsigned ref) -> Real* { return slot_at<Real>(st.lb, ref); };
        Real tb[C > 0 ? C : 1];
#pragma unroll
        for (int a = 0; a < C; ++a) tb[a] = *at(rt[a]);
        const Real b0 = *at(r0), b1 = *at(r1);
        *at(r0) = Real(0);
        *at(r1) = Real(0);
        const Real fk = av[0] + b0;
        const Real fk1 = fma(-av[2], fk, av[K1] + b1);
        Real* u = st.U + static_cast<size_t>(st.up) * kLanes;
        stream_store(&u[0], fk);
        stream_store(&u[static_cast<size_t>(3 + C) * kLanes], fk1);
#pragma unroll
        for (int a = 0; a < C; ++a) *at(rt[a]) = fma(-av[K1 + 2 + a], fk1, fma(-av[3 + a], fk, tb[a]));
        st.up += 5 + 2 * C;
        st.ap += NV;
        st.p += NW;
        return static_cast<int>(static_cast<unsigned>(__builtin_amdgcn_readfirstlane(head)) >> 16);
    }
    const int ap0 = st.ap;
    Real sv[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) sv[q] = *at(1 + q);
    Real tv[kAll && C > 0 ? C * C : 1], tb[C > 0 ? C : 1];
    if constexpr (kAll) {
#pragma unroll
        for (int i = 0; i < C * C; ++i) tv[i] = *at(NS + i);
    }
#pragma unroll
    for (int a = 0; a < C; ++a) tb[a] = *at(NS + C * C + a);
#pragma unroll
    for (int q = 0; q < NV; ++q) *at(1 + q) = Real(0);
    // A-stream order per pivot is [F, diag, ...] and source order [diag, b, ...]
    Real x[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        const int sq = (q == 0 || q == K1) ? q + 1 : (q == 1 || q == K1 + 1) ? q - 1 : q;
        x[q] = av[q] + sv[sq];
    }
    // pivot k
    const Real fk = x[0];
    Real d1 = x[1];
    const Real e = x[2];
    const Real b = x[3 + C];
    if (fabs(d1) < st.eps) {
        d1 = d1 < Real(0) ? -st.eps : st.eps;
        ++st.clamped;
    }
    const Real inv1 = pivot_inv(d1);
    const Real l = e * inv1;
    Real mk[C > 0 ? C : 1], uk[C > 0 ? C : 1], mk1[C > 0 ? C : 1], uk1[C > 0 ? C : 1];
#pragma unroll
    for (int i = 0; i < C; ++i) mk[i] = x[3 + i] * inv1;
#pragma unroll
    for (int j = 0; j < C; ++j) uk[j] = x[4 + C + j];
    Real* u = st.U + static_cast<size_t>(st.up) * kLanes;
    stream_store(&u[0], fk);
    stream_store(&u[kLanes], inv1);
    stream_store(&u[2 * kLanes], b);
#pragma unroll
    for (int j = 0; j < C; ++j) stream_store(&u[(3 + j) * kLanes], uk[j]);
    // pivot k+1, updated by pivot k
    const Real fk1 = fma(-l, fk, x[K1]);
    Real d2 = fma(-l, b, x[K1 + 1]);
    if (fabs(d2) < st.eps) {
        d2 = d2 < Real(0) ? -st.eps : st.eps;
        ++st.clamped;
    }
    const Real inv2 = pivot_inv(d2);
#pragma unroll
    for (int i = 0; i < C; ++i) mk1[i] = fma(-mk[i], b, x[K1 + 2 + i]) * inv2;
#pragma unroll
    for (int j = 0; j < C; ++j) uk1[j] = fma(-l, uk[j], x[K1 + 2 + C + j]);
    if constexpr (MODE == kFactorStoreL) {
        Real* aw = st.Aw + static_cast<size_t>(ap0) * kLanes;
        stream_store(&aw[2 * kLanes], l);
#pragma unroll
        for (int i = 0; i < C; ++i) stream_store(&aw[static_cast<size_t>(3 + i) * kLanes], mk[i]);
#pragma unroll
        for (int i = 0; i < C; ++i) stream_store(&aw[static_cast<size_t>(K1 + 2 + i) * kLanes], mk1[i]);
    }
    Real* u1 = u + static_cast<size_t>(3 + C) * kLanes;
    stream_store(&u1[0], fk1);
    stream_store(&u1[kLanes], inv2);
#pragma unroll
    for (int j = 0; j < C; ++j) stream_store(&u1[(2 + j) * kLanes], uk1[j]);
    st.up += 5 + 2 * C;
    st.ap += NV;
    // rank-2 update of S x S and b(S)
    if constexpr (kAll) {
#pragma unroll
        for (int a = 0; a < C; ++a)
#pragma unroll
            for (int c2 = 0; c2 < C; ++c2)
                *at(NS + a * C + c2) = fma(-mk1[a], uk1[c2], fma(-mk[a], uk[c2], tv[a * C + c2]));
    } else {
        Real t0[C], t1[C];
#pragma unroll
        for (int q = 0; q < C; ++q) t0[q] = *at(NS + q);
#pragma unroll
        for (int a = 0; a < C; ++a) {
            Real(&cur)[C] = (a & 1) ? t1 : t0;
            Real(&nxt)[C] = (a & 1) ? t0 : t1;
            if (a + 1 < C) {
#pragma unroll
                for (int q = 0; q < C; ++q) nxt[q] = *at(NS + (a + 1) * C + q);
            }
#pragma unroll
            for (int c2 = 0; c2 < C; ++c2) *at(NS + a * C + c2) = fma(-mk1[a], uk1[c2], fma(-mk[a], uk[c2], cur[c2]));
        }
    }
#pragma unroll
    for (int a = 0; a < C; ++a) *at(NS + C * C + a) = fma(-mk1[a], fk1, fma(-mk[a], fk, tb[a]));
    st.p += NW;
    return static_cast<int>(static_cast<unsigned>(__builtin_amdgcn_readfirstlane(word(w, 0))) >> 16);
}

template <int C, int MODE, typename Real, typename AR>
__device__ __forceinline__ int factor_dispatch2(int c, AR& A, ProgReader& P, FactorState<Real>& st)
{
    if constexpr (C > kFastC2) {
        return -1;
    } else {
        if (c == C) return factor_pivot2<C, MODE>(A, P, st);
        return factor_dispatch2<C + 1, MODE>(c, A, P, st);
    }
}

// Any column length (up to kMaxC), one word at a time: the rare long columns.
template <int MODE, typename Real, typename AR>
__device__ __forceinline__ int factor_pivot_any(int c, AR& A, ProgReader& P, FactorState<Real>& st, Real* lsc, Real* usc)
{
    const int ns = frec_src_words(c), rw = frec_row_words(c), nw = frec_words(c);
    P.window(st.p, st.p + nw - 1);
    auto ref = [&](int q) -> unsigned { return static_cast<unsigned>(P.at(st.p + q)); };
    auto take = [&](int q) -> Real {
        Real* s = slot_at<Real>(st.lb, ref(1 + q));
        const Real v = *s;
        *s = Real(0);
        return v;
    };
    A.window(st.ap, st.ap + 1 + 2 * c);
    auto aval = [&](int i) { return A.at(st.ap + i); };
    if constexpr (MODE == kForwardOnly) {
        const Real fk = aval(0) + take(1);
        stream_store(&st.U[static_cast<size_t>(st.up) * kLanes], fk);
        for (int a = 0; a < c; ++a) {
            Real* t = slot_at<Real>(st.lb, ref(ns + c * rw + a));
            *t = fma(aval(2 + a), fk, *t);
        }
        const int c_next = __builtin_amdgcn_readfirstlane(static_cast<int>(ref(0) >> 16));
        st.up += 2 + c;
        st.ap += 2 + 2 * c;
        st.p += nw;
        return c_next;
    }
    const int ap0 = st.ap;
    Real dv = aval(1) + take(0);
    const Real fk = aval(0) + take(1);
    for (int q = 0; q < c; ++q) lsc[q * kLanes] = aval(2 + q) + take(2 + q);
    for (int q = 0; q < c; ++q) usc[q * kLanes] = aval(2 + c + q) + take(2 + c + q);
    st.ap += 2 + 2 * c;
    if (fabs(dv) < st.eps) {
        dv = dv < Real(0) ? -st.eps : st.eps;
        ++st.clamped;
    }
    const Real inv = pivot_inv(dv);
    for (int q = 0; q < c; ++q) lsc[q * kLanes] *= -inv;
    if constexpr (MODE == kFactorStoreL) {
        for (int q = 0; q < c; ++q) stream_store(&st.Aw[static_cast<size_t>(ap0 + 2 + q) * kLanes], lsc[q * kLanes]);
    }
    Real* u = st.U + static_cast<size_t>(st.up) * kLanes;
    stream_store(&u[0], fk);
    stream_store(&u[kLanes], inv);
    for (int q = 0; q < c; ++q) stream_store(&u[(2 + q) * kLanes], usc[q * kLanes]);
    st.up += 2 + c;
    for (int a = 0; a < c; ++a)
        for (int b = 0; b < c; ++b) {
            Real* t = slot_at<Real>(st.lb, ref(ns + a * rw + b));
            *t = fma(lsc[a * kLanes], usc[b * kLanes], *t);
        }
    for (int a = 0; a < c; ++a) {
        Real* t = slot_at<Real>(st.lb, ref(ns + c * rw + a));
        *t = fma(lsc[a * kLanes], fk, *t);
    }
    const int c_next = __builtin_amdgcn_readfirstlane(static_cast<int>(ref(0) >> 16));
    st.p += nw;
    return c_next;
}

// the pivot of column length c; returns the next record's c (-1: no fast path)
template <int C, int MODE, typename Real, typename AR>
__device__ __forceinline__ int factor_dispatch(int c, AR& A, ProgReader& P, FactorState<Real>& st)
{
    if constexpr (C > kFastC) {
        return -1;
    } else {
        if (c == C) return factor_pivot<C, MODE>(A, P, st);
        return factor_dispatch<C + 1, MODE>(c, A, P, st);
    }
}

// cross-walker flags (LDS): walker w > 0 has finished its first phase; the
// top part of the backward substitution is done; each walker's clamp counts
struct WalkFlags {
    int done[kMaxWalkers];
    int top_done;
    int pad;
    int clamped[kMaxWalkers][kLanes];
};
static_assert(sizeof(WalkFlags) <= kWalkFlagBytes, "walker flags");

// bounded wait for an LDS flag another walker sets (false: gave up)
__device__ __forceinline__ bool wait_flag(int* flag)
{
    int spins = 0;
    while (!lds_load_acquire(flag)) {
        if (++spins > kSpinLimit) return false;
        __builtin_amdgcn_s_sleep(2);
    }
    return true;
}

// A merge record [kMergeHead | next << 16, n, (dst, src) x n]: dst += src,
// src = 0 (walker 0, after the other walkers' first phase); returns the
// next record's head
template <typename Real>
__device__ __forceinline__ int factor_merge(ProgReader& P, FactorState<Real>& st)
{
    P.window(st.p, st.p + 1);
    const int n = __builtin_amdgcn_readfirstlane(P.at(st.p + 1));
    const int nw = mrec_words(n > kMergePairs ? kMergePairs : (n < 0 ? 0 : n));
    P.window(st.p, st.p + nw - 1);
    for (int i = 0; i < n && i < kMergePairs; ++i) {
        const unsigned dst = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(P.at(st.p + 2 + 2 * i)));
        const unsigned src = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(P.at(st.p + 3 + 2 * i)));
        Real* pd = slot_at<Real>(st.lb, dst);
        Real* ps = slot_at<Real>(st.lb, src);
        const Real v = *ps;
        *ps = Real(0);
        *pd += v;
    }
    const int next = static_cast<int>(static_cast<unsigned>(__builtin_amdgcn_readfirstlane(P.at(st.p))) >> 16);
    st.p += nw;
    return next;
}

// split walk: dump [kDumpHead | next << 16, n, first, ref x n]: pend[first + i] =
// slot ref_i (then cleared); gather [kGatherHead | next << 16, n, (dst ref,
// index) x n]: dst += pend[index].  Both return the next record's head.
template <typename Real>
__device__ __forceinline__ int factor_dump(ProgReader& P, FactorState<Real>& st, Real* pend_lane, int pend_stride)
{
    P.window(st.p, st.p + 2);
    const int n = __builtin_amdgcn_readfirstlane(P.at(st.p + 1));
    const int first = __builtin_amdgcn_readfirstlane(P.at(st.p + 2));
    const int m = n > kSplitRecItems ? kSplitRecItems : (n < 0 ? 0 : n);
    const int nw = dump_words(m);
    P.window(st.p, st.p + nw - 1);
    // every slot is read before any is cleared or stored, so the record pays
    // one LDS round trip instead of one per item
    // (items past m read the first item's slot again and are dropped)
    unsigned ref[kSplitRecItems];
    Real v[kSplitRecItems];
#pragma unroll
    for (int i = 0; i < kSplitRecItems; ++i) {
        ref[i] = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(P.at(st.p + 3 + (i < m ? i : 0))));
        v[i] = *slot_at<Real>(st.lb, ref[i]);
    }
#pragma unroll
    for (int i = 0; i < kSplitRecItems; ++i) {
        if (i < m) {
            const int q = first + i;
            if (static_cast<unsigned>(q) < static_cast<unsigned>(pend_stride))
                pend_lane[static_cast<size_t>(q) * kLanes] = v[i];
            *slot_at<Real>(st.lb, ref[i]) = Real(0);
        }
    }
    const int next = static_cast<int>(static_cast<unsigned>(__builtin_amdgcn_readfirstlane(P.at(st.p))) >> 16);
    st.p += nw;
    return next;
}

template <typename Real>
__device__ __forceinline__ int factor_gather(ProgReader& P, FactorState<Real>& st, const Real* pend_lane, int pend_stride)
{
    P.window(st.p, st.p + 1);
    const int n = __builtin_amdgcn_readfirstlane(P.at(st.p + 1));
    const int m = n > kSplitRecItems ? kSplitRecItems : (n < 0 ? 0 : n);
    const int nw = pair_rec_words(m);
    P.window(st.p, st.p + nw - 1);
    // Every pending value is loaded, and every target slot read, before any
    // is used: one global-memory and one LDS round trip per record instead of
    // one each per item.  Items of one target are consecutive (the planner
    // emits a target's contributions together) and distinct targets are
    // distinct slots, so each run of equal targets sums its values onto the
    // slot's old value in item order and writes once.
    // (items past m repeat the first item's reads and are dropped)
    unsigned dst[kSplitRecItems];
    Real v[kSplitRecItems], t[kSplitRecItems];
#pragma unroll
    for (int i = 0; i < kSplitRecItems; ++i) {
        const int w = st.p + 2 + 2 * (i < m ? i : 0);
        dst[i] = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(P.at(w)));
        const int q = __builtin_amdgcn_readfirstlane(P.at(w + 1));
        const bool ok = static_cast<unsigned>(q) < static_cast<unsigned>(pend_stride);
        v[i] = pend_lane[static_cast<size_t>(ok ? q : 0) * kLanes];
        if (!ok) v[i] = Real(0);
        t[i] = *slot_at<Real>(st.lb, dst[i]);
    }
    Real acc = Real(0);
#pragma unroll
    for (int i = 0; i < kSplitRecItems; ++i) {
        if (i < m) {
            if (i == 0 || dst[i] != dst[i - 1]) acc = t[i];
            acc += v[i];
            const bool last = i + 1 >= m || (i + 1 < kSplitRecItems && dst[i + 1] != dst[i]);
            if (last) *slot_at<Real>(st.lb, dst[i]) = acc;
        }
    }
    const int next = static_cast<int>(static_cast<unsigned>(__builtin_amdgcn_readfirstlane(P.at(st.p))) >> 16);
    st.p += nw;
    return next;
}

__device__ __forceinline__ bool head_ok(int h, int max_c)
{
    return h == kMergeHead || h == kDumpHead || h == kGatherHead || (h & 0x7FFF) <= max_c;
}

// Workgroup over one tile of 64 rows: W walker wavefronts walk their
// program segments (LDS only, plus the U-stream stores); factor_value_producers(W)
// waves stream the walkers' A segments into their LDS rings (split evenly
// between the walkers) and W waves stream the program segments.
// SPLIT: one walker per workgroup over one program of the split walk (the
// launch covers programs seg0 .., T tiles each, program-major), smaller rings,
// dump and gather records; clamp counts are added per row.
template <int W, bool SPLIT, int MODE>
__host__ __device__ constexpr int factor_waves()
{
    return SPLIT ? 2 + kSplitProducers : factor_walk_waves(W, MODE);
}

template <typename Real, int W, bool SPLIT, int RG, int MODE>
__global__ __launch_bounds__((kLanes * factor_waves<W, SPLIT, MODE>())) void k_factor(FactorArgs<Real> a)
{
    static_assert(!SPLIT || W == 1, "a split-walk workgroup has one walker");
    static_assert(!SPLIT || MODE == kFactorFull, "the split walk factorizes in full");
    constexpr int RW = SPLIT ? kSplitRingF : RG > 0 ? RG : W;  // ring geometry (RG: another walker count's)
    constexpr int kCh = value_chunk(sizeof(Real), RW), kNch = value_ring_slots(sizeof(Real), RW);
    static_assert((kNch - 2) * kCh >= 6 + 4 * kFastC2 - 1, "the ring window spans a two-column step's values");
    using ARing = Ring<Real, kCh, kNch, false, true>;
    constexpr int kNP = SPLIT ? kSplitProducers : factor_value_producers(W, MODE), kP = kNP / W;
    static_assert(kP * W == kNP, "producers split evenly");
    static_assert(sizeof(typename ARing::Ctl) + sizeof(ProgRing::Ctl) <= kWalkerCtlBytes, "control blocks");
    const int lane = threadIdx.x & (kLanes - 1);
    // wave-uniform (scalar) role, so the segment table is read with scalar loads
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x / kLanes));
    const size_t tile = SPLIT ? blockIdx.x % a.T : blockIdx.x;
    const int seg_base = SPLIT ? a.seg0 + static_cast<int>(blockIdx.x / a.T) : 0;  // this workgroup's program
    if (a.tile_active[tile] <= 0) return;
    extern __shared__ __align__(16) unsigned char g2g_smem[];
    auto ctlA = [&](int w) { return reinterpret_cast<typename ARing::Ctl*>(g2g_smem + w * kWalkerCtlBytes); };
    auto ctlP = [&](int w) {
        return reinterpret_cast<ProgRing::Ctl*>(g2g_smem + w * kWalkerCtlBytes + sizeof(typename ARing::Ctl));
    };
    auto* flags = reinterpret_cast<WalkFlags*>(g2g_smem + W * kWalkerCtlBytes);
    unsigned char* rings = g2g_smem + W * kWalkerCtlBytes + kWalkFlagBytes;
    auto aring = [&](int w) { return reinterpret_cast<Real*>(rings + w * ARing::kLdsBytes); };
    auto pring = [&](int w) {
        return reinterpret_cast<int*>(rings + W * ARing::kLdsBytes + w * ProgRing::kLdsBytes);
    };
    Real* sl = reinterpret_cast<Real*>(rings + W * (ARing::kLdsBytes + ProgRing::kLdsBytes));
    if (wave == 0) {
        for (int w = 0; w < W; ++w) {
            ARing::init_ctl(ctlA(w), lane);
            ProgRing::init_ctl(ctlP(w), lane);
        }
        if (lane < kMaxWalkers) flags->done[lane] = 0;
        if (lane == 0) flags->top_done = 0;
    }
    // every slot (and the zero and trash slots) starts at zero
    for (int i = threadIdx.x; i < (a.slots + 2) * kLanes; i += blockDim.x) sl[i] = Real(0);
    __syncthreads();
    const Real* Atile = a.A + tile * static_cast<size_t>(a.a_stride) * kLanes;
    if (wave >= W && wave < W + kNP) {
        const int pw = (wave - W) / kP;
        const int* sg = a.seg + 9 * (seg_base + pw);
        ARing::produce(Atile + static_cast<size_t>(sg[4]) * kLanes, (sg[5] + kCh - 1) / kCh, aring(pw), ctlA(pw),
                       lane, (wave - W) % kP, kP);
        return;
    }
    if (wave >= W + kNP) {
        const int pw = wave - W - kNP;
        const int* sg = a.seg + 9 * (seg_base + pw);
        ProgRing::produce(a.prog + sg[0], sg[1] / kProgChunk, pring(pw), ctlP(pw), lane, 0, 1);
        return;
    }
    // walker `wave`: its program segment
    const int wk = wave;
    const int* sg = a.seg + 9 * (seg_base + wk);
    const int n_piv = sg[8];
    const int row = static_cast<int>(tile) * kLanes + lane;
    // long-column scratch rows of this walker (long columns only)
    Real* lsc = sl + static_cast<size_t>(a.slots + 2 + 2 * wk * a.max_c) * kLanes;
    Real* usc = lsc + static_cast<size_t>(a.max_c) * kLanes;
    typename ARing::Reader A;
    A.init(aring(wk), ctlA(wk), (sg[5] + kCh - 1) / kCh, lane);
    ProgReader P;
    P.init(pring(wk), ctlP(wk), sg[1] / kProgChunk, lane);
    FactorState<Real> st;
    st.lb = reinterpret_cast<unsigned char*>(sl + lane);
    st.U = a.U + tile * static_cast<size_t>(a.u_stride) * kLanes + static_cast<size_t>(sg[6]) * kLanes + lane;
    st.Aw = MODE == kFactorStoreL
                ? a.Lw + (tile * static_cast<size_t>(a.a_stride) + static_cast<size_t>(sg[4])) * kLanes + lane
                : nullptr;
    st.eps = static_cast<Real>(a.pivot_eps) * static_cast<Real>(a.jmax ? a.jmax[row] : 1.0);
    st.clamped = 0;
    st.p = st.ap = st.up = 0;
    Real* pend_lane = SPLIT ? a.pend + tile * static_cast<size_t>(a.pend_stride) * kLanes + lane : nullptr;
    int h = (n_piv > 0 || SPLIT) && sg[1] > 0 ? prog_word(P, 0) & 0xFFFF : 0;  // an empty segment has no record
    bool bad = !head_ok(h, a.max_c);
    bool waited = W == 1 || wk != 0;
    for (int k = 0; k < n_piv && !bad;) {
        if (SPLIT && h == kGatherHead) {
            h = factor_gather(P, st, pend_lane, a.pend_stride) & 0xFFFF;
            if (P.failed || !head_ok(h, a.max_c)) bad = true;
            continue;
        }
        if (h == kMergeHead) {
            // the other walkers' first phase is over: their pending
            // contributions to the top entries are final
            if (!waited) {
                waited = true;
                for (int w = 1; w < W; ++w)
                    if (!wait_flag(&flags->done[w])) bad = true;
                if (bad) break;
            }
            h = factor_merge(P, st) & 0xFFFF;
            if (P.failed || !head_ok(h, a.max_c)) bad = true;
            continue;
        }
        // each record's header announces the next record's column length and
        // width; a walk whose stream wait gave up stops before it trusts
        // another record
        const int c = h & 0x7FFF;
        int h_next;
        if (h >> 15) {
            h_next = factor_dispatch2<0, MODE>(c, A, P, st);
            if (h_next < 0) {  // never emitted: the host pairs columns up to kFastC2 only
                bad = true;
                break;
            }
            k += 2;
        } else {
            h_next = factor_dispatch<0, MODE>(c, A, P, st);
            if (h_next < 0) h_next = factor_pivot_any<MODE>(c, A, P, st, lsc + lane, usc + lane);
            k += 1;
        }
        h = h_next & 0xFFFF;
        if (A.failed || P.failed || (!head_ok(h, a.max_c) && k < n_piv)) {
            bad = true;
            break;
        }
    }
    if (SPLIT) {  // the pending update targets left by this program
        while (!bad && h == kDumpHead) {
            h = factor_dump(P, st, pend_lane, a.pend_stride) & 0xFFFF;
            if (P.failed) bad = true;
        }
    }
    if (bad) {  // release the producers at once
        if (lane == 0) atomicOr(a.error, 1);
        lds_store_release(&ctlA(wk)->abort, 1);
        lds_store_release(&ctlP(wk)->abort, 1);
    }
    if (W > 1) {
        if (wk != 0) {
            flags->clamped[wk][lane] = st.clamped;
            lds_store_release(&flags->done[wk], 1);  // also after a failure: walker 0 must not wait
            return;
        }
        if (!waited)
            for (int w = 1; w < W; ++w)
                if (!wait_flag(&flags->done[w]) && lane == 0) atomicOr(a.error, 1);
        // per row: each lane adds the other walkers' counts of its own row
        for (int w = 1; w < W; ++w) st.clamped += lds_load_lane(&flags->clamped[w][lane]);
    }
    if (a.clamped && row < a.rows) {
        if (SPLIT) {  // every program of the walk adds its count
            if (st.clamped) atomicAdd(&a.clamped[row], st.clamped);
        } else {
            a.clamped[row] = st.clamped;
        }
    }
}

// ---------------------------------------------------------------------------
// k_backward: U x = y in reverse postorder; x written in input labels
// (same three-role layout: walker, U-stream producers, program producer)
// ---------------------------------------------------------------------------

template <typename Real>
struct BackState {
    unsigned char* lb;
    Real* X;
    int p, up, d;
};

template <int C, typename Real, typename UR>
__device__ __forceinline__ int backward_pivot(UR& U, ProgReader& P, BackState<Real>& st)
{
    constexpr int NW = brec_words(C);
    P.window(st.p, st.p + NW - 1);
    v4u w[NW / 4];
    read_words<NW, 3 + C>(P, st.p, w);
    st.up -= C + 2;
    U.window(st.up + C + 1, st.up);  // the record is consumed from its top
    Real uv[C + 2];
    load_values<C + 2>(U, st.up, uv);
    Real xj[C > 0 ? C : 1];
#pragma unroll
    for (int q = 0; q < C; ++q) xj[q] = *slot_at<Real>(st.lb, word(w, 2 + q));
    Real s = uv[0];
#pragma unroll
    for (int q = 0; q < C; ++q) s = fma(-uv[2 + q], xj[q], s);
    const Real x = s * uv[1];  // U(k,k) is stored as its reciprocal
    const int perm_k = __builtin_amdgcn_readfirstlane(static_cast<int>(word(w, 1)));
    if (static_cast<unsigned>(perm_k) < static_cast<unsigned>(st.d)) st.X[static_cast<size_t>(perm_k) * kLanes] = x;
    *slot_at<Real>(st.lb, word(w, 2 + C)) = x;
    st.p += NW;
    return static_cast<int>(static_cast<unsigned>(__builtin_amdgcn_readfirstlane(word(w, 0))) >> 16);
}

// Backward 2-column supernode (pivots k, k+1; |S| = C): x(k+1) from its U row,
// then x(k) with x(k+1) in registers; the U records of the pair are adjacent:
//   [y(k), 1/U(k,k), U(k,k+1), U(k,S)[C], y(k+1), 1/U(k+1,k+1), U(k+1,S)[C]]
template <int C, typename Real, typename UR>
__device__ __forceinline__ int backward_pivot2(UR& U, ProgReader& P, BackState<Real>& st)
{
    constexpr int NW = brec2_words(C), NU = 5 + 2 * C;
    P.window(st.p, st.p + NW - 1);
    v4u w[NW / 4];
    read_words<NW, 5 + C>(P, st.p, w);
    st.up -= NU;
    U.window(st.up + NU - 1, st.up);
    Real uv[NU];
    load_values<NU>(U, st.up, uv);
    Real xj[C > 0 ? C : 1];
#pragma unroll
    for (int q = 0; q < C; ++q) xj[q] = *slot_at<Real>(st.lb, word(w, 3 + q));
    Real s2 = uv[3 + C];
#pragma unroll
    for (int q = 0; q < C; ++q) s2 = fma(-uv[5 + C + q], xj[q], s2);
    const Real x2 = s2 * uv[4 + C];
    Real s1 = fma(-uv[2], x2, uv[0]);
#pragma unroll
    for (int q = 0; q < C; ++q) s1 = fma(-uv[3 + q], xj[q], s1);
    const Real x1 = s1 * uv[1];
    const int perm_k1 = __builtin_amdgcn_readfirstlane(static_cast<int>(word(w, 1)));
    const int perm_k = __builtin_amdgcn_readfirstlane(static_cast<int>(word(w, 2)));
    if (static_cast<unsigned>(perm_k1) < static_cast<unsigned>(st.d)) st.X[static_cast<size_t>(perm_k1) * kLanes] = x2;
    if (static_cast<unsigned>(perm_k) < static_cast<unsigned>(st.d)) st.X[static_cast<size_t>(perm_k) * kLanes] = x1;
    *slot_at<Real>(st.lb, word(w, 3 + C)) = x2;
    *slot_at<Real>(st.lb, word(w, 4 + C)) = x1;
    st.p += NW;
    return static_cast<int>(static_cast<unsigned>(__builtin_amdgcn_readfirstlane(word(w, 0))) >> 16);
}

template <int C, typename Real, typename UR>
__device__ __forceinline__ int backward_dispatch2(int c, UR& U, ProgReader& P, BackState<Real>& st)
{
    if constexpr (C > kFastC2) {
        return -1;
    } else {
        if (c == C) return backward_pivot2<C>(U, P, st);
        return backward_dispatch2<C + 1>(c, U, P, st);
    }
}

template <typename Real, typename UR>
__device__ __forceinline__ int backward_pivot_any(int c, UR& U, ProgReader& P, BackState<Real>& st)
{
    const int nw = brec_words(c);
    P.window(st.p, st.p + nw - 1);
    auto ref = [&](int q) -> unsigned { return static_cast<unsigned>(P.at(st.p + q)); };
    st.up -= c + 2;
    U.window(st.up + c + 1, st.up);
    auto uval = [&](int i) { return U.at(st.up + i); };
    Real s = uval(0);
    for (int q = 0; q < c; ++q) s = fma(-uval(2 + q), *slot_at<Real>(st.lb, ref(2 + q)), s);
    const Real x = s * uval(1);
    const int perm_k = P.at(st.p + 1);
    if (static_cast<unsigned>(perm_k) < static_cast<unsigned>(st.d)) st.X[static_cast<size_t>(perm_k) * kLanes] = x;
    *slot_at<Real>(st.lb, ref(2 + c)) = x;
    const int c_next = __builtin_amdgcn_readfirstlane(static_cast<int>(ref(0) >> 16));
    st.p += nw;
    return c_next;
}

template <int C, typename Real, typename UR>
__device__ __forceinline__ int backward_dispatch(int c, UR& U, ProgReader& P, BackState<Real>& st)
{
    if constexpr (C > kFastC) {
        return -1;
    } else {
        if (c == C) return backward_pivot<C>(U, P, st);
        return backward_dispatch<C + 1>(c, U, P, st);
    }
}

// split walk: preload [kPreloadHead | next << 16, n, (dst ref, x index) x n]:
// dst = x[index], a higher program's solution; returns the next head
template <typename Real>
__device__ __forceinline__ int backward_preload(ProgReader& P, BackState<Real>& st)
{
    P.window(st.p, st.p + 1);
    const int n = __builtin_amdgcn_readfirstlane(P.at(st.p + 1));
    const int m = n > kSplitRecItems ? kSplitRecItems : (n < 0 ? 0 : n);
    const int nw = pair_rec_words(m);
    P.window(st.p, st.p + nw - 1);
    // every x is loaded before any is stored: one global-memory round trip
    // per record instead of one per item
    // (items past m repeat the first item's reads and are dropped)
    unsigned dst[kSplitRecItems];
    Real v[kSplitRecItems];
    bool ok[kSplitRecItems];
#pragma unroll
    for (int i = 0; i < kSplitRecItems; ++i) {
        const int w = st.p + 2 + 2 * (i < m ? i : 0);
        dst[i] = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(P.at(w)));
        const int q = __builtin_amdgcn_readfirstlane(P.at(w + 1));
        ok[i] = static_cast<unsigned>(q) < static_cast<unsigned>(st.d);
        v[i] = st.X[static_cast<size_t>(ok[i] ? q : 0) * kLanes];
    }
#pragma unroll
    for (int i = 0; i < kSplitRecItems; ++i)
        if (i < m && ok[i]) *slot_at<Real>(st.lb, dst[i]) = v[i];
    const int next = static_cast<int>(static_cast<unsigned>(__builtin_amdgcn_readfirstlane(P.at(st.p))) >> 16);
    st.p += nw;
    return next;
}

// Same roles as k_factor.  Walker 0 walks the top part first and then
// signals the others, which walk their subtrees once the top's x are in
// their slots.  SPLIT: one walker per workgroup over one program of the
// split walk, whose preload records bring the x of higher programs.
template <int W, bool SPLIT>
__host__ __device__ constexpr int backward_waves()
{
    return SPLIT ? 2 + kSplitProducers : walk_waves(W);
}

template <typename Real, int W, bool SPLIT>
__global__ __launch_bounds__((kLanes * backward_waves<W, SPLIT>())) void k_backward(BackwardArgs<Real> a)
{
    static_assert(!SPLIT || W == 1, "a split-walk workgroup has one walker");
    constexpr int RW = SPLIT ? kSplitRingB : W;
    constexpr int kCh = value_chunk(sizeof(Real), RW), kNch = value_ring_slots(sizeof(Real), RW);
    static_assert((kNch - 2) * kCh >= 5 + 2 * kFastC2 - 1, "the ring window spans a two-column record");
    using URing = Ring<Real, kCh, kNch, true, true>;
    constexpr int kNP = SPLIT ? kSplitProducers : value_producers(W), kP = kNP / W;
    static_assert(kP * W == kNP, "producers split evenly");
    static_assert(sizeof(typename URing::Ctl) + sizeof(ProgRing::Ctl) <= kWalkerCtlBytes, "control blocks");
    const int lane = threadIdx.x & (kLanes - 1);
    // wave-uniform (scalar) role, so the segment table is read with scalar loads
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x / kLanes));
    const size_t tile = SPLIT ? blockIdx.x % a.T : blockIdx.x;
    const int seg_base = SPLIT ? a.seg0 + static_cast<int>(blockIdx.x / a.T) : 0;
    if (a.tile_active[tile] <= 0) return;
    extern __shared__ __align__(16) unsigned char g2g_smem[];
    auto ctlU = [&](int w) { return reinterpret_cast<typename URing::Ctl*>(g2g_smem + w * kWalkerCtlBytes); };
    auto ctlP = [&](int w) {
        return reinterpret_cast<ProgRing::Ctl*>(g2g_smem + w * kWalkerCtlBytes + sizeof(typename URing::Ctl));
    };
    auto* flags = reinterpret_cast<WalkFlags*>(g2g_smem + W * kWalkerCtlBytes);
    unsigned char* rings = g2g_smem + W * kWalkerCtlBytes + kWalkFlagBytes;
    auto uring = [&](int w) { return reinterpret_cast<Real*>(rings + w * URing::kLdsBytes); };
    auto pring = [&](int w) {
        return reinterpret_cast<int*>(rings + W * URing::kLdsBytes + w * ProgRing::kLdsBytes);
    };
    Real* sl = reinterpret_cast<Real*>(rings + W * (URing::kLdsBytes + ProgRing::kLdsBytes));
    if (wave == 0) {
        for (int w = 0; w < W; ++w) {
            URing::init_ctl(ctlU(w), lane);
            ProgRing::init_ctl(ctlP(w), lane);
        }
        if (lane == 0) flags->top_done = 0;
    }
    __syncthreads();
    const Real* Utile = a.U + tile * static_cast<size_t>(a.u_stride) * kLanes;
    if (wave >= W && wave < W + kNP) {
        const int pw = (wave - W) / kP;
        const int* sg = a.seg + 9 * (seg_base + pw);
        URing::produce(Utile + static_cast<size_t>(sg[6]) * kLanes, (sg[7] + kCh - 1) / kCh, uring(pw), ctlU(pw),
                       lane, (wave - W) % kP, kP);
        return;
    }
    if (wave >= W + kNP) {
        const int pw = wave - W - kNP;
        const int* sg = a.seg + 9 * (seg_base + pw);
        ProgRing::produce(a.prog + sg[2], sg[3] / kProgChunk, pring(pw), ctlP(pw), lane, 0, 1);
        return;
    }
    const int wk = wave;
    const int* sg = a.seg + 9 * (seg_base + wk);
    const int n_piv = sg[8];
    typename URing::Reader U;
    U.init(uring(wk), ctlU(wk), (sg[7] + kCh - 1) / kCh, lane);
    ProgReader P;
    P.init(pring(wk), ctlP(wk), sg[3] / kProgChunk, lane);
    BackState<Real> st;
    st.lb = reinterpret_cast<unsigned char*>(sl + lane);
    st.X = a.X + tile * static_cast<size_t>(a.d) * kLanes + lane;
    st.p = 0;
    st.up = sg[7];
    st.d = a.d;
    bool bad = false;
    bool signaled = W == 1 || wk != 0;
    if (W > 1 && wk != 0 && !wait_flag(&flags->top_done)) bad = true;
    int h = n_piv > 0 && !bad ? prog_word(P, 0) & 0xFFFF : 0;  // an empty segment has no record
    if ((h & 0x7FFF) > a.max_c && !(SPLIT && h == kPreloadHead)) bad = true;
    for (int q = 0; q < n_piv && !bad;) {
        if (SPLIT && h == kPreloadHead) {
            h = backward_preload(P, st) & 0xFFFF;
            if (P.failed || ((h & 0x7FFF) > a.max_c && h != kPreloadHead)) bad = true;
            continue;
        }
        if (!signaled && q >= a.n_top) {
            signaled = true;
            lds_store_release(&flags->top_done, 1);
        }
        const int c = h & 0x7FFF;
        int h_next;
        if (h >> 15) {
            h_next = backward_dispatch2<0>(c, U, P, st);
            if (h_next < 0) {
                bad = true;
                break;
            }
            q += 2;
        } else {
            h_next = backward_dispatch<0>(c, U, P, st);
            if (h_next < 0) h_next = backward_pivot_any(c, U, P, st);
            q += 1;
        }
        h = h_next & 0xFFFF;
        if (U.failed || P.failed || ((h & 0x7FFF) > a.max_c && q < n_piv)) {
            bad = true;
            break;
        }
    }
    if (!signaled) lds_store_release(&flags->top_done, 1);  // also after a failure: the others must not wait
    if (bad) {
        if (lane == 0) atomicOr(a.error, 1);
        lds_store_release(&ctlU(wk)->abort, 1);
        lds_store_release(&ctlP(wk)->abort, 1);
    }
}

// ---------------------------------------------------------------------------
// k_update: polar V update of every bus from the solved step (one lane per row)
//   V <- V ((|V| + dVm) / |V|) exp(j dVa)   (acpf_nr_kernels.cu:728-788)
// ---------------------------------------------------------------------------

template <typename Real>
__global__ __launch_bounds__(kLanes* kMisWaves) void k_update(UpdateArgs<Real> a)
{
    const int lane = threadIdx.x & (kLanes - 1);
    // wave-uniform, so the bus tables are read with scalar loads
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x / kLanes));
    // splits workgroups per tile, split-major, each a contiguous bus range
    // (step scaling needs the whole row and runs unsplit)
    const int splits = a.splits > 1 ? a.splits : 1;
    const int T = static_cast<int>(gridDim.x) / splits;
    const int split = static_cast<int>(blockIdx.x) / T;
    const size_t tile = blockIdx.x - static_cast<size_t>(split) * T;
    if (a.tile_active[tile] <= 0) return;
    const int row = static_cast<int>(tile) * kLanes + lane;
    const bool live = row < a.rows && a.status[row] == kActive;
    const int part = split * kMisWaves + wave, parts = splits * kMisWaves;
    const int b0 = static_cast<int>((static_cast<long long>(a.nbus) * part) / parts);
    const int b1 = static_cast<int>((static_cast<long long>(a.nbus) * (part + 1)) / parts);
    // the step of unknown k (with the refinement's last correction added in Real)
    auto step = [&](int k) -> Real {
        const size_t q = tidx(tile, a.d, k, lane);
        return a.Xadd ? a.X[q] + a.Xadd[q] : a.X[q];
    };
    // Step scaling (acpf_nr_kernels.cu:677-723): alpha = min(1, max_dva /
    // max|dVa|, max_dvm / max|dVm|) per row, the maxima NaN-sticky so that a
    // non-finite step is left unscaled for the status rule to catch.
    double alpha = 1.0;
    if (a.max_dva > 0.0 && splits == 1) {
        __shared__ double s_mt[kMisWaves][kLanes];
        __shared__ double s_mv[kMisWaves][kLanes];
        double mt = 0.0, mv = 0.0;
        for (int bb = b0; bb < b1; ++bb) {
            const int kt = a.upd[3 * bb + 1], kv = a.upd[3 * bb + 2];
            const double at = fabs(static_cast<double>(step(kt)));
            if (!(mt >= at)) mt = at;
            if (kv >= 0) {
                const double av = fabs(static_cast<double>(step(kv)));
                if (!(mv >= av)) mv = av;
            }
        }
        s_mt[wave][lane] = mt;
        s_mv[wave][lane] = mv;
        __syncthreads();
        mt = 0.0;
        mv = 0.0;
        for (int w = 0; w < kMisWaves; ++w) {
            if (!(mt >= s_mt[w][lane])) mt = s_mt[w][lane];
            if (!(mv >= s_mv[w][lane])) mv = s_mv[w][lane];
        }
        if (mt > a.max_dva) alpha = fmin(alpha, a.max_dva / mt);
        if (mv > a.max_dvm) alpha = fmin(alpha, a.max_dvm / mv);
    }
    // buses in groups of kUpdBatch, every group's loads issued before its stores
    constexpr int kUpdBatch = 4;
    for (int bq = b0; bq < b1; bq += kUpdBatch) {
        int bus[kUpdBatch];
        double dva[kUpdBatch], dvm[kUpdBatch];
        double2 v[kUpdBatch];
#pragma unroll
        for (int q = 0; q < kUpdBatch; ++q) {
            const int bb = bq + q < b1 ? bq + q : b1 - 1;
            const int kt = a.upd[3 * bb + 1], kv = a.upd[3 * bb + 2];
            bus[q] = bq + q < b1 ? a.upd[3 * bb] : -1;
            dva[q] = static_cast<double>(step(kt));
            dvm[q] = kv >= 0 ? static_cast<double>(step(kv)) : 0.0;
            v[q] = a.V[tidx(tile, a.n, a.upd[3 * bb], lane)];
        }
#pragma unroll
        for (int q = 0; q < kUpdBatch; ++q) {
            if (bus[q] < 0) continue;
            if (alpha != 1.0) {
                dva[q] *= alpha;
                dvm[q] *= alpha;
            }
            const double m = sqrt(fma(v[q].x, v[q].x, v[q].y * v[q].y));
            const double sc = (m + dvm[q]) / m;
            double sn, cs;
            sincos(dva[q], &sn, &cs);
            if (live)
                a.V[tidx(tile, a.n, bus[q], lane)] =
                    make_double2(sc * fma(v[q].x, cs, -v[q].y * sn), sc * fma(v[q].x, sn, v[q].y * cs));
        }
    }
}

// ---------------------------------------------------------------------------
// Iterative refinement of the Newton step (own_lu/batched_lu.cu:239-260,391-411):
//   r = b - J x  written over the A stream's right-hand side, so that the next
//   k_factor / k_backward pass (the same factors, recomputed) solves J d = r;
//   then x += d.  J is read from the A stream through the CSR's positions; x is
//   the step in unknown labels.  On the first step b is saved from the A
//   stream, x is copied to Xs on every step.
// ---------------------------------------------------------------------------

// k_ir_residual: splits workgroups per tile (split-major, split s taking
// every splits-th group of kMismatchWaves buses, its waves one bus each), as
// a tile's buses are a latency-bound chain per wave; also keeps a copy of
// the step (the refinement adds its correction to it).
template <typename Real, bool PerRowY>
__global__ __launch_bounds__(kLanes* kMismatchWaves) void k_ir_residual(IrArgs<Real> a)
{
    const int lane = threadIdx.x & (kLanes - 1);
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x / kLanes));
    const int splits = a.splits > 1 ? a.splits : 1;
    const int T = static_cast<int>(gridDim.x) / splits;
    const int split = static_cast<int>(blockIdx.x) / T;
    const size_t tile = blockIdx.x - static_cast<size_t>(split) * T;
    if (a.tile_active[tile] <= 0) return;
    const size_t m = static_cast<size_t>(a.d) * kLanes;
    for (size_t k = static_cast<size_t>(split) * blockDim.x + threadIdx.x; k < m;
         k += static_cast<size_t>(blockDim.x) * splits)
        a.Xs[tile * m + k] = a.X[tile * m + k];
    // dV of bus b at V = v: V (j dtheta + dVm / |V|)
    auto dv = [&](int bus, double2 v) {
        const int kt = a.colmap[2 * bus], kv = a.colmap[2 * bus + 1];
        const double th = kt >= 0 ? static_cast<double>(a.X[tidx(tile, a.d, kt, lane)]) : 0.0;
        const double vm = kv >= 0 ? static_cast<double>(a.X[tidx(tile, a.d, kv, lane)]) : 0.0;
        const double s = vm / sqrt(fma(v.x, v.x, v.y * v.y));
        return make_double2(fma(v.x, s, -v.y * th), fma(v.y, s, v.x * th));
    };
    for (int bb = split * kMismatchWaves + wave; bb < a.nbus; bb += kMismatchWaves * splits) {
        const int* bus = a.k1_bus + kK1BusWords * bb;
        const int i = bus[0], e0 = bus[1], e1 = bus[2], fp = bus[3], fq = bus[4];
        const double2 vi = a.V[tidx(tile, a.n, i, lane)];
        const double2 sb = a.S[tidx(tile, a.n, i, lane)];
        const int ned = PerRowY && a.ed.flag[tile * a.n + i] ? a.ed.n_ed[tile * kLanes + lane] : 0;
        double2 I = make_double2(0.0, 0.0), dI = make_double2(0.0, 0.0);
        for (int e = e0; e < e1; ++e) {
            const int* en = a.k1_ent + 6 * e;
            const int j = en[0];
            double2 y = a.Y[en[1]];
            if (PerRowY && ned > 0) y = edited_y(a.ed, tile, lane, ned, en[1], y);
            const double2 vj = j == i ? vi : a.V[tidx(tile, a.n, j, lane)];
            const double2 t = cmul(y, vj), u = cmul(y, dv(j, vj));
            I.x += t.x;
            I.y += t.y;
            dI.x += u.x;
            dI.y += u.y;
        }
        double rp, rq;
        if (a.mask && a.mask[tidx(tile, a.n, i, lane)]) {
            // an islanded bus: identity rows, zero mismatch
            const int kt = a.colmap[2 * i], kv = a.colmap[2 * i + 1];
            rp = kt >= 0 ? -static_cast<double>(a.X[tidx(tile, a.d, kt, lane)]) : 0.0;
            rq = kv >= 0 ? -static_cast<double>(a.X[tidx(tile, a.d, kv, lane)]) : 0.0;
        } else {
            const double2 sc = cmulc(vi, I);  // V conj(I)
            const double2 d1 = cmulc(dv(i, vi), I), d2 = cmulc(vi, dI);
            rp = -(sc.x - sb.x) - (d1.x + d2.x);
            rq = -(sc.y - sb.y) - (d1.y + d2.y);
        }
        a.A[tidx(tile, a.a_stride, fp, lane)] = static_cast<Real>(rp);
        if (fq >= 0) a.A[tidx(tile, a.a_stride, fq, lane)] = static_cast<Real>(rq);
    }
}

template <typename Real>
__global__ __launch_bounds__(kLanes* kMisWaves) void k_lin_residual(RefineArgs<Real> a)
{
    const int lane = threadIdx.x & (kLanes - 1);
    // wave-uniform, so the CSR is read with scalar loads
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x / kLanes));
    const size_t tile = blockIdx.x;
    if (a.tile_active[tile] <= 0) return;
    const int i0 = static_cast<int>((static_cast<long long>(a.d) * wave) / kMisWaves);
    const int i1 = static_cast<int>((static_cast<long long>(a.d) * (wave + 1)) / kMisWaves);
    Real* A = a.A + tile * static_cast<size_t>(a.a_stride) * kLanes + lane;
    const Real* X = a.X + tile * static_cast<size_t>(a.d) * kLanes + lane;
    Real* Bs = a.Bs + tile * static_cast<size_t>(a.d) * kLanes + lane;
    Real* Xs = a.Xs + tile * static_cast<size_t>(a.d) * kLanes + lane;
    for (int i = i0; i < i1; ++i) {
        const size_t fp = static_cast<size_t>(a.fpos[i]) * kLanes;
        Real s;
        if (a.first) {
            s = A[fp];
            Bs[static_cast<size_t>(i) * kLanes] = s;
        } else {
            s = Bs[static_cast<size_t>(i) * kLanes];
        }
        for (int e = a.Jp[i]; e < a.Jp[i + 1]; ++e) {
            const int ap = a.apos[e];
            if (ap >= 0) s = fma(-A[static_cast<size_t>(ap) * kLanes], X[static_cast<size_t>(a.Ji[e]) * kLanes], s);
        }
        A[fp] = s;
        Xs[static_cast<size_t>(i) * kLanes] = X[static_cast<size_t>(i) * kLanes];
    }
}

template <typename Real>
__global__ void k_axpy_tiles(const Real* __restrict__ Xs, Real* __restrict__ X, const int* __restrict__ tile_active,
                             int d)
{
    const size_t tile = blockIdx.y;
    if (tile_active[tile] <= 0) return;
    const size_t m = static_cast<size_t>(d) * kLanes;
    const size_t base = tile * m;
    for (size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < m;
         k += static_cast<size_t>(gridDim.x) * blockDim.x)
        X[base + k] += Xs[base + k];
}

template <typename T>
__global__ void k_rows_to_tiles(const T* __restrict__ in, long long ld, int rows, int m, int row_offset,
                                T* __restrict__ out, int broadcast)
{
    // rows fastest: consecutive threads fill consecutive lanes of the tiled output
    const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long total = static_cast<long long>(rows) * m;
    if (gid >= total) return;
    const int k = static_cast<int>(gid / rows);
    const int r = static_cast<int>(gid - static_cast<long long>(k) * rows);
    const int dst = r + row_offset;
    const T v = in[(broadcast ? 0 : static_cast<long long>(r) * ld) + k];
    out[tidx(static_cast<size_t>(dst / kLanes), m, k, dst % kLanes)] = v;
}

template <typename T>
__global__ void k_tiles_to_rows(const T* __restrict__ in, int rows, int m, T* __restrict__ out, long long ld)
{
    const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long total = static_cast<long long>(rows) * m;
    if (gid >= total) return;
    const int k = static_cast<int>(gid / rows);
    const int r = static_cast<int>(gid - static_cast<long long>(k) * rows);
    out[static_cast<long long>(r) * ld + k] = in[tidx(static_cast<size_t>(r / kLanes), m, k, r % kLanes)];
}

// tiled rows [row0, row0 + rows) of [T][m][64] into row-major rows [rows][m]
// (row stride ld)
__global__ void k_slots_range_to_rows(const double2* __restrict__ in, int row0, int rows, int m,
                                      double2* __restrict__ out, long long ld)
{
    const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long total = static_cast<long long>(rows) * m;
    if (gid >= total) return;
    const int k = static_cast<int>(gid / rows);
    const int r = static_cast<int>(gid - static_cast<long long>(k) * rows);
    const int b = row0 + r;
    out[static_cast<long long>(r) * ld + k] = in[tidx(static_cast<size_t>(b / kLanes), m, k, b % kLanes)];
}

// tiled [T][m][64] values of the first `rows` rows into row-major FP64 rows
// [rows][m] with row stride ld
template <typename Real>
__global__ void k_tiles_to_rows_f64(const Real* __restrict__ in, int rows, int m, double* __restrict__ out,
                                    long long ld)
{
    const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long total = static_cast<long long>(rows) * m;
    if (gid >= total) return;
    const int k = static_cast<int>(gid / rows);
    const int r = static_cast<int>(gid - static_cast<long long>(k) * rows);
    out[static_cast<long long>(r) * ld + k] =
        static_cast<double>(in[tidx(static_cast<size_t>(r / kLanes), m, k, r % kLanes)]);
}

// Row edits (outage batches; the reference's apply_contingencies_kernel,
// acpf_nr_kernels.cu:328-348, kept as edits instead of a per-row Ybus copy).
// Rows slots[r] (slots nullable: row r) take row r's edits of the CSR lists.
__global__ void k_set_row_edits(const int* __restrict__ row_ptr, const int* __restrict__ ent,
                                const double2* __restrict__ delta, int R, const int* __restrict__ slots, int cap,
                                int* __restrict__ n_ed, int* __restrict__ ed_ent, double2* __restrict__ ed_val)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    const int b = slots ? slots[r] : r;
    const size_t tile = static_cast<size_t>(b / kLanes);
    const int lane = b % kLanes;
    const int q0 = row_ptr ? row_ptr[r] : 0, cnt = row_ptr ? row_ptr[r + 1] - q0 : 0;
    n_ed[b] = cnt;
    for (int t = 0; t < cnt; ++t) {
        ed_ent[tidx(tile, cap, t, lane)] = ent[q0 + t];
        ed_val[tidx(tile, cap, t, lane)] = delta[q0 + t];
    }
}

// flag[tile][bus] = 1 for the Ybus rows the edits of rows [0, rows) touch
// (entry_row: each Ybus entry's row); the flags start at zero
__global__ void k_edit_flags(const int* __restrict__ n_ed, const int* __restrict__ ed_ent, int cap, int rows,
                             const int* __restrict__ entry_row, int n, int* __restrict__ flag)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const size_t tile = static_cast<size_t>(r / kLanes);
    const int lane = r % kLanes;
    for (int t = 0; t < n_ed[r]; ++t) flag[tile * n + entry_row[ed_ent[tidx(tile, cap, t, lane)]]] = 1;
}

// edit lists of capacity cap_old into capacity cap_new (same rows)
template <typename T>
__global__ void k_regrow_edits(const T* __restrict__ src, int cap_old, T* __restrict__ dst, int cap_new, int tiles)
{
    const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= static_cast<long long>(tiles) * cap_old * kLanes) return;
    const int lane = static_cast<int>(gid % kLanes);
    const long long q = gid / kLanes;
    const int t = static_cast<int>(q % cap_old);
    const size_t tile = static_cast<size_t>(q / cap_old);
    dst[tidx(tile, cap_new, t, lane)] = src[tidx(tile, cap_old, t, lane)];
}

// row-major rows [R][m] (row stride ld; ld = 0 broadcasts row 0) into the
// tiled rows slots[0..R) of dst [T][m][64]
template <typename T>
__global__ void k_rows_to_slots(const T* __restrict__ src, long long ld, int R, int m, const int* __restrict__ slots,
                                T* __restrict__ dst)
{
    const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= static_cast<long long>(R) * m) return;
    const int r = static_cast<int>(gid / m), j = static_cast<int>(gid - static_cast<long long>(r) * m);
    const int b = slots[r];
    dst[tidx(static_cast<size_t>(b / kLanes), m, j, b % kLanes)] = src[static_cast<long long>(r) * ld + j];
}

// tiled rows slots[0..R) of src [T][m][64] into row-major rows [R][m] (stride ld)
template <typename T>
__global__ void k_slots_to_rows(const T* __restrict__ src, int R, int m, const int* __restrict__ slots,
                                T* __restrict__ dst, long long ld)
{
    const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= static_cast<long long>(R) * m) return;
    const int r = static_cast<int>(gid / m), j = static_cast<int>(gid - static_cast<long long>(r) * m);
    const int b = slots[r];
    dst[static_cast<long long>(r) * ld + j] = src[tidx(static_cast<size_t>(b / kLanes), m, j, b % kLanes)];
}

// islanded buses' voltages to NaN (mask_V_nan_kernel, acpf_nr_kernels.cu:289-307)
__global__ void k_mask_v_nan(const unsigned char* __restrict__ mask, long long total, double2* __restrict__ V)
{
    const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid < total && mask[gid]) V[gid] = make_double2(__builtin_nan(""), __builtin_nan(""));
}

// Rows marked for a restart (status kRejected while on their learned start)
// take their fallback start V0 back, bus by bus ...
__global__ __launch_bounds__(256) void k_restart_v(const int* __restrict__ status,
                                                   const unsigned char* __restrict__ warm, int n,
                                                   const double2* __restrict__ V0, int v0_row,
                                                   double2* __restrict__ V)
{
    // V0: tiled like V, or (v0_row) one row [n] for every row
    // workgroup (tile, chunk): a chunk of the tile's buses, for the lanes
    // whose row restarts; a tile with none leaves at once
    const size_t tile = blockIdx.x;
    const int lane = threadIdx.x & (kLanes - 1);
    const int q = threadIdx.x / kLanes, nq = blockDim.x / kLanes;
    const int row = static_cast<int>(tile) * kLanes + lane;
    const bool mine = status[row] == kRejected && warm[row];
    __shared__ int any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    if (mine) any = 1;
    __syncthreads();
    if (!any || !mine) return;
    const int chunk = (n + static_cast<int>(gridDim.y) - 1) / static_cast<int>(gridDim.y);
    const int b0 = static_cast<int>(blockIdx.y) * chunk;
    const int b1 = b0 + chunk < n ? b0 + chunk : n;
    for (int bus = b0 + q; bus < b1; bus += nq)
        V[tidx(tile, n, bus, lane)] = v0_row ? V0[bus] : V0[tidx(tile, n, bus, lane)];
}

// *differs = 1 if some row of [0, rows) has a V other than row 0's
__global__ __launch_bounds__(256) void k_v_uniform(const double2* __restrict__ V, int rows, int n,
                                                   int* __restrict__ differs)
{
    const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long tiles = (rows + kLanes - 1) / kLanes;
    if (gid >= tiles * n * kLanes) return;
    // gid enumerates (tile, bus, lane) in the tiled layout's order
    const int lane = static_cast<int>(gid % kLanes);
    const long long tb = gid / kLanes;
    const size_t tile = static_cast<size_t>(tb / n);
    const int bus = static_cast<int>(tb % n);
    if (static_cast<int>(tile) * kLanes + lane >= rows) return;
    const double2 v = V[tidx(tile, n, bus, lane)], r = V[tidx(0, n, bus, 0)];
    if (v.x != r.x || v.y != r.y) *differs = 1;
}

// ... then continue from it in the same iteration: the fallback gets max_iter
// updates on top of the ones the row had, and the row counts as re-queued.
__global__ void k_restart_rows(int* __restrict__ status, unsigned char* __restrict__ warm, const int* __restrict__ it0,
                               int* __restrict__ budget, int rows, int iter, int max_iter, int* __restrict__ requeued,
                               int* __restrict__ tile_restart)
{
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= rows || status[row] != kRejected || !warm[row]) return;
    budget[row] = (iter - it0[row]) + max_iter;
    warm[row] = 0;
    status[row] = kActive;
    atomicAdd(requeued, 1);
    tile_restart[row / kLanes] = 1;  // the evaluation at the new start covers this tile
}

// *p += v (one thread)
__global__ void k_add_int(int* p, int v)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *p += v;
}

// per-slot scalars: v[slots[r]] = value
__global__ void k_set_slots(const int* __restrict__ slots, int R, int* __restrict__ v, int value)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < R) v[slots[r]] = value;
}

// Gathers CSR values [rows][nnz] (and a rhs [rows][d]) into the A stream of
// the raw factorization.
template <typename Real>
__global__ void k_gather_astream(const double* __restrict__ vals, long long ldv, const double* __restrict__ rhs,
                                 long long ldb, int rows, int nnz, int d, const int* __restrict__ a_pos,
                                 const int* __restrict__ f_pos, int a_stride, Real* __restrict__ A)
{
    const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long per = static_cast<long long>(nnz) + d;
    const long long total = static_cast<long long>(rows) * per;
    if (gid >= total) return;
    const int r = static_cast<int>(gid / per);
    const int k = static_cast<int>(gid - static_cast<long long>(r) * per);
    const size_t tile = static_cast<size_t>(r / kLanes);
    const int lane = r % kLanes;
    if (k < nnz) {
        const int pos = a_pos[k];
        if (pos >= 0) A[tidx(tile, a_stride, pos, lane)] = static_cast<Real>(vals[static_cast<long long>(r) * ldv + k]);
    } else {
        const int e = k - nnz;
        A[tidx(tile, a_stride, f_pos[e], lane)] = static_cast<Real>(rhs[static_cast<long long>(r) * ldb + e]);
    }
}

// per-row max |value| of CSR rows (the pivot clamp scale of the raw path)
__global__ void k_row_absmax(const double* __restrict__ vals, long long ldv, int rows, int nnz, double* __restrict__ out)
{
    const int r = blockIdx.x;
    if (r >= rows) return;
    double m = 0.0;
    for (int k = threadIdx.x; k < nnz; k += blockDim.x) m = fmax(m, fabs(vals[static_cast<long long>(r) * ldv + k]));
    __shared__ double s[256];
    s[threadIdx.x] = m;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[r] = s[0];
}

__global__ void k_fill_int(int* __restrict__ p, int v, long long n)
{
    const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid < n) p[gid] = v;
}

// tile_active[t] = rows of tile t still iterating
__global__ void k_count_tiles(const int* __restrict__ status, int rows_pad, int* __restrict__ tile_active)
{
    const int t = blockIdx.x;
    const int lane = threadIdx.x;
    const int row = t * kLanes + lane;
    const unsigned long long m = __ballot(row < rows_pad && status[row] == kActive);
    if (lane == 0) tile_active[t] = __popcll(m);
}

// Row moves between a batch's tiled arrays [T][m][64] and a dense set of
// rows: gather dst row i <- src row map[i]; scatter dst row map[i] <- src row i.
// One thread per (position, i), i fastest, so the dense side is coalesced.
template <typename T>
__global__ void k_move_rows(const T* __restrict__ src, T* __restrict__ dst, const int* __restrict__ map, int R,
                            int m, int gather)
{
    const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= static_cast<long long>(R) * m) return;
    const int i = static_cast<int>(gid % R);
    const int pos = static_cast<int>(gid / R);
    const int r = map[i];
    const size_t a = tidx(static_cast<size_t>(r >> 6), static_cast<size_t>(m), static_cast<size_t>(pos), r & 63);
    const size_t d = tidx(static_cast<size_t>(i >> 6), static_cast<size_t>(m), static_cast<size_t>(pos), i & 63);
    if (gather) dst[d] = src[a];
    else dst[a] = src[d];
}

// per-row values: gather dst[i] = src[map[i]]; scatter dst[map[i]] = src[i] (+ dst[map[i]] when add)
template <typename T>
__global__ void k_move_vals(const T* __restrict__ src, T* __restrict__ dst, const int* __restrict__ map, int R,
                            int gather, int add)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= R) return;
    if (gather) dst[i] = src[map[i]];
    else dst[map[i]] = add ? dst[map[i]] + src[i] : src[i];
}

// Branch terminal currents on converged V (acpf_nr_kernels.cu:122-182):
//   I_or = yff V_f + yft V_t,  I_ex = ytf V_f + ytt V_t,  amps = |I| * base current
// One thread per (tile, branch, lane); outputs tiled [T][m][64].  A side with
// bus -1 (no terminal) reports 0.
__global__ void k_branch_flows(const double2* __restrict__ V, int n, const int* __restrict__ from,
                               const int* __restrict__ to, const double2* __restrict__ y4,
                               const double* __restrict__ base_or, const double* __restrict__ base_ex, int m,
                               int T, double* __restrict__ or_t, double* __restrict__ ex_t)
{
    const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid >= static_cast<long long>(T) * m * kLanes) return;
    const int lane = static_cast<int>(gid & (kLanes - 1));
    const long long tl = gid / kLanes;
    const int l = static_cast<int>(tl % m);
    const size_t tile = static_cast<size_t>(tl / m);
    const int f = from[l], t = to[l];
    const double2 z = make_double2(0.0, 0.0);
    const double2 vf = f >= 0 ? V[tidx(tile, n, f, lane)] : z;
    const double2 vt = t >= 0 ? V[tidx(tile, n, t, lane)] : z;
    const double2* y = y4 + 4 * static_cast<size_t>(l);
    double2 io = z, ie = z;
    if (f >= 0) {
        const double2 a = cmul(y[0], vf), b = cmul(y[1], vt);
        io = make_double2(a.x + b.x, a.y + b.y);
    }
    if (t >= 0) {
        const double2 a = cmul(y[2], vf), b = cmul(y[3], vt);
        ie = make_double2(a.x + b.x, a.y + b.y);
    }
    or_t[gid] = sqrt(fma(io.x, io.x, io.y * io.y)) * base_or[l];
    ex_t[gid] = sqrt(fma(ie.x, ie.x, ie.y * ie.y)) * base_ex[l];
}

// ---------------------------------------------------------------------------
// k_limit_violations: limit checks of solved rows (check_limit_violations_kernel,
// contingency/violation_kernels.cuh:80-160 semantics).  One wavefront per row:
// a row that did not converge gets one DIVERGENCE record; otherwise every
// branch terminal's current (kA) and every bus voltage (kV) is checked
// against its limit, the true count of each violation type is kept, and of
// each type the K most severe records (severity |value/limit - 1|, ties to
// the element scanned first) are written most severe first, CURRENT then
// LOW_VOLTAGE then HIGH_VOLTAGE.
//
// Scan order: branch l's two sides (keys 2l, 2l + 1), then bus b (key 2m + b).
// The wave walks the elements 64 at a time: a first pass counts and keeps the
// first K records of each type; a type with more than K finds the K-th
// largest severity by an 8-bit-digit radix select over its IEEE bits (eight
// passes; a severity is >= 0 so its bits order like the values) and a last
// pass keeps the records above it plus the first ties in scan order.  The
// kept records are ranked by (severity desc, key asc) in LDS.
// ---------------------------------------------------------------------------

constexpr int kGrpCurrent = 0, kGrpLow = 1, kGrpHigh = 2;
// LimitViolationType / ViolationElementType codes (limit_violation_types.hpp:46-52)
constexpr int kVtLow = 0, kVtHigh = 1, kVtCurrent = 2, kVtDivergence = 4;
constexpr int kEtBus = 0, kEtLine = 1, kEtTrafo = 2, kEtGrid = 3;

__device__ inline unsigned long long sev_key(double value, double limit)
{
    return static_cast<unsigned long long>(__double_as_longlong(fabs(value / limit - 1.0)));
}

// The records of element e: msk bit 0 = record 0 in CURRENT, bit 1 = record
// 1 in CURRENT, bit 2 = record 0 in LOW, bit 3 = record 0 in HIGH.  The
// arithmetic is not contracted, so each pass that re-evaluates an element
// gets the same severity to the last bit.
struct ViolRec {
    int msk;
    double val[2], lim[2];
    unsigned long long key[2];
};

__device__ __forceinline__ ViolRec viol_eval(const ViolationArgs& a, const double2* Vr, const int* trip, int nt,
                                             int e)
{
#pragma clang fp contract(off)
    ViolRec o;
    o.msk = 0;
    if (e < a.m) {
        for (int q = 0; q < nt; ++q)
            if (trip[q] == e) return o;  // the pi-model of a tripped branch no longer applies
        const double l1 = a.lim1[e], l2 = a.lim2[e];
        if (isnan(l1) && isnan(l2)) return o;
        const int f = a.from[e], t = a.to[e];
        const double2 vf = f >= 0 ? Vr[static_cast<size_t>(f) * kLanes] : make_double2(0.0, 0.0);
        const double2 vt = t >= 0 ? Vr[static_cast<size_t>(t) * kLanes] : make_double2(0.0, 0.0);
        const double2* y = a.y4 + 4 * static_cast<size_t>(e);
        const double2 yff = y[0], yft = y[1], ytf = y[2], ytt = y[3];
        double ior = 0.0, ioi = 0.0, ier = 0.0, iei = 0.0;
        if (f >= 0) {
            ior = (yff.x * vf.x - yff.y * vf.y) + (yft.x * vt.x - yft.y * vt.y);
            ioi = (yff.x * vf.y + yff.y * vf.x) + (yft.x * vt.y + yft.y * vt.x);
        }
        if (t >= 0) {
            ier = (ytf.x * vf.x - ytf.y * vf.y) + (ytt.x * vt.x - ytt.y * vt.y);
            iei = (ytf.x * vf.y + ytf.y * vf.x) + (ytt.x * vt.y + ytt.y * vt.x);
        }
        const double ko = sqrt(ior * ior + ioi * ioi) * a.base_or[e] * 0.001;
        const double ke = sqrt(ier * ier + iei * iei) * a.base_ex[e] * 0.001;
        if (!isnan(l1) && ko > l1 * a.up) {
            o.val[0] = ko;
            o.lim[0] = l1;
            o.key[0] = sev_key(ko, l1);
            o.msk |= 1;
        }
        if (!isnan(l2) && ke > l2 * a.up) {
            o.val[1] = ke;
            o.lim[1] = l2;
            o.key[1] = sev_key(ke, l2);
            o.msk |= 2;
        }
        return o;
    }
    const int b = e - a.m;
    const double vmin = a.vmin[b], vmax = a.vmax[b];
    if (isnan(vmin) && isnan(vmax)) return o;
    const double2 v = Vr[static_cast<size_t>(b) * kLanes];
    const double vm = sqrt(v.x * v.x + v.y * v.y) * a.vn_kv[b];  // NaN (islanded bus): no record
    if (!isnan(vmin) && vm < vmin * a.dn) {
        o.val[0] = vm;
        o.lim[0] = vmin;
        o.key[0] = sev_key(vm, vmin);
        o.msk = 4;
    } else if (!isnan(vmax) && vm > vmax * a.up) {
        o.val[0] = vm;
        o.lim[0] = vmax;
        o.key[0] = sev_key(vm, vmax);
        o.msk = 8;
    }
    return o;
}

// the group of record q of an element with record mask msk (-1: none)
__device__ inline int viol_group(int msk, int q)
{
    if (q == 0) return (msk & 1) ? kGrpCurrent : (msk & 4) ? kGrpLow : (msk & 8) ? kGrpHigh : -1;
    return (msk & 2) ? kGrpCurrent : -1;
}

__global__ __launch_bounds__(kLanes) void k_limit_violations(ViolationArgs a)
{
    extern __shared__ __align__(16) unsigned char g2g_smem[];
    const int lane = threadIdx.x;
    const int r = blockIdx.x;
    const int K = a.K;
    unsigned* hist = reinterpret_cast<unsigned*>(g2g_smem);                       // [3][256]
    unsigned long long* lkey = reinterpret_cast<unsigned long long*>(hist + 3 * 256);  // [3][K]
    double* lval = reinterpret_cast<double*>(lkey + 3 * K);                        // [3][K]
    double* llim = lval + 3 * K;                                                   // [3][K]
    int* lord = reinterpret_cast<int*>(llim + 3 * K);                              // [3][K] scan keys
    __shared__ unsigned long long s_sel[3];
    __shared__ int s_need[3];

    const size_t ob = static_cast<size_t>(r) * 3 * K;
    if (a.status[r] == kNotSimulated) {
        if (lane == 0) {
            a.count[r] = -1;
            a.truncated[r] = 0;
            for (int g = 0; g < 3; ++g) a.n_type[3 * r + g] = -1;
        }
        return;
    }
    const double res = a.res[r];
    if (isnan(res) || res > a.tol) {
        if (lane == 0) {
            a.etype[ob] = kEtGrid;
            a.eid[ob] = -1;
            a.side[ob] = 0;
            a.vtype[ob] = kVtDivergence;
            a.value[ob] = res;
            a.limit[ob] = a.tol;
            a.count[r] = 1;
            a.truncated[r] = 0;
            for (int g = 0; g < 3; ++g) a.n_type[3 * r + g] = 0;
        }
        return;
    }
    const double2* Vr = a.V + static_cast<size_t>(r / kLanes) * a.n * kLanes + (r % kLanes);
    const int* trip = nullptr;
    int nt = 0;
    if (a.trip_ptr) {
        trip = a.trip + a.trip_ptr[r];
        nt = a.trip_ptr[r + 1] - a.trip_ptr[r];
    }
    const int E = a.m + a.n;
    const unsigned long long lt = __lanemask_lt();
    auto scan_key = [&](int e, int q) { return e < a.m ? 2 * e + q : 2 * a.m + (e - a.m); };
    auto put = [&](int g, int slot, unsigned long long k, double v, double l, int order) {
        lkey[g * K + slot] = k;
        lval[g * K + slot] = v;
        llim[g * K + slot] = l;
        lord[g * K + slot] = order;
    };

    // pass 1: true counts, and the first K records of each type in scan order
    int cnt[3] = {0, 0, 0};
    for (int e0 = 0; e0 < E; e0 += kLanes) {
        const int e = e0 + lane;
        ViolRec v;
        v.msk = 0;
        if (e < E) v = viol_eval(a, Vr, trip, nt, e);
        const int g0 = viol_group(v.msk, 0), g1 = viol_group(v.msk, 1);
#pragma unroll
        for (int g = 0; g < 3; ++g) {
            const unsigned long long b0 = __ballot(g0 == g), b1 = __ballot(g1 == g);
            if ((b0 | b1) == 0) continue;
            const int s0 = cnt[g] + __popcll(b0 & lt) + __popcll(b1 & lt);
            const int s1 = s0 + (g0 == g ? 1 : 0);
            if (g0 == g && s0 < K) put(g, s0, v.key[0], v.val[0], v.lim[0], scan_key(e, 0));
            if (g1 == g && s1 < K) put(g, s1, v.key[1], v.val[1], v.lim[1], scan_key(e, 1));
            cnt[g] += __popcll(b0) + __popcll(b1);
        }
    }
    const bool over0 = cnt[0] > K, over1 = cnt[1] > K, over2 = cnt[2] > K;
    if (over0 || over1 || over2) {
        const bool over[3] = {over0, over1, over2};
        // radix select of the K-th largest severity of each overflowing type
        unsigned long long prefix[3] = {0, 0, 0}, pmask[3] = {0, 0, 0};
        int need[3] = {K, K, K};
        for (int pass = 0; pass < 8; ++pass) {
            const int shift = 56 - 8 * pass;
            for (int i = lane; i < 3 * 256; i += kLanes) hist[i] = 0;
            __syncthreads();
            for (int e0 = 0; e0 < E; e0 += kLanes) {
                const int e = e0 + lane;
                ViolRec v;
                v.msk = 0;
                if (e < E) v = viol_eval(a, Vr, trip, nt, e);
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    const int g = viol_group(v.msk, q);
                    if (g >= 0 && over[g] && (v.key[q] & pmask[g]) == prefix[g])
                        atomicAdd(&hist[g * 256 + static_cast<int>((v.key[q] >> shift) & 255)], 1u);
                }
            }
            __syncthreads();
            if (lane < 3 && over[lane]) {
                const int g = lane;
                unsigned acc = 0;
                int d = 255;
                for (; d > 0; --d) {
                    const unsigned c = hist[g * 256 + d];
                    if (acc + c >= static_cast<unsigned>(need[g])) break;
                    acc += c;
                }
                s_sel[g] = static_cast<unsigned long long>(d);
                s_need[g] = need[g] - static_cast<int>(acc);
            }
            __syncthreads();
#pragma unroll
            for (int g = 0; g < 3; ++g)
                if (over[g]) {
                    prefix[g] |= s_sel[g] << shift;
                    pmask[g] |= 255ull << shift;
                    need[g] = s_need[g];
                }
            __syncthreads();
        }
        // collect: above the threshold, plus the first `need` ties in scan order
        int kept[3] = {0, 0, 0}, ties[3] = {0, 0, 0};
        for (int e0 = 0; e0 < E; e0 += kLanes) {
            const int e = e0 + lane;
            ViolRec v;
            v.msk = 0;
            if (e < E) v = viol_eval(a, Vr, trip, nt, e);
            const int g0 = viol_group(v.msk, 0), g1 = viol_group(v.msk, 1);
#pragma unroll
            for (int g = 0; g < 3; ++g) {
                if (!over[g]) continue;
                const bool in0 = g0 == g, in1 = g1 == g;
                const bool eq0 = in0 && v.key[0] == prefix[g], eq1 = in1 && v.key[1] == prefix[g];
                const unsigned long long t0 = __ballot(eq0), t1 = __ballot(eq1);
                const int r0 = ties[g] + __popcll(t0 & lt) + __popcll(t1 & lt);
                const int r1 = r0 + (eq0 ? 1 : 0);
                const bool k0 = in0 && (v.key[0] > prefix[g] || (eq0 && r0 < need[g]));
                const bool k1 = in1 && (v.key[1] > prefix[g] || (eq1 && r1 < need[g]));
                const unsigned long long c0 = __ballot(k0), c1 = __ballot(k1);
                const int s0 = kept[g] + __popcll(c0 & lt) + __popcll(c1 & lt);
                const int s1 = s0 + (k0 ? 1 : 0);
                if (k0) put(g, s0, v.key[0], v.val[0], v.lim[0], scan_key(e, 0));
                if (k1) put(g, s1, v.key[1], v.val[1], v.lim[1], scan_key(e, 1));
                kept[g] += __popcll(c0) + __popcll(c1);
                ties[g] += __popcll(t0) + __popcll(t1);
            }
        }
    }
    __syncthreads();
    // rank the kept records of each type and write them out in type order
    const int k0 = min(cnt[0], K), k1 = min(cnt[1], K), k2 = min(cnt[2], K);
    const int kk[3] = {k0, k1, k2};
    const int off[3] = {0, k0, k0 + k1};
#pragma unroll
    for (int g = 0; g < 3; ++g) {
        const int ng = kk[g];
        for (int i = lane; i < ng; i += kLanes) {
            const unsigned long long ki = lkey[g * K + i];
            const int oi = lord[g * K + i];
            int rank = 0;
            for (int j = 0; j < ng; ++j) {
                const unsigned long long kj = lkey[g * K + j];
                rank += (kj > ki || (kj == ki && lord[g * K + j] < oi)) ? 1 : 0;
            }
            const size_t o = ob + off[g] + rank;
            if (oi < 2 * a.m) {
                const int l = oi >> 1;
                a.etype[o] = l < a.n_lines ? kEtLine : kEtTrafo;
                a.eid[o] = l < a.n_lines ? l : l - a.n_lines;
                a.side[o] = (oi & 1) + 1;
                a.vtype[o] = kVtCurrent;
            } else {
                a.etype[o] = kEtBus;
                a.eid[o] = oi - 2 * a.m;
                a.side[o] = 0;
                a.vtype[o] = g == kGrpLow ? kVtLow : kVtHigh;
            }
            a.value[o] = lval[g * K + i];
            a.limit[o] = llim[g * K + i];
        }
    }
    if (lane == 0) {
        a.count[r] = k0 + k1 + k2;
        a.truncated[r] = (cnt[0] > K || cnt[1] > K || cnt[2] > K) ? 1 : 0;
        a.n_type[3 * r + 0] = cnt[kGrpLow];
        a.n_type[3 * r + 1] = cnt[kGrpHigh];
        a.n_type[3 * r + 2] = cnt[kGrpCurrent];
    }
}

size_t limit_violations_lds(int K)
{
    return 3 * 256 * sizeof(unsigned) + 3 * static_cast<size_t>(K) * (8 + 8 + 8 + 4);
}

// ---------------------------------------------------------------------------
// explicit instantiations used by the C ABI
// ---------------------------------------------------------------------------

template __global__ void k_mismatch<double, false, true>(MismatchArgs<double>);
template __global__ void k_mismatch<double, true, true>(MismatchArgs<double>);
template __global__ void k_mismatch<double, false, false>(MismatchArgs<double>);
template __global__ void k_mismatch<double, true, false>(MismatchArgs<double>);
template __global__ void k_mismatch<float, false, true>(MismatchArgs<float>);
template __global__ void k_mismatch<float, true, true>(MismatchArgs<float>);
template __global__ void k_mismatch<float, false, false>(MismatchArgs<float>);
template __global__ void k_mismatch<float, true, false>(MismatchArgs<float>);
template __global__ void k_move_rows<double2>(const double2*, double2*, const int*, int, int, int);
template __global__ void k_move_rows<int>(const int*, int*, const int*, int, int, int);
template __global__ void k_regrow_edits<int>(const int*, int, int*, int, int);
template __global__ void k_regrow_edits<double2>(const double2*, int, double2*, int, int);
template __global__ void k_move_rows<unsigned char>(const unsigned char*, unsigned char*, const int*, int, int, int);
template __global__ void k_rows_to_slots<double2>(const double2*, long long, int, int, const int*, double2*);
template __global__ void k_slots_to_rows<double2>(const double2*, int, int, const int*, double2*, long long);
template __global__ void k_move_vals<int>(const int*, int*, const int*, int, int, int);
template __global__ void k_move_vals<double>(const double*, double*, const int*, int, int, int);
template __global__ void k_move_vals<unsigned char>(const unsigned char*, unsigned char*, const int*, int, int, int);
template __global__ void k_factor<double, 1, false, 0, kFactorFull>(FactorArgs<double>);
template __global__ void k_factor<float, 1, false, 0, kFactorFull>(FactorArgs<float>);
template __global__ void k_factor<double, 2, false, 0, kFactorFull>(FactorArgs<double>);
template __global__ void k_factor<float, 2, false, 0, kFactorFull>(FactorArgs<float>);
template __global__ void k_factor<float, 2, false, kHalfRingGeometry, kFactorFull>(FactorArgs<float>);
template __global__ void k_factor<double, 1, true, 0, kFactorFull>(FactorArgs<double>);
template __global__ void k_factor<float, 1, true, 0, kFactorFull>(FactorArgs<float>);
template __global__ void k_factor<float, 1, false, 0, kFactorStoreL>(FactorArgs<float>);
template __global__ void k_factor<float, 2, false, 0, kFactorStoreL>(FactorArgs<float>);
template __global__ void k_factor<float, 2, false, kHalfRingGeometry, kFactorStoreL>(FactorArgs<float>);
template __global__ void k_factor<float, 1, false, 0, kForwardOnly>(FactorArgs<float>);
template __global__ void k_factor<float, 2, false, 0, kForwardOnly>(FactorArgs<float>);
template __global__ void k_factor<float, 2, false, kHalfRingGeometry, kForwardOnly>(FactorArgs<float>);
template __global__ void k_backward<double, 1, false>(BackwardArgs<double>);
template __global__ void k_backward<float, 1, false>(BackwardArgs<float>);
template __global__ void k_backward<double, 2, false>(BackwardArgs<double>);
template __global__ void k_backward<float, 2, false>(BackwardArgs<float>);
template __global__ void k_backward<double, 4, false>(BackwardArgs<double>);
template __global__ void k_backward<float, 4, false>(BackwardArgs<float>);
template __global__ void k_backward<double, 1, true>(BackwardArgs<double>);
template __global__ void k_backward<float, 1, true>(BackwardArgs<float>);
template __global__ void k_ext_pre<double>(ExtArgs<double>);
template __global__ void k_ext_pre<float>(ExtArgs<float>);
template __global__ void k_ext_post<double>(ExtArgs<double>);
template __global__ void k_ext_post<float>(ExtArgs<float>);
template __global__ void k_ext_update<double>(ExtArgs<double>);
template __global__ void k_ext_update<float>(ExtArgs<float>);
template __global__ void k_ext_init<double>(ExtArgs<double>);
template __global__ void k_ext_init<float>(ExtArgs<float>);
template __global__ void k_move_rows<double>(const double*, double*, const int*, int, int, int);
template __global__ void k_mis_status<double, true>(MismatchArgs<double>);
template __global__ void k_mis_status<double, false>(MismatchArgs<double>);
template __global__ void k_mis_status<float, true>(MismatchArgs<float>);
template __global__ void k_mis_status<float, false>(MismatchArgs<float>);
template __global__ void k_update<double>(UpdateArgs<double>);
template __global__ void k_update<float>(UpdateArgs<float>);
template __global__ void k_tiles_to_rows_f64<double>(const double*, int, int, double*, long long);
template __global__ void k_tiles_to_rows_f64<float>(const float*, int, int, double*, long long);
template __global__ void k_lin_residual<double>(RefineArgs<double>);
template __global__ void k_lin_residual<float>(RefineArgs<float>);
template __global__ void k_ir_residual<double, false>(IrArgs<double>);
template __global__ void k_ir_residual<double, true>(IrArgs<double>);
template __global__ void k_ir_residual<float, false>(IrArgs<float>);
template __global__ void k_ir_residual<float, true>(IrArgs<float>);
template __global__ void k_axpy_tiles<double>(const double*, double*, const int*, int);
template __global__ void k_axpy_tiles<float>(const float*, float*, const int*, int);
template __global__ void k_rows_to_tiles<double2>(const double2*, long long, int, int, int, double2*, int);
template __global__ void k_rows_to_tiles<double>(const double*, long long, int, int, int, double*, int);
template __global__ void k_tiles_to_rows<double2>(const double2*, int, int, double2*, long long);
template __global__ void k_tiles_to_rows<double>(const double*, int, int, double*, long long);
template __global__ void k_gather_astream<double>(const double*, long long, const double*, long long, int, int, int,
                                                  const int*, const int*, int, double*);
template __global__ void k_gather_astream<float>(const double*, long long, const double*, long long, int, int, int,
                                                 const int*, const int*, int, float*);

}  // namespace g2g
