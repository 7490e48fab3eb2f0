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
//   k_backward      U x = y in reverse postorder with x(j) kept in LDS slots;
//                   the polar V update is fused (acpf_nr_kernels.cu:728-788).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_common.hpp"

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

__device__ __forceinline__ double nan_max(double m, double v)
{
    // NaN-sticky max of |v| (acpf_nr_kernels.cu:1101-1114)
    const double a = fabs(v);
    return (isnan(m) || isnan(a)) ? __builtin_nan("") : (a > m ? a : m);
}

// ---------------------------------------------------------------------------
// k_mismatch: one workgroup per tile, kMisWaves waves split the buses
// ---------------------------------------------------------------------------

template <typename Real, bool PerRowY, bool WithJ>
__global__ __launch_bounds__(kLanes* kMisWaves) void k_mismatch(MismatchArgs<Real> a)
{
    const int lane = threadIdx.x & (kLanes - 1);
    const int wave = threadIdx.x / kLanes;
    const size_t tile = blockIdx.x;
    const int row = static_cast<int>(tile) * kLanes + lane;
    __shared__ double s_res[kMisWaves][kLanes];
    __shared__ double s_jmax[kMisWaves][kLanes];

    const int st0 = a.status[row];
    const bool active = st0 == kActive;
    double rmax = 0.0, jmax = 0.0;
    if (a.tile_active[tile] > 0) {
        const int b0 = static_cast<int>((static_cast<long long>(a.nbus) * wave) / kMisWaves);
        const int b1 = static_cast<int>((static_cast<long long>(a.nbus) * (wave + 1)) / kMisWaves);
        const double2* V = a.V;
        const double2* S = a.S;
        for (int bb = b0; bb < b1; ++bb) {
            const int* bus = a.k1_bus + 6 * bb;
            const int i = bus[0], e0 = bus[1], e1 = bus[2], fp = bus[3], fq = bus[4], dk = bus[5];
            const double2 vi = V[tidx(tile, a.n, i, lane)];
            double2 I = make_double2(0.0, 0.0);
            double dr = 0.0, di = 0.0;  // diagonal admittance
            int dent = -1;
            for (int e = e0; e < e1; ++e) {
                const int* en = a.k1_ent + 6 * e;
                const int j = en[0], yk = en[1];
                const double2 y = PerRowY ? a.Yrow[tidx(tile, a.nY, yk, lane)] : a.Y[yk];
                if (j == i) {
                    dr = y.x;
                    di = y.y;
                    dent = e;
                    I.x = fma(y.x, vi.x, fma(-y.y, vi.y, I.x));
                    I.y = fma(y.x, vi.y, fma(y.y, vi.x, I.y));
                    continue;
                }
                const double2 vj = V[tidx(tile, a.n, j, lane)];
                const double2 t = cmul(y, vj);
                I.x += t.x;
                I.y += t.y;
                if (WithJ) {
                    const double2 w = cmulc(vi, t);  // V_i conj(Y_ij V_j)
                    const double inv = 1.0 / sqrt(fma(vj.x, vj.x, vj.y * vj.y));
                    const int a11 = en[2], a12 = en[3], a21 = en[4], a22 = en[5];
                    const double j11 = w.y, j21 = -w.x, j12 = w.x * inv, j22 = w.y * inv;
                    if (a11 >= 0) a.A[tidx(tile, a.a_len, a11, lane)] = static_cast<Real>(j11);
                    if (a12 >= 0) a.A[tidx(tile, a.a_len, a12, lane)] = static_cast<Real>(j12);
                    if (a21 >= 0) a.A[tidx(tile, a.a_len, a21, lane)] = static_cast<Real>(j21);
                    if (a22 >= 0) a.A[tidx(tile, a.a_len, a22, lane)] = static_cast<Real>(j22);
                    jmax = fmax(jmax, fmax(fmax(fabs(j11), fabs(j12)), fmax(fabs(j21), fabs(j22))));
                }
            }
            (void)dk;
            const double2 sc = cmulc(vi, I);  // V conj(I)
            const double2 sb = S[tidx(tile, a.n, i, lane)];
            const double fpv = -(sc.x - sb.x);
            rmax = nan_max(rmax, fpv);
            a.A[tidx(tile, a.a_len, fp, lane)] = static_cast<Real>(fpv);
            if (fq >= 0) {
                const double fqv = -(sc.y - sb.y);
                rmax = nan_max(rmax, fqv);
                a.A[tidx(tile, a.a_len, fq, lane)] = static_cast<Real>(fqv);
            }
            if (WithJ) {
                const int* en = a.k1_ent + 6 * dent;
                const double2 t = make_double2(fma(dr, vi.x, -di * vi.y), fma(dr, vi.y, di * vi.x));
                // dS/dVa = j V_i conj(I - t); dS/dVm = V_i conj(t + I) / |V_i|
                const double2 z = cmulc(vi, make_double2(I.x - t.x, I.y - t.y));
                const double2 u = cmulc(vi, make_double2(I.x + t.x, I.y + t.y));
                const double inv = 1.0 / sqrt(fma(vi.x, vi.x, vi.y * vi.y));
                const double j11 = -z.y, j21 = z.x, j12 = u.x * inv, j22 = u.y * inv;
                const int a11 = en[2], a12 = en[3], a21 = en[4], a22 = en[5];
                if (a11 >= 0) a.A[tidx(tile, a.a_len, a11, lane)] = static_cast<Real>(j11);
                if (a12 >= 0) a.A[tidx(tile, a.a_len, a12, lane)] = static_cast<Real>(j12);
                if (a21 >= 0) a.A[tidx(tile, a.a_len, a21, lane)] = static_cast<Real>(j21);
                if (a22 >= 0) a.A[tidx(tile, a.a_len, a22, lane)] = static_cast<Real>(j22);
                jmax = fmax(jmax, fmax(fmax(fabs(j11), fabs(j12)), fmax(fabs(j21), fabs(j22))));
            }
        }
    }
    s_res[wave][lane] = rmax;
    s_jmax[wave][lane] = jmax;
    __syncthreads();
    if (wave != 0) return;
    if (a.tile_active[tile] <= 0) return;
    double r = 0.0, jm = 0.0;
    for (int w = 0; w < kMisWaves; ++w) {
        const double v = s_res[w][lane];
        r = (isnan(r) || isnan(v)) ? __builtin_nan("") : fmax(r, v);
        jm = fmax(jm, s_jmax[w][lane]);
    }
    int still = 0;
    if (active && row < a.rows) {
        a.res[row] = r;
        if (WithJ) a.jmax[row] = jm;
        int st = kActive;
        if (!isfinite(r)) st = kDiverged;
        else if (a.freeze && r < a.tol) st = kConverged;
        else if (a.iter >= a.max_iter) st = r < a.tol ? kConverged : kMaxIter;
        if (st != kActive) {
            a.status[row] = st;
            a.iters[row] = a.iter;
        }
        still = st == kActive;
    }
    // wave-level count of rows still iterating
    const unsigned long long m = __ballot(still);
    if (lane == 0) {
        const int cnt = __popcll(m);
        a.tile_active[tile] = cnt;
        if (cnt) atomicAdd(a.n_active, cnt);
    }
}

// ---------------------------------------------------------------------------
// k_factor: right-looking static-pivot LU + forward substitution
// ---------------------------------------------------------------------------

template <typename Real>
__global__ __launch_bounds__(kLanes) void k_factor(FactorArgs<Real> a)
{
    const int lane = threadIdx.x;
    const size_t tile = blockIdx.x;
    if (a.tile_active[tile] <= 0) return;
    const int row = static_cast<int>(tile) * kLanes + lane;
    extern __shared__ unsigned char g2g_smem[];
    Real* sl = reinterpret_cast<Real*>(g2g_smem);          // [slots][64]
    Real* lsc = sl + static_cast<size_t>(a.slots) * kLanes;  // [kMaxC][64] L column scratch
    const Real* A = a.A + tile * static_cast<size_t>(a.a_len) * kLanes + lane;
    Real* U = a.U + tile * static_cast<size_t>(a.u_len) * kLanes + lane;
    const int* prog = a.prog;
    const Real eps = static_cast<Real>(a.pivot_eps) * static_cast<Real>(a.jmax ? a.jmax[row] : 1.0);
    int clamped = 0;
    int p = 0, ap = 0, up = 0;
    for (int k = 0; k < a.d; ++k) {
        const int c = prog[p];
        Real fk = A[static_cast<size_t>(ap++) * kLanes];
        const int bref = prog[p + 1];
        if (bref != kNoSlot) fk += sl[bref * kLanes + lane];
        const int dref = prog[p + 2];
        Real dv = Real(0);
        if (dref & kInA) dv = A[static_cast<size_t>(ap++) * kLanes];
        if ((dref & kSlotMask) != kNoSlot) dv += sl[(dref & kSlotMask) * kLanes + lane];
        if (fabs(dv) < eps) {  // static-pivot clamp (own_lu/batched_lu.cu:45-53); NaN passes
            dv = dv < Real(0) ? -eps : eps;
            ++clamped;
        }
        const Real inv = Real(1) / dv;
        p += 3;
        Real u[kMaxC];
#pragma unroll
        for (int q = 0; q < kMaxC; ++q) {
            if (q < c) {
                const int ref = prog[p + q];
                Real v = Real(0);
                if (ref & kInA) v = A[static_cast<size_t>(ap++) * kLanes];
                if ((ref & kSlotMask) != kNoSlot) v += sl[(ref & kSlotMask) * kLanes + lane];
                lsc[q * kLanes + lane] = v * inv;
            }
        }
        p += c;
#pragma unroll
        for (int q = 0; q < kMaxC; ++q) {
            u[q] = Real(0);
            if (q < c) {
                const int ref = prog[p + q];
                Real v = Real(0);
                if (ref & kInA) v = A[static_cast<size_t>(ap++) * kLanes];
                if ((ref & kSlotMask) != kNoSlot) v += sl[(ref & kSlotMask) * kLanes + lane];
                u[q] = v;
            }
        }
        p += c;
        // U record: y(k), U(k,k), U(k,j)...
        U[static_cast<size_t>(up++) * kLanes] = fk;
        U[static_cast<size_t>(up++) * kLanes] = dv;
#pragma unroll
        for (int q = 0; q < kMaxC; ++q)
            if (q < c) U[static_cast<size_t>(up++) * kLanes] = u[q];
        // rank-1 update of the trailing entries, then of the right-hand side
        for (int r = 0; r < c; ++r) {
            const Real l = lsc[r * kLanes + lane];
#pragma unroll
            for (int q = 0; q < kMaxC; ++q) {
                if (q < c) {
                    const int ref = prog[p + q];
                    Real* t = sl + (ref & kSlotMask) * kLanes + lane;
                    const Real v = -l * u[q];
                    if (ref & kFirst) *t = v;
                    else *t += v;
                }
            }
            p += c;
        }
        for (int r = 0; r < c; ++r) {
            const int ref = prog[p + r];
            Real* t = sl + (ref & kSlotMask) * kLanes + lane;
            const Real v = -lsc[r * kLanes + lane] * fk;
            if (ref & kFirst) *t = v;
            else *t += v;
        }
        p += c;
    }
    if (a.clamped && row < a.rows) a.clamped[row] = clamped;
}

// ---------------------------------------------------------------------------
// k_backward: U x = y, reverse postorder; fused polar update or raw output
// ---------------------------------------------------------------------------

template <typename Real, bool Update>
__global__ __launch_bounds__(kLanes) void k_backward(BackwardArgs<Real> a)
{
    const int lane = threadIdx.x;
    const size_t tile = blockIdx.x;
    if (a.tile_active[tile] <= 0) return;
    const int row = static_cast<int>(tile) * kLanes + lane;
    const bool live = row < a.rows && (!Update || a.status[row] == kActive);
    extern __shared__ unsigned char g2g_smem[];
    Real* sl = reinterpret_cast<Real*>(g2g_smem);  // [slots][64]
    const Real* U = a.U + tile * static_cast<size_t>(a.u_len) * kLanes + lane;
    const int* prog = a.prog;
    int p = 0, up = a.u_len;
    for (int k = a.d - 1; k >= 0; --k) {
        const int c = prog[p];
        up -= c + 2;
        Real s = U[static_cast<size_t>(up) * kLanes];
        const Real piv = U[static_cast<size_t>(up + 1) * kLanes];
        for (int q = 0; q < c; ++q) {
            const Real xj = sl[prog[p + 1 + q] * kLanes + lane];
            s = fma(-U[static_cast<size_t>(up + 2 + q) * kLanes], xj, s);
        }
        p += 1 + c;
        const Real x = s / piv;
        const int out = prog[p], op = prog[p + 1], bus = prog[p + 2], ps = prog[p + 3];
        p += 4;
        if (Update) {
            if (op != 0) {
                double dva, dvm = 0.0;
                if (op == kOpThetaSelf) {
                    dva = static_cast<double>(x);
                } else if (op == kOpThetaSelfVmSlot) {
                    dva = static_cast<double>(x);
                    dvm = static_cast<double>(sl[ps * kLanes + lane]);
                } else {
                    dvm = static_cast<double>(x);
                    dva = static_cast<double>(sl[ps * kLanes + lane]);
                }
                if (live) {
                    double2* vp = a.V + tidx(tile, a.n, bus, lane);
                    const double2 v = *vp;
                    const double m = sqrt(fma(v.x, v.x, v.y * v.y));
                    const double sc = (m + dvm) / m;
                    double sn, cs;
                    sincos(dva, &sn, &cs);
                    *vp = make_double2(sc * fma(v.x, cs, -v.y * sn), sc * fma(v.x, sn, v.y * cs));
                }
            }
        } else {
            if (row < a.rows) a.X[tidx(tile, a.d, a.perm[k], lane)] = x;
        }
        if (out != kNoSlot) sl[out * kLanes + lane] = x;
    }
}

// ---------------------------------------------------------------------------
// layout transforms and value edits
// ---------------------------------------------------------------------------

// rows-major [rows][m] (row stride ld elements) <-> tiled [T][m][64]
template <typename T>
__global__ void k_rows_to_tiles(const T* __restrict__ in, long long ld, int rows, int m, int row_offset,
                                T* __restrict__ out, int broadcast)
{
    const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long total = static_cast<long long>(rows) * m;
    if (gid >= total) return;
    const int r = static_cast<int>(gid / m);
    const int k = static_cast<int>(gid - static_cast<long long>(r) * m);
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
    const int r = static_cast<int>(gid / m);
    const int k = static_cast<int>(gid - static_cast<long long>(r) * m);
    out[static_cast<long long>(r) * ld + k] = in[tidx(static_cast<size_t>(r / kLanes), m, k, r % kLanes)];
}

// Per-row Ybus values: broadcast the shared values to every row, then subtract
// each row's outage deltas (the reference's apply_contingencies_kernel,
// acpf_nr_kernels.cu:328-348, on the tiled layout).
__global__ void k_ybus_broadcast(const double2* __restrict__ Y, int nY, int rows_pad, double2* __restrict__ Yrow)
{
    const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long total = static_cast<long long>(rows_pad) * nY;
    if (gid >= total) return;
    const int lane = static_cast<int>(gid & (kLanes - 1));
    const long long rest = gid >> 6;
    const int k = static_cast<int>(rest % nY);
    const long long tile = rest / nY;
    Yrow[tidx(static_cast<size_t>(tile), nY, k, lane)] = Y[k];
}

__global__ void k_ybus_edits(const int* __restrict__ row_ptr, const int* __restrict__ ent,
                             const double2* __restrict__ delta, int rows, int nY, double2* __restrict__ Yrow)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    for (int q = row_ptr[r]; q < row_ptr[r + 1]; ++q) {
        double2* y = Yrow + tidx(static_cast<size_t>(r / kLanes), nY, ent[q], r % kLanes);
        const double2 v = *y;
        *y = make_double2(v.x - delta[q].x, v.y - delta[q].y);
    }
}

// Gathers CSR values [rows][nnz] (and a rhs [rows][d]) into the A stream of
// the raw factorization.
template <typename Real>
__global__ void k_gather_astream(const double* __restrict__ vals, long long ldv, const double* __restrict__ rhs,
                                 long long ldb, int rows, int nnz, int d, const int* __restrict__ a_pos,
                                 const int* __restrict__ f_pos, int a_len, Real* __restrict__ A)
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
        if (pos >= 0) A[tidx(tile, a_len, pos, lane)] = static_cast<Real>(vals[static_cast<long long>(r) * ldv + k]);
    } else {
        const int e = k - nnz;
        A[tidx(tile, a_len, f_pos[e], lane)] = static_cast<Real>(rhs[static_cast<long long>(r) * ldb + e]);
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
template __global__ void k_factor<double>(FactorArgs<double>);
template __global__ void k_factor<float>(FactorArgs<float>);
template __global__ void k_backward<double, true>(BackwardArgs<double>);
template __global__ void k_backward<double, false>(BackwardArgs<double>);
template __global__ void k_backward<float, true>(BackwardArgs<float>);
template __global__ void k_backward<float, false>(BackwardArgs<float>);
template __global__ void k_rows_to_tiles<double2>(const double2*, long long, int, int, int, double2*, int);
template __global__ void k_rows_to_tiles<double>(const double*, long long, int, int, int, double*, int);
template __global__ void k_tiles_to_rows<double2>(const double2*, int, int, double2*, long long);
template __global__ void k_tiles_to_rows<double>(const double*, int, int, double*, long long);
template __global__ void k_gather_astream<double>(const double*, long long, const double*, long long, int, int, int,
                                                  const int*, const int*, int, double*);
template __global__ void k_gather_astream<float>(const double*, long long, const double*, long long, int, int, int,
                                                 const int*, const int*, int, float*);

}  // namespace g2g
