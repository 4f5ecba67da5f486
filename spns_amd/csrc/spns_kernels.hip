// HIP kernels for the pseudospectral HIT step on MI355X (gfx950).
//
// Data layout in HBM (k3 contiguous everywhere):
//   state u_c, R_c      : FP64 compact retained box [i2][i1][i3], nk2 x nk1 x nk3,
//                          the reference's Fourier-dump order (src/HIT_LES_output.cpp:46-53)
//   X1 (after the y pass): T complex [y][i1][i3], My x nk1 x nk3 per field
//   X2 (after the x pass): T complex [y][x][i3],  My x Mx  x nk3 per field
// Passes of one step (DNS; LES adds six SGS fields to the forward passes):
//   inv_y  : u_c -> {F_c = FFT_y u_c, G_c = FFT_y(i k2 u_c)}          (strided, pruned in)
//   inv_x  : {F_c, G_c} -> {u_c, d_x u_c, d_y u_c}                   (strided, pruned in)
//   z pass : c2r of 12 fields (d_z from i k3), advective products,
//            Smagorinsky nu_t and tau, CFL max, skewness, r2c          (pencil, pruned in/out)
//   fwd_x, fwd_y : forward transforms back to the compact box         (strided, pruned out)
//   epilogue: normalise, R, rotation, forcing, AB2 / kappa, projection,
//            E(K) and enstrophy partials                              (elementwise, FP64)
//   finalize: fixed-order sums, CFL time step, time, iteration count
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "spns_device.hpp"
#include "spns_internal.hpp"

#ifndef SPNS_PART
#define SPNS_PART 0
#endif

namespace spns {

// ---------------------------------------------------------------------------
// shell index K = round(|k|) computed exactly in integers (src/HIT_LES.cpp:1166)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int shell_of(int k2)
{
    int m = (int)sqrt((double)k2);
    while (m * m > k2) m--;
    while ((m + 1) * (m + 1) <= k2) m++;
    return (k2 > m * m + m) ? m + 1 : m;
}

__device__ __forceinline__ int wrap_k(int i, int n)  // compact index -> wavenumber
{
    return (i <= n / 2) ? i : i - n;
}

// i k v, formed in the input precision (src/HIT_LES.cpp:37) and rounded to
// the transform type T.  The products are rounded before they enter a
// butterfly: with contraction off, the compiler cannot fuse them into the
// butterflies' additions in one loader and not in another, so every pass
// that forms a derivative on load (the strided inverse passes, and the next
// step's inverse y pass inside the fused forward y kernel) has the same
// rounding.
template <typename T, typename TIN>
__device__ __forceinline__ cx<T> times_ik(TIN kk, cx<TIN> v)
{
#pragma clang fp contract(off)
    const TIN re = -kk * v.y, im = kk * v.x;
    return cx<T>{(T)re, (T)im};
}

// ---------------------------------------------------------------------------
// Strided batched pass: transforms along a row dimension of [A][rows][B]
// (B contiguous), TB consecutive columns per block.
// ---------------------------------------------------------------------------
template <typename TIN, typename TOUT, typename T, int L, int TB, int NT, int DIR>
__global__ __launch_bounds__(NT) void k_strided(StridedArgs a)
{
    __shared__ cx<T> lds[L * TB];
    if (a.skip && *a.skip) return;
    // Workgroups are dispatched round-robin over the 8 XCDs, each with its own
    // L2.  Remap the (padded, multiple-of-8) 1-D grid so every XCD owns a
    // contiguous run of tiles: neighbouring column tiles, which share cache
    // lines, then meet in the same L2.
    const int ntiles = (a.B + TB - 1) / TB;
    const int per = gridDim.x / 8;
    const int lin = (blockIdx.x % 8) * per + blockIdx.x / 8;
    const int64_t ai = lin / ntiles;
    if (ai >= a.A) return;
    const int b0 = (lin % ntiles) * TB;
    const cx<T>* tw = (const cx<T>*)a.tw;
    const int hin = a.nin / 2, hout = a.nout / 2;
    for (int jb = 0; jb < a.njobs; jb++) {
        const cx<TIN>* __restrict__ in = (const cx<TIN>*)a.job[jb].in + ai * a.in_a_stride;
        cx<TOUT>* __restrict__ out = (cx<TOUT>*)a.job[jb].out + ai * a.out_a_stride;
        const int deriv = a.job[jb].deriv;
        auto load = [&](int b, int p) -> cx<T> {
            const int bb = b0 + b;
            int i, k;
            if (a.nin == L) { i = p; k = (p <= L / 2) ? p : p - L; }
            else if (p <= hin) { i = p; k = p; }
            else if (p >= L - hin) { i = a.nin - (L - p); k = p - L; }
            else return cx<T>{(T)0, (T)0};
            if (bb >= a.B) return cx<T>{(T)0, (T)0};
            cx<TIN> v = in[(int64_t)i * a.in_row_stride + bb];
            if (deriv) return times_ik<T>((TIN)k, v);
            return cx<T>{(T)v.x, (T)v.y};
        };
        auto store = [&](int b, int p, cx<T> v) {
            const int bb = b0 + b;
            int i;
            if (a.nout == L || p <= hout) i = p;
            else if (p >= L - hout) i = a.nout - (L - p);
            else return;
            if (bb >= a.B) return;
            out[(int64_t)i * a.out_row_stride + bb] = cx<TOUT>{(TOUT)v.x, (TOUT)v.y};
        };
        Fft<T, L, TB, NT, TB, DIR>::run(lds, tw, load, store);
    }
}

// ---------------------------------------------------------------------------
// Strided batched pass with the next tile's input in flight (the default
// form of k_strided).  Persistent blocks walk work items (a, job, column
// tile); item order puts neighbouring column tiles of one row on the same
// XCD at the same time (they share cache lines).  Per item:
//   stage 0 reads the staged input tile [row][TB] (TIN) from LDS, with the
//   pruning and the i k derivative of k_strided; after its barrier the
//   block issues the asynchronous global->LDS copy (dma16) of the next
//   item's tile; the remaining stages run on fbuf; the last stage writes
//   fbuf; then each wave waits for its own copies and the block writes the
//   output tile from fbuf to global memory, so the wait never covers the
//   item's fresh stores.
// Requires the column tile rounded up to TB to stay inside the input
// buffers (StridedArgs::padded_ok, checked by the host).
// ---------------------------------------------------------------------------
template <int L, int DIR>
__host__ __device__ constexpr int sstage_rows()
{
    return DIR > 0 ? (2 * L) / 3 + 2 : L;  // pruned inverse input: N + 1 rows of M = 3N/2
}

template <int L, int TB>
__host__ __device__ constexpr int sstage_nt()
{
    int m = 0;
    for (int s = 0; s < Plan<L>::n; s++) {
        const int b = (L / Plan<L>::r(s)) * TB;
        m = b > m ? b : m;
    }
    m = (m + 63) / 64 * 64;
    return m > 1024 ? 1024 : (m < 64 ? 64 : m);
}

// NBM = 2: the transform holds 2 TB columns (the combined dual job below).
template <typename TIN, typename T, int L, int TB, int DIR, int NBM = 1>
__host__ __device__ constexpr size_t sstage_lds_bytes()
{
    return sizeof(cx<T>) * ((size_t)L * TB * NBM + fft_twiddle_span<L, sizeof(T) == 8 ? 1 : 2>()) +
           sizeof(cx<TIN>) * (size_t)sstage_rows<L, DIR>() * TB;
}

template <typename TIN, typename T, int L, int TB, int DIR, int NBM = 1>
__host__ __device__ constexpr bool sstage_fits()
{
    return sstage_lds_bytes<TIN, T, L, TB, DIR, NBM>() <= 160 * 1024 - 512 && (TB * sizeof(cx<TIN>)) % 16 == 0;
}

// STDP: inverse pass at the standard 3/2 padding (nin = 2 (L/3) + 1 stored
// rows) with L/R0 = 64 or 128 and one stage-0 butterfly per thread; the zone
// of each stage-0 input (stored low rows, stored high rows, zero padding) is
// then fixed at compile time.
// JSEL picks the jobs a launch takes: 0 all of them; 1 the dual jobs, each
// item as one transform of 2 TB columns (columns b < TB plain, b >= TB the
// i k transform of column b - TB, both formed from one staged tile), so the
// next item's copy is issued after the one shared stage 0 and stays in
// flight across both transforms; 2 the single jobs.
template <typename TIN, typename TOUT, typename T, int L, int TB, int NT, int DIR, bool STDP = false, int JSEL = 0>
__device__ __forceinline__ void sstage_body(const StridedArgs& a)
{
    constexpr int NB = JSEL == 1 ? 2 * TB : TB;  // columns of the transform (and LDS row pitch of fbuf)
    static_assert((TB & (TB - 1)) == 0, "column tile must be a power of two");
    constexpr int SROWS = sstage_rows<L, DIR>();
    constexpr int SPR = TB * (int)sizeof(cx<TIN>) / 16;  // 16-byte copy slots per staged row
    constexpr int EPS = 16 / (int)sizeof(cx<TIN>);        // input elements per slot
    constexpr int NW = NT / 64;
    constexpr int TWM = sizeof(T) == 8 ? 1 : 2;
    constexpr int TWN = fft_twiddle_span<L, TWM>();  // the table entries the transform reads
    __shared__ __align__(16) cx<T> fbuf[L * NB];
    __shared__ __align__(16) cx<TIN> stg[SROWS * TB];
    __shared__ __align__(16) cx<T> twl[TWN];
    if (a.skip && *a.skip) return;  // uniform: every thread reads the same flag
    const int t = threadIdx.x;
    const int hin = a.nin / 2, hout = a.nout / 2;
    const int ntiles = (a.B + TB - 1) / TB;
    // the launch's jobs as 4-bit job indices (at most 9 jobs)
    uint64_t jsel = 0;
    int nsel = 0;
    for (int j = 0; j < a.njobs; j++) {
        const bool d = a.job[j].out2 != nullptr;
        if (JSEL == 0 || (JSEL == 1) == d) {
            jsel |= (uint64_t)j << (4 * nsel);
            nsel++;
        }
    }
    const int64_t per_a = (int64_t)nsel * ntiles;
    const int64_t nitems = (int64_t)a.A * per_a;
    const int G = gridDim.x;  // a multiple of 8
    const int64_t lin = (int64_t)(blockIdx.x % 8) * (G / 8) + blockIdx.x / 8;

    for (int i = t; i < TWN; i += NT) twl[i] = ((const cx<T>*)a.tw)[i];
    __syncthreads();

    const int wave = __builtin_amdgcn_readfirstlane(t >> 6), lane = t & 63;
    const uint32_t stg_base = lds_addr(stg);
    const int nslot = a.nin * SPR;
    auto issue = [&](int64_t it) {
        const int64_t ai = it / per_a;
        const int r = (int)(it - ai * per_a);
        const int js = r / ntiles, b0 = (r - js * ntiles) * TB;
        const int jb = (int)((jsel >> (4 * js)) & 15);
        const cx<TIN>* in = (const cx<TIN>*)a.job[jb].in + ai * a.in_a_stride + b0;
        for (int m = wave; m * 64 < nslot; m += NW) {
            const int s = m * 64 + lane;
            const int row = s / SPR, part = s - row * SPR;
            if (s < nslot) dma16(in + (int64_t)row * a.in_row_stride + part * EPS, stg_base + (uint32_t)m * 1024u);
        }
    };

    int64_t it = lin;
    if (it < nitems) issue(it);
    dma_wait();
    lds_barrier();
    for (; it < nitems; it += G) {
        const int64_t ai = it / per_a;
        const int r = (int)(it - ai * per_a);
        const int js = r / ntiles, b0 = (r - js * ntiles) * TB;
        const int jb = (int)((jsel >> (4 * js)) & 15);
        const bool dual = JSEL == 1 || (JSEL == 0 && a.job[jb].out2 != nullptr);
        int deriv = dual ? 0 : a.job[jb].deriv;
        const int64_t nxt = it + G;
        auto load = [&](int b, int p) -> cx<T> {
            int i, k;
            if (a.nin == L) { i = p; k = (p <= L / 2) ? p : p - L; }
            else if (p <= hin) { i = p; k = p; }
            else if (p >= L - hin) { i = a.nin - (L - p); k = p - L; }
            else return cx<T>{(T)0, (T)0};
            const cx<TIN> v = stg[i * TB + (b & (TB - 1))];
            if (JSEL == 1 ? b >= TB : deriv) return times_ik<T>((TIN)k, v);
            return cx<T>{(T)v.x, (T)v.y};
        };
        auto load_std = [&](int b, int j, auto rc) -> cx<T> {
            constexpr int r = decltype(rc)::value;
            constexpr int LR0 = L / Plan<L>::r(0);
            constexpr int p0 = LR0 * r;
            constexpr int H = L / 3, NIN = 2 * H + 1;
            constexpr bool LO = p0 + LR0 - 1 <= H, HI = p0 >= L - H, ZERO = p0 > H && p0 + LR0 - 1 < L - H;
            if constexpr (ZERO) {
                return cx<T>{(T)0, (T)0};
            } else {
                const int p = p0 + j;
                bool lo = LO, hi = HI;
                if constexpr (!LO && !HI) {
                    lo = p <= H;
                    hi = p >= L - H;
                    if (!(lo || hi)) return cx<T>{(T)0, (T)0};
                }
                const int i = lo ? p : p - (L - NIN);
                const cx<TIN> v = stg[i * TB + (b & (TB - 1))];
                if (JSEL == 1 ? b >= TB : deriv) return times_ik<T>((TIN)(lo ? p : p - L), v);
                return cx<T>{(T)v.x, (T)v.y};
            }
        };
        auto store_lds = [&](int b, int p, cx<T> v) { fbuf[p * NB + b] = v; };
        auto prefetch = [&]() {
            if (nxt < nitems) issue(nxt);
        };
        // output tile from fbuf columns [0, TB) (and [TB, 2 TB) into dst2)
        auto write_tile = [&](void* dst, void* dst2) {
            cx<TOUT>* out = (cx<TOUT>*)dst + ai * a.out_a_stride + b0;
            cx<TOUT>* out2 = (JSEL == 1) ? (cx<TOUT>*)dst2 + ai * a.out_a_stride + b0 : out;
            for (int idx = t; idx < a.nout * NB; idx += NT) {
                const int i = idx / NB, bq = idx - i * NB, b = bq & (TB - 1);
                const int p = (a.nout == L || i <= hout) ? i : L - (a.nout - i);
                if (b0 + b < a.B) {
                    const cx<T> v = fbuf[p * NB + bq];
                    (bq >= TB ? out2 : out)[(int64_t)i * a.out_row_stride + b] = cx<TOUT>{(TOUT)v.x, (TOUT)v.y};
                }
            }
        };
        auto fft = [&](auto lead, auto& hook) {
            constexpr bool LEAD = decltype(lead)::value;
            if constexpr (STDP)
                FftS<T, L, NB, NT, NB, DIR, false, TWM>::template run<LEAD>(fbuf, twl, load_std, store_lds, hook, t);
            else
                FftS<T, L, NB, NT, NB, DIR, false, TWM>::template run<LEAD>(fbuf, twl, load, store_lds, hook, t);
        };
        if constexpr (JSEL == 1) {
            // both transforms of the dual job at once; the copy of the next
            // item's tile is in flight from the end of stage 0 on
            fft(std::false_type{}, prefetch);
        } else if (dual) {
            // plain transform first, from the staged tile; the tile stays
            // staged (no prefetch yet) for the i k transform below
            auto none = [] {};
            fft(std::false_type{}, none);
            lds_barrier();
            write_tile(a.job[jb].out, nullptr);
            deriv = 1;
            fft(std::true_type{}, prefetch);
        } else {
            fft(std::false_type{}, prefetch);
        }
        lds_barrier();
        dma_wait();  // the next item's tile (this wave's copies) has landed
        if constexpr (JSEL == 1) write_tile(a.job[jb].out, a.job[jb].out2);
        else write_tile(dual ? a.job[jb].out2 : a.job[jb].out, nullptr);
        lds_barrier();  // every wave's copies are in; fbuf is free again
    }
}

template <typename TIN, typename TOUT, typename T, int L, int TB, int NT, int DIR, bool STDP = false, int JSEL = 0>
__global__ __launch_bounds__(NT) void k_sstage(StridedArgs a)
{
    sstage_body<TIN, TOUT, T, L, TB, NT, DIR, STDP, JSEL>(a);
}

template <typename TIN, typename TOUT, typename T, int L, int TB, int NT, int DIR, bool STDP = false, int JSEL = 0>
__global__ __launch_bounds__(NT) void k_sstage_b(const StridedArgs* __restrict__ arr, const int* __restrict__ left)
{
    if (left[blockIdx.y] <= 0) return;
    sstage_body<TIN, TOUT, T, L, TB, NT, DIR, STDP, JSEL>(arr[blockIdx.y]);
}

// ---------------------------------------------------------------------------
// Block reductions
// ---------------------------------------------------------------------------
// Largest power of two <= n.
__host__ __device__ constexpr int pow2_floor(int n)
{
    int p = 1;
    while (2 * p <= n) p *= 2;
    return p;
}

// Fixed-order block reductions for any NT: the NT - P top values (P the
// largest power of two <= NT) are folded onto the first ones, then a tree.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* sh)
{
    constexpr int P = pow2_floor(NT);
    const int t = threadIdx.x;
    __syncthreads();
    sh[t] = v;
    __syncthreads();
    if constexpr (P < NT) {
        if (t < NT - P) sh[t] = sh[t] + sh[t + P];
        __syncthreads();
    }
#pragma unroll
    for (int s = P / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] = sh[t] + sh[t + s];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

template <int NT>
__device__ __forceinline__ double block_max(double v, double* sh)
{
    constexpr int P = pow2_floor(NT);
    const int t = threadIdx.x;
    __syncthreads();
    sh[t] = v;
    __syncthreads();
    if constexpr (P < NT) {
        if (t < NT - P) sh[t] = fmax(sh[t], sh[t + P]);
        __syncthreads();
    }
#pragma unroll
    for (int s = P / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] = fmax(sh[t], sh[t + s]);
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------------
// z pass: one (y,x) pencil per block iteration.
// Six paired c2r transforms of length L hold the twelve physical fields
//   (u,v) (w,ux) (uy,uz) (vx,vy) (vz,wx) (wy,wz)
// in LDS as [z][pair]; products are formed in place, then the forward
// transforms of (cx,cy) (cz,-) [DNS] or (cx,cy) (cz,t11) (t21,t22) (t31,t32)
// (t33,-) [LES] are unpacked to half spectra and written over the inputs.
// MODE 0: full step; MODE 1: velocity only (u,v,w -> CFL, optional export).
// ---------------------------------------------------------------------------
template <typename T, int L, int NT, int MODE>
__global__ __launch_bounds__(NT) void k_zpass(ZArgs a)
{
    constexpr int LD = 6;
    __shared__ cx<T> lds[L * LD];
    __shared__ double red[NT];
    const cx<T>* tw = (const cx<T>*)a.tw;
    const int nk = a.nk3;
    const int t = threadIdx.x;
    T mu = (T)0, mv = (T)0, mw = (T)0;
    double vmax_sa = 0.0, s2 = 0.0, s3 = 0.0;
    const double iAx = a.Ax, iAy = a.Ay, iAz = a.Az;
    const bool sa = a.sa;

    // Source of each logical field f (0 u,1 v,2 w | 3 ux,4 uy,5 uz | 6 vx,7 vy,8 vz
    // | 9 wx,10 wy,11 wz): X2 slot and whether it is d_z (i k3 times a
    // velocity spectrum).  Resolved once per block so the loads below are
    // branch-free and use no per-lane kernel-argument indexing.
    __shared__ const cx<T>* srcp[12];
    __shared__ int sderiv[12];
    if (t < 12) {
        const int f = t;
        const cx<T>* p = nullptr;
        int dz = 0;
        if (f < 3) p = (const cx<T>*)a.f[f];
        else {
            const int c = (f - 3) / 3, d = (f - 3) % 3;
            if (d == 0) p = (const cx<T>*)a.f[3 + c];
            else if (d == 1) p = (const cx<T>*)a.f[6 + c];
            else { p = (const cx<T>*)a.f[c]; dz = 1; }
        }
        if (MODE == 1 && f >= 3) p = nullptr;  // velocity mode: pairs (u,v), (w,0)
        srcp[f] = p;
        sderiv[f] = dz;
    }
    __syncthreads();

    for (int64_t pen = blockIdx.x; pen < a.npencil; pen += gridDim.x) {
        const int64_t off = pen * a.pitch;
        // -------- inverse: build paired Hermitian spectra and transform.
        // Z[p] = A[p] + i B[p] for 0 <= p < nk, conj(A[L-p]) + i conj(B[L-p])
        // for L-p < nk; only Re of the p = 0 terms enters (as in a c2r).
        constexpr int NPI = (MODE == 0) ? 6 : 2;
        auto load_inv = [&](int b, int p) -> cx<T> {
            const bool lo = p < nk, hi = p > L - nk;
            if (!(lo || hi)) return cx<T>{(T)0, (T)0};
            const int k = lo ? p : L - p;
            const T kk = (T)k;
            const int fa = 2 * b, fb = 2 * b + 1;
            const cx<T>* pa = srcp[fa];
            const cx<T>* pb = srcp[fb];
            cx<T> A = pa[off + k];
            if (sderiv[fa]) A = cx<T>{-kk * A.y, kk * A.x};
            cx<T> B = cx<T>{(T)0, (T)0};
            if (pb) {
                B = pb[off + k];
                if (sderiv[fb]) B = cx<T>{-kk * B.y, kk * B.x};
            }
            const T s = lo ? (T)1 : (T)-1;
            const T ay = (p == 0) ? (T)0 : s * A.y;
            const T by = (p == 0) ? (T)0 : s * B.y;
            return cx<T>{A.x - by, ay + B.x};
        };
        auto store_lds = [&](int b, int p, cx<T> v) { lds[p * LD + b] = v; };
        Fft<T, L, NPI, NT, LD, +1>::run(lds, tw, load_inv, store_lds);
        __syncthreads();

        // -------- physical space
        for (int n = t; n < L; n += NT) {
            cx<T>* row = &lds[n * LD];
            const T u = row[0].x, v = row[0].y, w = row[1].x;
            mu = fmax(mu, fabs(u));
            mv = fmax(mv, fabs(v));
            mw = fmax(mw, fabs(w));
            if (sa) vmax_sa = fmax(vmax_sa, fabs((double)u) / iAx + fabs((double)v) / iAy + fabs((double)w) / iAz);  // src/HIT_LES.cpp:377
            if constexpr (MODE == 1) {
                s2 += (double)u * (double)u + (double)v * (double)v + (double)w * (double)w;  // src/HIT_LES.cpp:1236-1239
                if (a.phys) {
                    // export physical velocity as [c][y][x][z]
                    const int64_t P = a.npencil * L;
                    T* o = (T*)a.phys;
                    o[pen * L + n] = u;
                    o[P + pen * L + n] = v;
                    o[2 * P + pen * L + n] = w;
                }
            } else {
                const T ux = row[1].y, uy = row[2].x, uz = row[2].y;
                const T vx = row[3].x, vy = row[3].y, vz = row[4].x;
                const T wx = row[4].y, wy = row[5].x, wz = row[5].y;
                s2 += (double)ux * (double)ux;
                s3 += (double)ux * (double)ux * (double)ux;
                // advective form (src/HIT_LES.cpp:118-120)
                const T cxv = u * ux + v * uy + w * uz;
                const T cyv = u * vx + v * vy + w * vz;
                const T czv = u * wx + v * wy + w * wz;
                if (a.les) {
                    // Smagorinsky (src/HIT_LES.cpp:130-154)
                    const T ct = (T)a.ct;
                    T ne = (T)2.0 * ux * ux;
                    ne += (T)2.0 * vy * vy;
                    ne += (T)2.0 * wz * wz;
                    ne += (uy + vx) * (uy + vx);
                    ne += (uz + wx) * (uz + wx);
                    ne += (vz + wy) * (vz + wy);
                    ne = ct * sqrt(ne);
                    const T t11 = (T)(-2.0) * ne * ux;
                    const T t21 = (T)(-1.0) * ne * (uy + vx);
                    const T t22 = (T)(-2.0) * ne * vy;
                    const T t31 = (T)(-1.0) * ne * (uz + wx);
                    const T t32 = (T)(-1.0) * ne * (vz + wy);
                    const T t33 = (T)(-2.0) * ne * wz;
                    row[0] = cx<T>{cxv, cyv};
                    row[1] = cx<T>{czv, t11};
                    row[2] = cx<T>{t21, t22};
                    row[3] = cx<T>{t31, t32};
                    row[4] = cx<T>{t33, (T)0};
                } else {
                    row[0] = cx<T>{cxv, cyv};
                    row[1] = cx<T>{czv, (T)0};
                }
            }
        }
        if constexpr (MODE == 0) {
            // -------- forward transforms of the products (in place in LDS)
            auto load_lds = [&](int b, int p) -> cx<T> { return lds[p * LD + b]; };
            if (a.les) Fft<T, L, 5, NT, LD, -1, true>::run(lds, tw, load_lds, store_lds);
            else Fft<T, L, 2, NT, LD, -1, true>::run(lds, tw, load_lds, store_lds);
            __syncthreads();
            // -------- unpack pairs to half spectra, keep k3 < nk
            const int npairs = a.les ? 5 : 2;
            const int nfields = a.les ? 9 : 3;
            for (int idx = t; idx < npairs * nk; idx += NT) {
                const int q = idx / nk, k = idx % nk;
                const cx<T> Z = lds[k * LD + q];
                const cx<T> Zc = lds[((L - k) % L) * LD + q];
                const T h = (T)0.5;
                const cx<T> A = cx<T>{h * (Z.x + Zc.x), h * (Z.y - Zc.y)};
                const cx<T> B = cx<T>{h * (Z.y + Zc.y), h * (Zc.x - Z.x)};
                const int fa = 2 * q, fb = 2 * q + 1;
                ((cx<T>*)a.f[fa])[off + k] = A;
                if (fb < nfields) ((cx<T>*)a.f[fb])[off + k] = B;
            }
        }
        __syncthreads();
    }
    // -------- block reductions -> partials / atomics
    const double bu = block_max<NT>((double)mu, red);
    const double bv = block_max<NT>((double)mv, red);
    const double bw = block_max<NT>((double)mw, red);
    const double vmax = fmax(bu / iAx, fmax(bv / iAy, bw / iAz));  // src/HIT_LES.cpp:393
    if (sa) vmax_sa = block_max<NT>(vmax_sa, red);
    if (MODE <= 1) {
        s2 = block_sum<NT>(s2, red);
        s3 = block_sum<NT>(s3, red);
    }
    if (t == 0) {
        atomicMax(a.maxvel_bits, (unsigned long long)__double_as_longlong(vmax));
        atomicMax(a.maxvel_sa_bits, (unsigned long long)__double_as_longlong(vmax_sa));
        if (MODE <= 1) {
            a.skew_partial[2 * blockIdx.x] = s2;
            a.skew_partial[2 * blockIdx.x + 1] = s3;
        }
    }
}


// ---------------------------------------------------------------------------
// Velocity-gradient invariants of g[3 i + j] = d u_i / d x_j, in the order
// QG QS RG RS V2 (src/tensor_algebra.cpp:250-303, calc_QO :356):
//   S = 0.5 G + 0.5 G^T, Omega = G - S, Q(T) = 0.5((tr T)^2 - tr(T T)),
//   R(T) = det T (calc_determinant's term order), V2 = 4(tr(S^2 Omega^2) - 2 QS QO),
//   QO = QG - QS.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void mul3(const T a[9], const T b[9], T ab[9])
{
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            T s = (T)0;
#pragma unroll
            for (int k = 0; k < 3; k++) s += a[3 * i + k] * b[3 * k + j];
            ab[3 * i + j] = s;
        }
}

template <typename T>
__device__ __forceinline__ T second_inv(const T a[9])
{
    T a2[9];
    mul3(a, a, a2);
    const T tr = a[0] + a[4] + a[8], tr2 = a2[0] + a2[4] + a2[8];
    return (T)0.5 * (tr * tr - tr2);
}

template <typename T>
__device__ __forceinline__ T det3(const T t[9])
{
    T d = (T)0;
    d += t[0] * t[4] * t[8];
    d += t[1] * t[5] * t[6];
    d += t[2] * t[3] * t[7];
    d -= t[2] * t[4] * t[6];
    d -= t[1] * t[3] * t[8];
    d -= t[0] * t[5] * t[7];
    return d;
}

template <typename T>
__device__ __forceinline__ void grad_invariants(const T g[9], T out[5])
{
    T S[9], O[9], S2[9], O2[9], S2O2[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) S[3 * i + j] = (T)0.5 * g[3 * i + j] + (T)0.5 * g[3 * j + i];
#pragma unroll
    for (int i = 0; i < 9; i++) O[i] = g[i] - S[i];
    const T QG = second_inv(g), QS = second_inv(S);
    mul3(S, S, S2);
    mul3(O, O, O2);
    mul3(S2, O2, S2O2);
    out[0] = QG;
    out[1] = QS;
    out[2] = det3(g);
    out[3] = det3(S);
    out[4] = (T)4.0 * ((S2O2[0] + S2O2[4] + S2O2[8]) - (T)2.0 * QS * (QG - QS));
}

// ---------------------------------------------------------------------------
// z pass with the next pencil's inputs in flight (the default form).
// Same arithmetic as k_zpass, restructured around HBM latency:
//   * the nine half-spectrum source fields of a pencil are staged into LDS
//     as [field][k] (pitch SP = pitch + pad) by asynchronous global->LDS
//     copies (dma16), issued for pencil n+1 as soon as pencil n's first
//     inverse stage has consumed the staging area, so the copy overlaps the
//     rest of pencil n's transforms and products;
//   * all barriers inside the loop are LDS-only (lds_barrier);
//   * twiddles are read from an LDS copy of the table (FftS);
//   * field pointers come from a small LDS table, so no per-lane indexing
//     of kernel arguments.
// Staging capacity is a compile-time maximum (zstage_spc); the launcher
// checks pitch + pad against it and the LDS total against the CU's 160 KB.
// ---------------------------------------------------------------------------
template <int L>
__host__ __device__ constexpr int zstage_spc()
{
    return (L / 3 + 1 + 15) / 16 * 16 + 2;
}

template <typename T>
__host__ __device__ constexpr int zstage_pad()
{
    return sizeof(T) == 8 ? 1 : 2;  // odd pitch (FP64) / even pitch for 16-byte pairs (FP32)
}

template <int MODE>
__host__ __device__ constexpr int zstage_ld()
{
    // LDS row pitch of the pair buffer, in complex elements: one more than the
    // pairs per point, so row-strided accesses (the physical-space loop, the
    // unpack) spread over all banks
    return MODE == 1 ? 3 : 7;
}

template <typename T, int L, int NT, int MODE, int LD = zstage_ld<MODE>()>
__host__ __device__ constexpr size_t zstage_lds_bytes()
{
    return sizeof(cx<T>) * ((size_t)L * LD + (size_t)(MODE == 1 ? 3 : 9) * zstage_spc<L>() + L) +
           sizeof(double) * NT + 2 * 9 * sizeof(void*);
}

template <typename T, int L, int NT, int MODE, int LD = zstage_ld<MODE>()>
__host__ __device__ constexpr bool zstage_fits()
{
    return zstage_lds_bytes<T, L, NT, MODE, LD>() <= 160 * 1024 - 512;
}

// ZOPAQUE: route the thread index through opaque() once per pencil, so the
// compiler recomputes the transforms' per-thread LDS addresses instead of
// holding them in registers across the pencil loop.  In FP32 that takes the
// kernel from 184 to 64-80 VGPRs and two blocks per CU (9.2-9.9 ms against
// 13.7 ms at 512^3); in FP64 the LDS allows one block per CU anyway and the
// recomputation only costs time, so it is off there.
// NKC > 0: the retained k3 count is the compile-time constant NKC (the
// standard 3/2 padding, nk3 = L/3 + 1), so the pruning of the first inverse
// stage resolves at compile time and its zero inputs fold away.
template <typename T, int L, int NT, int MODE, int LD = zstage_ld<MODE>(), int MINW = 1, bool ZOPAQUE = (sizeof(T) == 4),
          int NKC = 0, int TWM = (sizeof(T) == 8 ? 1 : 2)>
__device__ __forceinline__ void zstage_body(const ZArgs& a)
{
    static_assert(LD >= (MODE == 1 ? 2 : 6), "pair buffer pitch too small");
    constexpr int NSRC = (MODE == 1) ? 3 : 9;
    constexpr int SPC = zstage_spc<L>();
    constexpr int NW = NT / 64;
    constexpr int EPS = 16 / (int)sizeof(cx<T>);  // elements per 16-byte copy slot
    __shared__ __align__(16) cx<T> fbuf[L * LD];
    __shared__ __align__(16) cx<T> stg[NSRC * SPC];
    __shared__ __align__(16) cx<T> twl[L];
    __shared__ double red[NT];
    __shared__ const cx<T>* srcp[9];
    __shared__ cx<T>* dstp[9];
    const int t = threadIdx.x;
    const int nk = a.nk3, pitch = a.pitch;
    const int SP = pitch + zstage_pad<T>();
    const int nslot = NSRC * SP / EPS;
    T mu = (T)0, mv = (T)0, mw = (T)0;
    double vmax_sa = 0.0, s2 = 0.0, s3 = 0.0;
    const double iAx = a.Ax, iAy = a.Ay, iAz = a.Az;
    const bool sa = a.sa;

    if (t < 9) {
        srcp[t] = (const cx<T>*)a.f[t];
        dstp[t] = (cx<T>*)a.f[t];
    }
    for (int i = t; i < L; i += NT) twl[i] = ((const cx<T>*)a.tw)[i];
    __syncthreads();

    // copy slot s (16 bytes) holds staging elements s*EPS .. s*EPS+EPS-1,
    // element e = field f * SP + k; slots with k >= pitch (the pad) are skipped
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6), lane = t & 63;
    const uint32_t stg_base = lds_addr(stg);
    auto issue = [&](int64_t pen) {
        const int64_t off = pen * (int64_t)pitch;
        for (int m = wave; m * 64 < nslot; m += NW) {
            const int s = m * 64 + lane;
            const int e = s * EPS;
            const int f = e / SP, k = e - f * SP;
            if (s < nslot && k < pitch) dma16(srcp[f] + off + k, stg_base + (uint32_t)m * 1024u);
        }
    };

    // pair b of the inverse: logical fields (2b, 2b+1) of
    //   u v w ux uy uz vx vy vz wx wy wz
    // from source fields (u v w | ux vx wx | uy vy wy); d_z = i k3 * (u v w)
    //   A: u w uy vx vz wy  -> 0 2 6 4 1 8   (d_z for vz)
    //   B: v ux uz vy wx wz -> 1 3 0 7 5 2   (d_z for uz, wz)
    constexpr uint32_t SRC_A = 0x814620u, SRC_B = 0x257031u, DZ_A = 0x10u, DZ_B = 0x24u;

    int64_t pen = blockIdx.x;
    if (pen < a.npencil) issue(pen);
    dma_wait();
    lds_barrier();
    for (; pen < a.npencil; pen += gridDim.x) {
        const int64_t off = pen * (int64_t)pitch;
        const int64_t nxt = pen + gridDim.x;
        constexpr int NPI = (MODE == 1) ? 2 : 6;
        const int tl = ZOPAQUE ? opaque(t) : t;
        auto load_inv = [&](int b, int p) -> cx<T> {
            const int nkv = NKC > 0 ? NKC : nk;
            const bool lo = p < nkv, hi = p > L - nkv;
            if (!(lo || hi)) return cx<T>{(T)0, (T)0};
            const int k = lo ? p : L - p;
            const T kk = (T)k;
            const int fa = (SRC_A >> (4 * b)) & 15, fb = (SRC_B >> (4 * b)) & 15;
            cx<T> A = stg[fa * SP + k];
            if ((DZ_A >> b) & 1) A = cx<T>{-kk * A.y, kk * A.x};
            cx<T> B = cx<T>{(T)0, (T)0};
            if (MODE != 1 || b == 0) {
                B = stg[fb * SP + k];
                if ((DZ_B >> b) & 1) B = cx<T>{-kk * B.y, kk * B.x};
            }
            const T s = lo ? (T)1 : (T)-1;
            const T ay = (p == 0) ? (T)0 : s * A.y;
            const T by = (p == 0) ? (T)0 : s * B.y;
            return cx<T>{A.x - by, ay + B.x};
        };
        auto store_lds = [&](int b, int p, cx<T> v) { fbuf[p * LD + b] = v; };
        auto prefetch = [&]() {
            if (nxt < a.npencil) issue(nxt);
        };
        auto none = [&]() {};
        FftS<T, L, NPI, NT, LD, +1, false, TWM, sizeof(T) == 4>::template run<false>(fbuf, twl, load_inv, store_lds, prefetch, tl);
        lds_barrier();

        // -------- physical space
        for (int n = t; n < L; n += NT) {
            cx<T>* row = &fbuf[n * LD];
            if constexpr (MODE == 2) {
                // velocity-gradient invariants QG QS RG RS V2 (src/HIT_LES.cpp:1297-1319)
                const T g[9] = {row[1].y, row[2].x, row[2].y, row[3].x, row[3].y, row[4].x, row[4].y, row[5].x, row[5].y};
                T q[5];
                grad_invariants(g, q);
                row[0] = cx<T>{q[0], q[1]};
                row[1] = cx<T>{q[2], q[3]};
                row[2] = cx<T>{q[4], (T)0};
                continue;
            }
            const T u = row[0].x, v = row[0].y, w = row[1].x;
            mu = fmax(mu, fabs(u));
            mv = fmax(mv, fabs(v));
            mw = fmax(mw, fabs(w));
            if (sa) vmax_sa = fmax(vmax_sa, fabs((double)u) / iAx + fabs((double)v) / iAy + fabs((double)w) / iAz);  // src/HIT_LES.cpp:377
            if constexpr (MODE == 1) {
                s2 += (double)u * (double)u + (double)v * (double)v + (double)w * (double)w;  // src/HIT_LES.cpp:1236-1239
                if (a.phys) {
                    const int64_t P = a.npencil * L;
                    T* o = (T*)a.phys;
                    o[pen * L + n] = u;
                    o[P + pen * L + n] = v;
                    o[2 * P + pen * L + n] = w;
                }
            } else {
                const T ux = row[1].y, uy = row[2].x, uz = row[2].y;
                const T vx = row[3].x, vy = row[3].y, vz = row[4].x;
                const T wx = row[4].y, wy = row[5].x, wz = row[5].y;
                s2 += (double)ux * (double)ux;
                s3 += (double)ux * (double)ux * (double)ux;
                // advective form (src/HIT_LES.cpp:118-120)
                const T cxv = u * ux + v * uy + w * uz;
                const T cyv = u * vx + v * vy + w * vz;
                const T czv = u * wx + v * wy + w * wz;
                if (a.les) {
                    // Smagorinsky (src/HIT_LES.cpp:130-154)
                    const T ct = (T)a.ct;
                    T ne = (T)2.0 * ux * ux;
                    ne += (T)2.0 * vy * vy;
                    ne += (T)2.0 * wz * wz;
                    ne += (uy + vx) * (uy + vx);
                    ne += (uz + wx) * (uz + wx);
                    ne += (vz + wy) * (vz + wy);
                    ne = ct * sqrt(ne);
                    row[0] = cx<T>{cxv, cyv};
                    row[1] = cx<T>{czv, (T)(-2.0) * ne * ux};
                    row[2] = cx<T>{(T)(-1.0) * ne * (uy + vx), (T)(-2.0) * ne * vy};
                    row[3] = cx<T>{(T)(-1.0) * ne * (uz + wx), (T)(-1.0) * ne * (vz + wy)};
                    row[4] = cx<T>{(T)(-2.0) * ne * wz, (T)0};
                } else {
                    row[0] = cx<T>{cxv, cyv};
                    row[1] = cx<T>{czv, (T)0};
                }
            }
        }
        if constexpr (MODE == 2) {
            auto load_lds = [&](int b, int p) -> cx<T> { return fbuf[p * LD + b]; };
            FftS<T, L, 3, NT, LD, -1, true, TWM>::template run<true>(fbuf, twl, load_lds, store_lds, none, tl);
            lds_barrier();
            dma_wait();
            for (int idx = t; idx < 3 * nk; idx += NT) {
                const int q = idx / nk, k = idx - q * nk;
                const cx<T> Z = fbuf[k * LD + q];
                const cx<T> Zc = fbuf[((L - k) % L) * LD + q];
                const T h = (T)0.5;
                dstp[2 * q][off + k] = cx<T>{h * (Z.x + Zc.x), h * (Z.y - Zc.y)};
                if (q < 2) dstp[2 * q + 1][off + k] = cx<T>{h * (Z.y + Zc.y), h * (Zc.x - Z.x)};
            }
        } else if constexpr (MODE == 0) {
            auto load_lds = [&](int b, int p) -> cx<T> { return fbuf[p * LD + b]; };
            if (a.les) FftS<T, L, 5, NT, LD, -1, true, TWM>::template run<true>(fbuf, twl, load_lds, store_lds, none, tl);
            else FftS<T, L, 2, NT, LD, -1, true, TWM>::template run<true>(fbuf, twl, load_lds, store_lds, none, tl);
            lds_barrier();
            dma_wait();  // the next pencil's sources (this wave's copies) have landed
            const int npairs = a.les ? 5 : 2;
            const int nfields = a.les ? 9 : 3;
            for (int idx = t; idx < npairs * nk; idx += NT) {
                const int q = idx / nk, k = idx - q * nk;
                const cx<T> Z = fbuf[k * LD + q];
                const cx<T> Zc = fbuf[((L - k) % L) * LD + q];
                const T h = (T)0.5;
                const cx<T> A = cx<T>{h * (Z.x + Zc.x), h * (Z.y - Zc.y)};
                const cx<T> B = cx<T>{h * (Z.y + Zc.y), h * (Zc.x - Z.x)};
                dstp[2 * q][off + k] = A;
                if (2 * q + 1 < nfields) dstp[2 * q + 1][off + k] = B;
            }
        } else {
            dma_wait();
        }
        lds_barrier();  // every wave's copies are in; fbuf is free again
    }
    if constexpr (MODE == 2) return;  // a diagnostic pass: no CFL or skewness partials
    const double bu = block_max<NT>((double)mu, red);
    const double bv = block_max<NT>((double)mv, red);
    const double bw = block_max<NT>((double)mw, red);
    const double vmax = fmax(bu / iAx, fmax(bv / iAy, bw / iAz));  // src/HIT_LES.cpp:393
    if (sa) vmax_sa = block_max<NT>(vmax_sa, red);
    if (MODE <= 1) {
        s2 = block_sum<NT>(s2, red);
        s3 = block_sum<NT>(s3, red);
    }
    if (t == 0) {
        atomicMax(a.maxvel_bits, (unsigned long long)__double_as_longlong(vmax));
        atomicMax(a.maxvel_sa_bits, (unsigned long long)__double_as_longlong(vmax_sa));
        if (MODE <= 1) {
            a.skew_partial[2 * blockIdx.x] = s2;
            a.skew_partial[2 * blockIdx.x + 1] = s3;
        }
    }
}

template <typename T, int L, int NT, int MODE, int LD = zstage_ld<MODE>(), int MINW = 1, bool ZOPAQUE = (sizeof(T) == 4),
          int NKC = 0, int TWM = (sizeof(T) == 8 ? 1 : 2)>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(MINW))) void k_zstage(ZArgs a)
{
    zstage_body<T, L, NT, MODE, LD, MINW, ZOPAQUE, NKC, TWM>(a);
}

template <typename T, int L, int NT, int MODE, int LD = zstage_ld<MODE>(), int MINW = 1, bool ZOPAQUE = (sizeof(T) == 4),
          int NKC = 0, int TWM = (sizeof(T) == 8 ? 1 : 2)>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(MINW))) void k_zstage_b(const ZArgs* __restrict__ arr,
                                                                                           const int* __restrict__ left)
{
    if (left[blockIdx.y] <= 0) return;
    zstage_body<T, L, NT, MODE, LD, MINW, ZOPAQUE, NKC, TWM>(arr[blockIdx.y]);
}

// ---------------------------------------------------------------------------
// DNS z pass with the forward transforms of pencil n-1 folded into the
// inverse transforms of pencil n (the default DNS form).
// One Stockham transform of 8 columns per pencil step:
//   columns 0-5: the six paired c2r inputs of pencil n (from the staging
//                area, as in k_zstage);
//   columns 6-7: the products (cx, cy) and (cz, 0) of pencil n-1, written
//                by the previous step's physical-space loop.
// All columns run as inverse transforms; the forward ones through
// conj(inverse(conj x)), so one set of twiddles and butterflies serves both
// and every stage keeps all threads busy (768 = 64 x 12: 512 then 768
// butterflies at 768 threads).  Per step: the transform (the staging copy
// of pencil n+1 is issued after its first stage), then the unpack of
// pencil n-1's half spectra, then pencil n's products into columns 6-7.
// The loop runs one step past the block's last pencil to finish it.
// LDS row pitch 9 (odd), twiddle table trimmed to the entries read.  In FP32
// stage 0 maps threads column-major (one column per wave, so the loader runs
// without divergence: 8.96 -> 8.67 ms at 512^3); in FP64 the column-major
// stage-0 writes conflict in LDS and cost more than that saves.
// ---------------------------------------------------------------------------
// Pair buffer layout: the swizzled 8-column form of FftS where its index
// algebra applies and the compile-time stage-0 loader runs (L/R0 = 64 or 128:
// 768 and 1536), else pitch 9.
template <typename T, int L>
__host__ __device__ constexpr bool zfuse_swz()
{
    return fft_swz_ok<L, (int)sizeof(cx<T>)>() && (L / Plan<L>::r(0) == 64 || L / Plan<L>::r(0) == 128);
}

template <typename T, int L, int NT, int NKC, int TWM = (sizeof(T) == 8 ? 1 : 2)>
__host__ __device__ constexpr size_t zfuse_lds_bytes()
{
    return sizeof(cx<T>) * ((size_t)L * (zfuse_swz<T, L>() ? 8 : 9) + (size_t)9 * zstage_spc<L>() + fft_twiddle_span<L, TWM>()) +
           3 * 9 * sizeof(void*);
}

template <typename T, int L, int NT, int NKC = 0, int TWM = (sizeof(T) == 8 ? 1 : 2), bool ZOPAQUE = (sizeof(T) == 4)>
__device__ __forceinline__ void zfuse_body(const ZArgs& a)
{
    constexpr bool SWZ = zfuse_swz<T, L>();
    constexpr int LD = SWZ ? 8 : 9, NB = 8;
    constexpr int SPC = zstage_spc<L>();
    constexpr int NW = NT / 64;
    constexpr int EPS = 16 / (int)sizeof(cx<T>);
    constexpr int TWN = fft_twiddle_span<L, TWM>();
    static_assert(L * LD * sizeof(cx<T>) >= NT * sizeof(double), "reduction scratch aliases the pair buffer");
    __shared__ __align__(16) cx<T> fbuf[L * LD];
    __shared__ __align__(16) cx<T> stg[9 * SPC];
    __shared__ __align__(16) cx<T> twl[TWN];
    __shared__ const cx<T>* srcp[9];
    const int t = threadIdx.x;
    const int nk = a.nk3, pitch = a.pitch;
    const int SP = pitch + zstage_pad<T>();
    const int nslot = 9 * SP / EPS;
    T mu = (T)0, mv = (T)0, mw = (T)0;
    double vmax_sa = 0.0, s2 = 0.0, s3 = 0.0;
    const double iAx = a.Ax, iAy = a.Ay, iAz = a.Az;
    const bool sa = a.sa;
    // element (row, column) of the pair buffer
    auto fidx = [](int row, int col) -> int { return SWZ ? fft_swz<(int)sizeof(cx<T>)>(row, col) : row * LD + col; };

    if (t < 9) srcp[t] = (const cx<T>*)a.f[t];
    for (int i = t; i < TWN; i += NT) twl[i] = ((const cx<T>*)a.tw)[i];
    __syncthreads();

    const int wave = __builtin_amdgcn_readfirstlane(t >> 6), lane = t & 63;
    const uint32_t stg_base = lds_addr(stg);
    // Staging copies: slot m (16 bytes per lane) of wave m % NW.  Each lane's
    // source pointer (field f, k3 = k) is the same for every pencil up to the
    // pencil offset, so it is formed once here.
    constexpr int MAXM = (9 * (SPC + zstage_pad<T>()) / EPS + 64 * NW - 1) / (64 * NW);
    const cx<T>* gsrc[MAXM];
    bool gok[MAXM];
#pragma unroll
    for (int i = 0; i < MAXM; i++) {
        const int m = wave + i * NW;
        const int s = m * 64 + lane;
        const int e = s * EPS;
        const int f = e / SP, k = e - f * SP;
        gok[i] = s < nslot && k < pitch;
        gsrc[i] = srcp[gok[i] ? f : 0] + (gok[i] ? k : 0);
    }
    auto issue = [&](int64_t pen) {
        const int64_t off = pen * (int64_t)pitch;
#pragma unroll
        for (int i = 0; i < MAXM; i++) {
            const int m = wave + i * NW;
            if (m * 64 < nslot && gok[i]) dma16(gsrc[i] + off, stg_base + (uint32_t)m * 1024u);
        }
    };
    // inverse pair b (0..5): logical fields (2b, 2b+1) of u v w ux uy uz vx vy vz wx wy wz
    constexpr uint32_t SRC_A = 0x814620u, SRC_B = 0x257031u, DZ_A = 0x10u, DZ_B = 0x24u;
    // compile-time stage-0 loader: needs the compile-time k3 count, L/R0 = 64
    // or 128 and at most one stage-0 butterfly per thread
    constexpr int LR0 = L / Plan<L>::r(0);
    constexpr bool FAST = NKC > 0 && (LR0 == 64 || LR0 == 128) && NB * LR0 <= NT;
    constexpr bool S0COL = sizeof(T) == 4;  // column-major stage 0 in FP32 (its writes conflict in FP64)

    int64_t pen = blockIdx.x;
    if (pen < a.npencil) issue(pen);
    dma_wait();
    lds_barrier();
    int64_t prev = -1;  // pencil whose products sit in columns 6-7
    for (;; pen += gridDim.x) {
        const bool have = pen < a.npencil;
        const bool havep = prev >= 0;
        if (!have && !havep) break;
        const int tl = ZOPAQUE ? opaque(t) : t;
        const int64_t nxt = pen + gridDim.x;
        auto load = [&](int b, int p) -> cx<T> {
            if (b >= 6) {
                if (!havep) return cx<T>{(T)0, (T)0};
                const cx<T> v = fbuf[fidx(p, b)];
                return cx<T>{v.x, -v.y};  // forward = conj(inverse(conj x))
            }
            if (!have) return cx<T>{(T)0, (T)0};
            const int nkv = NKC > 0 ? NKC : nk;
            const bool lo = p < nkv, hi = p > L - nkv;
            if (!(lo || hi)) return cx<T>{(T)0, (T)0};
            const int k = lo ? p : L - p;
            const T kk = (T)k;
            const int fa = (SRC_A >> (4 * b)) & 15, fb = (SRC_B >> (4 * b)) & 15;
            cx<T> A = stg[fa * SP + k];
            if ((DZ_A >> b) & 1) A = cx<T>{-kk * A.y, kk * A.x};
            cx<T> B = stg[fb * SP + k];
            if ((DZ_B >> b) & 1) B = cx<T>{-kk * B.y, kk * B.x};
            const T s = lo ? (T)1 : (T)-1;
            const T ay = (p == 0) ? (T)0 : s * A.y;
            const T by = (p == 0) ? (T)0 : s * B.y;
            return cx<T>{A.x - by, ay + B.x};
        };
        // Fast loader.  With L/R0 = LR0 (64 or 128) and one stage-0 butterfly
        // per thread, input p = LR0 r + j with j < LR0, so the zone of p
        // (direct k3 = p, mirrored k3 = L - p, or zero padding) is fixed for
        // each r at compile time.  The thread's column is the wave index over
        // LR0 / 64 (a scalar) when stage 0 is column-major, else t mod 8.  Columns pair each velocity component
        // with its own z derivative, (f, i k f) -> (1 - k) f direct and
        // (1 + k) conj f mirrored (one read, no derivative product), then
        // (fx, fy) of u, v, w.
        const int col = S0COL ? wave / (LR0 / 64) : (tl & (NB - 1));
        const cx<T>* pa = stg + col * SP;        // f (columns 0-2) or fx (3-5)
        const cx<T>* pb = stg + (col + 3) * SP;  // fy (columns 3-5)
        auto load_fast = [&](int, int j, auto rc) -> cx<T> {
            constexpr int r = decltype(rc)::value;
            constexpr int p0 = LR0 * r;
            if (col >= 6) {
                if (!havep) return cx<T>{(T)0, (T)0};
                const cx<T> v = fbuf[fidx(p0 + j, col)];
                return cx<T>{v.x, -v.y};  // forward = conj(inverse(conj x))
            }
            if (!have) return cx<T>{(T)0, (T)0};
            constexpr bool LO = p0 + LR0 - 1 < NKC, HI = p0 > L - NKC, ZERO = p0 >= NKC && p0 + LR0 - 1 <= L - NKC;
            if constexpr (ZERO) {
                return cx<T>{(T)0, (T)0};
            } else {
                const int p = p0 + j;
                bool lo = LO, hi = HI;
                if constexpr (!LO && !HI) {
                    lo = p < NKC;
                    hi = p > L - NKC;
                    if (!(lo || hi)) return cx<T>{(T)0, (T)0};
                }
                const int k = lo ? p : L - p;
                const cx<T> A = pa[k];
                T ay = lo ? A.y : -A.y;
                if constexpr (r == 0) {
                    if (j == 0) ay = (T)0;  // the k3 = 0 imaginary part is dropped, as by a c2r
                }
                if (col < 3) {
                    const T sc = lo ? (T)1 - (T)k : (T)1 + (T)k;
                    return cx<T>{sc * A.x, sc * ay};
                }
                const cx<T> B = pb[k];
                T by = lo ? B.y : -B.y;
                if constexpr (r == 0) {
                    if (j == 0) by = (T)0;
                }
                return cx<T>{A.x - by, ay + B.x};
            }
        };
        auto store = [&](int b, int p, cx<T> v) { fbuf[p * LD + b] = (b >= 6) ? cx<T>{v.x, -v.y} : v; };
        auto store_swz = [&](int b, int, int li, cx<T> v) { fbuf[li] = (b >= 6) ? cx<T>{v.x, -v.y} : v; };
        auto prefetch = [&]() {
            if (have && nxt < a.npencil) issue(nxt);
        };
        if constexpr (SWZ) {
            static_assert(FAST, "the swizzled layout is only built with the compile-time loader");
            FftS<T, L, NB, NT, LD, +1, true, TWM, S0COL, true>::template run<false>(fbuf, twl, load_fast, store_swz, prefetch, tl);
        } else if constexpr (FAST) {
            FftS<T, L, NB, NT, LD, +1, true, TWM, S0COL>::template run<false>(fbuf, twl, load_fast, store, prefetch, tl);
        } else {
            FftS<T, L, NB, NT, LD, +1, true, TWM, S0COL>::template run<false>(fbuf, twl, load, store, prefetch, tl);
        }
        lds_barrier();
        dma_wait();  // pencil n+1's sources (this wave's copies) have landed
        if (havep) {
            // -------- unpack pencil n-1: (cx, cy) and (cz, -) half spectra
            const int64_t offp = prev * (int64_t)pitch;
            const int nkv = NKC > 0 ? NKC : nk;
            cx<T>* const d0 = (cx<T>*)a.f[0] + offp;
            cx<T>* const d1 = (cx<T>*)a.f[1] + offp;
            cx<T>* const d2 = (cx<T>*)a.f[2] + offp;
            for (int idx = t; idx < 2 * nkv; idx += NT) {
                const int q = idx >= nkv ? 1 : 0, k = idx - q * nkv;
                const cx<T> Z = fbuf[fidx(k, 6 + q)];
                const cx<T> Zc = fbuf[fidx(k == 0 ? 0 : L - k, 6 + q)];
                const T h = (T)0.5;
                (q ? d2 : d0)[k] = cx<T>{h * (Z.x + Zc.x), h * (Z.y - Zc.y)};
                if (q == 0) d1[k] = cx<T>{h * (Z.y + Zc.y), h * (Zc.x - Z.x)};
            }
            lds_barrier();
        }
        if (have) {
            // -------- physical space of pencil n: products into columns 6-7
            for (int n = t; n < L; n += NT) {
                // the row's 8 elements: 8n + c, XORed with the row's swizzle
                const int rb = SWZ ? 8 * n : n * LD, rh = SWZ ? fft_swz<(int)sizeof(cx<T>)>(n, 0) ^ (8 * n) : 0;
                auto row = [&](int c) -> cx<T>& { return fbuf[SWZ ? ((rb + c) ^ rh) : rb + c]; };
                T u, v, w, ux, uy, uz, vx, vy, vz, wx, wy, wz;
                if constexpr (FAST) {
                    u = row(0).x, uz = row(0).y, v = row(1).x, vz = row(1).y, w = row(2).x, wz = row(2).y;
                    ux = row(3).x, uy = row(3).y, vx = row(4).x, vy = row(4).y, wx = row(5).x, wy = row(5).y;
                } else {
                    u = row(0).x, v = row(0).y, w = row(1).x;
                    ux = row(1).y, uy = row(2).x, uz = row(2).y;
                    vx = row(3).x, vy = row(3).y, vz = row(4).x;
                    wx = row(4).y, wy = row(5).x, wz = row(5).y;
                }
                mu = fmax(mu, fabs(u));
                mv = fmax(mv, fabs(v));
                mw = fmax(mw, fabs(w));
                if (sa) vmax_sa = fmax(vmax_sa, fabs((double)u) / iAx + fabs((double)v) / iAy + fabs((double)w) / iAz);  // src/HIT_LES.cpp:377
                s2 += (double)ux * (double)ux;
                s3 += (double)ux * (double)ux * (double)ux;
                // advective form (src/HIT_LES.cpp:118-120)
                row(6) = cx<T>{u * ux + v * uy + w * uz, u * vx + v * vy + w * vz};
                row(7) = cx<T>{u * wx + v * wy + w * wz, (T)0};
            }
            lds_barrier();
        }
        prev = have ? pen : -1;
    }
    double* red = (double*)fbuf;
    const double bu = block_max<NT>((double)mu, red);
    const double bv = block_max<NT>((double)mv, red);
    const double bw = block_max<NT>((double)mw, red);
    const double vmax = fmax(bu / iAx, fmax(bv / iAy, bw / iAz));  // src/HIT_LES.cpp:393
    if (sa) vmax_sa = block_max<NT>(vmax_sa, red);
    s2 = block_sum<NT>(s2, red);
    s3 = block_sum<NT>(s3, red);
    if (t == 0) {
        atomicMax(a.maxvel_bits, (unsigned long long)__double_as_longlong(vmax));
        atomicMax(a.maxvel_sa_bits, (unsigned long long)__double_as_longlong(vmax_sa));
        a.skew_partial[2 * blockIdx.x] = s2;
        a.skew_partial[2 * blockIdx.x + 1] = s3;
    }
}

template <typename T, int L, int NT, int NKC = 0, int TWM = (sizeof(T) == 8 ? 1 : 2), bool ZOPAQUE = (sizeof(T) == 4)>
__global__ __launch_bounds__(NT) void k_zfuse(ZArgs a)
{
    zfuse_body<T, L, NT, NKC, TWM, ZOPAQUE>(a);
}

template <typename T, int L, int NT, int NKC = 0, int TWM = (sizeof(T) == 8 ? 1 : 2), bool ZOPAQUE = (sizeof(T) == 4)>
__global__ __launch_bounds__(NT) void k_zfuse_b(const ZArgs* __restrict__ arr, const int* __restrict__ left)
{
    if (left[blockIdx.y] <= 0) return;
    zfuse_body<T, L, NT, NKC, TWM, ZOPAQUE>(arr[blockIdx.y]);
}

// ---------------------------------------------------------------------------
// z pass, one wave per pencil (DNS and velocity-only modes, L = 64*E).
// Lane t owns the physical points z = t + 64 m; the twelve fields arrive pair
// by pair from WaveFft and the products u.grad(u) accumulate in registers in
// the reference's operand order ((u g1 + v g2) + w g3, src/HIT_LES.cpp:119).
// The forward transforms read their first stage straight from those
// registers; only the final unpack goes through the wave's LDS column.
// ---------------------------------------------------------------------------
template <typename T, int L>
__device__ __forceinline__ void wave_inv_pair(cx<T>* lds, const cx<T>* tw, const cx<T>* pa, int da, const cx<T>* pb, int db,
                                              int64_t off, int nk, T* outA, T* outB)
{
    auto load = [&](int p, int slot) -> cx<T> {
        (void)slot;
        const bool lo = p < nk, hi = p > L - nk;
        if (!(lo || hi)) return cx<T>{(T)0, (T)0};
        const int k = lo ? p : L - p;
        const T kk = (T)k;
        cx<T> A = pa[off + k];
        if (da) A = cx<T>{-kk * A.y, kk * A.x};
        cx<T> B = cx<T>{(T)0, (T)0};
        if (pb) {
            B = pb[off + k];
            if (db) B = cx<T>{-kk * B.y, kk * B.x};
        }
        const T sg = lo ? (T)1 : (T)-1;
        const T ay = (p == 0) ? (T)0 : sg * A.y;
        const T by = (p == 0) ? (T)0 : sg * B.y;
        return cx<T>{A.x - by, ay + B.x};
    };
    auto store = [&](int p, cx<T> v, int slot) {
        (void)p;
        outA[slot] = v.x;
        outB[slot] = v.y;
    };
    WaveFft<T, L, +1>::run(lds, tw, load, store);
}

template <typename T, int L>
__device__ __forceinline__ void wave_fwd_pair(cx<T>* lds, const cx<T>* tw, const T* inA, const T* inB, cx<T>* qa, cx<T>* qb,
                                              int64_t off, int nk)
{
    auto load = [&](int p, int slot) -> cx<T> {
        (void)p;
        return cx<T>{inA[slot], inB ? inB[slot] : (T)0};
    };
    auto store = [&](int p, cx<T> v, int slot) {
        (void)slot;
        lds[p] = v;
    };
    WaveFft<T, L, -1>::run(lds, tw, load, store);
    wave_sync();
    const int t = threadIdx.x & 63;
    for (int k = t; k < nk; k += 64) {
        const cx<T> Z = lds[k], Zc = lds[(L - k) % L];
        const T h = (T)0.5;
        qa[off + k] = cx<T>{h * (Z.x + Zc.x), h * (Z.y - Zc.y)};
        if (qb) qb[off + k] = cx<T>{h * (Z.y + Zc.y), h * (Zc.x - Z.x)};
    }
}

template <typename T, int L, int MODE>
__global__ __launch_bounds__(256) void k_zwave(ZArgs a)
{
    constexpr int E = L / 64, WPB = 4;
    __shared__ cx<T> ldsw[WPB][L];
    __shared__ double red[256];
    const int w = threadIdx.x >> 6, t = threadIdx.x & 63;
    cx<T>* lds = ldsw[w];
    const cx<T>* tw = (const cx<T>*)a.tw;
    const int nk = a.nk3;
    const cx<T>* F[9];
#pragma unroll
    for (int f = 0; f < 9; f++) F[f] = (const cx<T>*)a.f[f];
    // max over points of max(|u|/Ax, |v|/Ay, |w|/Az) equals the max of the
    // per-component maxima divided once (correctly rounded division by a
    // positive constant is monotonic); the sum form needs every point.
    T mu = (T)0, mv = (T)0, mw = (T)0;
    double vmax_sa = 0.0, s2 = 0.0, s3 = 0.0;
    const double iAx = a.Ax, iAy = a.Ay, iAz = a.Az;
    const bool sa = a.sa;

    for (int64_t pen = (int64_t)blockIdx.x * WPB + w; pen < a.npencil; pen += (int64_t)gridDim.x * WPB) {
        const int64_t off = pen * a.pitch;
        T U[E], V[E], W[E];
        wave_inv_pair<T, L>(lds, tw, F[0], 0, F[1], 0, off, nk, U, V);  // (u, v)
        if constexpr (MODE == 1) {
            T Z0[E];
            wave_inv_pair<T, L>(lds, tw, F[2], 0, nullptr, 0, off, nk, W, Z0);  // (w, -)
        }
        if constexpr (MODE == 1) {
#pragma unroll
            for (int m = 0; m < E; m++) {
                mu = fmax(mu, fabs(U[m]));
                mv = fmax(mv, fabs(V[m]));
                mw = fmax(mw, fabs(W[m]));
                s2 += (double)U[m] * (double)U[m] + (double)V[m] * (double)V[m] + (double)W[m] * (double)W[m];  // src/HIT_LES.cpp:1236-1239
            }
            if (sa) {
                for (int m = 0; m < E; m++)
                    vmax_sa = fmax(vmax_sa, fabs((double)U[m]) / iAx + fabs((double)V[m]) / iAy + fabs((double)W[m]) / iAz);
            }
            if (a.phys) {
                const int64_t P = a.npencil * L;
                T* o = (T*)a.phys;
#pragma unroll
                for (int m = 0; m < E; m++) {
                    const int64_t q = pen * L + t + 64 * m;
                    o[q] = U[m];
                    o[P + q] = V[m];
                    o[2 * P + q] = W[m];
                }
            }
        } else {
            T CX[E], CY[E], CZ[E], G1[E], G2[E];
            wave_inv_pair<T, L>(lds, tw, F[2], 0, F[3], 0, off, nk, W, G1);  // (w, ux)
            if (sa) {
                for (int m = 0; m < E; m++)
                    vmax_sa = fmax(vmax_sa, fabs((double)U[m]) / iAx + fabs((double)V[m]) / iAy + fabs((double)W[m]) / iAz);
            }
#pragma unroll
            for (int m = 0; m < E; m++) {
                mu = fmax(mu, fabs(U[m]));
                mv = fmax(mv, fabs(V[m]));
                mw = fmax(mw, fabs(W[m]));
                const double g = (double)G1[m];
                s2 += g * g;
                s3 += g * g * g;
                CX[m] = U[m] * G1[m];
            }
            wave_inv_pair<T, L>(lds, tw, F[6], 0, F[0], 1, off, nk, G1, G2);  // (uy, uz)
#pragma unroll
            for (int m = 0; m < E; m++) CX[m] = CX[m] + V[m] * G1[m] + W[m] * G2[m];
            wave_inv_pair<T, L>(lds, tw, F[4], 0, F[7], 0, off, nk, G1, G2);  // (vx, vy)
#pragma unroll
            for (int m = 0; m < E; m++) CY[m] = U[m] * G1[m] + V[m] * G2[m];
            wave_inv_pair<T, L>(lds, tw, F[1], 1, F[5], 0, off, nk, G1, G2);  // (vz, wx)
#pragma unroll
            for (int m = 0; m < E; m++) {
                CY[m] = CY[m] + W[m] * G1[m];
                CZ[m] = U[m] * G2[m];
            }
            wave_inv_pair<T, L>(lds, tw, F[8], 0, F[2], 1, off, nk, G1, G2);  // (wy, wz)
#pragma unroll
            for (int m = 0; m < E; m++) CZ[m] = CZ[m] + V[m] * G1[m] + W[m] * G2[m];
            wave_fwd_pair<T, L>(lds, tw, CX, CY, (cx<T>*)a.f[0], (cx<T>*)a.f[1], off, nk);
            wave_fwd_pair<T, L>(lds, tw, CZ, (const T*)nullptr, (cx<T>*)a.f[2], (cx<T>*)nullptr, off, nk);
        }
    }
    const double bu = block_max<256>((double)mu, red);
    const double bv = block_max<256>((double)mv, red);
    const double bw = block_max<256>((double)mw, red);
    const double vmax = fmax(bu / iAx, fmax(bv / iAy, bw / iAz));  // src/HIT_LES.cpp:393
    if (sa) vmax_sa = block_max<256>(vmax_sa, red);
    s2 = block_sum<256>(s2, red);
    s3 = block_sum<256>(s3, red);
    if (threadIdx.x == 0) {
        atomicMax(a.maxvel_bits, (unsigned long long)__double_as_longlong(vmax));
        atomicMax(a.maxvel_sa_bits, (unsigned long long)__double_as_longlong(vmax_sa));
        {
            a.skew_partial[2 * blockIdx.x] = s2;
            a.skew_partial[2 * blockIdx.x + 1] = s3;
        }
    }
}

// ---------------------------------------------------------------------------
// Plain z transforms for the FFT_package drop-in (reference layouts):
//   c2r: complex [y][x][nk] -> real [x][y][2nk] padded (src/util.h:75-80)
//   r2c: real [x][y][2nk] -> complex [y][x][nk]
// Two pencils per block iteration are paired into one complex transform.
// ---------------------------------------------------------------------------
template <typename T, int L, int NT, int DIR>
__global__ __launch_bounds__(NT) void k_zplain(ZPlainArgs a)
{
    __shared__ cx<T> lds[L];
    const cx<T>* tw = (const cx<T>*)a.tw;
    const int nk = L / 2 + 1;
    const int64_t np = (int64_t)a.Mx * a.My;
    for (int64_t pp = blockIdx.x; 2 * pp < np; pp += gridDim.x) {
        const int64_t p0 = 2 * pp, p1 = 2 * pp + 1;
        const bool has1 = p1 < np;
        // pencil index p -> (y = p / Mx, x = p % Mx)
        const int64_t y0 = p0 / a.Mx, x0 = p0 % a.Mx, y1 = p1 / a.Mx, x1 = p1 % a.Mx;
        if (DIR > 0) {
            const cx<T>* in = (const cx<T>*)a.in;
            auto ld = [&](int64_t y, int64_t x, int k) { return in[(y * a.Mx + x) * nk + k]; };
            auto load = [&](int b, int p) -> cx<T> {
                (void)b;
                cx<T> A, B = cx<T>{(T)0, (T)0};
                if (p == 0 || (p == L / 2)) {
                    A = ld(y0, x0, p);
                    if (has1) B = ld(y1, x1, p);
                    return cx<T>{A.x, B.x};
                }
                if (p < nk) {
                    A = ld(y0, x0, p);
                    if (has1) B = ld(y1, x1, p);
                    return cx<T>{A.x - B.y, A.y + B.x};
                }
                const int k = L - p;
                A = ld(y0, x0, k);
                if (has1) B = ld(y1, x1, k);
                return cx<T>{A.x + B.y, B.x - A.y};
            };
            T* out = (T*)a.out;
            auto store = [&](int b, int p, cx<T> v) {
                (void)b;
                out[(x0 * a.My + y0) * 2 * nk + p] = v.x;
                if (has1) out[(x1 * a.My + y1) * 2 * nk + p] = v.y;
            };
            Fft<T, L, 1, NT, 1, +1>::run(lds, tw, load, store);
        } else {
            const T* in = (const T*)a.in;
            auto load = [&](int b, int p) -> cx<T> {
                (void)b;
                T re = in[(x0 * a.My + y0) * 2 * nk + p];
                T im = has1 ? in[(x1 * a.My + y1) * 2 * nk + p] : (T)0;
                return cx<T>{re, im};
            };
            auto store = [&](int b, int p, cx<T> v) { (void)b; lds[p] = v; };
            Fft<T, L, 1, NT, 1, -1>::run(lds, tw, load, store);
            __syncthreads();
            cx<T>* out = (cx<T>*)a.out;
            for (int k = threadIdx.x; k < nk; k += NT) {
                const cx<T> Z = lds[k], Zc = lds[(L - k) % L];
                const T h = (T)0.5;
                out[(y0 * a.Mx + x0) * nk + k] = cx<T>{h * (Z.x + Zc.x), h * (Z.y - Zc.y)};
                if (has1) out[(y1 * a.Mx + x1) * nk + k] = cx<T>{h * (Z.y + Zc.y), h * (Zc.x - Z.x)};
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Spectral epilogue over the compact box (FP64 state).
// src/HIT_LES.cpp:205-242 (R), :243-278 (constant power), :295-339 (AB2 /
// kappa), :346-359 (projection); reductions :991-1003, :1137-1214.
// ---------------------------------------------------------------------------
// Step epilogue of one retained mode m = (k1, k2, k3) (src/HIT_LES.cpp:196-363):
// R1 from the normalised convective term C (and, LES, the SGS stress S:
// 11 21 22 31 32 33), viscosity, rotation, constant-power forcing; AB2 or the
// kappa scheme; projection; the state and history written back; then the
// shell-energy and enstrophy contributions of the new state.  With step
// false only the reductions of the current state.
struct EpiScalars {
    double dt, nu, alpha, kap, k05;
    bool first;
};

static __device__ __forceinline__ EpiScalars epi_scalars(const EpiArgs& a)
{
    EpiScalars q;
    q.dt = a.scal->dt;
    q.nu = a.scal->nu;
    q.alpha = a.scal->alpha;
    q.kap = a.scal->kappa;
    q.k05 = 1.0 / (q.kap + 0.5);
    q.first = a.scal->iter == 0;
    return q;
}

// RT: the storage type of the R history (the transform type of the form:
// the FP32-split form keeps R in FP32, whose convective part already carries
// the FP32 transforms' rounding; R is combined in FP64 either way).
template <typename RT>
static __device__ __forceinline__ void epi_mode(const EpiArgs& a, const EpiScalars& q, int64_t m, int k1, int k2, int k3,
                                         bool step, const cx<double>* C, const cx<double>* S, double* myh, double& om,
                                         double& ed, cx<double>* Nout = nullptr)
{
    const int ik2 = k1 * k1 + k2 * k2 + k3 * k3;
    const double rad2 = (double)ik2;
    const int K = shell_of(ik2);
    cx<double> U[3], R0[3], R1[3];
#pragma unroll
    for (int c = 0; c < 3; c++) U[c] = a.uk[c][m];
    if (step) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const cx<RT> r = ((const cx<RT>*)a.R[c])[m];
            R0[c] = cx<double>{(double)r.x, (double)r.y};
        }
        const double dt = q.dt, kap = q.kap, k05 = q.k05, alpha = q.alpha;
        const double ct = q.nu * rad2;
        const bool mean = !(rad2 > 0.0);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            R1[c].x = -(mean ? 0.0 : C[c].x) - ct * U[c].x;
            R1[c].y = -(mean ? 0.0 : C[c].y) - ct * U[c].y;
        }
        if (S) {
            // S: 11, 21, 22, 31, 32, 33 (src/HIT_LES.cpp:179-191)
            const double K1 = k1, K2 = k2, K3 = k3;
            const double dxr = -K1 * S[0].y - K2 * S[1].y - K3 * S[3].y;
            const double dxi = K1 * S[0].x + K2 * S[1].x + K3 * S[3].x;
            const double dyr = -K1 * S[1].y - K2 * S[2].y - K3 * S[4].y;
            const double dyi = K1 * S[1].x + K2 * S[2].x + K3 * S[4].x;
            const double dzr = -K1 * S[3].y - K2 * S[4].y - K3 * S[5].y;
            const double dzi = K1 * S[3].x + K2 * S[4].x + K3 * S[5].x;
            R1[0].x -= dxr; R1[0].y -= dxi;
            R1[1].x -= dyr; R1[1].y -= dyi;
            R1[2].x -= dzr; R1[2].y -= dzi;
        }
        if (a.rot) {
            const double ox = a.ox, oy = a.oy, oz = a.oz;
            R1[0].x += 2 * (oz * U[1].x - oy * U[2].x); R1[0].y += 2 * (oz * U[1].y - oy * U[2].y);
            R1[1].x += 2 * (ox * U[2].x - oz * U[0].x); R1[1].y += 2 * (ox * U[2].y - oz * U[0].y);
            R1[2].x += 2 * (oy * U[0].x - ox * U[1].x); R1[2].y += 2 * (oy * U[0].y - ox * U[1].y);
        }
        if (a.force_on) {
            if (a.kfmin <= K && K <= a.kfmax) {
#pragma unroll
                for (int c = 0; c < 3; c++) { R1[c].x += alpha * U[c].x; R1[c].y += alpha * U[c].y; }
            } else if (K == 0) {
#pragma unroll
                for (int c = 0; c < 3; c++) R1[c] = cx<double>{0.0, 0.0};
            }
        }
        cx<double> N[3];
        if (a.sa) {
            // self-adaptive kappa scheme; the oldest level at the first
            // step is the current field
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const cx<double> Up = a.ukprev[c][m];  // read unconditionally: a select of
                const cx<double> Ua = q.first ? U[c] : Up;  // pointers would push U to scratch
                a.ukprev[c][m] = U[c];
                N[c].x = k05 * (2.0 * kap * U[c].x + dt * ((1.0 + kap) * R1[c].x - kap * R0[c].x) - (kap - 0.5) * Ua.x);
                N[c].y = k05 * (2.0 * kap * U[c].y + dt * ((1.0 + kap) * R1[c].y - kap * R0[c].y) - (kap - 0.5) * Ua.y);
            }
        } else {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                if (a.keepprev) a.ukprev[c][m] = U[c];  // the start-of-step state
                N[c].x = U[c].x + dt * (1.5 * R1[c].x - 0.5 * R0[c].x);
                N[c].y = U[c].y + dt * (1.5 * R1[c].y - 0.5 * R0[c].y);
            }
        }
        {
            const double K1 = k1, K2 = k2, K3 = k3;
            const double ir2 = (rad2 > 0) ? 1.0 / rad2 : 0.0;
            const double axr = (K1 * N[0].x + K2 * N[1].x + K3 * N[2].x) * ir2;
            const double axi = (K1 * N[0].y + K2 * N[1].y + K3 * N[2].y) * ir2;
            N[0].x -= axr * K1; N[1].x -= axr * K2; N[2].x -= axr * K3;
            N[0].y -= axi * K1; N[1].y -= axi * K2; N[2].y -= axi * K3;
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            a.uk[c][m] = N[c];
            ((cx<RT>*)a.R[c])[m] = cx<RT>{(RT)R1[c].x, (RT)R1[c].y};
            U[c] = N[c];
            if (Nout) Nout[c] = N[c];
        }
    }
    // reductions on the (new) state
    const double ke = 0.5 * (U[0].x * U[0].x + U[0].y * U[0].y + U[1].x * U[1].x + U[1].y * U[1].y +
                             U[2].x * U[2].x + U[2].y * U[2].y);
    const bool conj = (k3 != 0) && !((a.Mz % 2 == 0) && (k3 == a.Mz / 2));
    atomicAdd(&myh[K], conj ? 2 * ke : ke);
    const double en = 2 * rad2 * ke;
    om += conj ? 2 * en : en;
    ed += conj ? 2 * ke : ke;
}

// Per-wave shell histograms [NW][nbins] summed in wave order into the
// block's partial, and the block's enstrophy partial.
template <int NT>
__device__ __forceinline__ void epi_flush(const EpiArgs& a, const double* hist, double om, double* red)
{
    constexpr int NW = NT / 64;
    const int nb = a.nbins;
    __syncthreads();
    for (int K = threadIdx.x; K < nb; K += NT) {
        double s = 0.0;
        for (int w = 0; w < NW; w++) s += hist[w * nb + K];
        a.hist_partial[(int64_t)blockIdx.x * nb + K] = s;
    }
    om = block_sum<NT>(om, red);
    if (threadIdx.x == 0) a.om_partial[blockIdx.x] = om;
}

template <typename T, int NT>
__device__ __forceinline__ void epilogue_body(const EpiArgs& a)
{
    constexpr int NW = NT / 64;
    extern __shared__ double hist[];  // NW * nbins
    __shared__ double red[NT];
    const int nb = a.nbins;
    for (int i = threadIdx.x; i < NW * nb; i += NT) hist[i] = 0.0;
    __syncthreads();
    double* myh = hist + (threadIdx.x / 64) * nb;
    const EpiScalars q = epi_scalars(a);
    double om = 0.0, ed = 0.0;
    const int nk1 = a.nk1, nk2 = a.nk2, nk3 = a.nk3, nk3p = a.nk3p;
    const double inv_nf = 1.0 / a.nf;
    const bool step = a.mode == EPI_STEP;
    auto mode = [&](int row, int i3) {
        const int i2 = row / nk1, i1 = row - i2 * nk1;
        const int64_t m = (int64_t)row * nk3p + i3;
        const int k1 = wrap_k(i1, nk1), k2 = wrap_k(i2, nk2), k3 = i3;
        if (a.notcubic && ((k1 % a.Lx) || (k2 % a.Ly) || (k3 % a.Lz))) return;  // src/HIT_LES.cpp:867-871
        cx<double> C[3], S[6];
        if (step) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const cx<T> cv = ((const cx<T>*)a.conv[c])[m];
                C[c] = cx<double>{(double)cv.x * inv_nf, (double)cv.y * inv_nf};  // Fourier_Normalization
            }
            if (a.les) {
#pragma unroll
                for (int s = 0; s < 6; s++) {
                    const cx<T> v = ((const cx<T>*)a.conv[3 + s])[m];
                    S[s] = cx<double>{(double)v.x * inv_nf, (double)v.y * inv_nf};
                }
            }
        }
        epi_mode<T>(a, q, m, k1, k2, k3, step, C, (step && a.les) ? S : nullptr, myh, om, ed);
    };
    // The first nfull = 64 floor(nk3 / 64) modes of each (i2, i1) row: one
    // wave per row, lanes along i3 (no per-mode index division, coalesced
    // row reads).  The rest of every row (nk3 mod 64 modes: 1 at the
    // 2^n-cubed grids, where a wave per row would run its last pass with one
    // lane) are spread over all threads of the grid, row-major.
    const int lane = threadIdx.x & 63;
    const int nrow = nk1 * nk2;
    const int nfull = nk3 & ~63, rem = nk3 - nfull;
    if (nfull > 0) {
        for (int row = blockIdx.x * NW + (int)(threadIdx.x >> 6); row < nrow; row += gridDim.x * NW)
            for (int i3 = lane; i3 < nfull; i3 += 64) mode(row, i3);
    }
    if (rem > 0) {
        // row = floor(f / rem) through FP64: (f + 1/2) / rem is never within
        // 1/(2 rem) of an integer, far above the product's rounding error
        const double inv_rem = 1.0 / rem;
        const int64_t ntail = (int64_t)nrow * rem;
        for (int64_t f = (int64_t)blockIdx.x * NT + threadIdx.x; f < ntail; f += (int64_t)gridDim.x * NT) {
            const int row = (int)(((double)f + 0.5) * inv_rem);
            mode(row, nfull + (int)(f - (int64_t)row * rem));
        }
    }
    epi_flush<NT>(a, hist, om, red);
    if (a.ed_partial) {
        // the direct sum of E (HIT::Check_PhysFour_Energy's EkFour, src/HIT_LES.cpp:1218-1230)
        ed = block_sum<NT>(ed, red);
        if (threadIdx.x == 0) a.ed_partial[blockIdx.x] = ed;
    }
}

template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_epilogue(EpiArgs a)
{
    epilogue_body<T, NT>(a);
}

template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_epilogue_b(const EpiArgs* __restrict__ arr, const int* __restrict__ left)
{
    if (left[blockIdx.y] <= 0) return;
    epilogue_body<T, NT>(arr[blockIdx.y]);
}

// ---------------------------------------------------------------------------
// Forward y pass of the three DNS convective fields fused with the step
// epilogue.  A persistent block walks items (k1 index, tile of TB k3
// columns); per item the three components run one after the other through
// the staged strided transform (input tile copied asynchronously, the next
// component's copy issued after stage 0), components 0 and 1 keep their
// retained outputs in registers, and after component 2 every retained mode
// of the item goes through epi_mode.  The convective spectra never reach
// HBM.  Shell histograms are per wave, as in k_epilogue.
// ---------------------------------------------------------------------------
// INV: after the epilogue of an item, the next step's inverse y pass of the
// item's new state (the same (k1, k3 tile) columns): component 0 goes into the
// staging tile (FP64, [2 (L/3) + 1][TB]) as the epilogue computes each mode;
// components 1 and 2 are staged back from the state this block has just
// written (through the cache hierarchy), each copy issued during the previous
// component's transforms (holding them in registers would spill in FP32).
// Each component is transformed twice, plain and i k2, into u_c and d/dy u_c
// (X1 slots c and 3 + c), with the same arithmetic as the strided inverse
// pass at the next step; the kernel sets *x1v, which lets that pass skip.  The item's X1 columns are the ones its
// forward transforms read, so the writes touch nothing another item still
// needs.  The staging tile grows to the FP64 state tile (at 512^3 FP32 the
// LDS is then 160,584 bytes with the 12 waves' histograms).
// Staging tile of k_yepi (elements of cx<T>): the forward pass's input
// tile, or with INV also a staged FP64 state tile [2 (L/3) + 1][TB].
template <typename T, int L, int TB, bool INV>
__host__ __device__ constexpr int yepi_stg_elems()
{
    constexpr size_t a = (size_t)L * TB, b = INV ? ((size_t)(2 * (L / 3) + 1) * TB * sizeof(cx<double>) + sizeof(cx<T>) - 1) / sizeof(cx<T>) : 0;
    return (int)(a > b ? a : b);
}

template <typename T, int L, int TB, int NT, bool INV = false>
__device__ __forceinline__ void yepi_body(const StridedArgs& a, const EpiArgs& e, const StridedArgs& ia, int* x1v)
{
    constexpr int SPR = TB * (int)sizeof(cx<T>) / 16;  // 16-byte copy slots per staged row
    constexpr int EPS = 16 / (int)sizeof(cx<T>);
    constexpr int NW = NT / 64;
    constexpr int TWM = sizeof(T) == 8 ? 1 : 2;
    constexpr int TWN = fft_twiddle_span<L, TWM>();
    constexpr int NOUTMAX = 2 * (L / 3) + 1;  // the launcher checks nout against it
    constexpr int Q = (NOUTMAX * TB + NT - 1) / NT;
    constexpr int STGN = yepi_stg_elems<T, L, TB, INV>();
    static_assert(L * TB * sizeof(cx<T>) >= NT * sizeof(double), "reduction scratch aliases fbuf");
    __shared__ __align__(16) cx<T> fbuf[L * TB];
    __shared__ __align__(16) cx<T> stg[STGN];
    __shared__ __align__(16) cx<T> twl[TWN];
    extern __shared__ double hist[];  // [NW][nbins]
    const int t = threadIdx.x;
    const int nb = e.nbins;
    if (INV && blockIdx.x == 0 && t == 0) *x1v = 1;  // read by the next step's inverse y pass only
    for (int i = t; i < NW * nb; i += NT) hist[i] = 0.0;
    for (int i = t; i < TWN; i += NT) twl[i] = ((const cx<T>*)a.tw)[i];
    __syncthreads();
    double* myh = hist + (t >> 6) * nb;
    const EpiScalars q = epi_scalars(e);
    double om = 0.0, ed_unused = 0.0;
    const double inv_nf = 1.0 / e.nf;
    const int nout = a.nout, hout = nout / 2;
    const int ntiles = (a.B + TB - 1) / TB;
    const int64_t nitems = (int64_t)a.A * ntiles;
    const int G = gridDim.x;  // a multiple of 8
    const int64_t lin = (int64_t)(blockIdx.x % 8) * (G / 8) + blockIdx.x / 8;

    const int wave = __builtin_amdgcn_readfirstlane(t >> 6), lane = t & 63;
    const uint32_t stg_base = lds_addr(stg);
    const int nslot = a.nin * SPR;
    auto issue = [&](int64_t it, int c) {
        const int64_t ai = it / ntiles;
        const int b0 = (int)(it - ai * ntiles) * TB;
        const cx<T>* in = (const cx<T>*)a.job[c].in + ai * a.in_a_stride + b0;
        for (int m = wave; m * 64 < nslot; m += NW) {
            const int s = m * 64 + lane;
            const int row = s / SPR, part = s - row * SPR;
            if (s < nslot) dma16(in + (int64_t)row * a.in_row_stride + part * EPS, stg_base + (uint32_t)m * 1024u);
        }
    };
    auto load = [&](int b, int p) -> cx<T> { return stg[p * TB + b]; };
    auto store_lds = [&](int b, int p, cx<T> v) { fbuf[p * TB + b] = v; };
    // INV: the FP64 state tile [nin][TB] of the component being transformed
    cx<double>* stgw = (cx<double>*)stg;
    auto issue_inv = [&](int64_t it, int c) {
        const int64_t ai = it / ntiles;
        const int b0 = (int)(it - ai * ntiles) * TB;
        const cx<double>* in = (const cx<double>*)ia.job[c].in + ai * ia.in_a_stride + b0;
        const int ns = ia.nin * TB;
        for (int m = wave; m * 64 < ns; m += NW) {
            const int s = m * 64 + lane;
            const int row = s / TB, col = s - row * TB;
            if (s < ns) dma16(in + (int64_t)row * ia.in_row_stride + col, stg_base + (uint32_t)m * 1024u);
        }
    };

    int64_t it = lin;
    if (it < nitems) issue(it, 0);
    dma_wait();
    lds_barrier();
    for (; it < nitems; it += G) {
        const int64_t ai = it / ntiles;
        const int b0 = (int)(it - ai * ntiles) * TB;
        const int i1 = (int)ai;
        const int k1 = wrap_k(i1, e.nk1);
        cx<T> keep[2][Q];
        auto comp = [&](auto cc) {
            constexpr int c = decltype(cc)::value;
            auto prefetch = [&]() {
                if (c < 2) issue(it, c + 1);
                else if (!INV && it + G < nitems) issue(it + G, 0);  // INV: the state parts come first
            };
            FftS<T, L, TB, NT, TB, -1, false, TWM>::template run<false>(fbuf, twl, load, store_lds, prefetch, t);
            lds_barrier();
            dma_wait();  // the next component's tile (this wave's copies) has landed
#pragma unroll
            for (int qq = 0; qq < Q; qq++) {
                const int idx = t + qq * NT;
                if (idx < nout * TB) {
                    const int i = idx / TB, b = idx - i * TB;
                    const int p = (i <= hout) ? i : L - (nout - i);
                    if constexpr (c < 2) {
                        keep[c][qq] = fbuf[p * TB + b];
                    } else {
                        const int k2 = wrap_k(i, e.nk2), k3 = b0 + b;
                        if (k3 < a.B && !(e.notcubic && ((k1 % e.Lx) || (k2 % e.Ly) || (k3 % e.Lz)))) {  // src/HIT_LES.cpp:867-871
                            const cx<T> v2 = fbuf[p * TB + b];
                            const cx<double> C[3] = {
                                cx<double>{(double)keep[0][qq].x * inv_nf, (double)keep[0][qq].y * inv_nf},
                                cx<double>{(double)keep[1][qq].x * inv_nf, (double)keep[1][qq].y * inv_nf},
                                cx<double>{(double)v2.x * inv_nf, (double)v2.y * inv_nf}};  // Fourier_Normalization
                            const int64_t m = ((int64_t)i * e.nk1 + i1) * e.nk3p + k3;
                            if constexpr (INV) {
                                cx<double> nn[3];
                                epi_mode<T>(e, q, m, k1, k2, k3, true, C, nullptr, myh, om, ed_unused, nn);
                                stgw[i * TB + b] = nn[0];
                            } else {
                                epi_mode<T>(e, q, m, k1, k2, k3, true, C, nullptr, myh, om, ed_unused);
                            }
                        } else if constexpr (INV) {
                            stgw[i * TB + b] = cx<double>{0.0, 0.0};  // padding columns: transformed, never written out
                        }
                    }
                }
            }
            lds_barrier();  // every wave's copies are in; fbuf is free again
        };
        static_for<3>(comp);
        if constexpr (INV) {
            // -------- the next step's inverse y pass of this item's new state
            // (component 0 is in the staging tile, published by the barrier
            // that ended the epilogue; the copies of components 1 and 2 read
            // the epilogue's stores, so every wave's stores complete first)
            dma_wait();
            lds_barrier();
            const int ihin = ia.nin / 2;
            // a runtime loop over the components, with the thread index through
            // an opaque copy: per-thread addresses are recomputed per component
            // rather than held across the item (register pressure)
            for (int c = 0; c < 3; c++) {
                const int tl = opaque(t);
                auto next = [&]() {
                    if (c < 2) issue_inv(it, c + 1);
                    else if (it + G < nitems) issue(it + G, 0);
                };
                auto none = [] {};
                int deriv = 0;  // 0 for the plain transform, 1 for the i k2 one
                // the inverse pass's loader (its input precision is FP64: i k u
                // formed in FP64, then rounded to T)
                auto load_inv = [&](int b, int p) -> cx<T> {
                    int i, k;
                    if (p <= ihin) { i = p; k = p; }
                    else if (p >= L - ihin) { i = ia.nin - (L - p); k = p - L; }
                    else return cx<T>{(T)0, (T)0};
                    const cx<double> v = stgw[i * TB + b];
                    if (deriv) return times_ik<T>((double)k, v);
                    return cx<T>{(T)v.x, (T)v.y};
                };
                auto write_inv = [&](void* dst) {
                    cx<T>* o = (cx<T>*)dst + ai * ia.out_a_stride + b0;
                    for (int idx = tl; idx < L * TB; idx += NT) {
                        const int i = idx / TB, b = idx - i * TB;
                        if (b0 + b < ia.B) o[(int64_t)i * ia.out_row_stride + b] = fbuf[i * TB + b];
                    }
                };
                FftS<T, L, TB, NT, TB, +1, false, TWM>::template run<false>(fbuf, twl, load_inv, store_lds, none, tl);
                lds_barrier();
                write_inv(ia.job[c].out);
                deriv = 1;
                FftS<T, L, TB, NT, TB, +1, false, TWM>::template run<true>(fbuf, twl, load_inv, store_lds, next, tl);
                lds_barrier();
                dma_wait();  // the next component's (or item's) tile has landed
                write_inv(ia.job[c].out2);
                lds_barrier();
            }
        }
    }
    epi_flush<NT>(e, hist, om, (double*)fbuf);
}

template <typename T, int L, int TB, int NT>
__global__ __launch_bounds__(NT) void k_yepi(StridedArgs a, EpiArgs e)
{
    yepi_body<T, L, TB, NT>(a, e, a, nullptr);
}

template <typename T, int L, int TB, int NT>
__global__ __launch_bounds__(NT) void k_yepi_inv(StridedArgs a, EpiArgs e, StridedArgs ia, int* x1v)
{
    yepi_body<T, L, TB, NT, true>(a, e, ia, x1v);
}

// batched k_yepi_inv: slot = blockIdx.y; each slot's flag is its inverse
// pass struct's setflag
template <typename T, int L, int TB, int NT>
__global__ __launch_bounds__(NT) void k_yepi_inv_b(const StridedArgs* __restrict__ arr, const EpiArgs* __restrict__ earr,
                                                   const StridedArgs* __restrict__ iarr, const int* __restrict__ left)
{
    if (left[blockIdx.y] <= 0) return;
    const StridedArgs& ia = iarr[blockIdx.y];
    yepi_body<T, L, TB, NT, true>(arr[blockIdx.y], earr[blockIdx.y], ia, ia.setflag);
}

template <typename T, int L, int TB, int NT>
__global__ __launch_bounds__(NT) void k_yepi_b(const StridedArgs* __restrict__ arr, const EpiArgs* __restrict__ earr,
                                               const int* __restrict__ left)
{
    if (left[blockIdx.y] <= 0) return;
    yepi_body<T, L, TB, NT>(arr[blockIdx.y], earr[blockIdx.y], arr[blockIdx.y], nullptr);
}

// ---------------------------------------------------------------------------
// Self-adaptive time step helpers (src/HIT_LES.h:249-269, src/HIT_LES.cpp:1049-1078)
// ---------------------------------------------------------------------------
static __device__ double FuL(double x, double x0, double x1, double f0, double f1) { return f0 + (x - x0) * (f1 - f0) / (x1 - x0); }
static __device__ double FuQ(double x, double x0, double x1) { return (x - x0) * (x - x1); }
static __device__ double FuG(double x, double a, double b, double c, double x0, double x1, double f0, double f1)
{
    return (a * x * x + b * x + c) * FuQ(x, x0, x1) + FuL(x, x0, x1, f0, f1);
}

// ---------------------------------------------------------------------------
// Finalize: fixed-order sums of the partials; time step (src/HIT_LES.cpp:365-403)
// ---------------------------------------------------------------------------
template <int NT>
__device__ __forceinline__ void finalize_body(const FinArgs& a)
{
    __shared__ double red[NT];
    DevScalars* s = a.scal;
    const int t = threadIdx.x;
    if (a.flags & FIN_REDUCE) {
        // E(K): bins are split over NT/QB thread groups; each thread sums a
        // fixed quarter of the block partials with 4 independent
        // accumulators, and the quarters are combined in a fixed order.
        constexpr int QB = 4;
        __shared__ double q[QB][NT / QB];
        const int lane = t % (NT / QB), quarter = t / (NT / QB);
        const int per = (a.nblk_epi + QB - 1) / QB;
        const int b0 = quarter * per, b1 = min(a.nblk_epi, b0 + per);
        for (int K0 = 0; K0 < a.nbins; K0 += NT / QB) {
            const int K = K0 + lane;
            double acc[4] = {0.0, 0.0, 0.0, 0.0};
            if (K < a.nbins) {
                int b = b0;
                for (; b + 3 < b1; b += 4) {
#pragma unroll
                    for (int u = 0; u < 4; u++) acc[u] += a.hist_partial[(int64_t)(b + u) * a.nbins + K];
                }
                for (; b < b1; b++) acc[0] += a.hist_partial[(int64_t)b * a.nbins + K];
            }
            q[quarter][lane] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
            __syncthreads();
            if (quarter == 0 && K < a.nbins) a.ek[K] = ((q[0][lane] + q[1][lane]) + q[2][lane]) + q[3][lane];
            __syncthreads();
        }
        double om = 0.0;
        for (int b = t; b < a.nblk_epi; b += NT) om += a.om_partial[b];
        om = block_sum<NT>(om, red);
        __syncthreads();
        if (t == 0) {
            double e = 0.0;
            for (int K = 0; K < a.nbins; K++) e += a.ek[K];
            s->ek = e;
            s->om = om;
            s->relambda = e / s->nu * sqrt(20.0 / 3.0 / om);
            // constant-power coefficient for the next step (src/HIT_LES.cpp:251-256)
            double ef = 0.0;
            for (int K = a.kfmin; K <= a.kfmax && K < a.nbins; K++) ef += a.ek[K];
            s->alpha = (ef > 1e-30) ? a.force_eps / (2.0 * ef) : 0.0;
        }
    }
    if (a.flags & FIN_SKEW) {
        double s2[4] = {0.0, 0.0, 0.0, 0.0}, s3[4] = {0.0, 0.0, 0.0, 0.0};
        int b = t;
        for (; b + 3 * NT < a.nblk_z; b += 4 * NT) {
#pragma unroll
            for (int u = 0; u < 4; u++) { s2[u] += a.skew_partial[2 * (b + u * NT)]; s3[u] += a.skew_partial[2 * (b + u * NT) + 1]; }
        }
        for (; b < a.nblk_z; b += NT) { s2[0] += a.skew_partial[2 * b]; s3[0] += a.skew_partial[2 * b + 1]; }
        double S2 = block_sum<NT>((s2[0] + s2[1]) + (s2[2] + s2[3]), red);
        double S3 = block_sum<NT>((s3[0] + s3[1]) + (s3[2] + s3[3]), red);
        if (t == 0) { s->skew2 = S2; s->skew3 = S3; }
    }
    if (a.flags & FIN_EFOUR) {
        double e = 0.0;
        for (int b = t; b < a.nblk_epi; b += NT) e += a.ed_partial[b];
        e = block_sum<NT>(e, red);
        if (t == 0) s->ek_four = e;
    }
    if (a.flags & FIN_EPHYS) {
        double e = 0.0;
        for (int b = t; b < a.nblk_z; b += NT) e += a.skew_partial[2 * b];
        e = block_sum<NT>(e, red);
        if (t == 0) s->ek_phys = e;
    }
    if (t == 0) {
        if (a.flags & FIN_MAXVEL) {
            s->maxvel = __longlong_as_double((long long)*a.maxvel_bits);
            s->maxvel_sa = __longlong_as_double((long long)*a.maxvel_sa_bits);
        }
        if (a.flags & FIN_DT) {
            const double nu = s->nu;
            if (a.sa) {
                const double MaxVel = s->maxvel_sa, Ah = a.Ah;
                const double phi0 = atan(MaxVel * Ah * Ah / (0.35 * nu * 12.0));
                double Topt, Kopt;
                const double PI = 3.141592653589793238462643;
                if (phi0 >= 0.0 && phi0 < a.phi[0]) Topt = FuG(phi0, 0.0, a.c[0], a.c[1], 0.0, a.phi[0], 4.0 / 3.0, a.t1);
                else Topt = FuG(phi0, a.c[2], a.c[3], a.c[4], a.phi[0], PI / 2.0, a.t1, 1.0);
                if (phi0 >= 0.0 && phi0 <= a.phi[0]) Kopt = 1.0;
                else if (phi0 > a.phi[0] && phi0 <= a.phi[1]) Kopt = FuG(phi0, a.c[5], a.c[6], a.c[7], a.phi[0], a.phi[1], 1.0, a.kk[0]);
                else if (phi0 > a.phi[1] && phi0 <= a.phi[2]) Kopt = FuG(phi0, a.c[8], a.c[9], a.c[10], a.phi[1], a.phi[2], a.kk[0], a.kk[1]);
                else Kopt = FuG(phi0, a.c[11], a.c[12], a.c[13], a.phi[2], PI / 2.0, a.kk[1], 0.0);
                double At = Topt / sqrt((MaxVel * MaxVel) / (0.35 * 0.35) + (144.0 * nu * nu / (Ah * Ah * Ah * Ah)));
                s->kappa = Kopt;
                s->dt = a.C_At * At;
            } else {
                const double ax2 = a.Ax * a.Ax, ay2 = a.Ay * a.Ay, az2 = a.Az * a.Az;
                const double C_visc = 0.2 * fmin(ax2, fmin(ay2, az2)) / nu;
                const double C_conv = 0.35 / s->maxvel;
                double At = fmin(C_visc, C_conv);
                s->dt = a.C_At * At;
                if (a.fixed_dt > 0.0) s->dt = a.fixed_dt;
            }
        }
        if (a.flags & FIN_ADVANCE) {
            s->time = s->time + s->dt;
            s->iter = s->iter + 1;
        }
    }
}

template <int NT>
__global__ __launch_bounds__(NT) void k_finalize(FinArgs a)
{
    finalize_body<NT>(a);
}

// Batched form: the slot's step ends here, so its steps-left counter drops
// by one (after every wave of the block has read it at entry).
template <int NT>
__global__ __launch_bounds__(NT) void k_finalize_b(const FinArgs* __restrict__ arr, int* __restrict__ left)
{
    if (left[blockIdx.y] <= 0) return;
    finalize_body<NT>(arr[blockIdx.y]);
    __syncthreads();
    if (threadIdx.x == 0) left[blockIdx.y] -= 1;
}

static __global__ void k_clear(unsigned long long* a, unsigned long long* b)
{
    if (threadIdx.x == 0) { *a = 0ull; *b = 0ull; }
}

static __global__ void k_clear_b(const ClearArgs* __restrict__ arr, const int* __restrict__ left)
{
    if (left[blockIdx.y] <= 0) return;
    if (threadIdx.x == 0) { *arr[blockIdx.y].a = 0ull; *arr[blockIdx.y].b = 0ull; }
}

// Streaming copy (the measured practical HBM ceiling beside the data-sheet
// peak): each thread moves U 16-byte vectors per iteration, all loads
// issued before the stores, non-temporal so the copy does not keep lines in
// the caches; XCD-interleaved block order is irrelevant for a linear sweep.
typedef float vf4 __attribute__((ext_vector_type(4)));
template <int U>
__global__ __launch_bounds__(256) void k_stream_copy(const vf4* __restrict__ in, vf4* __restrict__ out, int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * 256 * U;
    for (int64_t base = (int64_t)blockIdx.x * 256 * U + threadIdx.x; base < n; base += stride) {
        vf4 v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t i = base + (int64_t)u * 256;
            if (i < n) v[u] = __builtin_nontemporal_load(in + i);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t i = base + (int64_t)u * 256;
            if (i < n) __builtin_nontemporal_store(v[u], out + i);
        }
    }
}

// Steps-left counters of a batch for its next chunk, from the kernel's
// arguments (so the host may reuse its copy right after the launch).
static __global__ void k_set_left(int* left, SlotSteps s, int nslots)
{
    if ((int)threadIdx.x < nslots) left[threadIdx.x] = s.n[threadIdx.x];
}

// Spectrum imposition (src/HIT_LES.cpp:410-444): FT[K] from E(K), then scale.
static __global__ void k_forcing_ft(const double* ek, const double* forced, double* ft, int last_eff)
{
    for (int K = threadIdx.x; K <= last_eff; K += blockDim.x)
        ft[K] = ((ek[K] > 0.0) ? sqrt(forced[K] / ek[K]) : 0.0);
}

static __global__ void k_scale_shells(ShellArgs a)
{
    const int nk1 = a.nk1, nk2 = a.nk2, nk3 = a.nk3, nk3p = a.nk3p;
    for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m < a.nmodes; m += (int64_t)gridDim.x * blockDim.x) {
        const int i3 = (int)(m % nk3p);
        if (i3 >= nk3) continue;  // pitch padding
        const int64_t r12 = m / nk3p;
        const int i1 = (int)(r12 % nk1), i2 = (int)(r12 / nk1);
        const int k1 = wrap_k(i1, nk1), k2 = wrap_k(i2, nk2), k3 = i3;
        if (a.notcubic && ((k1 % a.Lx) || (k2 % a.Ly) || (k3 % a.Lz))) continue;
        const int K = shell_of(k1 * k1 + k2 * k2 + k3 * k3);
        for (int c = 0; c < 3; c++) {
            cx<double> v = a.uk[c][m];
            if (K <= a.last_eff) { v.x *= a.ft[K]; v.y *= a.ft[K]; }
            else if (a.nullify) v = cx<double>{0.0, 0.0};
            a.uk[c][m] = v;
        }
    }
}

// Complex conjugate correction of the k3 = 0 plane (src/HIT_LES.cpp:446-476)
static __global__ void k_ccc(ShellArgs a)
{
    const int nk1 = a.nk1, nk2 = a.nk2;
    const int n = nk1 * nk2;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const int i1 = q % nk1, i2 = q / nk1;
        const int k1 = wrap_k(i1, nk1), k2 = wrap_k(i2, nk2);
        if (k1 < 0 || (k1 == 0 && k2 < 0)) {
            const int j1 = (-k1 >= 0) ? -k1 : nk1 - k1;
            const int j2 = (-k2 >= 0) ? -k2 : nk2 - k2;
            const int64_t m = ((int64_t)i2 * nk1 + i1) * a.nk3p;
            const int64_t mc = ((int64_t)j2 * nk1 + j1) * a.nk3p;
            for (int c = 0; c < 3; c++) {
                cx<double> v = a.uk[c][mc];
                a.uk[c][m] = cx<double>{v.x, -v.y};
            }
        }
    }
}

// Normalisation of a reference-layout complex field on the retained set
// (src/FFTW_package.cpp:33-42).
template <typename T>
__global__ void k_normalize(NormArgs a)
{
    const int nk = a.Mz / 2 + 1;
    const int64_t n = (int64_t)a.My * a.Mx * nk;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int k3 = (int)(q % nk);
        const int64_t r = q / nk;
        const int ax = (int)(r % a.Mx), by = (int)(r / a.Mx);
        const int k1 = (ax <= a.Mx / 2) ? ax : ax - a.Mx, k2 = (by <= a.My / 2) ? by : by - a.My;
        bool rel = (abs(k1) <= a.Nx2 && abs(k2) <= a.Ny2 && k3 <= a.Nz2);
        if (a.notcubic && ((k1 % a.Lx) || (k2 % a.Ly) || (k3 % a.Lz))) rel = false;
        if (a.all || rel) {
            cx<T>* p = (cx<T>*)a.data;
            p[q].x = p[q].x / (T)a.nf;
            p[q].y = p[q].y / (T)a.nf;
        }
    }
}

// ---------------------------------------------------------------------------
// Launch tables
// ---------------------------------------------------------------------------
template <typename T, int L>
constexpr int strided_tb()
{
    // largest column tile with the FFT tile in <= 48 KB of LDS
    constexpr int bytes = (int)sizeof(cx<T>) * L;
    return (bytes * 16 <= 49152) ? 16 : (bytes * 8 <= 49152) ? 8 : (bytes * 4 <= 49152) ? 4 : (bytes * 2 <= 49152) ? 2 : 1;
}

template <int L>
constexpr int zpass_nt()
{
    return L >= 768 ? 512 : (L >= 288 ? 256 : (L >= 144 ? 128 : 64));
}

// One launch of the staged kernel over the jobs JSEL selects (njsel of them).
template <typename TIN, typename TOUT, typename T, int L, int TB, int DIR, int JSEL>
static hipError_t launch_sstage(const StridedArgs& b, int A, int njsel, hipStream_t st, const BatchLaunch* bl)
{
    constexpr int NB = JSEL == 1 ? 2 * TB : TB;
    constexpr int NT = sstage_nt<L, NB>();
    constexpr int LR0 = L / Plan<L>::r(0);
    constexpr bool STDP_OK = DIR > 0 && (LR0 == 64 || LR0 == 128) && NB * LR0 <= NT;
    const bool stdp = STDP_OK && b.nin == 2 * (L / 3) + 1;
    static const int occ = [] {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_sstage<TIN, TOUT, T, L, TB, NT, DIR, false, JSEL>, NT, 0) !=
            hipSuccess)
            n = 1;
        return n > 0 ? n : 1;
    }();
    int dev = 0, ncu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    const int64_t items = (int64_t)A * njsel * ((b.B + TB - 1) / TB);
    int64_t g = (int64_t)ncu * occ;
    if (items < g) g = items;
    g = (g + 7) / 8 * 8;
    if (bl) {
        // batched: each slot's struct (with A set) is in bl->args
        if constexpr (batch_length(L)) {
            const dim3 grid((unsigned)g, (unsigned)bl->nslots);
            const StridedArgs* arr = (const StridedArgs*)bl->args;
            if constexpr (STDP_OK) {
                if (stdp) {
                    hipLaunchKernelGGL((k_sstage_b<TIN, TOUT, T, L, TB, NT, DIR, true, JSEL>), grid, dim3(NT), 0, st, arr, bl->left);
                    return hipGetLastError();
                }
            }
            hipLaunchKernelGGL((k_sstage_b<TIN, TOUT, T, L, TB, NT, DIR, false, JSEL>), grid, dim3(NT), 0, st, arr, bl->left);
            return hipGetLastError();
        }
        return hipErrorNotSupported;
    }
    if constexpr (STDP_OK) {
        if (stdp) {
            hipLaunchKernelGGL((k_sstage<TIN, TOUT, T, L, TB, NT, DIR, true, JSEL>), dim3((unsigned)g), dim3(NT), 0, st, b);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((k_sstage<TIN, TOUT, T, L, TB, NT, DIR, false, JSEL>), dim3((unsigned)g), dim3(NT), 0, st, b);
    return hipGetLastError();
}

// The combined dual-job kernel (JSEL 1) for the inverse passes at L >= 512,
// unless SPNS_DUALC=0; the single jobs of the same pass then take a second
// launch.  At 512^3 it took FP64 inv_x from 8.00 to 7.59 ms and inv_y from
// 3.81 to 3.60 ms.  In FP32 it took inv_x from 4.01 to 3.76 ms, but with the
// pitch-9 z pass the z pass that follows ran 0.3 ms longer at the same cycle
// count (rocprofv3 GRBM_GUI_ACTIVE: a lower engine clock after the faster
// memory-bound kernel) and the step gained nothing; with the swizzled z pass
// the FP32 inv_x goes from 3.98 to 3.72 ms and the step to 16.76 ms on one
// box (16.81-17.03 ms without, on others), so both forms use it.
template <typename TIN, typename T, int L, int TB, int DIR>
constexpr bool dualc_applies()
{
    return DIR > 0 && L >= 512 && sstage_fits<TIN, T, L, TB, DIR, 2>();
}

static bool dualc_enabled()
{
    static const bool on = [] {
        const char* s = getenv("SPNS_DUALC");
        return !(s && s[0] == '0');
    }();
    return on;
}

template <typename TIN, typename TOUT, typename T, int L, int DIR>
static hipError_t launch_strided_L(const StridedArgs& a, int A, hipStream_t st, const BatchLaunch* bl)
{
    constexpr int TB = strided_tb<T, L>();
    StridedArgs b = a;
    b.A = A;
    if constexpr (sstage_fits<TIN, T, L, TB, DIR>()) {
        if (a.padded_ok && a.nin <= sstage_rows<L, DIR>()) {
            if constexpr (dualc_applies<TIN, T, L, TB, DIR>()) {
                int nd = 0;
                for (int j = 0; j < a.njobs; j++) nd += a.job[j].out2 != nullptr;
                if (nd > 0 && dualc_enabled()) {
                    hipError_t e = launch_sstage<TIN, TOUT, T, L, TB, DIR, 1>(b, A, nd, st, bl);
                    if (e == hipSuccess && nd < a.njobs) e = launch_sstage<TIN, TOUT, T, L, TB, DIR, 2>(b, A, a.njobs - nd, st, bl);
                    return e;
                }
            }
            return launch_sstage<TIN, TOUT, T, L, TB, DIR, 0>(b, A, a.njobs, st, bl);
        }
    }
    if (bl) return hipErrorNotSupported;  // no batched form of the direct-load kernel
    // the direct-load kernel takes single jobs: a dual job becomes two
    int n = 0;
    for (int j = 0; j < a.njobs; j++) {
        if (a.job[j].out2) {
            if (n + 2 > 9) return hipErrorInvalidValue;
            b.job[n++] = StridedJob{a.job[j].in, a.job[j].out, 0, 0, nullptr};
            b.job[n++] = StridedJob{a.job[j].in, a.job[j].out2, 1, 0, nullptr};
        } else {
            if (n + 1 > 9) return hipErrorInvalidValue;
            b.job[n++] = a.job[j];
        }
    }
    b.njobs = n;
    constexpr int NT = 256;
    const int64_t blocks = (int64_t)((a.B + TB - 1) / TB) * A;
    dim3 grid((unsigned)((blocks + 7) / 8 * 8), 1, 1);
    hipLaunchKernelGGL((k_strided<TIN, TOUT, T, L, TB, NT, DIR>), grid, dim3(NT), 0, st, b);
    return hipGetLastError();
}

#define SPNS_FOR_EACH_L(X) \
    X(16) X(24) X(32) X(36) X(48) X(64) X(72) X(96) X(128) X(144) X(192) X(256) X(288) X(384) X(512) X(576) X(768) X(1024) X(1152) X(1536)

#if SPNS_PART == 0 || SPNS_PART == 4
int strided_tile(int L, int prec)
{
#define X(LL) \
    if (L == LL) return prec == PREC_FP64 ? strided_tb<double, LL>() : strided_tb<float, LL>();
    SPNS_FOR_EACH_L(X)
#undef X
    return 0;
}

bool supported_length(int L)
{
#define X(LL) \
    if (L == LL) return true;
    SPNS_FOR_EACH_L(X)
#undef X
    return false;
}

#endif

template <typename TIN, typename TOUT, typename T, int DIR>
static hipError_t launch_strided_T(int L, const StridedArgs& a, int A, hipStream_t st, const BatchLaunch* bl)
{
#define X(LL) \
    if (L == LL) return launch_strided_L<TIN, TOUT, T, LL, DIR>(a, A, st, bl);
    SPNS_FOR_EACH_L(X)
#undef X
    return hipErrorInvalidValue;
}

#if SPNS_PART == 0 || SPNS_PART == 1
hipError_t launch_strided_fp64(int L, int dir, const StridedArgs& a, int A, hipStream_t st, const BatchLaunch* bl)
{
    return dir > 0 ? launch_strided_T<double, double, double, +1>(L, a, A, st, bl)
                   : launch_strided_T<double, double, double, -1>(L, a, A, st, bl);
}
#endif

#if SPNS_PART == 0 || SPNS_PART == 2
hipError_t launch_strided_fp32(int L, int in_fp64, int dir, const StridedArgs& a, int A, hipStream_t st, const BatchLaunch* bl)
{
    if (in_fp64) {
        return dir > 0 ? launch_strided_T<double, float, float, +1>(L, a, A, st, bl)
                       : launch_strided_T<double, float, float, -1>(L, a, A, st, bl);
    }
    return dir > 0 ? launch_strided_T<float, float, float, +1>(L, a, A, st, bl)
                   : launch_strided_T<float, float, float, -1>(L, a, A, st, bl);
}
#endif

#if SPNS_PART == 0 || SPNS_PART == 4
hipError_t launch_strided(int L, int prec, int in_fp64, int out_fp64, int dir, const StridedArgs& a, int A, hipStream_t st,
                          const BatchLaunch* bl)
{
    (void)out_fp64;
    if (prec == PREC_FP64) return launch_strided_fp64(L, dir, a, A, st, bl);
    return launch_strided_fp32(L, in_fp64, dir, a, A, st, bl);
}
#endif

#if SPNS_PART == 0 || SPNS_PART == 3
// Compile-time retained k3 count of the standard 3/2 padding (M = 3N/2,
// nk3 = N/2 + 1 = M/3 + 1), 0 where M is not a multiple of 3.
template <int L>
constexpr int zstage_nkc()
{
    return (L % 3 == 0) ? L / 3 + 1 : 0;
}

// Threads of the staged z pass.  FP32 at L = 768: 12 waves (measured 9.5 ms
// against 9.9 ms with 8 waves at 512^3); otherwise as the block kernel.
template <typename T, int L>
constexpr int zstage_nt()
{
    return (sizeof(T) == 4 && L == 768) ? 768 : zpass_nt<L>();
}

template <typename T, int L, int MODE>
static bool zstage_ok(const ZArgs& a)
{
    constexpr int NT = zstage_nt<T, L>();
    if constexpr (!zstage_fits<T, L, NT, MODE>()) return false;
    return a.pitch + zstage_pad<T>() <= zstage_spc<L>() && a.nk3 <= a.pitch;
}

// z-pass kernel variants and the choice among them
enum { ZV_BLOCK, ZV_WAVE, ZV_STAGE, ZV_STAGE_NKC, ZV_FUSE_NKC };

template <typename T, int L>
constexpr int zfuse_nt()
{
    return sstage_nt<L, 8>();
}

template <typename T, int L>
constexpr bool zfuse_fits()
{
    return zstage_nkc<L>() > 0 && zfuse_lds_bytes<T, L, zfuse_nt<T, L>(), zstage_nkc<L>()>() <= 160 * 1024 - 512;
}

template <typename T, int L, int MODE>
static int zpass_pick(const ZArgs& a)
{
    if constexpr (MODE != 2 && has_wave_plan<L>::value) {
        if ((!a.les || MODE == 1) && a.kernel == ZK_WAVE) return ZV_WAVE;
    }
    if (a.kernel == ZK_STAGED || a.kernel == ZK_STAGED_NOFUSE) {
        const bool pitch_ok = a.pitch + zstage_pad<T>() <= zstage_spc<L>() && a.nk3 <= a.pitch;
        const bool nkc = zstage_nkc<L>() > 0 && a.nk3 == zstage_nkc<L>();
        if constexpr (MODE == 0 && zfuse_fits<T, L>()) {
            if (a.kernel == ZK_STAGED && !a.les && pitch_ok && nkc) return ZV_FUSE_NKC;
        }
        if constexpr (zstage_fits<T, L, zstage_nt<T, L>(), MODE>()) {
            if (pitch_ok) return nkc ? ZV_STAGE_NKC : ZV_STAGE;
        }
    }
    return ZV_BLOCK;
}

template <typename T, int L, int MODE>
static hipError_t launch_zpass_L(const ZArgs& a, int grid, hipStream_t st, const BatchLaunch* bl)
{
    constexpr int NTS = zstage_nt<T, L>();
    constexpr int NKC = zstage_nkc<L>();
    if (bl) {
        // batched step (MODE 0) over the slots' structs in bl->args
        if constexpr (MODE == 0 && batch_length(L)) {
            const dim3 g((unsigned)grid, (unsigned)bl->nslots);
            const ZArgs* arr = (const ZArgs*)bl->args;
            switch (zpass_pick<T, L, MODE>(a)) {
            case ZV_FUSE_NKC:
                if constexpr (zfuse_fits<T, L>()) {
                    constexpr int NTF = zfuse_nt<T, L>();
                    hipLaunchKernelGGL((k_zfuse_b<T, L, NTF, NKC>), g, dim3(NTF), 0, st, arr, bl->left);
                    return hipGetLastError();
                }
                break;
            case ZV_STAGE_NKC:
                if constexpr (NKC > 0 && zstage_fits<T, L, NTS, MODE>()) {
                    hipLaunchKernelGGL((k_zstage_b<T, L, NTS, MODE, zstage_ld<MODE>(), 1, sizeof(T) == 4, NKC>), g, dim3(NTS), 0, st,
                                       arr, bl->left);
                    return hipGetLastError();
                }
                break;
            case ZV_STAGE:
                if constexpr (zstage_fits<T, L, NTS, MODE>()) {
                    hipLaunchKernelGGL((k_zstage_b<T, L, NTS, MODE>), g, dim3(NTS), 0, st, arr, bl->left);
                    return hipGetLastError();
                }
                break;
            default:
                break;
            }
        }
        return hipErrorNotSupported;
    }
    switch (zpass_pick<T, L, MODE>(a)) {
    case ZV_WAVE:
        if constexpr (has_wave_plan<L>::value) {
            hipLaunchKernelGGL((k_zwave<T, L, MODE>), dim3(grid), dim3(256), 0, st, a);
            return hipGetLastError();
        }
        break;
    case ZV_FUSE_NKC:
        if constexpr (MODE == 0 && zfuse_fits<T, L>()) {
            constexpr int NTF = zfuse_nt<T, L>();
            hipLaunchKernelGGL((k_zfuse<T, L, NTF, NKC>), dim3(grid), dim3(NTF), 0, st, a);
            return hipGetLastError();
        }
        break;
    case ZV_STAGE_NKC:
        if constexpr (NKC > 0 && zstage_fits<T, L, NTS, MODE>()) {
            hipLaunchKernelGGL((k_zstage<T, L, NTS, MODE, zstage_ld<MODE>(), 1, sizeof(T) == 4, NKC>), dim3(grid), dim3(NTS),
                               0, st, a);
            return hipGetLastError();
        }
        break;
    case ZV_STAGE:
        if constexpr (zstage_fits<T, L, NTS, MODE>()) {
            hipLaunchKernelGGL((k_zstage<T, L, NTS, MODE>), dim3(grid), dim3(NTS), 0, st, a);
            return hipGetLastError();
        }
        break;
    default:
        break;
    }
    if constexpr (MODE == 2) {
        return hipErrorNotSupported;  // the invariants pass exists only in the staged kernel
    } else {
        constexpr int NT = zpass_nt<L>();
        hipLaunchKernelGGL((k_zpass<T, L, NT, MODE>), dim3(grid), dim3(NT), 0, st, a);
        return hipGetLastError();
    }
}

template <class K>
static int blocks_per_cu(K kernel, int nt)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, nt, 0) != hipSuccess) return -1;
    return nb > 0 ? nb : 1;
}

// Grid of the z pass: the staged kernels run persistent blocks (as many as
// fit on the device at once, counted for the full step's kernel), the others
// one block per pencil up to 65536.
template <typename T, int L>
static int zpass_grid_L(const ZArgs& a, int64_t npen)
{
    constexpr int NTS = zstage_nt<T, L>();
    constexpr int NKC = zstage_nkc<L>();
    int nb = 0;
    switch (zpass_pick<T, L, 0>(a)) {
    case ZV_FUSE_NKC:
        if constexpr (zfuse_fits<T, L>()) nb = blocks_per_cu(k_zfuse<T, L, zfuse_nt<T, L>(), NKC>, zfuse_nt<T, L>());
        break;
    case ZV_STAGE_NKC:
        if constexpr (NKC > 0 && zstage_fits<T, L, NTS, 0>())
            nb = blocks_per_cu(k_zstage<T, L, NTS, 0, zstage_ld<0>(), 1, sizeof(T) == 4, NKC>, NTS);
        break;
    case ZV_STAGE:
        if constexpr (zstage_fits<T, L, NTS, 0>()) nb = blocks_per_cu(k_zstage<T, L, NTS, 0>, NTS);
        break;
    default:
        return (int)(npen < 65536 ? npen : 65536);
    }
    if (nb <= 0) return -1;
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
    const int64_t g = (int64_t)ncu * nb;
    return (int)(npen < g ? npen : g);
}

int zpass_grid(int L, int prec, const ZArgs& a, int64_t npen)
{
#define X(LL)                                                                                          \
    if (L == LL) return prec == PREC_FP64 ? zpass_grid_L<double, LL>(a, npen) : zpass_grid_L<float, LL>(a, npen);
    SPNS_FOR_EACH_L(X)
#undef X
    return -1;
}

// Kernel a z pass with these arguments runs (ZK_*; the staged forms report
// ZK_STAGED): falls back to ZK_BLOCK where the staged kernels' LDS does not fit.
template <typename T, int L>
static int zpass_kernel_L(const ZArgs& a)
{
    switch (zpass_pick<T, L, 0>(a)) {
    case ZV_WAVE: return ZK_WAVE;
    case ZV_BLOCK: return ZK_BLOCK;
    default: return ZK_STAGED;
    }
}

int zpass_kernel(int L, int prec, const ZArgs& a)
{
#define X(LL)                                                                                        \
    if (L == LL) return prec == PREC_FP64 ? zpass_kernel_L<double, LL>(a) : zpass_kernel_L<float, LL>(a);
    SPNS_FOR_EACH_L(X)
#undef X
    return -1;
}

hipError_t launch_zpass(int L, int prec, int mode, const ZArgs& a, int grid, hipStream_t st, const BatchLaunch* bl)
{
#define X(LL)                                                                                  \
    if (L == LL) {                                                                             \
        if (prec == PREC_FP64)                                                                 \
            return mode == 0   ? launch_zpass_L<double, LL, 0>(a, grid, st, bl)                \
                   : mode == 1 ? launch_zpass_L<double, LL, 1>(a, grid, st, bl)                \
                               : launch_zpass_L<double, LL, 2>(a, grid, st, bl);               \
        return mode == 0   ? launch_zpass_L<float, LL, 0>(a, grid, st, bl)                     \
               : mode == 1 ? launch_zpass_L<float, LL, 1>(a, grid, st, bl)                     \
                           : launch_zpass_L<float, LL, 2>(a, grid, st, bl);                    \
    }
    SPNS_FOR_EACH_L(X)
#undef X
    return hipErrorInvalidValue;
}

#endif

#if SPNS_PART == 0 || SPNS_PART == 4
template <typename T, int L, int DIR>
static hipError_t launch_zplain_L(const ZPlainArgs& a, int grid, hipStream_t st)
{
    constexpr int NT = 64;
    hipLaunchKernelGGL((k_zplain<T, L, NT, DIR>), dim3(grid), dim3(NT), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_zplain(int L, int prec, int dir, const ZPlainArgs& a, int grid, hipStream_t st)
{
#define X(LL)                                                                                        \
    if (L == LL) {                                                                                   \
        if (prec == PREC_FP64) return dir > 0 ? launch_zplain_L<double, LL, +1>(a, grid, st)         \
                                              : launch_zplain_L<double, LL, -1>(a, grid, st);        \
        return dir > 0 ? launch_zplain_L<float, LL, +1>(a, grid, st) : launch_zplain_L<float, LL, -1>(a, grid, st); \
    }
    SPNS_FOR_EACH_L(X)
#undef X
    return hipErrorInvalidValue;
}


// ---------------------------------------------------------------------------
// Shell spectra of the five velocity-gradient invariants (src/HIT_LES.cpp:
// 1297-1319 with the integrands |I_k|^2 of :1006-1025 and the shell sums of
// Integrate_Field :1137-1214): one wave per block walks (i2, i1) rows of the
// compact box, lanes along i3; per-block LDS histograms [5][nbins] (fixed
// order within the wave), then a fixed-order sum over the blocks.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(64) void k_inv_spectra(InvSpecArgs a)
{
    // One wave per block, rows (k1, k2) grid-strided.  Along a row the shell
    // K = round(|k|) never decreases with k3, so each 64-mode chunk splits
    // into runs of equal K: the first lane of a run sums it in order and adds
    // it to the block's bins.  Runs of one chunk hit distinct bins and the
    // chunks follow in program order, so the sums are deterministic (no
    // atomics); k_inv_sum adds the block partials in block order.
    extern __shared__ double hs[];  // [5][nbins] bins, [5][64] values, [64] shells
    const int nb = a.nbins, lane = threadIdx.x;
    double* vals = hs + 5 * nb;
    int* kb = (int*)(vals + 5 * 64);
    for (int i = lane; i < 5 * nb; i += 64) hs[i] = 0.0;
    __syncthreads();
    const int nrow = a.nk1 * a.nk2;
    for (int row = blockIdx.x; row < nrow; row += gridDim.x) {
        const int i2 = row / a.nk1, i1 = row - i2 * a.nk1;
        const int k1 = wrap_k(i1, a.nk1), k2 = wrap_k(i2, a.nk2);
        for (int base = 0; base < a.nk3; base += 64) {
            const int i3 = base + lane;
            int K = -1;
            if (i3 < a.nk3) {
                const int k3 = i3;
                K = shell_of(k1 * k1 + k2 * k2 + k3 * k3);
                const bool skip = a.notcubic && ((k1 % a.Lx) || (k2 % a.Ly) || (k3 % a.Lz));  // src/HIT_LES.cpp:867-871
                const bool conj = (k3 != 0) && !((a.Mz % 2 == 0) && (k3 == a.Mz / 2));       // src/HIT_LES.cpp:880
                const double w = skip ? 0.0 : (conj ? 2.0 : 1.0) * a.inv_nf2;
                const int64_t m = (int64_t)row * a.nk3p + i3;
#pragma unroll
                for (int f = 0; f < 5; f++) {
                    const cx<T> c = ((const cx<T>*)a.f[f])[m];
                    vals[f * 64 + lane] = w * ((double)c.x * (double)c.x + (double)c.y * (double)c.y);
                }
            }
            kb[lane] = K;
            __syncthreads();
            if (K >= 0 && (lane == 0 || kb[lane - 1] != K)) {
                int e = lane + 1;
                while (e < 64 && kb[e] == K) e++;
#pragma unroll
                for (int f = 0; f < 5; f++) {
                    double sum = 0.0;
                    for (int j = lane; j < e; j++) sum += vals[f * 64 + j];
                    hs[f * nb + K] += sum;
                }
            }
            __syncthreads();
        }
    }
    for (int i = lane; i < 5 * nb; i += 64) a.partial[(int64_t)blockIdx.x * 5 * nb + i] = hs[i];
}

static __global__ void k_inv_sum(const double* partial, int nblk, int n, double* out)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int b = 0; b < nblk; b++) s += partial[(int64_t)b * n + i];
        out[i] = s;
    }
}

hipError_t launch_inv_spectra(int prec, const InvSpecArgs& a, int grid, double* out, hipStream_t st)
{
    const size_t shm = sizeof(double) * (5 * (size_t)a.nbins + 5 * 64) + sizeof(int) * 64;
    if (prec == PREC_FP64) hipLaunchKernelGGL((k_inv_spectra<double>), dim3(grid), dim3(64), shm, st, a);
    else hipLaunchKernelGGL((k_inv_spectra<float>), dim3(grid), dim3(64), shm, st, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int n = 5 * a.nbins;
    hipLaunchKernelGGL(k_inv_sum, dim3((n + 255) / 256), dim3(256), 0, st, a.partial, grid, n, out);
    return hipGetLastError();
}

// Fused forward y pass + step epilogue (DNS): grid for A x tiles items with
// nbins shell bins, or 0 where the kernel does not apply (staging or LDS
// does not fit, or more retained rows than the 3/2 padding leaves).
// Block size of k_yepi: that of the staged strided pass, except
//   * for the longest transforms, where half of it leaves LDS for the
//     per-wave shell histograms;
//   * where that block has 1024 threads (L = 192 and 384 with 16-column
//     tiles): 16 waves leave 128 VGPRs a thread, and the fused kernel
//     spills there; 8 waves leave 256.
// The transforms then take up to two butterflies per thread.
template <int L, int TB>
__host__ __device__ constexpr int yepi_nt()
{
    return (L > 1152 || sstage_nt<L, TB>() >= 1024) ? sstage_nt<L, TB>() / 2 : sstage_nt<L, TB>();
}

// LDS check and occupancy of k_yepi (INV false) or k_yepi_inv (true)
template <typename T, int L, int TB, int NT, bool INV, class K>
static int yepi_blocks_per_cu(K kern, int nbins, int max_spill)
{
    const size_t shm = sizeof(double) * (NT / 64) * (size_t)nbins;
    const size_t lds = sizeof(cx<T>) * ((size_t)L * TB + yepi_stg_elems<T, L, TB, INV>() +
                                        fft_twiddle_span<L, sizeof(T) == 8 ? 1 : 2>());
    if (lds + shm > 160 * 1024 - 512) return 0;
    hipFuncAttributes fa;
    if (hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(kern)) != hipSuccess || (int)fa.localSizeBytes > max_spill) return 0;
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kern, NT, shm) != hipSuccess || n < 1) return 0;
    return n;
}

template <typename T, int L>
static int yepi_grid_L(int nbins, int nout, int64_t items, int max_grid, bool inv)
{
    constexpr int TB = strided_tb<T, L>();
    if constexpr (TB <= 0 || (TB * sizeof(cx<T>)) % 16 != 0) {
        return 0;
    } else {
        if (nout > 2 * (L / 3) + 1) return 0;
        constexpr int NT = yepi_nt<L, TB>();
        // a variant that spills registers (the 1024-thread tiles) is slower
        // than the two separate kernels; k_yepi_inv may keep a few
        // loop-invariant values in scratch (FP32 at L = 768: 20 bytes, three
        // reloads per item outside the transforms)
        const int n = inv ? yepi_blocks_per_cu<T, L, TB, NT, true>(k_yepi_inv<T, L, TB, NT>, nbins, 32)
                          : yepi_blocks_per_cu<T, L, TB, NT, false>(k_yepi<T, L, TB, NT>, nbins, 0);
        if (n < 1) return 0;
        int dev = 0, ncu = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return 0;
        int64_t g = (int64_t)ncu * n;
        if (items < g) g = items;
        g = (g + 7) / 8 * 8;
        if (g > max_grid) g = max_grid / 8 * 8;
        return (int)g;
    }
}

int yepi_grid(int L, int prec, int nbins, int nout, int64_t items, int max_grid, bool inv)
{
#define X(LL)                                                                                   \
    if (L == LL)                                                                                \
        return prec == PREC_FP64 ? yepi_grid_L<double, LL>(nbins, nout, items, max_grid, inv) \
                                 : yepi_grid_L<float, LL>(nbins, nout, items, max_grid, inv);
    SPNS_FOR_EACH_L(X)
#undef X
    return 0;
}

template <typename T, int L>
static hipError_t launch_yepi_L(const StridedArgs& a, int A, const EpiArgs& e, int grid, hipStream_t st, const BatchLaunch* bl,
                                const StridedArgs* ia, int* x1v)
{
    constexpr int TB = strided_tb<T, L>();
    if constexpr (TB <= 0 || (TB * sizeof(cx<T>)) % 16 != 0) {
        return hipErrorNotSupported;
    } else {
        StridedArgs b = a;
        b.A = A;
        if (!a.padded_ok || a.nin != L || a.njobs != 3 || a.nout > 2 * (L / 3) + 1 || grid < 8 || grid % 8) return hipErrorInvalidValue;
        if (bl) {
            constexpr int NT = yepi_nt<L, TB>();
            const size_t shm = sizeof(double) * (NT / 64) * (size_t)e.nbins;
            if constexpr (batch_length(L)) {
                if (ia) {
                    // with the next step's inverse y pass: ia is a slot's struct (all
                    // slots share the shape class), the slots' structs are in args3
                    bool ok = bl->args3 && ia->setflag && ia->padded_ok && ia->njobs == 3 && ia->nout == L && ia->nin == a.nout &&
                              ia->B == a.B;
                    for (int j = 0; ok && j < 3; j++) ok = ia->job[j].in && ia->job[j].out && ia->job[j].out2;
                    if (!ok) return hipErrorInvalidValue;
                    hipLaunchKernelGGL((k_yepi_inv_b<T, L, TB, NT>), dim3(grid, bl->nslots), dim3(NT), shm, st,
                                       (const StridedArgs*)bl->args, (const EpiArgs*)bl->args2, (const StridedArgs*)bl->args3, bl->left);
                    return hipGetLastError();
                }
                hipLaunchKernelGGL((k_yepi_b<T, L, TB, NT>), dim3(grid, bl->nslots), dim3(NT), shm, st, (const StridedArgs*)bl->args,
                                   (const EpiArgs*)bl->args2, bl->left);
                return hipGetLastError();
            }
            return hipErrorNotSupported;
        }
        if (ia) {
            // the inverse pass's items must be the forward pass's: same
            // outer extent and column tiles, all rows of the transform out
            bool ok = x1v && ia->padded_ok && ia->njobs == 3 && ia->nout == L && ia->nin == a.nout && ia->B == a.B;
            for (int j = 0; ok && j < 3; j++) ok = ia->job[j].in && ia->job[j].out && ia->job[j].out2;
            if (!ok) return hipErrorInvalidValue;
            StridedArgs ib = *ia;
            ib.A = A;
            constexpr int NT = yepi_nt<L, TB>();
            const size_t shm = sizeof(double) * (NT / 64) * (size_t)e.nbins;
            hipLaunchKernelGGL((k_yepi_inv<T, L, TB, NT>), dim3(grid), dim3(NT), shm, st, b, e, ib, x1v);
            return hipGetLastError();
        }
        constexpr int NT = yepi_nt<L, TB>();
        const size_t shm = sizeof(double) * (NT / 64) * (size_t)e.nbins;
        hipLaunchKernelGGL((k_yepi<T, L, TB, NT>), dim3(grid), dim3(NT), shm, st, b, e);
        return hipGetLastError();
    }
}

hipError_t launch_yepi(int L, int prec, const StridedArgs& a, int A, const EpiArgs& e, int grid, hipStream_t st, const BatchLaunch* bl,
                       const StridedArgs* ia, int* x1v)
{
#define X(LL)                                                                                        \
    if (L == LL)                                                                                     \
        return prec == PREC_FP64 ? launch_yepi_L<double, LL>(a, A, e, grid, st, bl, ia, x1v) \
                                 : launch_yepi_L<float, LL>(a, A, e, grid, st, bl, ia, x1v);
    SPNS_FOR_EACH_L(X)
#undef X
    return hipErrorInvalidValue;
}

// Sets an int flag in device memory (stream-ordered, no memset node).
static __global__ void k_set_int(int* p, int v)
{
    *p = v;
}

hipError_t launch_set_int(int* p, int v, hipStream_t st)
{
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, st, p, v);
    return hipGetLastError();
}

hipError_t launch_epilogue(int prec, const EpiArgs& a, int grid, hipStream_t st, const BatchLaunch* bl)
{
    constexpr int NT = 256;
    const size_t shm = sizeof(double) * (NT / 64) * a.nbins;
    if (bl) {
        const EpiArgs* arr = (const EpiArgs*)bl->args;
        if (prec == PREC_FP64) hipLaunchKernelGGL((k_epilogue_b<double, NT>), dim3(grid, bl->nslots), dim3(NT), shm, st, arr, bl->left);
        else hipLaunchKernelGGL((k_epilogue_b<float, NT>), dim3(grid, bl->nslots), dim3(NT), shm, st, arr, bl->left);
        return hipGetLastError();
    }
    if (prec == PREC_FP64) hipLaunchKernelGGL((k_epilogue<double, NT>), dim3(grid), dim3(NT), shm, st, a);
    else hipLaunchKernelGGL((k_epilogue<float, NT>), dim3(grid), dim3(NT), shm, st, a);
    return hipGetLastError();
}

hipError_t launch_finalize(const FinArgs& a, hipStream_t st, const BatchLaunch* bl)
{
    if (bl) {
        hipLaunchKernelGGL((k_finalize_b<1024>), dim3(1, bl->nslots), dim3(1024), 0, st, (const FinArgs*)bl->args, bl->left);
        return hipGetLastError();
    }
    hipLaunchKernelGGL((k_finalize<1024>), dim3(1), dim3(1024), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_clear(unsigned long long* a, unsigned long long* b, hipStream_t st, const BatchLaunch* bl)
{
    if (bl) {
        hipLaunchKernelGGL(k_clear_b, dim3(1, bl->nslots), dim3(64), 0, st, (const ClearArgs*)bl->args, bl->left);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_clear, dim3(1), dim3(64), 0, st, a, b);
    return hipGetLastError();
}

hipError_t launch_stream_copy(const void* in, void* out, int64_t bytes, int blocks, hipStream_t st)
{
    if (bytes % 16 || blocks < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_stream_copy<4>), dim3(blocks), dim3(256), 0, st, (const vf4*)in, (vf4*)out, bytes / 16);
    return hipGetLastError();
}

hipError_t launch_set_left(int* left, const SlotSteps& s, int nslots, hipStream_t st)
{
    if (nslots < 1 || nslots > kMaxBatchSlots) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_set_left, dim3(1), dim3(kMaxBatchSlots), 0, st, left, s, nslots);
    return hipGetLastError();
}

hipError_t launch_forcing(const double* ek, const double* forced, double* ft, int last_eff, const ShellArgs& a, hipStream_t st)
{
    hipLaunchKernelGGL(k_forcing_ft, dim3(1), dim3(256), 0, st, ek, forced, ft, last_eff);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    int grid = (int)((a.nmodes + 255) / 256);
    if (grid > 4096) grid = 4096;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(k_scale_shells, dim3(grid), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_ccc(const ShellArgs& a, hipStream_t st)
{
    int n = a.nk1 * a.nk2;
    int grid = (n + 255) / 256;
    hipLaunchKernelGGL(k_ccc, dim3(grid), dim3(256), 0, st, a);
    return hipGetLastError();
}


// ---------------------------------------------------------------------------
// Synthetic initial field: random-phase, solenoidal, E(k) ~ k^4 exp(-2(k/k0)^2)
// (unscaled).  Counter-based: each mode's value depends only on
// (seed, k1, k2, k3), so the field is the same for any launch shape.  The
// k3 = 0 plane is made Hermitian by taking the mirror of the canonical half.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t splitmix(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__device__ __forceinline__ double unif(uint64_t seed, int k1, int k2, int k3, int d)
{
    uint64_t key = splitmix(seed) ^ splitmix(((uint64_t)(uint32_t)(k1 + 65536) << 40) ^
                                             ((uint64_t)(uint32_t)(k2 + 65536) << 20) ^ (uint64_t)(uint32_t)k3 ^
                                             ((uint64_t)d << 58));
    return ((double)(splitmix(key) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

static __global__ void k_init_random(RandArgs a)
{
    const int nk1 = a.nk1, nk2 = a.nk2, nk3 = a.nk3, nk3p = a.nk3p;
    const double PI = 3.141592653589793238462643;
    for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m < a.nmodes; m += (int64_t)gridDim.x * blockDim.x) {
        const int i3 = (int)(m % nk3p);
        if (i3 >= nk3) continue;  // pitch padding
        const int64_t r12 = m / nk3p;
        const int i1 = (int)(r12 % nk1), i2 = (int)(r12 / nk1);
        int k1 = wrap_k(i1, nk1), k2 = wrap_k(i2, nk2);
        const int k3 = i3;
        bool masked = a.notcubic && ((k1 % a.Lx) || (k2 % a.Ly) || (k3 % a.Lz));
        const bool canon = k3 > 0 || k1 > 0 || (k1 == 0 && k2 > 0);
        const bool mirror = (k3 == 0) && !canon && !(k1 == 0 && k2 == 0);
        cx<double> out[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
        if (!masked && (canon || mirror)) {
            if (mirror) { k1 = -k1; k2 = -k2; }
            const double K1 = k1, K2 = k2, K3 = k3;
            const double r2 = K1 * K1 + K2 * K2 + K3 * K3;
            double gr[3], gi[3];
            for (int c = 0; c < 3; c++) {
                const double u1 = unif(a.seed, k1, k2, k3, 2 * c), u2 = unif(a.seed, k1, k2, k3, 2 * c + 1);
                const double rr = sqrt(-2.0 * log(u1));
                gr[c] = rr * cos(2 * PI * u2);
                gi[c] = rr * sin(2 * PI * u2);
            }
            const double dr = (K1 * gr[0] + K2 * gr[1] + K3 * gr[2]) / r2;
            const double di = (K1 * gi[0] + K2 * gi[1] + K3 * gi[2]) / r2;
            const double kk[3] = {K1, K2, K3};
            double nrm = 0.0;
            for (int c = 0; c < 3; c++) {
                gr[c] -= dr * kk[c];
                gi[c] -= di * kk[c];
                nrm += gr[c] * gr[c] + gi[c] * gi[c];
            }
            if (nrm > 0.0) {
                const double km = sqrt(r2);
                const double Ek = km * km * km * km * exp(-2.0 * (km / a.k0) * (km / a.k0));
                const double amp = sqrt(Ek / (4.0 * PI * r2) / nrm);
                for (int c = 0; c < 3; c++) out[c] = cx<double>{gr[c] * amp, (mirror ? -1.0 : 1.0) * gi[c] * amp};
            }
        }
        for (int c = 0; c < 3; c++) a.uk[c][m] = out[c];
    }
}

static __global__ void k_scale_state(ShellArgs a, double sc)
{
    for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m < a.nmodes; m += (int64_t)gridDim.x * blockDim.x)
        for (int c = 0; c < 3; c++) {
            cx<double> v = a.uk[c][m];
            a.uk[c][m] = cx<double>{v.x * sc, v.y * sc};
        }
}

hipError_t launch_init_random(const RandArgs& a, hipStream_t st)
{
    int grid = (int)((a.nmodes + 255) / 256);
    if (grid > 16384) grid = 16384;
    hipLaunchKernelGGL(k_init_random, dim3(grid), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_scale_state(const ShellArgs& a, double sc, hipStream_t st)
{
    int grid = (int)((a.nmodes + 255) / 256);
    if (grid > 16384) grid = 16384;
    hipLaunchKernelGGL(k_scale_state, dim3(grid), dim3(256), 0, st, a, sc);
    return hipGetLastError();
}

hipError_t launch_normalize(int prec, const NormArgs& a, hipStream_t st)
{
    const int64_t n = (int64_t)a.My * a.Mx * (a.Mz / 2 + 1);
    int grid = (int)((n + 255) / 256);
    if (grid > 8192) grid = 8192;
    if (prec == PREC_FP64) hipLaunchKernelGGL((k_normalize<double>), dim3(grid), dim3(256), 0, st, a);
    else hipLaunchKernelGGL((k_normalize<float>), dim3(grid), dim3(256), 0, st, a);
    return hipGetLastError();
}

#endif
}  // namespace spns
