// Device building blocks for the MI355X (gfx950) pseudospectral step:
// complex arithmetic, small in-register DFTs and a batched Stockham FFT that
// runs its stages through LDS.
//
// The FFT is the Stockham autosort formulation (Govindaraju et al., "High
// performance discrete Fourier transforms on graphics processors", SC08):
// stage s with radix R and span Ns reads x[j + r*L/R], multiplies by
// W_{Ns*R}^{r*(j mod Ns)}, does an R-point DFT and writes
// y[(j/Ns)*Ns*R + (j mod Ns) + r*Ns].  A block transforms NB independent
// columns of length L that live in LDS as [row][ld] (ld >= NB).  The first
// stage reads its inputs through a caller functor (global memory, pruned
// zero padding) and the last stage hands its outputs to another functor, so
// a transform costs (stages-1) LDS round trips.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>
#include <utility>

#include "spns_internal.hpp"

namespace spns {

// f(std::integral_constant<int, i>) for i = 0 .. N-1, unrolled at compile time.
template <class F, int... I>
__device__ __forceinline__ void static_for_seq(F& f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f)
{
    static_for_seq(f, std::make_integer_sequence<int, N>{});
}

template <typename T>
__device__ __forceinline__ cx<T> mk(T a, T b) { return cx<T>{a, b}; }
template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return cx<T>{a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return cx<T>{a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) { return cx<T>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
template <typename T>
__device__ __forceinline__ cx<T> cscale(cx<T> a, T s) { return cx<T>{a.x * s, a.y * s}; }
// multiply by -i (FWD) or +i (INV): dir = -1 forward, +1 inverse
template <int DIR, typename T>
__device__ __forceinline__ cx<T> mul_i(cx<T> a) { return DIR < 0 ? cx<T>{a.y, -a.x} : cx<T>{-a.y, a.x}; }

// ---------------------------------------------------------------------------
// Small DFTs in registers, y_k = sum_n x_n exp(DIR*2*pi*i*n*k/R).
// ---------------------------------------------------------------------------
template <int DIR, typename T>
__device__ __forceinline__ void dft2(cx<T>& a, cx<T>& b)
{
    cx<T> t = a - b;
    a = a + b;
    b = t;
}

template <int DIR, typename T>
__device__ __forceinline__ void dft3(cx<T>& x0, cx<T>& x1, cx<T>& x2)
{
    const T s = (T)0.86602540378443864676;  // sqrt(3)/2
    cx<T> t1 = x1 + x2;
    cx<T> t2 = x0 - cscale(t1, (T)0.5);
    cx<T> t3 = cscale(mul_i<DIR>(x1 - x2), s);
    x0 = x0 + t1;
    x1 = t2 + t3;
    x2 = t2 - t3;
}

template <int DIR, typename T>
__device__ __forceinline__ void dft4(cx<T>& x0, cx<T>& x1, cx<T>& x2, cx<T>& x3)
{
    cx<T> a = x0 + x2, b = x0 - x2, c = x1 + x3, d = mul_i<DIR>(x1 - x3);
    x0 = a + c;
    x2 = a - c;
    x1 = b + d;
    x3 = b - d;
}

// W_R^k as (cos, DIR*sin) for the constants the composite radices need
template <int DIR, typename T>
__device__ __forceinline__ cx<T> wk(T c, T s) { return cx<T>{c, (T)DIR * s}; }

template <int DIR, typename T>
__device__ __forceinline__ void dft8(cx<T>* x)
{
    const T h = (T)0.70710678118654752440;
    cx<T> e0 = x[0], e1 = x[2], e2 = x[4], e3 = x[6];
    cx<T> o0 = x[1], o1 = x[3], o2 = x[5], o3 = x[7];
    dft4<DIR>(e0, e1, e2, e3);
    dft4<DIR>(o0, o1, o2, o3);
    o1 = cmul(o1, wk<DIR>(h, h));
    o2 = mul_i<DIR>(o2);
    o3 = cmul(o3, wk<DIR>(-h, h));
    x[0] = e0 + o0; x[4] = e0 - o0;
    x[1] = e1 + o1; x[5] = e1 - o1;
    x[2] = e2 + o2; x[6] = e2 - o2;
    x[3] = e3 + o3; x[7] = e3 - o3;
}

// 6 = 2 x 3 and 12 = 3 x 4 by the Good-Thomas prime-factor algorithm (the
// factors are coprime, so the index maps n = (N2 n1 + N1 n2) mod N and
// k = (N2 (N2^-1 mod N1) k1 + N1 (N1^-1 mod N2) k2) mod N leave no twiddles):
//   6:  n = (3 n1 + 2 n2) mod 6,  k = (3 k1 + 4 k2) mod 6;
//   12: n = (4 n1 + 3 n2) mod 12, k = (4 k1 + 9 k2) mod 12.
template <int DIR, typename T>
__device__ __forceinline__ void dft6(cx<T>* x)
{
    cx<T> y[2][3];
#pragma unroll
    for (int n1 = 0; n1 < 2; n1++) {
#pragma unroll
        for (int n2 = 0; n2 < 3; n2++) y[n1][n2] = x[(3 * n1 + 2 * n2) % 6];
        dft3<DIR>(y[n1][0], y[n1][1], y[n1][2]);
    }
#pragma unroll
    for (int k2 = 0; k2 < 3; k2++) {
        cx<T> a = y[0][k2], b = y[1][k2];
        dft2<DIR>(a, b);
        x[(4 * k2) % 6] = a;
        x[(3 + 4 * k2) % 6] = b;
    }
}

template <int DIR, typename T>
__device__ __forceinline__ void dft12(cx<T>* x)
{
    cx<T> y[3][4];
#pragma unroll
    for (int n1 = 0; n1 < 3; n1++) {
#pragma unroll
        for (int n2 = 0; n2 < 4; n2++) y[n1][n2] = x[(4 * n1 + 3 * n2) % 12];
        dft4<DIR>(y[n1][0], y[n1][1], y[n1][2], y[n1][3]);
    }
#pragma unroll
    for (int k2 = 0; k2 < 4; k2++) {
        cx<T> a0 = y[0][k2], a1 = y[1][k2], a2 = y[2][k2];
        dft3<DIR>(a0, a1, a2);
        x[(9 * k2) % 12] = a0;
        x[(4 + 9 * k2) % 12] = a1;
        x[(8 + 9 * k2) % 12] = a2;
    }
}

template <int DIR, typename T>
__device__ __forceinline__ void dft16(cx<T>* x)
{
    const T c1 = (T)0.92387953251128675613, s1 = (T)0.38268343236508977173, h = (T)0.70710678118654752440;
    cx<T> e[8], o[8];
#pragma unroll
    for (int k = 0; k < 8; k++) { e[k] = x[2 * k]; o[k] = x[2 * k + 1]; }
    dft8<DIR>(e);
    dft8<DIR>(o);
    o[1] = cmul(o[1], wk<DIR>(c1, s1));
    o[2] = cmul(o[2], wk<DIR>(h, h));
    o[3] = cmul(o[3], wk<DIR>(s1, c1));
    o[4] = mul_i<DIR>(o[4]);
    o[5] = cmul(o[5], wk<DIR>(-s1, c1));
    o[6] = cmul(o[6], wk<DIR>(-h, h));
    o[7] = cmul(o[7], wk<DIR>(-c1, s1));
#pragma unroll
    for (int k = 0; k < 8; k++) {
        x[k] = e[k] + o[k];
        x[k + 8] = e[k] - o[k];
    }
}

// 24 = 3 x 8: three 8-point DFTs of the decimated sequences, twiddles W24^(n2 k),
// then 3-point DFTs across n2.
template <int DIR, typename T>
__device__ __forceinline__ void dft24(cx<T>* x)
{
    // cos/sin of k*15 degrees, k = 0..7
    const T C[8] = {(T)1.0, (T)0.96592582628906828675, (T)0.86602540378443864676, (T)0.70710678118654752440,
                    (T)0.5, (T)0.25881904510252076235, (T)0.0, (T)-0.25881904510252076235};
    const T S[8] = {(T)0.0, (T)0.25881904510252076235, (T)0.5, (T)0.70710678118654752440,
                    (T)0.86602540378443864676, (T)0.96592582628906828675, (T)1.0, (T)0.96592582628906828675};
    cx<T> d[3][8];
#pragma unroll
    for (int n2 = 0; n2 < 3; n2++) {
#pragma unroll
        for (int n1 = 0; n1 < 8; n1++) d[n2][n1] = x[3 * n1 + n2];
        dft8<DIR>(d[n2]);
    }
    // W24^(2k): cos/sin of k*30 degrees, k = 0..7
    const T C2[8] = {(T)1.0, (T)0.86602540378443864676, (T)0.5, (T)0.0, (T)-0.5, (T)-0.86602540378443864676, (T)-1.0, (T)-0.86602540378443864676};
    const T S2[8] = {(T)0.0, (T)0.5, (T)0.86602540378443864676, (T)1.0, (T)0.86602540378443864676, (T)0.5, (T)0.0, (T)-0.5};
#pragma unroll
    for (int k = 1; k < 8; k++) {
        d[1][k] = cmul(d[1][k], wk<DIR>(C[k], S[k]));
        d[2][k] = cmul(d[2][k], wk<DIR>(C2[k], S2[k]));
    }
#pragma unroll
    for (int k = 0; k < 8; k++) {
        cx<T> a0 = d[0][k], a1 = d[1][k], a2 = d[2][k];
        dft3<DIR>(a0, a1, a2);
        x[k] = a0;
        x[k + 8] = a1;
        x[k + 16] = a2;
    }
}

template <int R, int DIR, typename T>
__device__ __forceinline__ void dft(cx<T>* x)
{
    if constexpr (R == 2) dft2<DIR>(x[0], x[1]);
    else if constexpr (R == 3) dft3<DIR>(x[0], x[1], x[2]);
    else if constexpr (R == 4) dft4<DIR>(x[0], x[1], x[2], x[3]);
    else if constexpr (R == 6) dft6<DIR>(x);
    else if constexpr (R == 8) dft8<DIR>(x);
    else if constexpr (R == 12) dft12<DIR>(x);
    else if constexpr (R == 16) dft16<DIR>(x);
    else if constexpr (R == 24) dft24<DIR>(x);
    else static_assert(R == 2, "unsupported radix");
}

// ---------------------------------------------------------------------------
// Radix plans for the supported transform lengths (2^a 3^b, b <= 2).
// ---------------------------------------------------------------------------
template <int L>
struct Plan;
#define SPNS_PLAN(LEN, ...)                                                \
    template <>                                                           \
    struct Plan<LEN> {                                                    \
        static constexpr int n = sizeof((int[]){__VA_ARGS__}) / sizeof(int); \
        __host__ __device__ static constexpr int r(int s)                 \
        {                                                                 \
            constexpr int a[] = {__VA_ARGS__};                            \
            return a[s];                                                  \
        }                                                                 \
    };
SPNS_PLAN(16, 4, 4)
SPNS_PLAN(24, 3, 8)
SPNS_PLAN(32, 4, 8)
SPNS_PLAN(36, 3, 12)
SPNS_PLAN(48, 6, 8)
SPNS_PLAN(64, 8, 8)
SPNS_PLAN(72, 6, 12)
SPNS_PLAN(96, 12, 8)
SPNS_PLAN(128, 2, 8, 8)
SPNS_PLAN(144, 12, 12)
SPNS_PLAN(192, 3, 8, 8)
SPNS_PLAN(256, 4, 8, 8)
SPNS_PLAN(288, 6, 6, 8)
SPNS_PLAN(384, 6, 8, 8)
SPNS_PLAN(512, 8, 8, 8)
SPNS_PLAN(576, 12, 6, 8)
SPNS_PLAN(768, 12, 8, 8)
SPNS_PLAN(1024, 16, 8, 8)
SPNS_PLAN(1152, 12, 12, 8)
SPNS_PLAN(1536, 12, 16, 8)
#undef SPNS_PLAN

template <int L, int S>
__host__ __device__ constexpr int plan_span()  // product of the radices before stage S
{
    int p = 1;
    for (int s = 0; s < S; s++) p *= Plan<L>::r(s);
    return p;
}

// ---------------------------------------------------------------------------
// Batched Stockham FFT through LDS.
//   T: compute type; L: length; NB: columns; NT: threads; LD: LDS row pitch
//   DIR: -1 forward exp(-i..), +1 inverse; SYNC0: the first stage's loader
//   reads LDS (barrier between its reads and its writes).
// lds must hold L*LD elements.  tw[m] = exp(-2 pi i m / L) (forward table).
// Thread mapping: consecutive threads take consecutive columns b, so the
// loader's global reads are coalesced when columns are contiguous.
// ---------------------------------------------------------------------------
template <typename T, int L, int NB, int NT, int LD, int DIR, bool SYNC0 = false>
struct Fft {
    using C = cx<T>;
    static constexpr int NS = Plan<L>::n;

    template <int S, class Load, class Store>
    __device__ static __forceinline__ void stage(C* __restrict__ lds, const C* __restrict__ tw, Load& load, Store& store)
    {
        constexpr int R = Plan<L>::r(S);
        constexpr int Ns = plan_span<L, S>();
        constexpr int LR = L / R;
        constexpr int NBF = LR * NB;               // butterflies in this stage
        constexpr int BPT = (NBF + NT - 1) / NT;   // per thread
        constexpr int TWS = L / (Ns * R);          // twiddle index stride
        const int tid = threadIdx.x;
        C v[BPT][R];
#pragma unroll
        for (int q = 0; q < BPT; q++) {
            const int idx = tid + q * NT;
            if (NBF % NT == 0 || idx < NBF) {
                const int b = idx % NB, j = idx / NB;
#pragma unroll
                for (int r = 0; r < R; r++) {
                    if constexpr (S == 0) v[q][r] = load(b, j + r * LR);
                    else v[q][r] = lds[(j + r * LR) * LD + b];
                }
            }
        }
        if constexpr (S > 0 || SYNC0) __syncthreads();
#pragma unroll
        for (int q = 0; q < BPT; q++) {
            const int idx = tid + q * NT;
            if (NBF % NT == 0 || idx < NBF) {
                const int b = idx % NB, j = idx / NB;
                const int k = j % Ns;
                if constexpr (S > 0) {
#pragma unroll
                    for (int r = 1; r < R; r++) {
                        C w = tw[r * k * TWS];
                        if (DIR > 0) w.y = -w.y;
                        v[q][r] = cmul(v[q][r], w);
                    }
                }
                dft<R, DIR>(v[q]);
                const int base = (j / Ns) * Ns * R + k;
#pragma unroll
                for (int r = 0; r < R; r++) {
                    if constexpr (S == NS - 1) store(b, base + r * Ns, v[q][r]);
                    else lds[(base + r * Ns) * LD + b] = v[q][r];
                }
            }
        }
        if constexpr (S < NS - 1) {
            __syncthreads();
            stage<S + 1>(lds, tw, load, store);
        }
    }

    template <class Load, class Store>
    __device__ static __forceinline__ void run(C* __restrict__ lds, const C* __restrict__ tw, Load load, Store store)
    {
        __syncthreads();  // LDS may still be read by a previous use
        stage<0>(lds, tw, load, store);
    }
};


// ---------------------------------------------------------------------------
// LDS-only block barrier: waits for this wave's LDS operations, then meets
// the other waves of the block.  Unlike __syncthreads it does not wait for
// outstanding global-memory operations, so asynchronous global->LDS copies
// (dma16) issued earlier keep flowing across it.  The memory clobber keeps
// the compiler from moving memory accesses across the barrier.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Asynchronous copy of 16 bytes per lane from global memory to LDS
// (gfx950 global_load_lds_dwordx4): active lane i's bytes land at LDS byte
// address lds_base + 16*i; lds_base must be wave-uniform.  The copy is
// counted by vmcnt only: dma_wait() then a barrier publish it to the block.
// Issued from inline assembly so the compiler does not serialise later LDS
// reads of unrelated buffers behind it.  M0 is reserved to the compiler, so
// the sequence saves and restores it; s_nop gives the one wait state the
// LDS-DMA instruction needs after an M0 write.
__device__ __forceinline__ void dma16(const void* g, uint32_t lds_base)
{
    uint32_t saved;
    asm volatile(
        "s_mov_b32 %0, m0\n\ts_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %2, off\n\ts_mov_b32 m0, %0"
        : "=&s"(saved)
        : "s"(lds_base), "v"(g)
        : "memory");
}

// An opaque copy of v: the compiler cannot see through it, so values derived
// from it are not hoisted out of the enclosing loop.
__device__ __forceinline__ int opaque(int v)
{
    asm volatile("" : "+v"(v));
    return v;
}

__device__ __forceinline__ void dma_wait()
{
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

template <typename P>
__device__ __forceinline__ uint32_t lds_addr(P* p)
{
    return (uint32_t)(uintptr_t)(__attribute__((address_space(3))) void*)(p);
}

// ---------------------------------------------------------------------------
// Batched Stockham FFT like Fft, for kernels that overlap asynchronous
// global->LDS copies with the transform:
//   * stages meet at lds_barrier() (no wait on global memory);
//   * twiddles come from an LDS copy of the forward table; in FP64 each
//     butterfly reads only w = tw[k*TWS] and forms w^r by repeated complex
//     multiplication (error a few ulp, LDS traffic 1/(R-1) of a full read);
//   * hook() runs once, right after the barrier that ends stage 0, when
//     every wave has finished reading stage 0's inputs.
// ---------------------------------------------------------------------------
// Number of leading entries of the forward twiddle table that FftS with
// twiddle mode TWM (below) reads for length L.
template <int L, int TWM>
__host__ __device__ constexpr int fft_twiddle_span()
{
    int m = 1;
    for (int S = 1; S < Plan<L>::n; S++) {
        const int R = Plan<L>::r(S);
        int Ns = 1;
        for (int s = 0; s < S; s++) Ns *= Plan<L>::r(s);
        const int TWS = L / (Ns * R);
        int rmax = R - 1;
        if (TWM == 1) rmax = 1;
        if (TWM == 2) {
            rmax = 1;
            while (2 * rmax <= R - 1) rmax *= 2;
        }
        const int n = rmax * (Ns - 1) * TWS + 1;
        m = n > m ? n : m;
    }
    return m;
}

// Swizzled 8-column layout of FftS (SWZ) for elements of ES bytes: rows are
// grouped in 256-byte blocks of RPB = 256 / (8 ES) rows (4 in FP32, 2 in
// FP64), and element (row, b) sits at
// (8 row + b) ^ ((row / RPB) mod 8 RPB), inside its own block.
template <int ES>
__host__ __device__ constexpr int fft_swz_rpb()
{
    return 256 / (8 * ES);
}

template <int ES>
__host__ __device__ __forceinline__ constexpr int fft_swz(int row, int b)
{
    constexpr int RPB = fft_swz_rpb<ES>(), SH = RPB == 4 ? 2 : 1;
    return (8 * row + b) ^ ((row >> SH) & (8 * RPB - 1));
}

// The swizzled stages index rows j + r*(L/R) (reads) and base + r*Ns
// (writes) without per-access shifts: every L/R and every Ns after stage 0
// a multiple of RPB, and the first radix too (stage-0 rows R j + r).
template <int L, int ES>
__host__ __device__ constexpr bool fft_swz_ok()
{
    constexpr int RPB = fft_swz_rpb<ES>();
    if (RPB != 2 && RPB != 4) return false;
    if (Plan<L>::r(0) % RPB != 0) return false;
    int ns = 1;
    for (int s = 0; s < Plan<L>::n; s++) {
        const int R = Plan<L>::r(s);
        if (s > 0 && ((L / R) % RPB != 0 || ns % RPB != 0)) return false;
        ns *= R;
    }
    return L % RPB == 0;
}

// TWM, the twiddle source of a butterfly's R-1 factors W^r:
//   0: R-1 reads of the table;
//   1: one read (W) and W^r by repeated multiplication (R-2 products, error
//      up to R-2 roundings);
//   2: reads of the power-of-two exponents and one product for each other r
//      (W^r = W^h W^(r-h), h the highest power of two below r; at most
//      popcount(r)-1 roundings).
// S0COL: stage 0 maps consecutive threads to consecutive rows of one column
// (b = idx / (L/R0), j = idx % (L/R0)) instead of across columns, so with
// L/R0 a multiple of 64 every wave of stage 0 works on a single column and a
// column-dependent loader runs without divergence.
// Stage 0 calls load(b, j, std::integral_constant<int, r>) for input
// j + r*(L/R0) when the functor accepts that form, else load(b, j + r*(L/R0)).
// SWZ: the LDS buffer holds 8 columns with element (row, b) at
// fft_swz<sizeof(C)>(row, b) instead of row*LD + b.  The XOR stays inside
// aligned 256-byte blocks, so the buffer is exactly L*8 elements.  The
// transform's access patterns (8 columns x consecutive rows, the stage-0
// writes of rows R j + r across lanes j) and a thread-per-row access of one
// column then spread over all 64 banks, where the plain pitch-9 layout
// conflicts two- to four-way.  The last stage passes the element's index to
// store(b, p, index, v).  Needs fft_swz_ok<L, sizeof(C)>() (the index algebra
// below).
template <typename T, int L, int NB, int NT, int LD, int DIR, bool SYNC0 = false, int TWM = (sizeof(T) == 8 ? 1 : 2),
          bool S0COL = false, bool SWZ = false>
struct FftS {
    using C = cx<T>;
    static constexpr int NS = Plan<L>::n;
    static_assert(NS >= 2, "FftS needs at least two stages");
    static_assert(!SWZ || (NB == 8 && LD == 8 && fft_swz_ok<L, (int)sizeof(C)>()), "swizzled layout: 8 columns, pitch 8");
    static constexpr int RPB = fft_swz_rpb<(int)sizeof(C)>();  // rows per 256-byte block (SWZ)
    static constexpr int SH = RPB == 4 ? 2 : 1, SM = 8 * RPB - 1;

    template <int S, class Load, class Store, class Hook>
    __device__ static __forceinline__ void stage(C* __restrict__ lds, const C* __restrict__ twl, Load& load, Store& store,
                                                 Hook& hook, const int tid)
    {
        constexpr int R = Plan<L>::r(S);
        constexpr int Ns = plan_span<L, S>();
        constexpr int LR = L / R;
        constexpr int NBF = LR * NB;
        constexpr int BPT = (NBF + NT - 1) / NT;
        constexpr int TWS = L / (Ns * R);
        C v[BPT][R];
#pragma unroll
        for (int q = 0; q < BPT; q++) {
            const int idx = tid + q * NT;
            if (NBF % NT == 0 || idx < NBF) {
                const int b = (S == 0 && S0COL) ? idx / LR : idx % NB;
                const int j = (S == 0 && S0COL) ? idx % LR : idx / NB;
                __builtin_assume(j >= 0 && j < LR);  // idx < NBF = LR * NB
                __builtin_assume(b >= 0 && b < NB);
                if constexpr (S == 0 && std::is_invocable_v<Load&, int, int, std::integral_constant<int, 0>>) {
                    // load(b, j, r as a compile-time constant): the loader can
                    // resolve per-input structure (zero padding, mirroring) statically
                    auto ld = [&](auto rc) { v[q][decltype(rc)::value] = load(b, j, rc); };
                    static_for<R>(ld);
                } else {
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        if constexpr (S == 0) {
                            v[q][r] = load(b, j + r * LR);
                        } else if constexpr (SWZ) {
                            // row j + r LR: row / RPB = j / RPB + r LR / RPB, and 8 LR r is a
                            // multiple of the block, outside the bits the XOR touches
                            const int h = ((j >> SH) + ((r * (LR / RPB)) & SM)) & SM;
                            v[q][r] = lds[((8 * j + b) ^ h) + 8 * LR * r];
                        } else {
                            v[q][r] = lds[(j + r * LR) * LD + b];
                        }
                    }
                }
            }
        }
        if constexpr (S > 0 || SYNC0) lds_barrier();
#pragma unroll
        for (int q = 0; q < BPT; q++) {
            const int idx = tid + q * NT;
            if (NBF % NT == 0 || idx < NBF) {
                const int b = (S == 0 && S0COL) ? idx / LR : idx % NB;
                const int j = (S == 0 && S0COL) ? idx % LR : idx / NB;
                __builtin_assume(j >= 0 && j < LR);  // idx < NBF = LR * NB
                __builtin_assume(b >= 0 && b < NB);
                const int k = j % Ns;
                if constexpr (S > 0) {
                    if constexpr (TWM == 1) {
                        C w1 = twl[k * TWS];
                        if (DIR > 0) w1.y = -w1.y;
                        C w = w1;
#pragma unroll
                        for (int r = 1; r < R; r++) {
                            v[q][r] = cmul(v[q][r], w);
                            if (r + 1 < R) w = cmul(w, w1);
                        }
                    } else if constexpr (TWM == 2) {
                        C w[R];
#pragma unroll
                        for (int r = 1; r < R; r++) {
                            if ((r & (r - 1)) == 0) {
                                w[r] = twl[r * k * TWS];
                                if (DIR > 0) w[r].y = -w[r].y;
                            } else {
                                int h = 1;
                                while (2 * h < r) h *= 2;
                                w[r] = cmul(w[h], w[r - h]);
                            }
                            v[q][r] = cmul(v[q][r], w[r]);
                        }
                    } else {
#pragma unroll
                        for (int r = 1; r < R; r++) {
                            C w = twl[r * k * TWS];
                            if (DIR > 0) w.y = -w.y;
                            v[q][r] = cmul(v[q][r], w);
                        }
                    }
                }
                dft<R, DIR>(v[q]);
                const int base = (j / Ns) * Ns * R + k;
                // swizzled index of row base + r Ns (SWZ): for Ns a multiple of
                // RPB as for the reads; at stage 0 (Ns = 1, base = R j with R a
                // multiple of RPB) the rows are base + r, row / RPB = base / RPB + r / RPB
                auto widx = [&](int r) -> int {
                    if constexpr (Ns % RPB == 0) {
                        return ((8 * base + b) ^ (((base >> SH) + ((r * (Ns / RPB)) & SM)) & SM)) + 8 * Ns * r;
                    } else {
                        static_assert(!SWZ || Ns == 1, "swizzled stage rows");
                        return 8 * base + ((8 * r + b) ^ (((base >> SH) + (r >> SH)) & SM));
                    }
                };
#pragma unroll
                for (int r = 0; r < R; r++) {
                    if constexpr (S == NS - 1) {
                        if constexpr (SWZ) store(b, base + r * Ns, widx(r), v[q][r]);
                        else store(b, base + r * Ns, v[q][r]);
                    } else if constexpr (SWZ) {
                        lds[widx(r)] = v[q][r];
                    } else {
                        lds[(base + r * Ns) * LD + b] = v[q][r];
                    }
                }
            }
        }
        if constexpr (S < NS - 1) {
            lds_barrier();
            if constexpr (S == 0) hook();
            stage<S + 1>(lds, twl, load, store, hook, tid);
        }
    }

    // LEAD: start with a barrier (the buffer may still be read by a previous
    // use).  tid is the thread's index in the block; callers that loop may
    // pass it through an opaque copy so the per-thread index arithmetic is
    // recomputed per call rather than held in registers across the loop.
    template <bool LEAD, class Load, class Store, class Hook>
    __device__ static __forceinline__ void run(C* __restrict__ lds, const C* __restrict__ twl, Load load, Store store,
                                               Hook hook, const int tid)
    {
        if constexpr (LEAD) lds_barrier();
        stage<0>(lds, twl, load, store, hook, tid);
    }
};

// ---------------------------------------------------------------------------
// One-wave FFT of a single length-L column (L = 64*E).  The first radix is E,
// so stage 0 reads points t + 64 r (r < E) of lane t: one coalesced load per
// r.  Every later radix divides E, so the last stage leaves lane t holding
// exactly the points t + 64 m (m < E); the store functor receives m as a
// compile-time slot and can keep the values in registers.  Exchanges use the
// wave's own LDS column with wave-level ordering only (no block barrier).
// ---------------------------------------------------------------------------
template <int L>
struct WavePlan;
#define SPNS_WPLAN(LEN, ...)                                               \
    template <>                                                           \
    struct WavePlan<LEN> {                                                \
        static constexpr int n = sizeof((int[]){__VA_ARGS__}) / sizeof(int); \
        __host__ __device__ static constexpr int r(int s)                 \
        {                                                                 \
            constexpr int a[] = {__VA_ARGS__};                            \
            return a[s];                                                  \
        }                                                                 \
    };
SPNS_WPLAN(512, 8, 8, 8)
SPNS_WPLAN(768, 12, 4, 4, 4)
SPNS_WPLAN(1024, 16, 8, 8)
SPNS_WPLAN(1536, 24, 8, 8)
#undef SPNS_WPLAN

template <int L>
struct has_wave_plan {
    static constexpr bool value = (L == 512 || L == 768 || L == 1024 || L == 1536);
};

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T, int L, int DIR>
struct WaveFft {
    using C = cx<T>;
    static constexpr int E = L / 64;
    static constexpr int NS = WavePlan<L>::n;

    template <int S>
    __host__ __device__ static constexpr int span()
    {
        int p = 1;
        for (int s = 0; s < S; s++) p *= WavePlan<L>::r(s);
        return p;
    }

    template <int S, class Load, class Store>
    __device__ static __forceinline__ void stage(C* __restrict__ lds, const C* tw, Load& load, Store& store)
    {
        constexpr int R = WavePlan<L>::r(S);
        constexpr int Ns = span<S>();
        constexpr int LR = L / R;
        constexpr int BPT = LR / 64;
        constexpr int TWS = L / (Ns * R);
        static_assert(LR % 64 == 0, "wave plan: every radix must divide L/64");
        const int t = threadIdx.x & 63;
        C v[BPT][R];
#pragma unroll
        for (int q = 0; q < BPT; q++) {
            const int j = t + 64 * q;
#pragma unroll
            for (int r = 0; r < R; r++) {
                if constexpr (S == 0) v[q][r] = load(j + r * LR, r);
                else v[q][r] = lds[j + r * LR];
            }
        }
        if constexpr (S > 0) wave_sync();
#pragma unroll
        for (int q = 0; q < BPT; q++) {
            const int j = t + 64 * q;
            const int k = j % Ns;
            if constexpr (S > 0) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    C w = tw[r * k * TWS];
                    if (DIR > 0) w.y = -w.y;
                    v[q][r] = cmul(v[q][r], w);
                }
            }
            dft<R, DIR>(v[q]);
            const int base = (j / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) {
                if constexpr (S == NS - 1) store(base + r * Ns, v[q][r], q + r * (E / R));
                else lds[base + r * Ns] = v[q][r];
            }
        }
        if constexpr (S < NS - 1) {
            wave_sync();
            stage<S + 1>(lds, tw, load, store);
        }
    }

    template <class Load, class Store>
    __device__ static __forceinline__ void run(C* __restrict__ lds, const C* tw, Load load, Store store)
    {
        // The twiddles depend only on the lane; without this compiler barrier
        // they are hoisted out of the callers' loops and stay live in
        // registers for the whole kernel.
        asm volatile("" ::: "memory");
        wave_sync();
        stage<0>(lds, tw, load, store);
    }
};

}  // namespace spns
