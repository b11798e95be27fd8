// MI355X (gfx950) device layer: upload of the H2 representation into flat
// arrays, the execution plan, the HIP kernels of the matvec and of the
// Coulomb uncontract/contract, and the hot-path C-ABI entry points.
//
// Semantics follow H2P-ERI's AOT matvec (src/H2ERI_matvec.c:882-957):
//   upward sweep   y0[n] = U_n^T x_n (leaf) or U_n^T [y0 of children]   :93-175
//   coupling       three level cases of each admissible pair            :299-407
//   downward sweep t = U_n y1[n]; leaf: y += t, else push to children  :663-757
//   near field     diagonal y_n += D_nn x_n (no transpose half);
//                  inadmissible y_0 += D x_1, y_1 += D^T x_0           :572-660
//   low-rank block out0 += U (V in0), out1 += V^T (U^T in1)             :225-264
// and Coulomb: src/H2ERI_build_Coulomb.c:26-152.
//
// Device design (independent of the reference's OpenMP structure):
//   * every matrix block (U, dense B/D, both halves of low-rank B/D) lives in
//     one flat storage array per precision, row stride padded to the vector
//     width; the storage is uploaded once;
//   * work is a list of tiles ("bi-GEMV tasks"): tile x in0 -> out0 (row
//     dots), tile^T x in1 -> out1 (column sums), one workgroup each; the tiles
//     of one stage run in one launch;
//   * every tile output goes to a private slot; every intermediate vector
//     (y0, y1, the k-vectors of low-rank blocks, the output segments) is
//     "virtual": a fixed list of slots that the consuming tile sums once into
//     LDS.  Where a vector has many early slots and several consumers, a fold
//     task (a fixed-order sum into one slot) runs in the launch before its
//     first consumer, beside that launch's tiles; the output y is written by
//     one final gather launch.  All sums run in a fixed order: results are
//     bit-reproducible and no atomics are used;
//   * the near-field tiles depend on x only and are balanced over the launches
//     of the sweeps' critical path;
//   * low-rank blocks take 1.5 passes: the smaller half twice, the other half
//     once with a fused bi-GEMV;
//   * NV vectors (1, 2, 4, 8, 16) are processed per pass over the matrix data
//     (at 16 the tiles are FP64 MFMA products);
//   * U and B may be stored in FP32 (accumulation stays FP64), D stays FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <numeric>
#include <mutex>
#include <queue>
#include <type_traits>

#include "h2eri_internal.h"

namespace h2e {

// ---------------------------------------------------------------------------
// Descriptors
// ---------------------------------------------------------------------------
// Vector sources (vsrc entries): e >= 0 is an internal vector at element
// offset e of the workspace (NV-interleaved: element i of vector v lives at
// Wv[(e + i) * NV + v]); e < 0 is the external input at element offset
// -e - 1 (vector v at X[v * ldx + ...]); an external source is always the
// only source of its vector.
struct BTask {
    long long mat;   // element offset into the storage array of its precision
    int nrow, ncol, ld;
    int lpr;         // lanes per row: 4, 8, 16, 32 or 64
    int v0b, v0e, v0off;   // in0 (length ncol): sum of vsrc[v0b..v0e) at element offset v0off
    int v1b, v1e, v1off;   // in1 (length nrow)
    int out0, out1;  // slot offsets (-1: unused)
    int pad_[2];
};
static_assert(sizeof(BTask) == 64, "BTask layout");

// Reduce task: dst[i] = sum over rsrc[src_beg..src_end) of Wv[src + soff + i].
struct RTask {
    long long dst;   // element offset: external output (flags & 1) or workspace
    int len;         // elements: <= the plan's reduce chunk (len * NV <= 1024)
    int src_beg, src_end;
    int soff;        // element offset of this chunk inside every source slot
    int flags;       // 1: external output
    int pad_;
};
static_assert(sizeof(RTask) == 32, "RTask layout");

// Task descriptors are the same for every lane of a wave.  Reading them
// through v_readfirstlane lets the compiler keep them in scalar registers
// and treat branches on them as uniform; a branch it must assume divergent
// keeps the registers of both sides live at once.
__device__ __forceinline__ int ufl(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ long long ufl(long long v) {
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(v & 0xffffffff));
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(v >> 32));
    return (long long)(((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ BTask uniform(const BTask &a) {
    BTask t;
    t.mat = ufl(a.mat);
    t.nrow = ufl(a.nrow); t.ncol = ufl(a.ncol); t.ld = ufl(a.ld); t.lpr = ufl(a.lpr);
    t.v0b = ufl(a.v0b); t.v0e = ufl(a.v0e); t.v0off = ufl(a.v0off);
    t.v1b = ufl(a.v1b); t.v1e = ufl(a.v1e); t.v1off = ufl(a.v1off);
    t.out0 = ufl(a.out0); t.out1 = ufl(a.out1);
    t.pad_[0] = t.pad_[1] = 0;
    return t;
}
__device__ __forceinline__ RTask uniform(const RTask &a) {
    RTask t;
    t.dst = ufl(a.dst);
    t.len = ufl(a.len); t.src_beg = ufl(a.src_beg); t.src_end = ufl(a.src_end);
    t.soff = ufl(a.soff); t.flags = ufl(a.flags); t.pad_ = 0;
    return t;
}

static const int WG = 256;       // threads per workgroup (4 waves)
// Output elements per reduce task: 1024 / NV values fill the task's 256
// threads four deep.  Both executors use the same chunks, so their plans and
// summation orders are identical.
static inline int rchunk_for(int nv) { return 1024 / nv; }
static const int LDS_DOUBLES = 4096;   // 32 KB of LDS per workgroup
static const int XIN_DOUBLES = 1024;   // in0 staging (ncol * NV <= 1024)
static const int YIN_DOUBLES = LDS_DOUBLES - XIN_DOUBLES;   // in1 staging (nrow * NV)

// Per (storage type, NV) tile configuration: W elements per vector load, KP
// loads per row and lane (columns per lane CPL = W * KP), RU rows per block
// (two blocks of RU*KP loads are in flight).  CPL * NV <= 16 keeps the
// cross-wave reduction in 32 KB of LDS; the sizes keep the kernel within 128
// VGPRs (four waves per SIMD).
template <typename T, int NV> struct Cfg;
template <> struct Cfg<double, 1> { typedef double2 V; static const int W = 2, KP = 4, RU = 2; };
template <> struct Cfg<double, 2> { typedef double2 V; static const int W = 2, KP = 4, RU = 1; };
template <> struct Cfg<double, 4> { typedef double2 V; static const int W = 2, KP = 2, RU = 2; };
template <> struct Cfg<double, 8> { typedef double2 V; static const int W = 2, KP = 1, RU = 2; };
template <> struct Cfg<float, 1> { typedef float4 V; static const int W = 4, KP = 2, RU = 4; };
template <> struct Cfg<float, 2> { typedef float4 V; static const int W = 4, KP = 2, RU = 2; };
template <> struct Cfg<float, 4> { typedef float4 V; static const int W = 4, KP = 1, RU = 2; };
template <> struct Cfg<float, 8> { typedef float2 V; static const int W = 2, KP = 1, RU = 2; };

// columns per lane for a batch width (same for both storage types)
static inline int cpl_of(int nv) { return nv <= 2 ? 8 : (nv == 4 ? 4 : (nv == 8 ? 2 : 1)); }

__device__ __forceinline__ double vget(const double2 &v, int i) { return i == 0 ? v.x : v.y; }
__device__ __forceinline__ double vget(const float2 &v, int i) { return (double)(i == 0 ? v.x : v.y); }
__device__ __forceinline__ double vget(const float4 &v, int i) {
    return (double)(i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w)));
}

// Workspace accesses.  COH = true (persistent executor, where a producer and
// its consumer may run on different XCDs within one launch) makes them
// agent-scope relaxed atomics: coherent at device scope without any L2
// writeback or invalidation; the ordering against the dependency counters
// comes from waiting for the stores' completion before a counter moves.
template <bool COH>
__device__ __forceinline__ double wld(const double *p) {
    if constexpr (COH) return __hip_atomic_load(const_cast<double *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else return *p;
}
template <bool COH>
__device__ __forceinline__ void wst(double *p, double v) {
    if constexpr (COH) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else *p = v;
}

// Stages n elements (NV values each) of the virtual vector vsrc[b..e) at
// element offset `off` into dst[n * NV] (LDS).  From the workspace the
// threads take values q = tid, tid + WG, ... (contiguous in its
// NV-interleaved layout), so every thread works for any NV, and each sums
// the slots in list order with eight loads in flight.  From the external
// input the threads take elements and read the NV vectors.
template <int NV, bool COH, int NT, int XS = NV>
__device__ __forceinline__ void stage_vec(const int *__restrict__ vsrc, int b, int e, int off, int n,
                                          const double *__restrict__ Wv, const double *__restrict__ X, long long ldx,
                                          double *dst, int tid) {
    const int s0 = vsrc[b];
    if (s0 < 0 && ldx < 0) {   // the external input as an element-major copy (NV values per element)
        const double *p = X + (long long)(-s0 - 1 + off) * NV;
        for (int q = tid; q < n * NV; q += NT) dst[XS == NV ? q : (q / NV) * XS + q % NV] = p[q];
        return;
    }
    if (s0 < 0) {
        const double *p = X + (long long)(-s0 - 1 + off);
        for (int i = tid; i < n; i += NT) {
            #pragma unroll
            for (int v = 0; v < NV; v++) dst[i * XS + v] = p[(long long)v * ldx + i];
        }
        return;
    }
    const int nq = n * NV;
    for (int q = tid; q < nq; q += NT) {
        const double *base = Wv + (long long)off * NV + q;
        double sum = 0.0;
        int j = b;
        for (; j + 7 < e; j += 8) {
            double a[8];
            #pragma unroll
            for (int u = 0; u < 8; u++) a[u] = wld<COH>(base + (long long)vsrc[j + u] * NV);
            #pragma unroll
            for (int u = 0; u < 8; u++) sum += a[u];
        }
        for (; j < e; j++) sum += wld<COH>(base + (long long)vsrc[j] * NV);
        dst[XS == NV ? q : (q / NV) * XS + q % NV] = sum;
    }
}

// Synchronisation of the threads of one tile: the workgroup, or one wave
// (a wave's LDS operations complete in order; the fences keep the compiler
// from moving memory operations across the point).
template <int NT>
__device__ __forceinline__ void tile_sync() {
    if constexpr (NT == WG) {
        __syncthreads();
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// A tile of NT threads: the workgroup (NT = WG) or one wave (NT = 64, a
// "wave tile": four of them run side by side in a workgroup, each with a
// quarter of the LDS).
template <int NT> struct TileLDS {
    static const int X = NT == WG ? XIN_DOUBLES : XIN_DOUBLES / 4;      // in0 staging (ncol * NV)
    static const int Y = NT == WG ? YIN_DOUBLES : YIN_DOUBLES / 4;      // in1 staging (nrow * NV)
};

// Row-side-only tile at NV = 8 (y = M x with no transposed product: the
// down sweep and the first pass of a low-rank block): RD_L = 4 lanes share a
// row and each holds RB = 2 rows, so a lane keeps 2 x 8 sums and reads its x
// values from LDS per column instead of holding a register copy of its
// columns (what limits the general tile at NV = 8).  Two register buffers
// take the steps alternately, so the loads of the next two steps are in
// flight while a step is consumed (a step is short, and a rotating buffer
// would wait for its loads at every copy).  The four partial sums of a row
// are combined by a butterfly; lane q % RD_L stores value q of the row pair.
static const int RD_L = 4;
template <typename T, int NV, bool COH, int XS, int W>
__device__ void rowdot_tile(const BTask &t, const T *__restrict__ S, double *__restrict__ Wv, const double *xin) {
    constexpr int RB = 2;
    // W columns per lane and step, as NL loads of E elements
    constexpr int E = (W * sizeof(T) >= 16) ? 16 / (int)sizeof(T) : W;
    constexpr int NL = W / E;
    typedef typename std::conditional<sizeof(T) == 8, double2, typename std::conditional<E == 4, float4, float2>::type>::type V;
    constexpr int NG = WG / RD_L;
    const int tid = threadIdx.x, lr = tid % RD_L, rg = tid / RD_L;
    const int R = t.nrow, C = t.ncol;
    const int nstep = (C + RD_L * W - 1) / (RD_L * W);
    const int cmax = ((C - 1) / E) * E;   // last load position in a row
    const T *base = S + t.mat;
    for (int r0 = 0; r0 < R; r0 += NG * RB) {
        const T *rowp[RB];
        #pragma unroll
        for (int b = 0; b < RB; b++) rowp[b] = base + (long long)min(r0 + rg + NG * b, R - 1) * t.ld;
        double d[RB][NV];
        #pragma unroll
        for (int b = 0; b < RB; b++)
            #pragma unroll
            for (int v = 0; v < NV; v++) d[b][v] = 0.0;
        // steps past the end re-read the last load position
        auto load = [&](V (&buf)[RB][NL], int k) {
            const int c0 = (k * RD_L + lr) * W;
            #pragma unroll
            for (int b = 0; b < RB; b++)
                #pragma unroll
                for (int l = 0; l < NL; l++) buf[b][l] = *reinterpret_cast<const V *>(rowp[b] + min(c0 + l * E, cmax));
        };
        auto consume = [&](const V (&m)[RB][NL], int k) {
            const int c0 = (k * RD_L + lr) * W;
            #pragma unroll
            for (int w = 0; w < W; w++) {
                const int col = c0 + w;
                const bool in = col < C;
                const double *xc = xin + min(col, C - 1) * XS;
                double xv[NV];
                #pragma unroll
                for (int v = 0; v < NV; v++) xv[v] = xc[v];
                #pragma unroll
                for (int b = 0; b < RB; b++) {
                    const double mv = in ? vget(m[b][w / E], w % E) : 0.0;
                    #pragma unroll
                    for (int v = 0; v < NV; v++) d[b][v] = fma(mv, xv[v], d[b][v]);
                }
            }
        };
        // The loop body is straight-line: every load is issued and every step
        // consumed (a step past the end meets zero matrix values), and the
        // scheduling barriers keep the two buffers' loads in the same order
        // before and inside the loop, so the compiler's wait for one buffer's
        // loads does not also wait for the other's.
        V ba[RB][NL], bb[RB][NL];
        load(ba, 0);
        __builtin_amdgcn_sched_barrier(0);
        load(bb, 1);
        for (int k = 0; k < nstep; k += 2) {
            consume(ba, k);
            load(ba, k + 2);
            __builtin_amdgcn_sched_barrier(0);
            consume(bb, k + 1);
            load(bb, k + 3);
            __builtin_amdgcn_sched_barrier(0);
        }
        #pragma unroll
        for (int b = 0; b < RB; b++)
            #pragma unroll
            for (int v = 0; v < NV; v++)
                #pragma unroll
                for (int sft = 1; sft < RD_L; sft <<= 1) d[b][v] += __shfl_xor(d[b][v], sft, RD_L);
        #pragma unroll
        for (int q = 0; q < RB * NV; q++) {
            const int b = q / NV, v = q - b * NV;
            const int r = r0 + rg + NG * b;
            if (q % RD_L == lr && r < R) wst<COH>(Wv + (long long)(t.out0 + r) * NV + v, d[b][v]);
        }
    }
}

// Column-side-only tile at NV = 8 (M^T y with no row product: the up sweep
// and the last pass of a low-rank block).  A lane owns W = 2 adjacent columns
// (the plan's lanes per row cover the tile's columns at two per lane) and
// sums over the rows of its row group; it loads RC rows per block, and two
// register buffers take the blocks alternately, so two blocks' loads are in
// flight while a block is consumed.  Rows past the tile are clamped to its
// last row and meet a zero matrix value.  The column sums are combined across
// the row groups of a wave by shuffles and across the four waves through LDS,
// in a fixed order.
template <typename T, int NV, bool COH, int RC>
__device__ void coldot_tile(const BTask &t, const T *__restrict__ S, double *__restrict__ Wv, const double *yin,
                            double *lds) {
    constexpr int W = 2;
    typedef typename std::conditional<sizeof(T) == 8, double2, float2>::type V;
    const int lpr = t.lpr;
    const int tid = threadIdx.x, lane = tid & (lpr - 1), grp = tid / lpr, G = WG / lpr;
    const int R = t.nrow, C = t.ncol;
    const T *base = S + t.mat + min(lane * W, ((C - 1) / W) * W);
    double acc[W][NV];
    #pragma unroll
    for (int w = 0; w < W; w++)
        #pragma unroll
        for (int v = 0; v < NV; v++) acc[w][v] = 0.0;
    auto load = [&](V (&mm)[RC], int r0) {
        #pragma unroll
        for (int u = 0; u < RC; u++) mm[u] = *reinterpret_cast<const V *>(base + (long long)min(r0 + u * G, R - 1) * t.ld);
    };
    auto consume = [&](const V (&m)[RC], int r0) {
        #pragma unroll
        for (int u = 0; u < RC; u++) {
            const int r = r0 + u * G;
            const bool in = r < R;
            const double *yr = yin + min(r, R - 1) * NV;
            double y[NV];
            #pragma unroll
            for (int v = 0; v < NV; v++) y[v] = yr[v];
            #pragma unroll
            for (int w = 0; w < W; w++) {
                const double mv = in ? vget(m[u], w) : 0.0;
                #pragma unroll
                for (int v = 0; v < NV; v++) acc[w][v] = fma(mv, y[v], acc[w][v]);
            }
        }
    };
    // two buffers take the blocks alternately, with a straight-line loop
    // body (see rowdot_tile; rows past the end meet zero matrix values)
    V ba[RC], bb[RC];
    const int blk = RC * G;
    load(ba, grp);
    __builtin_amdgcn_sched_barrier(0);
    load(bb, grp + blk);
    for (int r0 = grp; r0 < R; r0 += 2 * blk) {
        consume(ba, r0);
        load(ba, r0 + 2 * blk);
        __builtin_amdgcn_sched_barrier(0);
        consume(bb, r0 + blk);
        load(bb, r0 + 3 * blk);
        __builtin_amdgcn_sched_barrier(0);
    }
    #pragma unroll
    for (int w = 0; w < W; w++)
        #pragma unroll
        for (int v = 0; v < NV; v++)
            for (int s = lpr; s < 64; s <<= 1) acc[w][v] += __shfl_xor(acc[w][v], s, 64);
    __syncthreads();   // the staged y is dead from here on
    double *red = lds;
    const int wave = tid >> 6, lw = tid & 63;
    constexpr int STRIDE = 64 * W * NV;   // doubles per wave
    static_assert(4 * STRIDE <= LDS_DOUBLES, "column sums fit the LDS");
    if (lw < lpr) {
        #pragma unroll
        for (int w = 0; w < W; w++)
            #pragma unroll
            for (int v = 0; v < NV; v++) red[wave * STRIDE + (lw * W + w) * NV + v] = acc[w][v];
    }
    __syncthreads();
    for (int q = tid; q < C * NV; q += WG) {
        const double s = (red[q] + red[STRIDE + q]) + (red[2 * STRIDE + q] + red[3 * STRIDE + q]);
        wst<COH>(Wv + (long long)t.out1 * NV + q, s);
    }
}

// NV = 8 paths: rows per block of the column tile (RC); the stride in
// doubles per element of the row-dot tile's staged x (XS, padded so that the
// four lanes of a row read different LDS banks) and its columns per lane and
// step (RW).
template <typename T> struct Dot8 { static const int RC = 4, XS = 10, RW = 4; };

template <typename T, int NV, bool COH, int NT = WG>
__device__ void bgemv_tile_valu(const BTask &task, const T *__restrict__ S, const int *__restrict__ vsrc,
                                double *__restrict__ Wv, const double *__restrict__ X, long long ldx, double *lds) {
    // At NV = 8 the descriptor stays as loaded: the register allocation of
    // the two tile paths together fits 128 VGPRs only that way (measured on
    // the compiler's output), and the path choice is made uniform below.
    const BTask t = NV == 8 ? task : uniform(task);
    typedef Cfg<T, NV> C_;
    typedef typename C_::V V;
    constexpr int W = C_::W, KP = C_::KP, RU = C_::RU, CPL = W * KP;
    const int lpr = t.lpr;
    const int tid = threadIdx.x & (NT - 1);
    const int lane = tid & (lpr - 1);
    const int grp = tid / lpr;
    const int G = NT / lpr;
    const int R = t.nrow, C = t.ncol;
    const bool do0 = t.out0 >= 0, do1 = t.out1 >= 0;
    double *xin = lds;
    double *yin = lds + TileLDS<NT>::X;
    if constexpr (NV == 8 && NT == WG) {
        // readfirstlane: a branch known to be uniform does not keep the
        // registers of both paths live at once
        if (__builtin_amdgcn_readfirstlane(t.out1) < 0) {
            stage_vec<NV, COH, NT, Dot8<T>::XS>(vsrc, t.v0b, t.v0e, t.v0off, C, Wv, X, ldx, lds, threadIdx.x);
            __syncthreads();
            rowdot_tile<T, NV, COH, Dot8<T>::XS, Dot8<T>::RW>(t, S, Wv, lds);
            return;
        }
        if (__builtin_amdgcn_readfirstlane(t.out0) < 0) {
            stage_vec<NV, COH, NT>(vsrc, t.v1b, t.v1e, t.v1off, R, Wv, X, ldx, yin, threadIdx.x);
            __syncthreads();
            coldot_tile<T, NV, COH, Dot8<T>::RC>(t, S, Wv, yin, lds);
            return;
        }
    }
    const T *base = S + t.mat;
    // Row blocks of RU rows; every lane issues RU*KP vector loads per block.
    // Loads are unconditional: rows past the tile are clamped to its last row
    // and column positions past it to its last vector position.  Such
    // entries meet a zero input (x0 = 0 for c >= C, y = 0 for r >= R) and
    // their outputs are not stored (the stored blocks hold finite values).
    const int cmax = ((C - 1) / W) * W;
    int coff[KP];
    #pragma unroll
    for (int k = 0; k < KP; k++) coff[k] = min((k * lpr + lane) * W, cmax);
    auto load_rows = [&](V (&mm)[RU][KP], int r0) {
        #pragma unroll
        for (int u = 0; u < RU; u++) {
            const T *row = base + (long long)min(r0 + u * G, R - 1) * t.ld;
            #pragma unroll
            for (int k = 0; k < KP; k++) mm[u][k] = *reinterpret_cast<const V *>(row + coff[k]);
        }
    };
    // Software pipeline: the next row block is always in flight while the
    // current one is consumed; the first is in flight during the staging.
    V nxt[RU][KP];
    load_rows(nxt, grp);

    // stage the input vectors once per tile (sums of their slots)
    if (do0) stage_vec<NV, COH, NT>(vsrc, t.v0b, t.v0e, t.v0off, C, Wv, X, ldx, xin, tid);
    if (do1) stage_vec<NV, COH, NT>(vsrc, t.v1b, t.v1e, t.v1off, R, Wv, X, ldx, yin, tid);
    tile_sync<NT>();

    double x0[KP][W][NV];
    double a1[KP][W][NV];
    #pragma unroll
    for (int k = 0; k < KP; k++)
        #pragma unroll
        for (int w = 0; w < W; w++) {
            const int c = (k * lpr + lane) * W + w;
            #pragma unroll
            for (int v = 0; v < NV; v++) {
                x0[k][w][v] = (do0 && c < C) ? xin[c * NV + v] : 0.0;
                a1[k][w][v] = 0.0;
            }
        }
    for (int r0 = grp; r0 < R; r0 += RU * G) {
        V m[RU][KP];
        #pragma unroll
        for (int u = 0; u < RU; u++)
            #pragma unroll
            for (int k = 0; k < KP; k++) m[u][k] = nxt[u][k];
        load_rows(nxt, r0 + RU * G);
        if (do1) {
            #pragma unroll
            for (int u = 0; u < RU; u++) {
                const int r = r0 + u * G;
                double y[NV];
                #pragma unroll
                for (int v = 0; v < NV; v++) y[v] = r < R ? yin[r * NV + v] : 0.0;
                #pragma unroll
                for (int k = 0; k < KP; k++)
                    #pragma unroll
                    for (int w = 0; w < W; w++) {
                        const double mv = vget(m[u][k], w);
                        #pragma unroll
                        for (int v = 0; v < NV; v++) a1[k][w][v] = fma(mv, y[v], a1[k][w][v]);
                    }
            }
        }
        if (do0) {
            double d[RU][NV];
            #pragma unroll
            for (int u = 0; u < RU; u++) {
                #pragma unroll
                for (int v = 0; v < NV; v++) d[u][v] = 0.0;
                #pragma unroll
                for (int k = 0; k < KP; k++)
                    #pragma unroll
                    for (int w = 0; w < W; w++) {
                        const double mv = vget(m[u][k], w);
                        #pragma unroll
                        for (int v = 0; v < NV; v++) d[u][v] = fma(mv, x0[k][w][v], d[u][v]);
                    }
            }
            // Reduce-scatter over the lpr lanes of the row group: each step
            // halves the values a lane holds (it keeps one half, sends the
            // other to its partner), so RU*NV sums cost about RU*NV shuffles
            // instead of RU*NV*log2(lpr); a plain butterfly finishes when
            // lanes outnumber values.
            constexpr int NVAL = RU * NV;
            double vals[NVAL];
            #pragma unroll
            for (int u = 0; u < RU; u++)
                #pragma unroll
                for (int v = 0; v < NV; v++) vals[u * NV + v] = d[u][v];
            int s = lpr >> 1, base = 0, held = NVAL;
            #pragma unroll
            for (int cnt = NVAL; cnt > 1; cnt >>= 1) {
                if (s > 0) {
                    const bool up = (lane & s) != 0;
                    const int half = cnt >> 1;
                    #pragma unroll
                    for (int i = 0; i < half; i++) {
                        const double send = up ? vals[i] : vals[i + half];
                        const double keep = up ? vals[i + half] : vals[i];
                        vals[i] = keep + __shfl_xor(send, s, lpr);
                    }
                    if (up) base += half;
                    held = half;
                    s >>= 1;
                }
            }
            const int wmask = s > 0 ? (s << 1) - 1 : 0;   // lanes that differ below this hold equal sums
            for (; s > 0; s >>= 1) vals[0] += __shfl_xor(vals[0], s, lpr);
            if ((lane & wmask) == 0) {
                #pragma unroll
                for (int i = 0; i < NVAL; i++) {
                    if (i < held) {
                        const int q = base + i;
                        const int u = q / NV, v = q - u * NV;
                        const int r = r0 + u * G;
                        if (r < R) wst<COH>(Wv + (long long)(t.out0 + r) * NV + v, vals[i]);
                    }
                }
            }
        }
    }
    if (do1) {
        // reduce across the row groups of a wave, then across the 4 waves
        #pragma unroll
        for (int k = 0; k < KP; k++)
            #pragma unroll
            for (int w = 0; w < W; w++)
                #pragma unroll
                for (int v = 0; v < NV; v++)
                    for (int s = lpr; s < 64; s <<= 1) a1[k][w][v] += __shfl_xor(a1[k][w][v], s, 64);
        if constexpr (NT != WG) {   // one wave: lanes below lpr hold the column sums
            if (lane == tid) {
                #pragma unroll
                for (int k = 0; k < KP; k++)
                    #pragma unroll
                    for (int w = 0; w < W; w++) {
                        const int c = (k * lpr + lane) * W + w;
                        if (c < C) {
                            #pragma unroll
                            for (int v = 0; v < NV; v++) wst<COH>(Wv + (long long)(t.out1 + c) * NV + v, a1[k][w][v]);
                        }
                    }
            }
            return;
        }
        __syncthreads();   // the staged in1 is dead from here on
        double *red = lds;
        const int wave = tid >> 6;
        const int lw = tid & 63;
        constexpr int STRIDE = 64 * CPL * NV;   // doubles per wave
        if (lw < lpr) {
            #pragma unroll
            for (int k = 0; k < KP; k++)
                #pragma unroll
                for (int w = 0; w < W; w++)
                    #pragma unroll
                    for (int v = 0; v < NV; v++) red[wave * STRIDE + ((k * lpr + lw) * W + w) * NV + v] = a1[k][w][v];
        }
        __syncthreads();
        for (int q = tid; q < C * NV; q += WG) {
            const double s = (red[q] + red[STRIDE + q]) + (red[2 * STRIDE + q] + red[3 * STRIDE + q]);
            wst<COH>(Wv + (long long)t.out1 * NV + q, s);
        }
    }
}

// ---------------------------------------------------------------------------
// 16 vectors per pass: the tile as FP64 MFMA products
// ---------------------------------------------------------------------------
// At 16 vectors a tile is a small matrix product, out0 (R x 16) = M X and
// out1 (C x 16) = M^T Y, every matrix element used 16 times per side.  Both
// run on v_mfma_f64_16x16x4f64 (lane l holds A[l%16][l/16] and B[l/16][l%16];
// output register r of lane l is C[l/16 + 4r][l%16]) with FP64 accumulation;
// FP32-stored elements are widened to FP64 in registers.
//
// Inputs are staged in LDS (an element's 16 values contiguous, the slots of
// a virtual vector summed once) when X and Y fit beside each other in the
// 4096 doubles of a workgroup.  A row product spans the whole width of its
// block, so that no partial sums over column tiles are left (each would be
// 16 values per output row for the consumers to read back); when its X does
// not fit the LDS it reads the operands from memory: the workspace is
// element-major, and at 16 vectors the pass's first launch makes an
// element-major copy of the external input, so the 16 lanes that need an
// element read one 128-byte run, and the tile's rows reuse it from the
// caches.
//  * Row products: a wave takes up to four blocks of 16 rows at a time (the
//    waves' blocks interleave, so a tile of 64 rows keeps all four busy) and
//    walks the columns in groups of 16.  For a group, lane l loads columns
//    4(l/16) .. 4(l/16)+3 of row l%16 of each block (one 16-byte load for
//    FP32, two for FP64): the A operands of four MFMAs j whose inner index
//    runs over the columns 4k + j of the group (a permutation of the inner
//    index, which leaves the sum alone).  The B operands (X at those
//    columns) serve all of the wave's blocks.
//  * Column products, out1^T = Y^T M, at most 128 columns: a span of 64
//    columns is shared by 2 or 4 waves, each taking blocks of 16 rows in
//    turn.  For a group of 4 rows, lane l loads columns
//    4(l%16) .. 4(l%16)+3 of row l/16 of the span: the B operands of four
//    MFMAs j, MFMA j covering the 16 columns 4n + j of the span; A = those
//    rows of Y.  Waves sharing a span add their sums in LDS in a fixed order.
//  * A tile with both outputs (mfma_tile16_2s) reads its rows once: the two
//    products need the matrix in transposed register layouts, and a wave
//    moves each 16 x 16 block from one to the other through 2 KB of LDS of
//    its own.
// Rows and columns past the tile are clamped to its last load position and
// meet zero operands; their outputs are not stored.
typedef double d4v __attribute__((ext_vector_type(4)));
static const int T16_MAXC = 128;   // columns of a tile with a column product (two spans of 64)
static const int T16_MAXC2 = 64;   // columns of a tile with both products (column sums in registers)
static const int T16_LDS = LDS_DOUBLES;   // staging doubles of a 16-vector tile (X and Y)

// Four consecutive elements from c0 (clamped so no load leaves the row's
// padded stride): FP32 one 16-byte load at min(c0, cmax), FP64 two at
// min(c0, cmax) and min(c0 + 2, cmax).
__device__ __forceinline__ void load_quad(const float *row, int c0, int cmax, double (&o)[4]) {
    const float4 v = *reinterpret_cast<const float4 *>(row + min(c0, cmax));
    o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
}
__device__ __forceinline__ void load_quad(const double *row, int c0, int cmax, double (&o)[4]) {
    const double2 a = *reinterpret_cast<const double2 *>(row + min(c0, cmax));
    const double2 b = *reinterpret_cast<const double2 *>(row + min(c0 + 2, cmax));
    o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
}

// A tile input at 16 vectors: vsrc[b..e) at element offset off.  Element i,
// vector v is the external input's element-major copy, or the sum of the
// listed workspace slots in list order.  The first slots' offsets are held
// in registers (uniform across the workgroup).
struct In16 {
    const int *vsrc;
    int b, e, off;
    int s[4];
    int ns;   // slots held in s (all of them when e - b <= 4), 0: external
};
__device__ __forceinline__ In16 in16_make(const int *__restrict__ vsrc, int b, int e, int off) {
    In16 r;
    r.vsrc = vsrc; r.b = b; r.e = e; r.off = off;
    const int s0 = ufl(vsrc[b]);
    r.ns = s0 < 0 ? 0 : min(e - b, 4);
    #pragma unroll
    for (int k = 0; k < 4; k++) r.s[k] = (k < r.ns) ? ufl(vsrc[b + k]) : s0;
    return r;
}
template <bool COH>
__device__ __forceinline__ double in16_get(const In16 &in, int i, int v, const double *__restrict__ Wv,
                                           const double *__restrict__ Xt) {
    if (in.ns == 0) return Xt[(long long)(-in.s[0] - 1 + in.off + i) * 16 + v];
    double sum = wld<COH>(Wv + (long long)(in.s[0] + in.off + i) * 16 + v);
    #pragma unroll
    for (int k = 1; k < 4; k++)
        if (k < in.ns) sum += wld<COH>(Wv + (long long)(in.s[k] + in.off + i) * 16 + v);
    for (int j = in.b + 4; j < in.e; j++) sum += wld<COH>(Wv + (long long)(in.vsrc[j] + in.off + i) * 16 + v);
    return sum;
}

// A tile with both products (at most 64 columns, any number of rows), in one
// walk over its rows.  Waves take blocks of 16 rows (block b of the tile on
// wave b % 4).  For a group of 16 columns, lane l loads columns
// 4(l/16) .. 4(l/16)+3 of row l%16 of the block: the row product's A
// operands (X staged in LDS supplies B, as in the row product above).  The
// lane writes them to the wave's 16 x 16 LDS tile and reads back row
// 4q + l/16, column l%16: the column product's B operand of MFMA q (rows
// 4q .. 4q+3 of the block), whose A operand is Y at those rows, read from
// memory once per block (element-major: 16 lanes read one 128-byte run).
// Each wave keeps its column sums in registers, one accumulator per group of
// 16 columns, until the end of the tile; then the four waves add them in LDS
// in a fixed order.
template <typename T, bool COH>
__device__ void mfma_tile16_2s(const BTask &t, const T *__restrict__ base, const int *__restrict__ vsrc,
                               double *__restrict__ Wv, const double *__restrict__ Xt, double *lds) {
    constexpr int NV = 16;
    constexpr int VW = sizeof(T) == 4 ? 4 : 2;   // elements per vector load
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int R = t.nrow, C = t.ncol;
    const int cmax = ((C - 1) / VW) * VW;
    const int Cp = (C + 15) & ~15, ng = Cp >> 4;   // groups of 16 columns, at most 4
    double *xin = lds;                                // Cp x 16, zero past C
    double *tr = lds + T16_MAXC * NV + 256 * w;       // this wave's 16 x 16 transposition tile
    stage_vec<NV, COH, WG>(vsrc, t.v0b, t.v0e, t.v0off, C, Wv, Xt, -1, xin, tid);
    for (int q = C * NV + tid; q < Cp * NV; q += WG) xin[q] = 0.0;
    __syncthreads();
    const In16 y = in16_make(vsrc, t.v1b, t.v1e, t.v1off);
    d4v acc1[T16_MAXC2 / 16];
    #pragma unroll
    for (int g = 0; g < T16_MAXC2 / 16; g++) acc1[g] = d4v{0.0, 0.0, 0.0, 0.0};
    for (int rb = 16 * w; rb < R; rb += 64) {
        double ya[4];   // Y at rows rb + 4q + l/16 (zero past the tile), vector l%16
        #pragma unroll
        for (int q = 0; q < 4; q++) {
            const int rr = rb + 4 * q + lq;
            ya[q] = rr < R ? in16_get<COH>(y, rr, l16, Wv, Xt) : 0.0;
        }
        const T *row = base + (long long)min(rb + l16, R - 1) * t.ld;
        d4v acc0 = d4v{0.0, 0.0, 0.0, 0.0};
        #pragma unroll
        for (int g = 0; g < T16_MAXC2 / 16; g++) {
            if (g < ng) {
                const int c0 = 16 * g + 4 * lq;
                double a[4];
                load_quad(row, c0, cmax, a);
                #pragma unroll
                for (int j = 0; j < 4; j++)
                    acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[j], xin[(c0 + j) * NV + l16], acc0, 0, 0, 0);
                // written and read back as doubles through the same pointer,
                // so the compiler keeps the reads after the writes
                #pragma unroll
                for (int j = 0; j < 4; j++) tr[l16 * 16 + 4 * lq + j] = a[j];
                #pragma unroll
                for (int q = 0; q < 4; q++)
                    acc1[g] = __builtin_amdgcn_mfma_f64_16x16x4f64(ya[q], tr[(4 * q + lq) * 16 + l16], acc1[g], 0, 0, 0);
            }
        }
        #pragma unroll
        for (int r = 0; r < 4; r++) {
            const int rr = rb + lq + 4 * r;
            if (rr < R) wst<COH>(Wv + (long long)(t.out0 + rr) * NV + l16, acc0[r]);
        }
    }
    __syncthreads();   // X and the transposition tiles are dead from here on
    double *red = lds;  // per wave: 64 columns x 16 vectors
    #pragma unroll
    for (int g4 = 0; g4 < T16_MAXC2 / 16; g4++)
        #pragma unroll
        for (int r = 0; r < 4; r++) red[w * 1024 + (16 * g4 + l16) * NV + lq + 4 * r] = acc1[g4][r];
    __syncthreads();
    for (int q = tid; q < C * NV; q += WG) {
        const double v = (red[q] + red[1024 + q]) + (red[2048 + q] + red[3072 + q]);
        wst<COH>(Wv + (long long)t.out1 * NV + q, v);
    }
}

template <typename T, bool COH>
__device__ void mfma_tile16(const BTask &task, const T *__restrict__ S, const int *__restrict__ vsrc,
                            double *__restrict__ Wv, const double *__restrict__ Xt, double *lds) {
    constexpr int NV = 16;
    constexpr int VW = sizeof(T) == 4 ? 4 : 2;   // elements per vector load
    const BTask t = uniform(task);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int R = t.nrow, C = t.ncol;
    const bool do0 = t.out0 >= 0, do1 = t.out1 >= 0;
    const T *base = S + t.mat;
    if (do0 && do1) { mfma_tile16_2s<T, COH>(t, base, vsrc, Wv, Xt, lds); return; }
    const int cmax = ((C - 1) / VW) * VW;
    // inputs staged in LDS (their slots summed once) when they fit beside
    // each other, padded with zeros to whole groups of 16; otherwise (a row
    // product wider than the LDS) the row product reads X from memory
    const int Cp = (C + 15) & ~15, Rp = (R + 15) & ~15;
    const int need1 = do1 ? Rp * NV : 0;
    const bool st0 = do0 && Cp * NV + need1 <= T16_LDS;
    double *xin = lds;
    double *yin = lds + (st0 ? Cp * NV : 0);
    if (st0) {
        stage_vec<NV, COH, WG>(vsrc, t.v0b, t.v0e, t.v0off, C, Wv, Xt, -1, xin, tid);
        for (int q = C * NV + tid; q < Cp * NV; q += WG) xin[q] = 0.0;
    }
    if (do1) {
        stage_vec<NV, COH, WG>(vsrc, t.v1b, t.v1e, t.v1off, R, Wv, Xt, -1, yin, tid);
        for (int q = R * NV + tid; q < Rp * NV; q += WG) yin[q] = 0.0;
    }
    __syncthreads();
    if (do0) {
        const In16 x = in16_make(vsrc, t.v0b, t.v0e, t.v0off);
        const int ng = Cp >> 4;   // groups of 16 columns
        // a step of the tile covers 256 rows: wave w takes its blocks of 16
        // rows w, w + 4, w + 8, w + 12 (fewer at the end of the tile)
        for (int rs = 0; rs < R; rs += 256) {
            const int rb0 = rs + 16 * w;
            const int nb = rb0 < R ? min(4, (R - rb0 + 63) >> 6) : 0;   // blocks of this wave in the step (wave-uniform)
            if (nb == 0) break;
            const T *row[4];
            #pragma unroll
            for (int k = 0; k < 4; k++) row[k] = base + (long long)min(rb0 + 64 * k + l16, R - 1) * t.ld;
            d4v acc[4];
            #pragma unroll
            for (int k = 0; k < 4; k++) acc[k] = d4v{0.0, 0.0, 0.0, 0.0};
            for (int g = 0; g < ng; g++) {
                const int c0 = 16 * g + 4 * lq;
                double a[4][4], bx[4];
                #pragma unroll
                for (int k = 0; k < 4; k++)
                    if (k < nb) load_quad(row[k], c0, cmax, a[k]);
                #pragma unroll
                for (int j = 0; j < 4; j++)
                    bx[j] = st0 ? xin[(c0 + j) * NV + l16] : (c0 + j < C ? in16_get<COH>(x, c0 + j, l16, Wv, Xt) : 0.0);
                #pragma unroll
                for (int k = 0; k < 4; k++) {
                    if (k >= nb) break;
                    #pragma unroll
                    for (int j = 0; j < 4; j++) acc[k] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[k][j], bx[j], acc[k], 0, 0, 0);
                }
            }
            #pragma unroll
            for (int k = 0; k < 4; k++) {
                if (k >= nb) break;
                #pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int rr = rb0 + 64 * k + lq + 4 * r;
                    if (rr < R) wst<COH>(Wv + (long long)(t.out0 + rr) * NV + l16, acc[k][r]);
                }
            }
        }
    }
    if (do1) {
        const int nspan = (C + 63) >> 6;   // spans of 64 columns (1 or 2)
        const int wps = 4 / nspan;         // waves per span
        const int span = w / wps, part = w - span * wps;
        const int c0 = 64 * span;
        d4v acc1[4];
        #pragma unroll
        for (int j = 0; j < 4; j++) acc1[j] = d4v{0.0, 0.0, 0.0, 0.0};
        for (int rb = 16 * part; rb < R; rb += 16 * wps) {
            #pragma unroll
            for (int q = 0; q < 4; q++) {
                const int rr = rb + 4 * q + lq;
                double b[4];
                load_quad(base + (long long)min(rr, R - 1) * t.ld, c0 + 4 * l16, cmax, b);
                const double ya = yin[rr * NV + l16];   // rr < Rp: zero past the tile
                #pragma unroll
                for (int j = 0; j < 4; j++) acc1[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ya, b[j], acc1[j], 0, 0, 0);
            }
        }
        __syncthreads();   // the staged inputs are dead from here on
        double *red = lds;   // per wave: 64 columns x 16 vectors
        #pragma unroll
        for (int j = 0; j < 4; j++)
            #pragma unroll
            for (int r = 0; r < 4; r++) red[w * 1024 + (4 * l16 + j) * NV + lq + 4 * r] = acc1[j][r];
        __syncthreads();
        for (int q = tid; q < C * NV; q += WG) {
            const int sp = q >> 10, loc = q & 1023;
            const double *rs = red + sp * wps * 1024 + loc;
            const double v = wps == 4 ? (rs[0] + rs[1024]) + (rs[2048] + rs[3072]) : rs[0] + rs[1024];
            wst<COH>(Wv + (long long)t.out1 * NV + q, v);
        }
    }
}

// A tile task: the MFMA products at 16 vectors per pass, the VALU tile below.
template <typename T, int NV, bool COH, int NT = WG>
__device__ __forceinline__ void bgemv_tile(const BTask &task, const T *__restrict__ S, const int *__restrict__ vsrc,
                                           double *__restrict__ Wv, const double *__restrict__ X, long long ldx,
                                           double *lds) {
    if constexpr (NV == 16) {
        // X is the element-major copy of the input (the pass's first launch
        // makes it); no wave tiles at 16
        if constexpr (NT == WG) mfma_tile16<T, COH>(task, S, vsrc, Wv, X, lds);
    } else {
        bgemv_tile_valu<T, NV, COH, NT>(task, S, vsrc, Wv, X, ldx, lds);
    }
}

// Fixed-order sum of slots into one chunk (n = len * NV <= 1024 values).
// K = 4, 2 or 1 thread groups (n <= 64, <= 128, larger) split the sources:
// group g sums sources g, g + K, ... in order, each thread up to four values
// with all their loads of four sources in flight; the group sums are added
// as (g0 + g1) + (g2 + g3), or g0 + g1.
template <int NV, bool COH>
__device__ void reduce_task(const RTask &task, const int *__restrict__ src, double *__restrict__ Wv,
                            double *__restrict__ Y, long long ldy, double *lds) {
    const RTask t = uniform(task);
    const int n = t.len * NV;
    const int K = n <= 64 ? 4 : (n <= 128 ? 2 : 1);
    const int gsz = WG / K;
    const int g = threadIdx.x / gsz, lt = threadIdx.x - g * gsz;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const double *base = Wv + (long long)t.soff * NV + lt;
    int j = t.src_beg + g;
    for (; j + 3 * K < t.src_end; j += 4 * K) {
        const long long o0 = (long long)src[j] * NV, o1 = (long long)src[j + K] * NV;
        const long long o2 = (long long)src[j + 2 * K] * NV, o3 = (long long)src[j + 3 * K] * NV;
        #pragma unroll
        for (int k = 0; k < 4; k++) {
            if (lt + k * gsz < n) {
                const double *p = base + k * gsz;
                const double a0 = wld<COH>(p + o0), a1 = wld<COH>(p + o1), a2 = wld<COH>(p + o2), a3 = wld<COH>(p + o3);
                acc[k] += a0;
                acc[k] += a1;
                acc[k] += a2;
                acc[k] += a3;
            }
        }
    }
    for (; j < t.src_end; j += K) {
        const long long o = (long long)src[j] * NV;
        #pragma unroll
        for (int k = 0; k < 4; k++)
            if (lt + k * gsz < n) acc[k] += wld<COH>(base + k * gsz + o);
    }
    auto put = [&](int q, double v) {
        if (t.flags & 1) {
            const int i = q / NV, w = q - i * NV;
            Y[(long long)w * ldy + t.dst + i] = v;
        } else {
            wst<COH>(Wv + t.dst * NV + q, v);
        }
    };
    if (K == 1) {
        #pragma unroll
        for (int k = 0; k < 4; k++)
            if (lt + k * WG < n) put(lt + k * WG, acc[k]);
        return;
    }
    double *red = lds;   // K x 1024
    if (lt < n) red[g * 1024 + lt] = acc[0];
    __syncthreads();
    for (int q = threadIdx.x; q < n; q += WG) {
        const double v = K == 4 ? (red[q] + red[1024 + q]) + (red[2048 + q] + red[3072 + q]) : red[q] + red[1024 + q];
        put(q, v);
    }
}

// One launch per stage: n64 FP64 tiles, then n32 FP32 tiles, then nr reduce
// tasks, then the wave tiles, four per workgroup (FP64 first), all
// independent of each other within the launch.
// Waves per SIMD the kernels are built for: 4 (at most 128 VGPRs).  A
// two-sided 16-vector tile keeps the column sums of its (at most 64)
// columns in registers; with 128 columns it needed 3 waves per SIMD and ran
// 6% slower (C100H202, 16 densities: 3.60 against 3.38 ms).
template <int NV> constexpr int waves_per_simd() { return 4; }

// Task b of a staged launch: tiles (FP64, FP32), reduce tasks, then groups of
// four wave tiles.
template <int NV>
__device__ __forceinline__ void stage_task(int b, const BTask *__restrict__ t64, int n64, const BTask *__restrict__ t32,
                                           int n32, const RTask *__restrict__ rt, int nr, const BTask *__restrict__ w64,
                                           int nw64, const BTask *__restrict__ w32, int nw32,
                                           const int *__restrict__ vsrc, const int *__restrict__ rsrc,
                                           const double *__restrict__ S64, const float *__restrict__ S32,
                                           double *__restrict__ Wv, const double *__restrict__ X,
                                           double *__restrict__ Y, long long ldx, double *lds) {
    if (b < n64) {
        const BTask t = t64[b];
        bgemv_tile<double, NV, false>(t, S64, vsrc, Wv, X, ldx, lds);
    } else if (b - n64 < n32) {
        const BTask t = t32[b - n64];
        bgemv_tile<float, NV, false>(t, S32, vsrc, Wv, X, ldx, lds);
    } else if (b - n64 - n32 < nr) {
        const RTask t = rt[b - n64 - n32];
        reduce_task<NV, false>(t, rsrc, Wv, Y, ldx, lds);
    } else {
        const int wave = threadIdx.x >> 6;
        const int i = 4 * (b - n64 - n32 - nr) + wave;   // wave tile of this wave
        double *wl = lds + wave * (LDS_DOUBLES / 4);
        const int j = i - ((nw64 + 3) & ~3);   // FP32 wave tiles start at a new workgroup
        if (i < nw64) {
            const BTask t = w64[i];
            bgemv_tile<double, NV, false, 64>(t, S64, vsrc, Wv, X, ldx, wl);
        } else if (j >= 0 && j < nw32) {
            const BTask t = w32[j];
            bgemv_tile<float, NV, false, 64>(t, S32, vsrc, Wv, X, ldx, wl);
        }
    }
}

// One launch per stage: one task per workgroup.
template <int NV>
__global__ __launch_bounds__(WG, waves_per_simd<NV>()) void stage_kernel(const BTask *__restrict__ t64, int n64, const BTask *__restrict__ t32,
                                                   int n32, const RTask *__restrict__ rt, int nr,
                                                   const BTask *__restrict__ w64, int nw64, const BTask *__restrict__ w32,
                                                   int nw32, const int *__restrict__ vsrc, const int *__restrict__ rsrc,
                                                   const double *__restrict__ S64, const float *__restrict__ S32,
                                                   double *__restrict__ Wv, const double *__restrict__ X,
                                                   double *__restrict__ Y, long long ldx) {
    __shared__ double lds[LDS_DOUBLES];
    stage_task<NV>(blockIdx.x, t64, n64, t32, n32, rt, nr, w64, nw64, w32, nw32, vsrc, rsrc, S64, S32, Wv, X, Y, ldx,
                   lds);
}

// A near-field launch with fewer workgroups than tasks (the sweeps' launches
// on the other stream keep CUs of their own): workgroup g takes tasks
// g, g + gridDim.x, ...
template <int NV>
__global__ __launch_bounds__(WG, waves_per_simd<NV>()) void stage_loop_kernel(const BTask *__restrict__ t64, int n64, const BTask *__restrict__ t32,
                                                        int n32, const RTask *__restrict__ rt, int nr,
                                                        const BTask *__restrict__ w64, int nw64, const BTask *__restrict__ w32,
                                                        int nw32, const int *__restrict__ vsrc, const int *__restrict__ rsrc,
                                                        const double *__restrict__ S64, const float *__restrict__ S32,
                                                        double *__restrict__ Wv, const double *__restrict__ X,
                                                        double *__restrict__ Y, long long ldx) {
    __shared__ double lds[LDS_DOUBLES];
    const int total = n64 + n32 + nr + (nw64 + 3) / 4 + (nw32 + 3) / 4;
    for (int b = blockIdx.x; b < total; b += gridDim.x) {
        if (b != (int)blockIdx.x) __syncthreads();   // the previous task's LDS
        stage_task<NV>(b, t64, n64, t32, n32, rt, nr, w64, nw64, w32, nw32, vsrc, rsrc, S64, S32, Wv, X, Y, ldx, lds);
    }
}

// Pointers of one representation, for launches over several molecules.
struct MolPtrs {
    const double *S64;
    const float *S32;
    double *Wv;
    const BTask *t64, *t32;
    const RTask *rt;
    const int *vsrc, *rsrc;
    const double *X;
    double *Y;
    long long ldx;
};
// ---------------------------------------------------------------------------
// Persistent executor: the whole pass in one launch.
// ---------------------------------------------------------------------------
// The tasks of every stage of a plan, in a topological order computed on the
// host.  Dependencies are per input list: a tile's in0/in1 or a reduce task's
// source list is ready once every task writing one of its slots has counted
// itself into the list's counter.  Every task is taken from one ticket
// counter, in chunks of QCHUNK consecutive tasks, the next chunk claimed
// while the current one runs; a workgroup holds a ticket only once it runs.
// A task only ever waits for tasks of lower ticket, and those were claimed
// earlier by workgroups that are running, so the lowest unfinished ticket
// can always proceed and the launch drains whatever the residency (work of
// other streams on the same CUs included).  Every wait is bounded: past
// `spin_limit` polls the workgroup raises the error flag and goes on, so
// every wave exits; the synchronous entry points read the flag and report
// the pass as failed.
//
// Data between workgroups goes through agent-coherent workspace accesses
// (wld/wst with COH); a task's counters move only after every wave's stores
// have completed, and a consumer reads only after its counters are complete.
struct QTask {
    int kind;         // 0: FP64 tile, 1: FP32 tile, 2: reduce task
    int idx;          // index into t64 / t32 / rt
    int in0, in1;     // input list counters to wait for (-1: none)
    int out0, out1;   // list counters this task counts into (-1: none)
    int pad_[2];
};
static_assert(sizeof(QTask) == 32, "QTask layout");

// control words of the counter array (then one counter per list)
static const int QC_TICKET = 0;    // next chunk to claim
static const int QC_ERROR = 32;    // set when a wait exceeded its bound
static const int QC_LISTS = 64;    // first list counter
static const int QCHUNK = 1;       // tasks per claim: consecutive tasks are often dependent

__device__ __forceinline__ void q_wait(int *ctr, int list, int need, int spin_limit) {
    if (list < 0) return;
    const int *c = ctr + QC_LISTS + list;
    int it = 0;
    while (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
        if (++it > spin_limit) {
            __hip_atomic_store(ctr + QC_ERROR, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// s_waitcnt vmcnt(0): every vector memory operation of this wave (stores
// included on gfx9) has completed.
__device__ __forceinline__ void wait_vm_stores() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// GROUP = false: one representation, pointers from the arguments; GROUP =
// true: QTask::pad_[0] names the molecule whose pointers come from mols.
template <int NV, bool GROUP>
__global__ __launch_bounds__(WG, waves_per_simd<NV>()) void queue_kernel(const QTask *__restrict__ tasks, int ntask,
                                                      const int *__restrict__ need, int *ctr, int spin_limit,
                                                      const MolPtrs *__restrict__ mols, const BTask *__restrict__ t64,
                                                      const BTask *__restrict__ t32, const RTask *__restrict__ rt,
                                                      const int *__restrict__ vsrc, const int *__restrict__ rsrc,
                                                      const double *__restrict__ S64, const float *__restrict__ S32,
                                                      double *__restrict__ Wv, const double *__restrict__ X,
                                                      double *__restrict__ Y, long long ldx) {
    __shared__ double lds[LDS_DOUBLES];
    __shared__ int s_next[2];
    if (threadIdx.x == 0) s_next[0] = __hip_atomic_fetch_add(ctr + QC_TICKET, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    int t0 = s_next[0] * QCHUNK;
    int t1 = min(t0 + QCHUNK, ntask);
    for (int par = 1; t0 < ntask; par ^= 1) {
        for (int t = t0; t < t1; t++) {
            QTask q = tasks[t];
            q.kind = ufl(q.kind); q.idx = ufl(q.idx); q.in0 = ufl(q.in0); q.in1 = ufl(q.in1);
            q.out0 = ufl(q.out0); q.out1 = ufl(q.out1); q.pad_[0] = ufl(q.pad_[0]);
            // the next chunk, claimed while this one runs: the ticket stays in
            // a register of thread 0 and reaches LDS after the task, so its
            // atomic's latency is not on the task's path
            int next_ticket = 0;
            if (threadIdx.x == 0) {
                if (t == t0) next_ticket = __hip_atomic_fetch_add(ctr + QC_TICKET, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                q_wait(ctr, q.in0, q.in0 >= 0 ? need[q.in0] : 0, spin_limit);
                q_wait(ctr, q.in1, q.in1 >= 0 ? need[q.in1] : 0, spin_limit);
            }
            __syncthreads();
            if constexpr (GROUP) {
                const MolPtrs m = mols[q.pad_[0]];
                if (q.kind == 0) bgemv_tile<double, NV, true>(m.t64[q.idx], m.S64, m.vsrc, m.Wv, m.X, m.ldx, lds);
                else if (q.kind == 1) bgemv_tile<float, NV, true>(m.t32[q.idx], m.S32, m.vsrc, m.Wv, m.X, m.ldx, lds);
                else reduce_task<NV, true>(m.rt[q.idx], m.rsrc, m.Wv, m.Y, m.ldx, lds);
            } else {
                if (q.kind == 0) bgemv_tile<double, NV, true>(t64[q.idx], S64, vsrc, Wv, X, ldx, lds);
                else if (q.kind == 1) bgemv_tile<float, NV, true>(t32[q.idx], S32, vsrc, Wv, X, ldx, lds);
                else reduce_task<NV, true>(rt[q.idx], rsrc, Wv, Y, ldx, lds);
            }
            if (threadIdx.x == 0 && t == t0) s_next[par] = next_ticket;
            wait_vm_stores();
            __syncthreads();
            if (threadIdx.x == 0) {
                if (q.out0 >= 0) __hip_atomic_fetch_add(ctr + QC_LISTS + q.out0, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (q.out1 >= 0) __hip_atomic_fetch_add(ctr + QC_LISTS + q.out1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        t0 = s_next[par] * QCHUNK;
        t1 = min(t0 + QCHUNK, ntask);
    }
}

// Clears n counters (every control word included).
__global__ void counters_clear_kernel(int *ctr, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ctr[i] = 0;
}

// Clears the ticket and list counters before a queue launch (the error flag
// stays until the host reads it).
__global__ void queue_reset_kernel(int *ctr, int nlist) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == QC_TICKET) ctr[i] = 0;
    if (i < nlist) ctr[QC_LISTS + i] = 0;
}

// x[i] = c_i * D[g_i >> 1], c_i = 2 when the low bit is set
// (src/H2ERI_build_Coulomb.c:26-62).
__global__ void uncontract_kernel(const int *__restrict__ gidx, const double *__restrict__ D, double *__restrict__ x,
                                  int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        int g = gidx[i];
        double v = D[g >> 1];
        x[i] = (g & 1) ? 2.0 * v : v;
    }
}

// Jh[g_i >> 1] = s_i * y[i], s_i = 1/2 when the low bit is clear (M == N)
// (src/H2ERI_build_Coulomb.c:90-116).  The map i -> g_i >> 1 is injective.
__global__ void contract_scatter_kernel(const int *__restrict__ gidx, const double *__restrict__ y,
                                        double *__restrict__ Jh, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        int g = gidx[i];
        Jh[g >> 1] = (g & 1) ? y[i] : 0.5 * y[i];
    }
}

// J = Jh + Jh^T through 32x32 LDS tiles (src/H2ERI_build_Coulomb.c:119-132).
__global__ void symmetrize_kernel(const double *__restrict__ Jh, double *__restrict__ J, int n) {
    __shared__ double tile[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8 threads
    for (int k = ty; k < 32; k += 8) {
        int r = bx + k, c = by + tx;   // read the transposed tile
        tile[k][tx] = (r < n && c < n) ? Jh[(size_t)r * n + c] : 0.0;
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        int r = by + k, c = bx + tx;
        if (r < n && c < n) J[(size_t)r * n + c] = Jh[(size_t)r * n + c] + tile[tx][k];
    }
}

// Element-major copy of 16 vectors (Xt[i * 16 + v] = X[v * ld + i]), the
// first launch of a 16-vector pass: both sides coalesced through a 16 x 256
// LDS tile.
__global__ void transpose16_kernel(const double *__restrict__ X, long long ld, double *__restrict__ Xt, int n) {
    __shared__ double tile[16][257];
    const int i0 = blockIdx.x * 256, tid = threadIdx.x;
    #pragma unroll
    for (int v = 0; v < 16; v++) tile[v][tid] = i0 + tid < n ? X[(long long)v * ld + i0 + tid] : 0.0;
    __syncthreads();
    #pragma unroll
    for (int k = 0; k < 16; k++) {
        const int q = k * 256 + tid, il = q >> 4, v = q & 15;
        if (i0 + il < n) Xt[(long long)(i0 + il) * 16 + v] = tile[v][il];
    }
}

__global__ void zero_kernel(double *__restrict__ p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0.0;
}

// ---------------------------------------------------------------------------
// Storage (uploaded once per precision) and plans (one per batch width)
// ---------------------------------------------------------------------------
enum Phase { PH_UP = 0, PH_B = 1, PH_Y1 = 2, PH_DOWN = 3, PH_D = 4, PH_GATHER = 5, PH_N = 6 };

struct MatRef {
    long long off = -1;
    int nrow = 0, ncol = 0, ld = 0;
    bool f32 = false;
};
struct BlkRef {
    bool lr = false;
    MatRef M;      // dense block
    MatRef U, V;   // low-rank halves: U (nrow x k), V^T (ncol x k)
};

struct Storage {
    std::vector<double> h64;
    std::vector<float> h32;
    std::vector<MatRef> U;
    std::vector<BlkRef> B, D;
    double bytes_U = 0, bytes_B = 0, bytes_D = 0;   // algorithmic (U counted twice)

    MatRef store(const double *src, int nrow, int ncol, bool fp32) {
        MatRef m;
        const int W = fp32 ? 4 : 2;
        m.nrow = nrow; m.ncol = ncol; m.f32 = fp32;
        m.ld = (ncol + W - 1) / W * W;
        const size_t n = (size_t)nrow * m.ld;
        if (fp32) {
            m.off = (long long)h32.size();
            h32.resize(h32.size() + ((n + 63) & ~(size_t)63), 0.0f);
            for (int r = 0; r < nrow; r++)
                for (int c = 0; c < ncol; c++) h32[m.off + (size_t)r * m.ld + c] = (float)src[(size_t)r * ncol + c];
        } else {
            m.off = (long long)h64.size();
            h64.resize(h64.size() + ((n + 31) & ~(size_t)31), 0.0);
            for (int r = 0; r < nrow; r++) memcpy(&h64[m.off + (size_t)r * m.ld], src + (size_t)r * ncol, sizeof(double) * ncol);
        }
        return m;
    }
    BlkRef store_blk(H2E_dense_mat_p b, bool fp32, double &bytes) {
        BlkRef r;
        if (!b || b->nrow == 0 || b->ncol == 0) return r;
        const double esz = fp32 ? 4.0 : 8.0;
        if (b->ld >= 0) {
            r.M = store(b->data, b->nrow, b->ncol, fp32);
            bytes += esz * b->nrow * b->ncol;
        } else {
            const int k = -b->ld;
            r.lr = true;
            r.U = store(b->data, b->nrow, k, fp32);
            r.V = store(b->data + (size_t)b->nrow * k, b->ncol, k, fp32);
            bytes += esz * (double)(b->nrow + b->ncol) * k;
        }
        return r;
    }
    void build(H2ERI_p h, bool ub32) {
        U.assign(h->n_node, MatRef());
        for (int i = 0; i < h->n_node; i++) {
            H2E_dense_mat_p u = h->U[i];
            if (!u || u->nrow == 0 || u->ncol == 0) continue;
            U[i] = store(u->data, u->nrow, u->ncol, ub32);
            bytes_U += 2.0 * (ub32 ? 4.0 : 8.0) * u->nrow * u->ncol;
        }
        B.resize(h->n_B);
        for (int i = 0; i < h->n_B; i++) B[i] = store_blk(h->c_B_blks[i], ub32, bytes_B);
        D.resize(h->n_D);
        for (int i = 0; i < h->n_D; i++) D[i] = store_blk(h->c_D_blks[i], false, bytes_D);
    }
};

struct Stage {
    int phase;
    int stream;         // 0: main stream, 1: near-field stream (joins before the gather)
    int t64_beg, t64_cnt, t32_beg, t32_cnt;
    int r_beg, r_cnt;
    int w64_beg, w64_cnt, w32_beg, w32_cnt;   // wave tiles (four per workgroup)
    double mat_bytes;   // matrix bytes streamed by this launch
};

struct Plan {
    int nv = 1;
    std::vector<BTask> t64, t32;
    std::vector<BTask> w64, w32;   // wave tiles
    std::vector<RTask> rt;
    std::vector<int> rsrc, vsrc;
    std::vector<Stage> stages;
    long long nW = 0;   // workspace elements per vector
    double bytes_vec = 0;
    int n_fold = 0;     // fold reduce tasks
    std::vector<std::pair<long long, int>> slots;   // workspace slots in allocation order: (offset, length)
};

// One timed launch of the last call (phase timing mode).
struct LaunchRec {
    int phase, blocks;
    double mat_bytes;
    float ms;
    float start;   // ms after the pass's base event (recorded before any of its launches)
};

struct DevPlan {
    Plan P;
    bool ready = false;
    double *Wv = nullptr;
    BTask *d_t64 = nullptr, *d_t32 = nullptr, *d_w64 = nullptr, *d_w32 = nullptr;
    RTask *d_rt = nullptr;
    int *d_rsrc = nullptr, *d_vsrc = nullptr;
    // persistent executor
    bool q_ready = false;
    int q_n = 0, q_static = 0, q_nlist = 0, q_grid = 0;
    double q_est_us = 0;
    QTask *d_q = nullptr;
    int *d_need = nullptr, *d_ctr = nullptr;
    // 16 vectors per pass: element-major copy of the pass's input (N x 16)
    double *Xt = nullptr;
};

static const int NWIDTH = 5;   // batch widths: vectors per pass over the matrix data
static const int NV_VARIANTS[NWIDTH] = {1, 2, 4, 8, 16};
static inline int nv_index(int nv) { return nv == 1 ? 0 : (nv == 2 ? 1 : (nv == 4 ? 2 : (nv == 8 ? 3 : 4))); }
static inline bool nv_valid(int nv) { return nv == 1 || nv == 2 || nv == 4 || nv == 8 || nv == 16; }
// widest pass for the vectors left: 16, 8, 4, 2 or 1
static inline int nv_pass(int rem) { return rem >= 16 ? 16 : (rem >= 8 ? 8 : (rem >= 4 ? 4 : (rem >= 2 ? 2 : 1))); }

// Tuning configuration of one batch width.
struct Tune {
    double tile_bytes = 128.0 * 1024.0;
    int nf_schedule = 3;      // near-field placement, see H2ERI_set_schedule (two groups on a second stream)
    int executor = 0;         // 0: one launch per stage, 1: one persistent launch per pass (H2ERI_set_executor)
    double wave_bytes = 32768.0;   // staged executor: matrix parts up to this size run as wave tiles
    int nf_grid = 0;          // staged executor: workgroups of a near-field launch (0: one per task)
};

struct DevState {
    int dev = -1;
    int ub_bits = 64;
    Tune cfg[NWIDTH];         // per batch width (nv_index); the setters change all of them, the autotuner one
    Storage st;
    bool st_ready = false;
    double *S64 = nullptr;
    float *S32 = nullptr;
    DevPlan plans[NWIDTH];
    // Coulomb and host-pointer staging
    int *d_gidx = nullptr;
    double *d_D = nullptr, *d_J = nullptr, *d_Jh = nullptr, *d_y = nullptr, *d_x = nullptr;
    size_t d_io_len = 0;
    hipStream_t side = nullptr;         // near-field stream (schedule 2)
    hipGraphExec_t graph = nullptr;     // captured hot-path sequence (H2ERI_graph_capture)
    int graph_kind = -1;                // its kind (0 matvec, 1 Coulomb, 2 exchange)
    bool capturing = false;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool phase_timing = false;
    std::vector<hipEvent_t> evs;
    hipEvent_t ev_base = nullptr;   // timing base of the last timed pass
    std::vector<LaunchRec> launches;
    bool collected = true;
    double phase_ms[7] = {0, 0, 0, 0, 0, 0, 0};
    size_t resident_bytes = 0;
    // bound of every dependency wait of the persistent executor, in polls
    // (about a second; H2ERI_set_spin_limit)
    int spin_limit = 1 << 22;
    // buffers of the captured Coulomb graph: only H2ERI_graph_capture
    // allocates them and only release_graph frees them, so no eager call
    // that grows its own buffers can leave the graph pointing at freed memory
    double *g_x = nullptr, *g_y = nullptr, *g_Jh = nullptr;
    // workspace of the pass being enqueued when it is not the plan's own
    // (the allocation-free entry point puts it in the file-scope scratch
    // array); null: the plan's Wv and counters
    double *ws_Wv = nullptr;
    int *ws_ctr = nullptr;
    int *last_noalloc_ctr = nullptr;   // counters of the last allocation-free persistent pass
    JitState *jit = nullptr;           // a BD_JIT handle's on-the-fly state (jit.hip)
};

static DevState *dstate(H2ERI_p h) { return (DevState *)h->dev; }

#define HIP_OK(expr)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            set_error(ERR_DEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                      __LINE__);                                                                \
            return false;                                                                       \
        }                                                                                       \
    } while (0)

static void free_dev(void *&p) {
    if (p) { (void)hipFree(p); p = nullptr; }
}

static void release_graph(DevState *d);
static void release_plan(DevState *d, int i) {
    release_graph(d);
    DevPlan &p = d->plans[i];
    void **ps[] = {(void **)&p.Wv, (void **)&p.d_t64, (void **)&p.d_t32, (void **)&p.d_w64, (void **)&p.d_w32,
                   (void **)&p.d_rt, (void **)&p.d_rsrc, (void **)&p.d_vsrc, (void **)&p.d_q, (void **)&p.d_need,
                   (void **)&p.d_ctr, (void **)&p.Xt};
    for (void **q : ps) free_dev(*q);
    p.ready = false;
    p.q_ready = false;
    p.P = Plan();
}

static void release_plans(DevState *d) {
    for (int i = 0; i < NWIDTH; i++) release_plan(d, i);
}

static void release_graph(DevState *d) {
    if (d->graph) { (void)hipGraphExecDestroy(d->graph); d->graph = nullptr; }
    d->graph_kind = -1;
    free_dev((void *&)d->g_x);
    free_dev((void *&)d->g_y);
    free_dev((void *&)d->g_Jh);
}

static void release_buffers(DevState *d) {
    release_graph(d);   // a graph refers to the buffers below
    release_plans(d);
    void **ps[] = {(void **)&d->S64, (void **)&d->S32, (void **)&d->d_gidx, (void **)&d->d_D, (void **)&d->d_J,
                   (void **)&d->d_Jh, (void **)&d->d_y, (void **)&d->d_x};
    for (void **p : ps) free_dev(*p);
    jit_release(d->jit);
    d->d_io_len = 0;
    d->st = Storage();
    d->st_ready = false;
    d->resident_bytes = 0;
}

void device_release(H2ERI_p h) {
    exchange_release(h);
    DevState *d = dstate(h);
    if (!d) return;
    release_buffers(d);
    for (hipEvent_t e : d->evs) (void)hipEventDestroy(e);
    if (d->ev_fork) (void)hipEventDestroy(d->ev_fork);
    if (d->ev_join) (void)hipEventDestroy(d->ev_join);
    if (d->ev_base) (void)hipEventDestroy(d->ev_base);
    if (d->side) (void)hipStreamDestroy(d->side);
    delete d;
    h->dev = nullptr;
}

void device_release_graph_kind(H2ERI_p h, int kind) {
    DevState *d = h ? dstate(h) : nullptr;
    if (d && d->graph && d->graph_kind == kind) release_graph(d);
}

void device_invalidate(H2ERI_p h) {
    DevState *d = dstate(h);
    if (d) release_buffers(d);
    exchange_release(h);
}

static DevState *dstate_get(H2ERI_p h) {
    if (!h->dev) h->dev = new DevState();
    return dstate(h);
}

namespace {

struct Src {
    int off;     // workspace element offset (or the encoded external input)
    int stage;   // launch that writes it (-1: external input)
};

// Virtual vector: element i = sum over srcs (in order) of the source at i.
struct VVec {
    int len = 0;
    std::vector<Src> srcs;
    int fc = INT_MAX;     // first consuming stage
    double reads = 0;     // elements read by its consumer tiles
    int lb = -1, le = -1; // range in vsrc once emitted
};

// Tile before emission: inputs are vector ids.
struct PTile {
    BTask t;
    int st;
    bool f32;
    bool wave;   // a wave tile
    int iv0, iv1;
};

struct Planner {
    H2ERI_p h;
    const Storage &S;
    Plan &P;
    int nv;
    double tile_bytes;
    int nf_schedule;
    int rchunk;  // output elements per reduce task
    double wave_bytes;   // matrix parts up to this size become wave tiles (0: none)
    bool split2;         // two-sided tiles become two one-sided tasks (staged executor, NV = 8)
    int maxc;   // max tile columns for this batch width
    std::vector<VVec> vv;
    std::vector<PTile> tiles;
    std::vector<std::vector<RTask>> SR;   // reduce tasks of each stage
    std::vector<int> sphase;
    std::vector<double> sload;
    int nst = 0, sG = 0, sS = 0;   // stages: main 0..sG, near-field stream sS..nst-1 (schedule 2)

    Planner(H2ERI_p h_, const Storage &s_, Plan &p_, int nv_, double tb, int nfs, int rch, double wb, bool sp)
        : h(h_), S(s_), P(p_), nv(nv_), tile_bytes(tb), nf_schedule(nfs), rchunk(rch), wave_bytes(nv_ == 16 ? 0.0 : wb),
          split2(sp && nv_ == 8), maxc(64 * cpl_of(nv_)) {}

    long long walloc(long long n) {
        long long o = P.nW;
        P.nW += (n + 7) & ~7LL;
        P.slots.push_back({o, (int)n});
        return o;
    }
    int new_vec(int len) {
        VVec v;
        v.len = len;
        vv.push_back(v);
        return (int)vv.size() - 1;
    }
    void use(int vec, int st, int len) {
        VVec &v = vv[vec];
        v.fc = std::min(v.fc, st);
        v.reads += len;
    }

    // Tiles of rows [r0, r0+nrow) of matrix m in stage st.  in0 (length ncol)
    // is vector v0 at element offset o0, in1 (length nrow) vector v1 at o1;
    // out0 feeds vector d0 (full length nrow), out1 vector d1 (full length ncol).
    void add_matrix(int st, const MatRef &m, int r0_, int nrow, int v0, int o0, int v1, int o1, int d0, int d1) {
        const int ncol = m.ncol;
        if (nrow <= 0 || ncol <= 0 || (d0 < 0 && d1 < 0)) return;
        const double esz = m.f32 ? 4.0 : 8.0;
        // a small part is one wave tile: one wave of a workgroup, a quarter of its LDS
        const bool wave = esz * nrow * ncol <= wave_bytes && ncol <= maxc &&
                          ncol * nv <= TileLDS<64>::X && (d1 < 0 || nrow * nv <= TileLDS<64>::Y);
        if (nv == 16) { add_matrix16(st, m, r0_, nrow, v0, o0, v1, o1, d0, d1); return; }
        const int nct = wave ? 1 : (ncol + maxc - 1) / maxc;
        const int ctw = wave ? ncol : ((ncol + nct - 1) / nct + 7) / 8 * 8;   // multiple of 8 (vector alignment), <= maxc
        int rows = wave ? nrow : (int)std::max(1.0, tile_bytes / (esz * ctw));
        if (d1 >= 0 && !wave) rows = std::min(rows, YIN_DOUBLES / nv);
        emit_tiles(st, m, r0_, nrow, nct, ctw, rows, wave, v0, o0, v1, o1, d0, d1);
    }

    // 16 vectors per pass: a row product spans whole rows and leaves no
    // partial sums over column tiles (a partial sum is 16 values per output
    // row); its X is staged in LDS when it fits and read from memory when
    // not.  Row products: row strips of the full width, about tile_bytes each
    // (at least 64 rows, one block of 16 per wave).  Column products: strips
    // of at most 128 columns (one or two spans of 64, so all four waves share
    // the work) by row strips of at most 256 rows (Y staged in LDS); their
    // partial sums over row strips are short (the strip's width).  A block
    // with both outputs wider than 256 columns is a set of row strips plus a
    // set of column tiles (the matrix read twice, no partial row sums); a
    // narrower one takes two-sided tiles of at most 64 columns by row strips
    // of about tile_bytes, which read their rows once.
    void add_matrix16(int st, const MatRef &m, int r0_, int nrow, int v0, int o0, int v1, int o1, int d0, int d1) {
        const int ncol = m.ncol;
        const double esz = m.f32 ? 4.0 : 8.0;
        if (d0 >= 0 && d1 >= 0 && ncol > 2 * T16_MAXC) {
            add_matrix16(st, m, r0_, nrow, v0, o0, -1, 0, d0, -1);
            add_matrix16(st, m, r0_, nrow, -1, 0, v1, o1, -1, d1);
            return;
        }
        int nct = 1, ctw = ncol;
        if (d1 >= 0) {
            const int mc = d0 >= 0 ? T16_MAXC2 : T16_MAXC;
            nct = (ncol + mc - 1) / mc;
            ctw = std::min(mc, ((ncol + nct - 1) / nct + 7) / 8 * 8);
        }
        int rows = std::max(64, (int)(tile_bytes / (esz * ctw)));
        // a column product alone stages Y in LDS (256 rows); a tile with both
        // products reads Y from memory
        if (d1 >= 0 && d0 < 0) rows = std::min(rows, T16_LDS / 16);
        emit_tiles(st, m, r0_, nrow, nct, ctw, rows, false, v0, o0, v1, o1, d0, d1);
    }

    // Tiles of nct column strips of ctw columns (the last narrower) by row
    // strips of about `rows` rows (balanced), with their output slots.
    void emit_tiles(int st, const MatRef &m, int r0_, int nrow, int nct, int ctw, int rows, bool wave, int v0, int o0,
                    int v1, int o1, int d0, int d1) {
        const int ncol = m.ncol;
        const double esz = m.f32 ? 4.0 : 8.0;
        const int nrt = (nrow + rows - 1) / rows;
        rows = (nrow + nrt - 1) / nrt;
        std::vector<long long> s0(nct, -1), s1(nrt, -1);
        if (d0 >= 0) for (int c = 0; c < nct; c++) { s0[c] = walloc(nrow); vv[d0].srcs.push_back(Src{(int)s0[c], st}); }
        if (d1 >= 0) for (int r = 0; r < nrt; r++) { s1[r] = walloc(ncol); vv[d1].srcs.push_back(Src{(int)s1[r], st}); }
        for (int rt = 0; rt < nrt; rt++) {
            const int ra = rt * rows, nr = std::min(rows, nrow - ra);
            for (int ct = 0; ct < nct; ct++) {
                const int c0 = ct * ctw, nc = std::min(ctw, ncol - c0);
                PTile p;
                memset(&p.t, 0, sizeof(p.t));
                BTask &t = p.t;
                t.mat = m.off + (long long)(r0_ + ra) * m.ld + c0;
                t.nrow = nr;
                t.ncol = nc;
                t.ld = m.ld;
                int lpr = 4;
                while (lpr * cpl_of(nv) < nc) lpr <<= 1;
                t.lpr = lpr;
                t.v0off = o0 + c0;
                t.v1off = o1 + ra;
                t.out0 = d0 >= 0 ? (int)(s0[ct] + ra) : -1;
                t.out1 = d1 >= 0 ? (int)(s1[rt] + c0) : -1;
                p.st = st;
                p.f32 = m.f32;
                p.wave = wave;
                p.iv0 = d0 >= 0 ? v0 : -1;
                p.iv1 = d1 >= 0 ? v1 : -1;
                if (p.iv0 >= 0) use(v0, st, nc);
                if (p.iv1 >= 0) use(v1, st, nr);
                tiles.push_back(p);
                sload[st] += esz * nr * nc;
            }
        }
    }

    // Dense or low-rank block with inputs (v0,o0) / (v1,o1) and outputs d0 / d1
    // in stages a < b < c (low-rank: smaller half in a and c, the other in b).
    void add_block(const BlkRef &blk, int v0, int o0, int v1, int o1, int d0, int d1, int a, int b, int c) {
        if (!blk.lr) {
            if (blk.M.off < 0) return;
            add_matrix(a, blk.M, 0, blk.M.nrow, v0, o0, v1, o1, d0, d1);
            return;
        }
        const int nr = blk.U.nrow, nc = blk.V.nrow, k = blk.U.ncol;
        const int vt = new_vec(k), vs = new_vec(k);
        if (d1 < 0) {
            add_matrix(a, blk.V, 0, nc, -1, 0, v0, o0, -1, vt);   // t = V in0
            add_matrix(b, blk.U, 0, nr, vt, 0, -1, 0, d0, -1);    // out0 = U t
            return;
        }
        if (nr <= nc) {   // U twice
            add_matrix(a, blk.U, 0, nr, -1, 0, v1, o1, -1, vs);   // s = U^T in1
            add_matrix(b, blk.V, 0, nc, vs, 0, v0, o0, d1, vt);   // out1 = V^T s, t = V in0
            add_matrix(c, blk.U, 0, nr, vt, 0, -1, 0, d0, -1);    // out0 = U t
        } else {          // V^T twice
            add_matrix(a, blk.V, 0, nc, -1, 0, v0, o0, -1, vt);   // t = V in0
            add_matrix(b, blk.U, 0, nr, vt, 0, v1, o1, d0, vs);   // out0 = U t, s = U^T in1
            add_matrix(c, blk.V, 0, nc, vs, 0, -1, 0, d1, -1);    // out1 = V^T s
        }
    }

    // Reduce tasks at stage st: slot dst (or external output) = sum of srcs.
    void emit_reduce(int st, long long dst, bool external, int len, const std::vector<Src> &srcs) {
        const int sb = (int)P.rsrc.size();
        for (const Src &s : srcs) P.rsrc.push_back(s.off);
        for (int e0 = 0; e0 < len; e0 += rchunk) {
            RTask r;
            memset(&r, 0, sizeof(r));
            r.dst = dst + e0;
            r.len = std::min(rchunk, len - e0);
            r.src_beg = sb;
            r.src_end = (int)P.rsrc.size();
            r.soff = e0;
            r.flags = external ? 1 : 0;
            SR[st].push_back(r);
        }
    }

    // Replaces the sources written before stage c - 1 by one fold slot written
    // in stage c - 1 when that saves reads (consumer tiles: `reads` elements
    // in total) or, for reduce consumers, when there are many of them.
    // Stages of one stream are [lo, c); a fold runs in stage c - 1 of that stream.
    std::vector<Src> fold(const std::vector<Src> &L, int c, double reads, int len, bool reducer, int depth) {
        const int lo = c > sG ? sS : 0;
        if (c - 1 < lo || depth > 8) return L;
        std::vector<Src> early, late;
        for (const Src &s : L) {
            if (s.stage >= lo && s.stage <= c - 2) early.push_back(s);
            else late.push_back(s);
        }
        const double m = (double)early.size();
        if (m < 2) return L;
        const bool worth = reducer ? m >= 4 : (m - 1) * reads > (m + 1) * (double)len;
        if (!worth) return L;
        std::vector<Src> sub = fold(early, c - 1, len, len, true, depth + 1);
        const long long slot = walloc(len);
        emit_reduce(c - 1, slot, false, len, sub);
        P.n_fold++;
        std::vector<Src> out;
        out.push_back(Src{(int)slot, c - 1});
        out.insert(out.end(), late.begin(), late.end());
        return out;
    }

    std::pair<int, int> emit_vec(int vec) {
        VVec &v = vv[vec];
        if (v.lb < 0) {
            v.lb = (int)P.vsrc.size();
            for (const Src &s : v.srcs) P.vsrc.push_back(s.off);
            v.le = (int)P.vsrc.size();
        }
        return {v.lb, v.le};
    }

    bool build() {
        const int n_node = h->n_node, nleaf = h->n_leaf_node, N = h->krnl_mat_size;
        P = Plan();
        P.nv = nv;
        // the external input: one vector whose only source is the encoded "-1"
        const int X = new_vec(N);
        vv[X].srcs.push_back(Src{-1, -1});
        std::vector<int> y0(n_node, -1), y1(n_node, -1), yv(n_node, -1);
        long long tot_r = 0;
        for (int i = 0; i < n_node; i++) {
            const int rk = S.U[i].off >= 0 ? S.U[i].ncol : 0;
            tot_r += rk;
            if (rk > 0) { y0[i] = new_vec(rk); y1[i] = new_vec(rk); }
            if (h->n_child[i] == 0) yv[i] = new_vec(h->mat_cluster[2 * i + 1] - h->mat_cluster[2 * i] + 1);
        }
        auto xo = [&](int node) { return h->mat_cluster[2 * node]; };
        // stages along the critical path
        const int minL = h->min_adm_level, maxL = h->max_level;
        const int nU = (minL >= 0 && minL <= maxL) ? (maxL - minL + 1) : 0;
        const int s0 = nf_schedule == 1 ? 3 : 0;
        const int sBA = s0 + nU, sBB = sBA + 1, sBC = sBA + 2;
        const int sD0 = sBA + 3;
        sG = sD0 + nU;
        sS = sG + 1;
        // schedule >= 2: near field on its own stream in g = schedule - 1
        // byte-balanced groups of blocks, software-pipelined over g + 2
        // launches: launch k runs the first pass of group k, the second of
        // group k - 1 and the third of group k - 2.  A low-rank block's
        // smaller half is then read again about two groups of bytes after
        // its first read, which a small enough group keeps within reach of
        // the last-level cache.
        const int ngroup = nf_schedule >= 2 ? nf_schedule - 1 : 0;
        std::vector<double> gload(std::max(ngroup, 1), 0.0);
        nst = ngroup > 0 ? sS + ngroup + 2 : sG + 1;
        SR.assign(nst, {});
        sload.assign(nst, 0.0);
        sphase.assign(nst, PH_GATHER);
        for (int q = 0; q < s0; q++) sphase[q] = PH_D;
        for (int q = 0; q < nU; q++) { sphase[s0 + q] = PH_UP; sphase[sD0 + q] = PH_DOWN; }
        sphase[sBA] = PH_B; sphase[sBB] = PH_B; sphase[sBC] = PH_Y1;
        for (int q = sS; q < nst; q++) sphase[q] = PH_D;
        // y0 of a node is needed when a coupling block reads it or its parent's is needed
        std::vector<char> need0(n_node, 0);
        for (int i = 0; i < h->n_B; i++) {
            const int n0 = h->r_adm_pairs[2 * i], n1 = h->r_adm_pairs[2 * i + 1];
            const int l0 = h->node_level[n0], l1 = h->node_level[n1];
            if (l0 >= l1) need0[n0] = 1;
            if (l1 >= l0) need0[n1] = 1;
        }
        for (int lv = 1; lv <= maxL; lv++)
            for (int j = 0; j < h->level_n_node[lv]; j++) {
                const int node = h->level_nodes[(size_t)lv * nleaf + j];
                if (h->parent[node] >= 0 && need0[h->parent[node]]) need0[node] = 1;
            }
        // ---- upward sweep: level maxL - q at stage s0 + q; inner nodes tile by child
        for (int lv = maxL; nU > 0 && lv >= minL; lv--) {
            const int st = s0 + maxL - lv;
            for (int j = 0; j < h->level_n_node[lv]; j++) {
                const int node = h->level_nodes[(size_t)lv * nleaf + j];
                if (y0[node] < 0 || !need0[node]) continue;
                const MatRef &U = S.U[node];
                if (h->n_child[node] == 0) {
                    add_matrix(st, U, 0, U.nrow, -1, 0, X, xo(node), -1, y0[node]);
                } else {
                    int row = 0;
                    for (int k = 0; k < h->n_child[node]; k++) {
                        const int ch = h->children[(size_t)node * h->max_child + k];
                        const int rc = S.U[ch].off >= 0 ? S.U[ch].ncol : 0;
                        if (rc > 0) add_matrix(st, U, row, rc, -1, 0, y0[ch], 0, -1, y0[node]);
                        row += rc;
                    }
                    if (row != U.nrow) { set_error(ERR_ARG, "device plan: U row count of node %d", node); return false; }
                }
            }
        }
        // ---- coupling blocks in stages sBA..sBC
        for (int i = 0; i < h->n_B; i++) {
            const int n0 = h->r_adm_pairs[2 * i], n1 = h->r_adm_pairs[2 * i + 1];
            const int l0 = h->node_level[n0], l1 = h->node_level[n1];
            int v0, o0, v1, o1, d0, d1;
            if (l0 == l1) { v0 = y0[n1]; o0 = 0; v1 = y0[n0]; o1 = 0; d0 = y1[n0]; d1 = y1[n1]; }
            else if (l0 > l1) { v0 = X; o0 = xo(n1); v1 = y0[n0]; o1 = 0; d0 = y1[n0]; d1 = yv[n1]; }
            else { v0 = y0[n1]; o0 = 0; v1 = X; o1 = xo(n0); d0 = yv[n0]; d1 = y1[n1]; }
            if (v0 < 0 || v1 < 0 || d0 < 0 || d1 < 0) continue;   // zero-rank side: no contribution
            add_block(S.B[i], v0, o0, v1, o1, d0, d1, sBA, sBB, sBC);
        }
        // ---- downward sweep: level minL + q at stage sD0 + q; inner nodes tile by child
        for (int lv = minL; nU > 0 && lv <= maxL; lv++) {
            const int st = sD0 + (lv - minL);
            for (int j = 0; j < h->level_n_node[lv]; j++) {
                const int node = h->level_nodes[(size_t)lv * nleaf + j];
                if (y1[node] < 0 || vv[y1[node]].srcs.empty()) continue;   // nothing to push down
                const MatRef &U = S.U[node];
                if (h->n_child[node] == 0) {
                    add_matrix(st, U, 0, U.nrow, y1[node], 0, -1, 0, yv[node], -1);
                } else {
                    int row = 0;
                    for (int k = 0; k < h->n_child[node]; k++) {
                        const int ch = h->children[(size_t)node * h->max_child + k];
                        const int rc = S.U[ch].off >= 0 ? S.U[ch].ncol : 0;
                        if (rc > 0) add_matrix(st, U, row, rc, y1[node], 0, -1, 0, y1[ch], -1);
                        row += rc;
                    }
                }
            }
        }
        // ---- near field: depends on x only, balanced over the tile launches
        std::vector<int> order(h->n_D);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return blk_elems(h->c_D_blks[a]) > blk_elems(h->c_D_blks[b]); });
        for (int i : order) {
            const BlkRef &blk = S.D[i];
            if (!blk.lr && blk.M.off < 0) continue;
            int v0 = X, o0, v1 = X, o1, d0, d1;
            if (i < nleaf) {
                const int node = h->height_nodes[i];
                o0 = xo(node); o1 = 0; d0 = yv[node]; d1 = -1; v1 = -1;
            } else {
                const int n0 = h->r_inadm_pairs[2 * (i - nleaf)], n1 = h->r_inadm_pairs[2 * (i - nleaf) + 1];
                o0 = xo(n1); o1 = xo(n0); d0 = yv[n0]; d1 = yv[n1];
            }
            if (d0 < 0 || (d1 < 0 && i >= nleaf)) { set_error(ERR_ARG, "device plan: near-field block %d on a non-leaf", i); return false; }
            const double bytes = 8.0 * (double)blk_elems(h->c_D_blks[i]);
            const int span = !blk.lr ? 1 : (d1 < 0 ? 2 : 3);
            int best = 0;
            if (ngroup > 0) {
                const int g = (int)(std::min_element(gload.begin(), gload.end()) - gload.begin());
                gload[g] += bytes;
                best = sS + g;
            } else if (nf_schedule == 0) {
                double bestv = 1e300;
                for (int a = 0; a + span - 1 < sG; a++) {
                    double v = 0;
                    for (int q = 0; q < span; q++) v = std::max(v, sload[a + q] + bytes / span);
                    if (v < bestv) { bestv = v; best = a; }
                }
            }
            add_block(blk, v0, o0, v1, o1, d0, d1, best, best + 1, best + 2);
        }
        if (has_error()) return false;
        // ---- folds of the tile inputs, then the final gather of every leaf
        for (size_t q = 0; q < vv.size(); q++) {
            VVec &v = vv[q];
            if ((int)q == X || v.fc == INT_MAX) continue;
            v.srcs = fold(v.srcs, v.fc, v.reads, v.len, false, 0);
        }
        for (int i = 0; i < n_node; i++) {
            if (yv[i] < 0) continue;
            const VVec &v = vv[yv[i]];
            if (v.srcs.empty()) { set_error(ERR_ARG, "device plan: leaf %d receives no contribution", i); return false; }
            std::vector<Src> mainS, sideS;
            for (const Src &q : v.srcs) (q.stage > sG ? sideS : mainS).push_back(q);
            std::vector<Src> all = fold(mainS, sG, v.len, v.len, true, 0);
            if (!sideS.empty()) {
                std::vector<Src> f = fold(sideS, nst, v.len, v.len, true, 0);
                all.insert(all.end(), f.begin(), f.end());
            }
            emit_reduce(sG, xo(i), true, v.len, all);
        }
        finish();
        P.bytes_vec = 8.0 * nv * (2.0 * N + 4.0 * (double)tot_r);
        if (P.nW * nv >= (1LL << 31)) { set_error(ERR_UNSUPPORTED, "workspace exceeds 2^31 doubles"); return false; }
        return validate();
    }

    void finish() {
        std::vector<std::vector<BTask>> S64(nst), S32(nst), W64(nst), W32(nst);
        for (PTile &p : tiles) {
            if (p.iv0 >= 0) { auto r = emit_vec(p.iv0); p.t.v0b = r.first; p.t.v0e = r.second; }
            if (p.iv1 >= 0) { auto r = emit_vec(p.iv1); p.t.v1b = r.first; p.t.v1e = r.second; }
            if (p.wave) (p.f32 ? W32 : W64)[p.st].push_back(p.t);
            else (p.f32 ? S32 : S64)[p.st].push_back(p.t);
        }
        auto key = [](const BTask &t) { return (long long)t.nrow * t.ncol; };
        // Staged executor at 8 densities per pass: a tile with both outputs
        // runs as two one-sided tasks (row-dot and column tile; together
        // faster than the two-sided tile at this width, whose row sums need a
        // reduction over up to 64 lanes per row).  The halves of eight
        // consecutive tiles are placed eight blocks apart: workgroups go to the
        // 8 XCDs round robin, so the two halves of a tile run on one XCD at
        // about the same time and the second read of the matrix is mostly
        // served by its L2.  (The persistent executor's claim order gives no
        // such placement, and there the split is slower.)
        auto split_two_sided = [&](std::vector<BTask> &L) {
            if (!split2) return;
            std::vector<BTask> out, ga, gb;
            out.reserve(2 * L.size());
            for (const BTask &t : L) {
                if (t.out0 < 0 || t.out1 < 0) { out.push_back(t); continue; }
                BTask a = t, b = t;
                a.out1 = -1;
                b.out0 = -1;
                ga.push_back(a);
                gb.push_back(b);
                if (ga.size() == 8) {
                    out.insert(out.end(), ga.begin(), ga.end());
                    out.insert(out.end(), gb.begin(), gb.end());
                    ga.clear();
                    gb.clear();
                }
            }
            for (size_t i = 0; i < ga.size(); i++) { out.push_back(ga[i]); out.push_back(gb[i]); }
            L.swap(out);
        };
        for (int st = 0; st < nst; st++) {
            std::stable_sort(S64[st].begin(), S64[st].end(), [&](const BTask &x, const BTask &y) { return key(x) > key(y); });
            std::stable_sort(S32[st].begin(), S32[st].end(), [&](const BTask &x, const BTask &y) { return key(x) > key(y); });
            split_two_sided(S64[st]);
            split_two_sided(S32[st]);
            std::stable_sort(W64[st].begin(), W64[st].end(), [&](const BTask &x, const BTask &y) { return key(x) > key(y); });
            std::stable_sort(W32[st].begin(), W32[st].end(), [&](const BTask &x, const BTask &y) { return key(x) > key(y); });
            Stage s;
            s.phase = sphase[st];
            s.stream = st > sG ? 1 : 0;
            s.mat_bytes = 0;
            for (const BTask &t : S64[st]) s.mat_bytes += 8.0 * t.nrow * t.ncol;
            for (const BTask &t : S32[st]) s.mat_bytes += 4.0 * t.nrow * t.ncol;
            for (const BTask &t : W64[st]) s.mat_bytes += 8.0 * t.nrow * t.ncol;
            for (const BTask &t : W32[st]) s.mat_bytes += 4.0 * t.nrow * t.ncol;
            s.t64_beg = (int)P.t64.size();
            s.t64_cnt = (int)S64[st].size();
            s.t32_beg = (int)P.t32.size();
            s.t32_cnt = (int)S32[st].size();
            P.t64.insert(P.t64.end(), S64[st].begin(), S64[st].end());
            P.t32.insert(P.t32.end(), S32[st].begin(), S32[st].end());
            s.w64_beg = (int)P.w64.size();
            s.w64_cnt = (int)W64[st].size();
            s.w32_beg = (int)P.w32.size();
            s.w32_cnt = (int)W32[st].size();
            P.w64.insert(P.w64.end(), W64[st].begin(), W64[st].end());
            P.w32.insert(P.w32.end(), W32[st].begin(), W32[st].end());
            s.r_beg = (int)P.rt.size();
            s.r_cnt = (int)SR[st].size();
            P.rt.insert(P.rt.end(), SR[st].begin(), SR[st].end());
            if (s.t64_cnt + s.t32_cnt + s.r_cnt + s.w64_cnt + s.w32_cnt > 0) P.stages.push_back(s);
        }
    }

    // Every access of every launch checked on the host before any launch.
    bool validate() const {
        const long long N = h->krnl_mat_size;
        auto src_ok = [&](int b, int e, long long off, long long len) {
            if (b < 0 || e > (int)P.vsrc.size() || b >= e) return false;
            if (P.vsrc[b] < 0) return e == b + 1 && -(long long)P.vsrc[b] - 1 + off >= 0 && -(long long)P.vsrc[b] - 1 + off + len <= N;
            for (int j = b; j < e; j++) {
                const long long s = P.vsrc[j];
                if (s < 0 || s + off < 0 || s + off + len > P.nW) return false;
            }
            return true;
        };
        auto chk = [&](const std::vector<BTask> &tl, bool f32, int xcap, int ycap) {
            const long long nstore = f32 ? (long long)S.h32.size() : (long long)S.h64.size();
            const int W = !f32 ? 2 : 4;   // widest vector load of any tile path (FP32: float4)
            for (const BTask &t : tl) {
                bool ok = t.nrow > 0 && t.ncol > 0 && t.ld >= t.ncol && t.ld % W == 0 && t.mat % W == 0 && t.mat >= 0 &&
                          (t.out0 >= 0 || t.out1 >= 0);
                if (nv == 16) {   // the MFMA tile: 64 columns with both products, 128 with a column product alone
                    const int Rp = (t.nrow + 15) & ~15;
                    ok = ok && (t.out1 < 0 || (t.out0 >= 0 ? t.ncol <= T16_MAXC2 : (t.ncol <= T16_MAXC && Rp * 16 <= T16_LDS)));
                } else {
                    ok = ok && t.ncol <= maxc && t.ncol * nv <= xcap && (t.lpr & (t.lpr - 1)) == 0 && t.lpr >= 4 &&
                         t.lpr <= 64 && t.lpr * cpl_of(nv) >= t.ncol && (t.out1 < 0 || t.nrow * nv <= ycap);
                }
                const long long last = t.mat + (long long)(t.nrow - 1) * t.ld + ((t.ncol - 1) / W) * W + W;
                ok = ok && last <= nstore;
                if (t.out0 >= 0) ok = ok && src_ok(t.v0b, t.v0e, t.v0off, t.ncol) && t.out0 + (long long)t.nrow <= P.nW;
                if (t.out1 >= 0) ok = ok && src_ok(t.v1b, t.v1e, t.v1off, t.nrow) && t.out1 + (long long)t.ncol <= P.nW;
                if (!ok) {
                    set_error(ERR_ARG,
                              "device plan check failed (nv %d): tile mat=%lld %dx%d ld=%d lpr=%d in0 [%d,%d)+%d "
                              "in1 [%d,%d)+%d out %d %d",
                              nv, t.mat, t.nrow, t.ncol, t.ld, t.lpr, t.v0b, t.v0e, t.v0off, t.v1b, t.v1e, t.v1off,
                              t.out0, t.out1);
                    return false;
                }
            }
            return true;
        };
        if (nv == 16 && (!P.w64.empty() || !P.w32.empty())) { set_error(ERR_ARG, "device plan check failed: wave tiles at 16 vectors"); return false; }
        if (!chk(P.t64, false, XIN_DOUBLES, YIN_DOUBLES) || !chk(P.t32, true, XIN_DOUBLES, YIN_DOUBLES) ||
            !chk(P.w64, false, TileLDS<64>::X, TileLDS<64>::Y) || !chk(P.w32, true, TileLDS<64>::X, TileLDS<64>::Y))
            return false;
        for (const RTask &r : P.rt) {
            bool ok = r.len > 0 && r.len <= rchunk && r.len * nv <= 1024 && r.dst >= 0 && r.soff >= 0 && r.src_beg >= 0 &&
                      r.src_end <= (int)P.rsrc.size() && r.src_beg < r.src_end;
            ok = ok && ((r.flags & 1) ? r.dst + r.len <= N : r.dst + r.len <= P.nW);
            for (int j = r.src_beg; ok && j < r.src_end; j++) ok = P.rsrc[j] >= 0 && (long long)P.rsrc[j] + r.soff + r.len <= P.nW;
            if (!ok) { set_error(ERR_ARG, "device plan check failed: reduce dst=%lld len=%d", r.dst, r.len); return false; }
        }
        for (const Stage &s : P.stages)
            if (s.t64_beg + s.t64_cnt > (int)P.t64.size() || s.t32_beg + s.t32_cnt > (int)P.t32.size() ||
                s.r_beg + s.r_cnt > (int)P.rt.size() || s.w64_beg + s.w64_cnt > (int)P.w64.size() ||
                s.w32_beg + s.w32_cnt > (int)P.w32.size()) {
                set_error(ERR_ARG, "device plan check failed: stage ranges");
                return false;
            }
        return true;
    }
};

}  // namespace

// Wave tiles (parts up to wave_bytes) and split two-sided tiles belong to the
// staged executor only.
static bool build_plan(H2ERI_p h, const Storage &S, Plan &P, int nv, double tb, int nfs, int executor,
                       double wave_bytes) {
    Planner pl(h, S, P, nv, tb, nfs, rchunk_for(nv), executor == 0 ? wave_bytes : 0.0, executor == 0);
    return pl.build();
}

// Queue form of a plan (persistent executor): the tasks of all its stages in
// one topological order, with their input and output lists.
struct QPlan {
    std::vector<QTask> tasks;   // ticket order
    std::vector<int> need;      // producer count of every list
    int nstatic = 0;            // leading tasks without inputs (at most one per workgroup; statistics)
    double est_us = 0;          // list-scheduling estimate of the pass
};

// Dependency graph of one plan's tasks for the persistent executor.  Lists
// are the input vectors of the tiles (vsrc ranges) and the source lists of
// the reduce tasks (rsrc ranges); a workspace slot belongs to at most one
// list, and a task counts itself into the list of every slot it writes.
struct QGraph {
    std::vector<QTask> q;                  // plan order; list ids local to the plan
    std::vector<int> nprod;                // producers per list
    std::vector<std::vector<int>> cons;    // consumers per list
    std::vector<double> dur;               // modelled duration per task
};

static bool queue_graph(const Plan &P, int nv, int G, QGraph &Gr) {
    Gr = QGraph();
    if (!P.w64.empty() || !P.w32.empty()) { set_error(ERR_ARG, "queue plan: wave tiles are not queued"); return false; }
    struct TRef { int kind, idx; };
    std::vector<TRef> T;
    for (const Stage &st : P.stages) {
        for (int i = 0; i < st.t64_cnt; i++) T.push_back({0, st.t64_beg + i});
        for (int i = 0; i < st.t32_cnt; i++) T.push_back({1, st.t32_beg + i});
        for (int i = 0; i < st.r_cnt; i++) T.push_back({2, st.r_beg + i});
    }
    const int n = (int)T.size();
    // lists, keyed by the begin index of their range
    std::vector<int> vlist(P.vsrc.size() + 1, -1), rlist(P.rsrc.size() + 1, -1);
    std::vector<std::pair<int, int>> lrange;   // [b, e) in vsrc (kind 0) or rsrc (kind 1)
    std::vector<char> lkind;
    auto list_of = [&](int kind, int b, int e) -> int {
        std::vector<int> &m = kind == 0 ? vlist : rlist;
        const std::vector<int> &src = kind == 0 ? P.vsrc : P.rsrc;
        if (b < 0 || e <= b || e > (int)src.size() || src[b] < 0) return -1;   // external input or nothing
        if (m[b] < 0) {
            m[b] = (int)lrange.size();
            lrange.push_back({b, e});
            lkind.push_back((char)kind);
        } else if (lrange[m[b]].second != e) {
            return -2;
        }
        return m[b];
    };
    std::vector<QTask> q(n);
    for (int i = 0; i < n; i++) {
        QTask &t = q[i];
        memset(&t, 0, sizeof(t));
        t.kind = T[i].kind;
        t.idx = T[i].idx;
        t.in0 = t.in1 = t.out0 = t.out1 = -1;
        if (t.kind < 2) {
            const BTask &b = (t.kind == 0 ? P.t64 : P.t32)[t.idx];
            if (b.out0 >= 0) t.in0 = list_of(0, b.v0b, b.v0e);
            if (b.out1 >= 0) t.in1 = list_of(0, b.v1b, b.v1e);
        } else {
            const RTask &r = P.rt[t.idx];
            t.in0 = list_of(1, r.src_beg, r.src_end);
        }
        if (t.in0 == -2 || t.in1 == -2) { set_error(ERR_ARG, "queue plan: overlapping source lists"); return false; }
    }
    const int nlist = (int)lrange.size();
    // slot -> list (slots are in increasing offset order)
    const std::vector<std::pair<long long, int>> &SL = P.slots;
    auto slot_at = [&](long long off) -> int {
        auto it = std::upper_bound(SL.begin(), SL.end(), std::make_pair(off, INT_MAX));
        if (it == SL.begin()) return -1;
        const int k = (int)(it - SL.begin()) - 1;
        return off < SL[k].first + SL[k].second ? k : -1;
    };
    std::vector<int> slot_list(SL.size(), -1);
    for (int L = 0; L < nlist; L++) {
        const std::vector<int> &src = lkind[L] == 0 ? P.vsrc : P.rsrc;
        for (int j = lrange[L].first; j < lrange[L].second; j++) {
            const int k = slot_at(src[j]);
            if (k < 0 || SL[k].first != src[j]) { set_error(ERR_ARG, "queue plan: list entry %d is not a slot", src[j]); return false; }
            if (slot_list[k] >= 0 && slot_list[k] != L) { set_error(ERR_ARG, "queue plan: slot in two lists"); return false; }
            slot_list[k] = L;
        }
    }
    // outputs -> lists
    std::vector<int> nprod(nlist, 0);
    auto out_list = [&](long long off, int len, int &out) -> bool {
        const int k = slot_at(off);
        if (k < 0 || off + len > SL[k].first + SL[k].second) { set_error(ERR_ARG, "queue plan: output outside its slot"); return false; }
        out = slot_list[k];
        return true;
    };
    for (int i = 0; i < n; i++) {
        QTask &t = q[i];
        if (t.kind < 2) {
            const BTask &b = (t.kind == 0 ? P.t64 : P.t32)[t.idx];
            if (b.out0 >= 0 && !out_list(b.out0, b.nrow, t.out0)) return false;
            if (b.out1 >= 0 && !out_list(b.out1, b.ncol, t.out1)) return false;
        } else {
            const RTask &r = P.rt[t.idx];
            if (!(r.flags & 1) && !out_list(r.dst, r.len, t.out0)) return false;
        }
        if (t.out1 >= 0 && t.out1 == t.out0) t.out1 = -1;
        if (t.out0 >= 0) nprod[t.out0]++;
        if (t.out1 >= 0) nprod[t.out1]++;
    }
    std::vector<std::vector<int>> cons(nlist);
    for (int i = 0; i < n; i++) {
        if (q[i].in1 == q[i].in0) q[i].in1 = -1;
        for (int L : {q[i].in0, q[i].in1}) {
            if (L < 0) continue;
            if (nprod[L] == 0) { set_error(ERR_ARG, "queue plan: a list without producers is read"); return false; }
            cons[L].push_back(i);
        }
    }
    // duration model: fixed cost (the counter wait, input staging and output
    // reductions of a task, as measured) plus bytes at the per-workgroup
    // share of HBM
    const double c0 = 6.0e-6, bw = 5.0e12 / std::max(G, 1);
    std::vector<double> dur(n);
    for (int i = 0; i < n; i++) {
        const QTask &t = q[i];
        double bytes;
        if (t.kind < 2) {
            const BTask &b = (t.kind == 0 ? P.t64 : P.t32)[t.idx];
            bytes = (t.kind == 0 ? 8.0 : 4.0) * b.nrow * b.ncol;
            if (b.out0 >= 0) bytes += 8.0 * nv * ((double)b.ncol * (b.v0e - b.v0b) + b.nrow);
            if (b.out1 >= 0) bytes += 8.0 * nv * ((double)b.nrow * (b.v1e - b.v1b) + b.ncol);
        } else {
            const RTask &r = P.rt[t.idx];
            bytes = 8.0 * nv * r.len * (double)(r.src_end - r.src_beg + 1);
        }
        dur[i] = c0 + bytes / bw;
    }
    Gr.q = std::move(q);
    Gr.nprod = std::move(nprod);
    Gr.cons = std::move(cons);
    Gr.dur = std::move(dur);
    return true;
}

// Orders the tasks of several graphs (disjoint lists: one per molecule of a
// group, or one) by list scheduling on G workgroups with the modelled
// durations, longest remaining path first, so the sweeps' critical paths run
// as soon as they can and the near field fills the gaps.  List ids become
// global (offset per graph) and QTask::pad_[0] names the graph.
static bool queue_schedule(const std::vector<QGraph> &gs, int G, QPlan &Q) {
    Q = QPlan();
    std::vector<QTask> q;
    std::vector<int> nprod;
    std::vector<std::vector<int>> cons;
    std::vector<double> dur;
    for (size_t gi = 0; gi < gs.size(); gi++) {
        const QGraph &g = gs[gi];
        const int tb = (int)q.size(), lb = (int)nprod.size();
        for (QTask t : g.q) {
            for (int *L : {&t.in0, &t.in1, &t.out0, &t.out1})
                if (*L >= 0) *L += lb;
            t.pad_[0] = (int)gi;
            q.push_back(t);
        }
        nprod.insert(nprod.end(), g.nprod.begin(), g.nprod.end());
        for (const std::vector<int> &c : g.cons) {
            cons.emplace_back();
            for (int i : c) cons.back().push_back(tb + i);
        }
        dur.insert(dur.end(), g.dur.begin(), g.dur.end());
    }
    const int n = (int)q.size(), nlist = (int)nprod.size();
    // topological order (Kahn), then the longest remaining path of every task
    std::vector<int> rem(n, 0), cnt(nlist, 0), topo;
    topo.reserve(n);
    for (int i = 0; i < n; i++) rem[i] = (q[i].in0 >= 0) + (q[i].in1 >= 0);
    for (int i = 0; i < n; i++) if (rem[i] == 0) topo.push_back(i);
    auto complete = [&](int i, std::vector<int> &cnt_, std::vector<int> &rem_, auto &&on_ready) {
        for (int L : {q[i].out0, q[i].out1}) {
            if (L < 0 || ++cnt_[L] != nprod[L]) continue;
            for (int c : cons[L])
                if (--rem_[c] == 0) on_ready(c);
        }
    };
    {
        std::vector<int> r2 = rem;
        for (size_t h = 0; h < topo.size(); h++) complete(topo[h], cnt, r2, [&](int c) { topo.push_back(c); });
    }
    if ((int)topo.size() != n) { set_error(ERR_ARG, "queue plan: dependency cycle"); return false; }
    std::vector<double> bl(n, 0.0), lbl(nlist, 0.0);
    for (int h = n - 1; h >= 0; h--) {
        const int i = topo[h];
        double m = 0.0;
        for (int L : {q[i].out0, q[i].out1}) if (L >= 0) m = std::max(m, lbl[L]);
        bl[i] = dur[i] + m;
        for (int L : {q[i].in0, q[i].in1}) if (L >= 0) lbl[L] = std::max(lbl[L], bl[i]);
    }
    // list scheduling on G workgroups, longest remaining path first
    std::priority_queue<std::pair<double, int>> ready;
    typedef std::pair<double, int> Ev;
    std::priority_queue<Ev, std::vector<Ev>, std::greater<Ev>> running;
    std::fill(cnt.begin(), cnt.end(), 0);
    for (int i = 0; i < n; i++) if (rem[i] == 0) ready.push({bl[i], i});
    std::vector<int> order;
    order.reserve(n);
    int free_slots = std::max(G, 1);
    double now = 0.0;
    while ((int)order.size() < n) {
        while (free_slots > 0 && !ready.empty()) {
            const int i = ready.top().second;
            ready.pop();
            order.push_back(i);
            running.push({now + dur[i], i});
            free_slots--;
        }
        if (running.empty()) { set_error(ERR_ARG, "queue plan: scheduling stalled"); return false; }
        const Ev e = running.top();
        running.pop();
        now = e.first;
        free_slots++;
        complete(e.second, cnt, rem, [&](int c) { ready.push({bl[c], c}); });
    }
    while (!running.empty()) { now = std::max(now, running.top().first); running.pop(); }
    Q.est_us = now * 1e6;
    // ticket order; every producer of a task's inputs precedes it
    std::vector<int> pos(n);
    for (int k = 0; k < n; k++) pos[order[k]] = k;
    std::vector<int> last_prod(nlist, -1);
    for (int i = 0; i < n; i++)
        for (int L : {q[i].out0, q[i].out1}) if (L >= 0) last_prod[L] = std::max(last_prod[L], pos[i]);
    Q.tasks.resize(n);
    for (int k = 0; k < n; k++) {
        const QTask &t = q[order[k]];
        for (int L : {t.in0, t.in1})
            if (L >= 0 && last_prod[L] >= k) { set_error(ERR_ARG, "queue plan: order is not topological"); return false; }
        Q.tasks[k] = t;
    }
    while (Q.nstatic < std::min(n, G) && Q.tasks[Q.nstatic].in0 < 0 && Q.tasks[Q.nstatic].in1 < 0) Q.nstatic++;
    Q.need = nprod;
    return true;
}

static bool build_queue(const Plan &P, int nv, int G, QPlan &Q) {
    std::vector<QGraph> gs(1);
    if (!queue_graph(P, nv, G, gs[0])) return false;
    return queue_schedule(gs, G, Q);
}

static bool ensure_events(DevState *d, size_t n) {
    while (d->evs.size() < n) {
        hipEvent_t e;
        HIP_OK(hipEventCreate(&e));
        d->evs.push_back(e);
    }
    return true;
}

static bool upload_storage(H2ERI_p h) {
    DevState *d = dstate_get(h);
    if (d->st_ready) return true;
    if (!h->U || !h->c_B_blks || !h->c_D_blks) { set_error(ERR_ARG, "no H2 representation built or loaded"); return false; }
    HIP_OK(hipGetDevice(&d->dev));
    release_buffers(d);
    if (h->BD_JIT) {   // U and the task lists; B and D are evaluated in each matvec
        if (!jit_upload(h, d->jit)) return false;
        d->resident_bytes += jit_resident_bytes(d->jit);
    } else {
        d->st.build(h, d->ub_bits == 32);
    }
    auto alloc_copy = [&](void **dst, const void *src, size_t bytes) -> bool {
        if (bytes == 0) bytes = 64;
        HIP_OK(hipMalloc(dst, bytes));
        if (src) HIP_OK(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
        d->resident_bytes += bytes;
        return true;
    };
    if (!alloc_copy((void **)&d->S64, d->st.h64.empty() ? nullptr : d->st.h64.data(), d->st.h64.size() * 8)) return false;
    if (!alloc_copy((void **)&d->S32, d->st.h32.empty() ? nullptr : d->st.h32.data(), d->st.h32.size() * 4)) return false;
    // Coulomb index map: x index -> (D element, 2x flag)
    if (h->num_bf > 0 && h->sp_shell_idx && h->shell_bf_sidx) {
        const int np = h->num_sp, nbf = h->num_bf, N = h->krnl_mat_size;
        if ((long long)nbf * nbf >= (1LL << 30)) { set_error(ERR_UNSUPPORTED, "num_bf too large for the index map"); return false; }
        std::vector<int> g(N);
        for (int i = 0; i < np; i++) {
            int s0 = h->sp_shell_idx[i], s1 = h->sp_shell_idx[i + np];
            if (s0 < 0 || s1 < 0 || s0 >= h->nshell || s1 >= h->nshell) {
                set_error(ERR_ARG, "Coulomb index map: shell pair %d names no shell", i);
                return false;
            }
            int srow = h->shell_bf_sidx[s0], nrow = h->shell_bf_sidx[s0 + 1] - srow;
            int scol = h->shell_bf_sidx[s1], ncol = h->shell_bf_sidx[s1 + 1] - scol;
            int two = (s0 != s1) ? 1 : 0;
            int xo = h->sp_bfp_sidx[i];
            if (xo < 0 || nrow < 0 || ncol < 0 || (long long)xo + (long long)nrow * ncol > N || srow < 0 || scol < 0 ||
                scol + ncol > nbf || srow + nrow > nbf) {
                set_error(ERR_ARG, "Coulomb index map: shell pair %d lies outside x or D", i);
                return false;
            }
            for (int j = 0; j < ncol; j++)
                for (int k = 0; k < nrow; k++) g[xo + j * nrow + k] = (((scol + j) * nbf + srow + k) << 1) | two;
        }
        if (!alloc_copy((void **)&d->d_gidx, g.data(), (size_t)N * sizeof(int))) return false;
    }
    d->st_ready = true;
    return true;
}

// Resident workgroups of the persistent kernel on the current device.
static int queue_grid(int nv) {
    int dev = 0, ncu = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    hipError_t e;
    switch (nv) {
    case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, queue_kernel<1, false>, WG, 0); break;
    case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, queue_kernel<2, false>, WG, 0); break;
    case 4: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, queue_kernel<4, false>, WG, 0); break;
    case 8: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, queue_kernel<8, false>, WG, 0); break;
    default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, queue_kernel<16, false>, WG, 0); break;
    }
    return e == hipSuccess ? ncu * std::max(per, 1) : 0;
}

// Queue form of an uploaded plan: built on the host, uploaded, counters zeroed.
static bool ensure_queue(DevState *d, DevPlan &dp) {
    if (dp.q_ready) return true;
    const int G = queue_grid(dp.P.nv);
    if (G <= 0) { set_error(ERR_DEVICE, "persistent executor: occupancy query failed"); return false; }
    QPlan Q;
    if (!build_queue(dp.P, dp.P.nv, G, Q)) return false;
    const size_t nq = std::max<size_t>(Q.tasks.size(), 1), nl = Q.need.size();
    HIP_OK(hipMalloc(&dp.d_q, nq * sizeof(QTask)));
    HIP_OK(hipMalloc(&dp.d_need, std::max<size_t>(nl, 1) * sizeof(int)));
    HIP_OK(hipMalloc(&dp.d_ctr, (QC_LISTS + nl) * sizeof(int)));
    if (!Q.tasks.empty()) HIP_OK(hipMemcpy(dp.d_q, Q.tasks.data(), Q.tasks.size() * sizeof(QTask), hipMemcpyHostToDevice));
    if (nl) HIP_OK(hipMemcpy(dp.d_need, Q.need.data(), nl * sizeof(int), hipMemcpyHostToDevice));
    HIP_OK(hipMemset(dp.d_ctr, 0, (QC_LISTS + nl) * sizeof(int)));
    HIP_OK(hipDeviceSynchronize());
    dp.q_n = (int)Q.tasks.size();
    dp.q_static = Q.nstatic;
    dp.q_nlist = (int)nl;
    dp.q_grid = std::min(G, std::max(dp.q_n, 1));
    dp.q_est_us = Q.est_us;
    dp.q_ready = true;
    (void)d;
    return true;
}

static DevPlan *get_plan(H2ERI_p h, int nv) {
    if (h->BD_JIT) {
        set_error(ERR_UNSUPPORTED, "a BD_JIT handle has no stored blocks to plan: plain and multi-vector matvec and "
                                   "Coulomb calls only");
        return nullptr;
    }
    if (!upload_storage(h)) return nullptr;
    DevState *d = dstate(h);
    DevPlan &dp = d->plans[nv_index(nv)];
    const Tune &c = d->cfg[nv_index(nv)];
    if (dp.ready) {
        if (c.executor == 1 && !ensure_queue(d, dp)) return nullptr;
        return &dp;
    }
    if (!build_plan(h, d->st, dp.P, nv, c.tile_bytes, c.nf_schedule, c.executor, c.wave_bytes)) return nullptr;
    const Plan &P = dp.P;
    auto alloc_copy = [&](void **dst, const void *src, size_t bytes) -> bool {
        if (bytes == 0) bytes = 64;
        hipError_t e = hipMalloc(dst, bytes);
        if (e == hipSuccess && src) e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) { set_error(ERR_DEVICE, "plan upload: %s", hipGetErrorString(e)); return false; }
        return true;
    };
    if (!alloc_copy((void **)&dp.Wv, nullptr, (size_t)P.nW * nv * 8)) return nullptr;
    if (hipMemset(dp.Wv, 0, (size_t)P.nW * nv * 8) != hipSuccess) { set_error(ERR_DEVICE, "workspace memset"); return nullptr; }
    if (!alloc_copy((void **)&dp.d_t64, P.t64.empty() ? nullptr : P.t64.data(), P.t64.size() * sizeof(BTask))) return nullptr;
    if (!alloc_copy((void **)&dp.d_t32, P.t32.empty() ? nullptr : P.t32.data(), P.t32.size() * sizeof(BTask))) return nullptr;
    if (!alloc_copy((void **)&dp.d_w64, P.w64.empty() ? nullptr : P.w64.data(), P.w64.size() * sizeof(BTask))) return nullptr;
    if (!alloc_copy((void **)&dp.d_w32, P.w32.empty() ? nullptr : P.w32.data(), P.w32.size() * sizeof(BTask))) return nullptr;
    if (!alloc_copy((void **)&dp.d_rt, P.rt.empty() ? nullptr : P.rt.data(), P.rt.size() * sizeof(RTask))) return nullptr;
    if (!alloc_copy((void **)&dp.d_rsrc, P.rsrc.empty() ? nullptr : P.rsrc.data(), P.rsrc.size() * sizeof(int))) return nullptr;
    if (!alloc_copy((void **)&dp.d_vsrc, P.vsrc.empty() ? nullptr : P.vsrc.data(), P.vsrc.size() * sizeof(int))) return nullptr;
    if (nv == 16 && !alloc_copy((void **)&dp.Xt, nullptr, (size_t)h->krnl_mat_size * 16 * sizeof(double))) return nullptr;
    dp.ready = true;
    if (c.executor == 1 && !ensure_queue(d, dp)) return nullptr;
    return &dp;
}

// Workgroups of one stage launch: a tile or reduce task each, four wave tiles each.
static inline int stage_blocks(const Stage &st) {
    return st.t64_cnt + st.t32_cnt + st.r_cnt + (st.w64_cnt + 3) / 4 + (st.w32_cnt + 3) / 4;
}

// Workspace and counters of the pass being enqueued.
static inline double *pass_Wv(const DevState *d, const DevPlan &dp) { return d->ws_Wv ? d->ws_Wv : dp.Wv; }
static inline int *pass_ctr(const DevState *d, const DevPlan &dp) { return d->ws_ctr ? d->ws_ctr : dp.d_ctr; }

// Workgroups of a launch: one per task; a near-field launch on the second
// stream takes at most the configured number, looping over its tasks, so the
// sweeps' launches beside it find free CUs.
static inline int launch_blocks(const Stage &st, const DevState *d, int nv) {
    const int nb = stage_blocks(st);
    const int cap = d->cfg[nv_index(nv)].nf_grid;
    return (st.stream == 1 && cap > 0) ? std::min(nb, cap) : nb;
}

template <int NV>
static void launch_stage(const Stage &st, const DevPlan &dp, DevState *d, const double *X, double *Y, long long ld,
                         hipStream_t s) {
    const int nb = stage_blocks(st), grid = launch_blocks(st, d, NV);
    hipLaunchKernelGGL(grid < nb ? stage_loop_kernel<NV> : stage_kernel<NV>, dim3(grid), dim3(WG), 0, s,
                       dp.d_t64 + st.t64_beg, st.t64_cnt, dp.d_t32 + st.t32_beg, st.t32_cnt, dp.d_rt + st.r_beg,
                       st.r_cnt, dp.d_w64 + st.w64_beg, st.w64_cnt, dp.d_w32 + st.w32_beg, st.w32_cnt, dp.d_vsrc,
                       dp.d_rsrc, d->S64, d->S32, pass_Wv(d, dp), NV == 16 ? dp.Xt : X, Y, ld);
}

static void collect_phase_ms(H2ERI_p h, DevState *d);

// Spin bound of one wait of a group's persistent kernel: about a second of
// polling, far beyond any dependency of a pass.
static const int QUEUE_SPIN_LIMIT = 1 << 22;

template <int NV>
static void launch_queue(const DevPlan &dp, DevState *d, const double *X, double *Y, long long ld, hipStream_t s) {
    hipLaunchKernelGGL((queue_kernel<NV, false>), dim3(dp.q_grid), dim3(WG), 0, s, dp.d_q, dp.q_n,
                       dp.d_need, pass_ctr(d, dp), d->spin_limit, (const MolPtrs *)nullptr, dp.d_t64, dp.d_t32, dp.d_rt,
                       dp.d_vsrc, dp.d_rsrc, d->S64, d->S32, pass_Wv(d, dp), NV == 16 ? dp.Xt : X, Y, ld);
}

// One pass of the persistent executor on stream s (reset + queue launch).
static bool run_queue(H2ERI_p h, DevState *d, DevPlan &dp, const double *X, double *Y, hipStream_t s) {
    if (!dp.q_ready) { set_error(ERR_DEVICE, "persistent executor: queue not built"); return false; }
    const long long ld = h->krnl_mat_size;
    const bool tim = d->phase_timing && !d->capturing;
    if (dp.P.nv == 16) {
        hipLaunchKernelGGL(transpose16_kernel, dim3((unsigned)((ld + 255) / 256)), dim3(256), 0, s, X, ld, dp.Xt, (int)ld);
        HIP_OK(hipGetLastError());
    }
    hipLaunchKernelGGL(queue_reset_kernel, dim3((std::max(dp.q_nlist, QC_LISTS) + 255) / 256), dim3(256), 0, s,
                       pass_ctr(d, dp), dp.q_nlist);
    HIP_OK(hipGetLastError());
    if (tim) {   // the event pair brackets the persistent launch only
        collect_phase_ms(h, d);
        if (!ensure_events(d, 2)) return false;
        d->launches.clear();
        d->collected = false;
        if (!d->ev_base) HIP_OK(hipEventCreate(&d->ev_base));
        HIP_OK(hipEventRecord(d->ev_base, s));
        HIP_OK(hipEventRecord(d->evs[0], s));
    }
    switch (dp.P.nv) {
    case 1: launch_queue<1>(dp, d, X, Y, ld, s); break;
    case 2: launch_queue<2>(dp, d, X, Y, ld, s); break;
    case 4: launch_queue<4>(dp, d, X, Y, ld, s); break;
    case 8: launch_queue<8>(dp, d, X, Y, ld, s); break;
    default: launch_queue<16>(dp, d, X, Y, ld, s); break;
    }
    HIP_OK(hipGetLastError());
    if (tim) {
        HIP_OK(hipEventRecord(d->evs[1], s));
        double mb = 0;
        for (const Stage &st : dp.P.stages) mb += st.mat_bytes;
        d->launches.push_back(LaunchRec{PH_N, dp.q_grid, mb, 0.f, 0.f});
    }
    return true;
}

// Enqueue one batched matvec of nv vectors: X (nv x N, vector-major) -> Y.
// In phase-timing mode an event pair brackets every launch.
static bool run_matvec(H2ERI_p h, DevState *d, DevPlan &dp, const double *X, double *Y, hipStream_t s) {
    if (d->cfg[nv_index(dp.P.nv)].executor == 1) return run_queue(h, d, dp, X, Y, s);
    const Plan &P = dp.P;
    const long long ld = h->krnl_mat_size;
    const bool tim = d->phase_timing && !d->capturing;
    bool has_side = false;
    for (const Stage &st : P.stages) has_side |= st.stream == 1;
    if (P.nv == 16) {   // element-major copy of the input, before the fork
        hipLaunchKernelGGL(transpose16_kernel, dim3((unsigned)((ld + 255) / 256)), dim3(256), 0, s, X, ld, dp.Xt, (int)ld);
        HIP_OK(hipGetLastError());
    }
    if (tim) {
        collect_phase_ms(h, d);   // a previous pass of the same call goes into the timers first
        if (!ensure_events(d, 2 * P.stages.size() + 2)) return false;
        d->launches.clear();
        d->collected = false;
        if (!d->ev_base) HIP_OK(hipEventCreate(&d->ev_base));
        HIP_OK(hipEventRecord(d->ev_base, s));   // before the fork: every launch starts after it
    }
    auto launch = [&](const Stage &st, hipStream_t q) -> bool {
        if (tim) HIP_OK(hipEventRecord(d->evs[2 * d->launches.size()], q));
        switch (P.nv) {
        case 1: launch_stage<1>(st, dp, d, X, Y, ld, q); break;
        case 2: launch_stage<2>(st, dp, d, X, Y, ld, q); break;
        case 4: launch_stage<4>(st, dp, d, X, Y, ld, q); break;
        case 8: launch_stage<8>(st, dp, d, X, Y, ld, q); break;
        default: launch_stage<16>(st, dp, d, X, Y, ld, q); break;
        }
        HIP_OK(hipGetLastError());
        if (tim) {
            HIP_OK(hipEventRecord(d->evs[2 * d->launches.size() + 1], q));
            d->launches.push_back(LaunchRec{st.phase, stage_blocks(st), st.mat_bytes, 0.f, 0.f});
        }
        return true;
    };
    if (has_side) {
        // near-field launches on their own stream, forked after everything
        // already enqueued on s and joined before the gather
        if (!d->side) HIP_OK(hipStreamCreateWithFlags(&d->side, hipStreamNonBlocking));
        if (!d->ev_fork) HIP_OK(hipEventCreateWithFlags(&d->ev_fork, hipEventDisableTiming));
        if (!d->ev_join) HIP_OK(hipEventCreateWithFlags(&d->ev_join, hipEventDisableTiming));
        HIP_OK(hipEventRecord(d->ev_fork, s));
        HIP_OK(hipStreamWaitEvent(d->side, d->ev_fork, 0));
        for (const Stage &st : P.stages)
            if (st.stream == 1 && !launch(st, d->side)) return false;
        HIP_OK(hipEventRecord(d->ev_join, d->side));
    }
    for (const Stage &st : P.stages) {
        if (st.stream != 0) continue;
        if (has_side && st.phase == PH_GATHER) HIP_OK(hipStreamWaitEvent(s, d->ev_join, 0));
        if (!launch(st, s)) return false;
    }
    return true;
}

// Per-launch and per-phase milliseconds of the last timed call.
static void collect_phase_ms(H2ERI_p h, DevState *d) {
    if (!d->phase_timing || d->launches.empty() || d->collected) return;
    d->collected = true;
    for (int p = 0; p < 7; p++) d->phase_ms[p] = 0.0;
    for (size_t i = 0; i < d->launches.size(); i++) {
        float ms = 0.f;
        if (hipEventSynchronize(d->evs[2 * i + 1]) != hipSuccess ||
            hipEventElapsedTime(&ms, d->evs[2 * i], d->evs[2 * i + 1]) != hipSuccess)
            ms = 0.f;
        d->launches[i].ms = ms;
        float t0 = 0.f;
        if (!d->ev_base || hipEventElapsedTime(&t0, d->ev_base, d->evs[2 * i]) != hipSuccess) t0 = 0.f;
        d->launches[i].start = t0;
        if (d->launches[i].phase < PH_N) d->phase_ms[d->launches[i].phase] += ms;
        else d->phase_ms[PH_D] += ms;   // one launch for the whole pass (persistent executor)
        d->phase_ms[6] += ms;
    }
    // reference timer slots (seconds): forward, intermediate, backward, dense, vector ops
    h->timers[MV_FWD_TIMER_IDX] += d->phase_ms[PH_UP] * 1e-3;
    h->timers[MV_MID_TIMER_IDX] += (d->phase_ms[PH_B] + d->phase_ms[PH_Y1]) * 1e-3;
    h->timers[MV_BWD_TIMER_IDX] += d->phase_ms[PH_DOWN] * 1e-3;
    h->timers[MV_DEN_TIMER_IDX] += d->phase_ms[PH_D] * 1e-3;
    h->timers[MV_VOP_TIMER_IDX] += d->phase_ms[PH_GATHER] * 1e-3;
}

// Batched device matvec: nvec vectors, processed in passes of 16, 8, 4, 2, 1.
static bool matvec_multi_dev_impl(H2ERI_p h, int nvec, const double *X, double *Y, hipStream_t s) {
    if (h->BD_JIT) {   // timed like the stored form: the reference's phase timer slots
        if (!upload_storage(h)) return false;
        DevState *d = dstate(h);
        const bool tim = d->phase_timing && !d->capturing;
        double ms5[5] = {0, 0, 0, 0, 0};
        if (!jit_matvec(h, d->jit, nvec, X, Y, s, tim ? ms5 : nullptr)) return false;
        if (tim) {
            h->timers[MV_FWD_TIMER_IDX] += ms5[0] * 1e-3;
            h->timers[MV_MID_TIMER_IDX] += ms5[1] * 1e-3;
            h->timers[MV_BWD_TIMER_IDX] += ms5[2] * 1e-3;
            h->timers[MV_DEN_TIMER_IDX] += ms5[3] * 1e-3;
            h->timers[MV_VOP_TIMER_IDX] += ms5[4] * 1e-3;
            const int ph[5] = {PH_UP, PH_B, PH_DOWN, PH_D, PH_GATHER};
            for (int p = 0; p < 7; p++) d->phase_ms[p] = 0.0;
            for (int i = 0; i < 5; i++) { d->phase_ms[ph[i]] = ms5[i]; d->phase_ms[6] += ms5[i]; }
        }
        return true;
    }
    const size_t N = (size_t)h->krnl_mat_size;
    int done = 0;
    while (done < nvec) {
        const int rem = nvec - done;
        const int nv = nv_pass(rem);
        DevPlan *dp = get_plan(h, nv);
        if (!dp) return false;
        if (!run_matvec(h, dstate(h), *dp, X + done * N, Y + done * N, s)) return false;
        done += nv;
    }
    return true;
}

static bool ensure_io(DevState *d, size_t n) {
    if (d->d_io_len >= n) return true;
    free_dev((void *&)d->d_x);
    free_dev((void *&)d->d_y);
    d->d_io_len = 0;
    HIP_OK(hipMalloc(&d->d_x, sizeof(double) * n));
    HIP_OK(hipMalloc(&d->d_y, sizeof(double) * n));
    d->d_io_len = n;
    return true;
}

// Host-pointer forms synchronise anyway, so they always time their launches
// and fill the reference's phase timers (src/H2ERI_matvec.c:890-954).
struct HostTiming {
    DevState *d;
    bool was;
    explicit HostTiming(DevState *d_) : d(d_), was(d_->phase_timing) { d->phase_timing = true; }
    ~HostTiming() { d->phase_timing = was; }
};

static int queue_flags(DevState *d);

static bool matvec_host_impl(H2ERI_p h, int nvec, const double *x, double *y) {
    if (!upload_storage(h)) return false;
    DevState *d = dstate(h);
    HostTiming tm(d);
    const size_t N = (size_t)h->krnl_mat_size;
    if (!ensure_io(d, N * nvec)) return false;
    HIP_OK(hipMemcpy(d->d_x, x, sizeof(double) * N * nvec, hipMemcpyHostToDevice));
    if (!matvec_multi_dev_impl(h, nvec, d->d_x, d->d_y, nullptr)) return false;
    HIP_OK(hipMemcpy(y, d->d_y, sizeof(double) * N * nvec, hipMemcpyDeviceToHost));
    collect_phase_ms(h, d);
    if (queue_flags(d) != 0) return false;
    h->n_matvec += nvec;
    return true;
}

// J_k = contract(A uncontract(D_k)) for ndens densities on the device, with
// x and y in xb / yb (N * ndens doubles each) and the half matrix in Jh
// (num_bf^2 doubles).
static bool coulomb_run(H2ERI_p h, int ndens, const double *D_dev, double *J_dev, hipStream_t s, double *xb,
                        double *yb, double *Jh) {
    DevState *d = dstate(h);
    const int N = h->krnl_mat_size, nbf = h->num_bf;
    const size_t nn = (size_t)nbf * nbf;
    for (int k = 0; k < ndens; k++) {
        hipLaunchKernelGGL(uncontract_kernel, dim3((N + 255) / 256), dim3(256), 0, s, d->d_gidx, D_dev + k * nn,
                           xb + (size_t)k * N, N);
        HIP_OK(hipGetLastError());
    }
    if (!matvec_multi_dev_impl(h, ndens, xb, yb, s)) return false;
    for (int k = 0; k < ndens; k++) {
        hipLaunchKernelGGL(zero_kernel, dim3((unsigned)std::min<size_t>((nn + 255) / 256, 4096)), dim3(256), 0, s,
                           Jh, nn);
        HIP_OK(hipGetLastError());
        hipLaunchKernelGGL(contract_scatter_kernel, dim3((N + 255) / 256), dim3(256), 0, s, d->d_gidx,
                           yb + (size_t)k * N, Jh, N);
        HIP_OK(hipGetLastError());
        dim3 grid((nbf + 31) / 32, (nbf + 31) / 32);
        hipLaunchKernelGGL(symmetrize_kernel, grid, dim3(256), 0, s, Jh, J_dev + k * nn, nbf);
        HIP_OK(hipGetLastError());
    }
    return true;
}

// Eager Coulomb build on the handle's own buffers (grown when needed).
static bool coulomb_dev_impl(H2ERI_p h, int ndens, const double *D_dev, double *J_dev, hipStream_t s) {
    if (!upload_storage(h)) return false;
    DevState *d = dstate(h);
    if (!d->d_gidx) { set_error(ERR_ARG, "Coulomb build needs shell-pair metadata"); return false; }
    const size_t nn = (size_t)h->num_bf * h->num_bf;
    if (!ensure_io(d, (size_t)h->krnl_mat_size * ndens)) return false;
    if (!d->d_Jh) HIP_OK(hipMalloc(&d->d_Jh, sizeof(double) * nn));
    return coulomb_run(h, ndens, D_dev, J_dev, s, d->d_x, d->d_y, d->d_Jh);
}

// Reads and clears the error flag of every persistent plan's counters, after
// the work that used them has completed.  Returns 1 with ERR_DEVICE set when
// a dependency wait exceeded its bound (that pass's results are invalid), 0
// when clean, -1 on a device error.
static int queue_flags(DevState *d) {
    int bad = 0;
    for (DevPlan &p : d->plans) {
        if (!p.q_ready) continue;
        int e = 0;
        if (hipMemcpy(&e, p.d_ctr + QC_ERROR, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
            set_error(ERR_DEVICE, "persistent executor: reading the error flag failed");
            return -1;
        }
        if (e) {
            bad = 1;
            const int zero = 0;
            if (hipMemcpy(p.d_ctr + QC_ERROR, &zero, sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
                set_error(ERR_DEVICE, "persistent executor: clearing the error flag failed");
                return -1;
            }
        }
    }
    if (d->last_noalloc_ctr) {   // the last allocation-free persistent pass (counters in the scratch array)
        int e = 0;
        if (hipMemcpy(&e, d->last_noalloc_ctr + QC_ERROR, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
            set_error(ERR_DEVICE, "persistent executor: reading the error flag failed");
            return -1;
        }
        d->last_noalloc_ctr = nullptr;
        if (e) bad = 1;
    }
    if (bad) set_error(ERR_DEVICE, "persistent executor: a dependency wait exceeded its bound; the results are invalid");
    return bad;
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free matvec
// ---------------------------------------------------------------------------
// The persistent executor's counters (when that executor runs), then the
// workspace of every intermediate vector of one single-vector pass.  Each GPU
// the library runs on holds its own copy until the process ends; every
// allocation-free call on a GPU shares it, so calls that overlap on one GPU
// race.  Host code never passes `scratch` itself (its name is a host-side
// stand-in): it passes the device address that hipGetSymbolAddress returns
// for the current device.
static const size_t NOALLOC_SCRATCH_BYTES = (size_t)128 << 20;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_BYTES / sizeof(double)];

static std::mutex g_scratch_mu;
static std::vector<double *> g_scratch_ptr;   // per device, filled once

// Device address of the scratch array on device `dev`, which must be current.
static double *noalloc_scratch(int dev) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (dev < 0) { set_error(ERR_DEVICE, "scratch: no current device"); return nullptr; }
    if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            set_error(ERR_DEVICE, "scratch: hipGetSymbolAddress failed: %s", hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[dev] = (double *)ptr;
    }
    return g_scratch_ptr[dev];
}

// One single-vector pass of a prepared handle with the workspace (and the
// persistent executor's counters) in the scratch array; allocates nothing.
static bool matvec_noalloc_impl(H2ERI_p h, const double *X, double *Y, hipStream_t s) {
    if (h->BD_JIT) { set_error(ERR_UNSUPPORTED, "H2ERI_matvec_dev_noalloc: not for BD_JIT handles"); return false; }
    DevState *d = h->dev ? dstate(h) : nullptr;
    DevPlan *dp = d ? &d->plans[0] : nullptr;
    if (!d || !d->st_ready || !dp->ready || !d->side || !d->ev_fork || !d->ev_join ||
        (d->cfg[0].executor == 1 && !dp->q_ready)) {
        set_error(ERR_ARG, "H2ERI_matvec_dev_noalloc: the handle is not prepared; call H2ERI_upload after the last "
                           "change of precision, executor, tile size or schedule");
        return false;
    }
    int dev = -1;
    HIP_OK(hipGetDevice(&dev));
    if (dev != d->dev) { set_error(ERR_ARG, "H2ERI_matvec_dev_noalloc: the handle's device %d is not current", d->dev); return false; }
    double *base = noalloc_scratch(dev);
    if (!base) return false;
    const bool queue = d->cfg[0].executor == 1;
    const size_t nctr = queue ? (size_t)QC_LISTS + dp->q_nlist : 0;
    const size_t ctr_bytes = (nctr * sizeof(int) + 255) & ~(size_t)255;
    const size_t need = ctr_bytes + (size_t)dp->P.nW * sizeof(double);
    if (need > NOALLOC_SCRATCH_BYTES) {
        set_error(ERR_ARG, "H2ERI_matvec_dev_noalloc: this representation needs %zu bytes of scratch, the maximum is %zu",
                  need, NOALLOC_SCRATCH_BYTES);
        return false;
    }
    int *ctr = (int *)base;
    if (queue) {
        // every counter, the error flag included, starts at zero (a kernel,
        // so that the call can be captured in a graph)
        hipLaunchKernelGGL(counters_clear_kernel, dim3((unsigned)((nctr + 255) / 256)), dim3(256), 0, s, ctr, (int)nctr);
        HIP_OK(hipGetLastError());
    }
    const bool was_timing = d->phase_timing;   // timing would create events
    d->phase_timing = false;
    d->ws_Wv = (double *)((char *)base + ctr_bytes);
    d->ws_ctr = queue ? ctr : nullptr;
    const bool ok = run_matvec(h, d, *dp, X, Y, s);
    d->ws_Wv = nullptr;
    d->ws_ctr = nullptr;
    d->phase_timing = was_timing;
    if (queue) d->last_noalloc_ctr = ctr;
    return ok;
}

}  // namespace h2e

using namespace h2e;

extern "C" {

void H2ERI_matvec(H2ERI_p h, const double *x, double *y) {
    if (!h || !x || !y) { set_error(ERR_ARG, "H2ERI_matvec: null argument"); return; }
    matvec_host_impl(h, 1, x, y);
}

void H2ERI_matvec_dev(H2ERI_p h, const double *x_dev, double *y_dev, void *stream) {
    if (!h || !x_dev || !y_dev) { set_error(ERR_ARG, "H2ERI_matvec_dev: null argument"); return; }
    matvec_multi_dev_impl(h, 1, x_dev, y_dev, (hipStream_t)stream);
}

// No counterpart in the reference: H2ERI_matvec_dev for callers that cannot
// allocate (graph capture, tight serving loops).
void H2ERI_matvec_dev_noalloc(H2ERI_p h, const double *x_dev, double *y_dev, void *stream) {
    if (!h || !x_dev || !y_dev) { set_error(ERR_ARG, "H2ERI_matvec_dev_noalloc: null argument"); return; }
    matvec_noalloc_impl(h, x_dev, y_dev, (hipStream_t)stream);
}

void H2ERI_matvec_multi(H2ERI_p h, const int nvec, const double *X, double *Y) {
    if (!h || nvec < 0 || (nvec > 0 && (!X || !Y))) { set_error(ERR_ARG, "H2ERI_matvec_multi: bad argument"); return; }
    if (nvec > 0) matvec_host_impl(h, nvec, X, Y);
}

void H2ERI_matvec_multi_dev(H2ERI_p h, const int nvec, const double *X_dev, double *Y_dev, void *stream) {
    if (!h || nvec < 0 || (nvec > 0 && (!X_dev || !Y_dev))) { set_error(ERR_ARG, "H2ERI_matvec_multi_dev: bad argument"); return; }
    if (nvec > 0) matvec_multi_dev_impl(h, nvec, X_dev, Y_dev, (hipStream_t)stream);
}

void H2ERI_build_Coulomb_multi(H2ERI_p h, const int ndens, const double *den_mats, double *J_mats) {
    if (!h || ndens < 0 || (ndens > 0 && (!den_mats || !J_mats))) { set_error(ERR_ARG, "H2ERI_build_Coulomb_multi: bad argument"); return; }
    if (ndens == 0 || !upload_storage(h)) return;
    DevState *d = dstate(h);
    HostTiming tm(d);
    const size_t nn = (size_t)h->num_bf * h->num_bf;
    free_dev((void *&)d->d_D);
    free_dev((void *&)d->d_J);
    if (hipMalloc(&d->d_D, nn * 8 * ndens) != hipSuccess || hipMalloc(&d->d_J, nn * 8 * ndens) != hipSuccess) {
        set_error(ERR_DEVICE, "H2ERI_build_Coulomb: device allocation of D/J failed");
        return;
    }
    if (hipMemcpy(d->d_D, den_mats, nn * 8 * ndens, hipMemcpyHostToDevice) != hipSuccess) { set_error(ERR_DEVICE, "copy D"); return; }
    if (!coulomb_dev_impl(h, ndens, d->d_D, d->d_J, nullptr)) return;
    if (hipMemcpy(J_mats, d->d_J, nn * 8 * ndens, hipMemcpyDeviceToHost) != hipSuccess) { set_error(ERR_DEVICE, "copy J"); return; }
    collect_phase_ms(h, d);
    if (queue_flags(d) != 0) return;
    h->n_matvec += ndens;
}

void H2ERI_build_Coulomb(H2ERI_p h, const double *den_mat, double *J_mat) {
    if (!h || !den_mat || !J_mat) { set_error(ERR_ARG, "H2ERI_build_Coulomb: null argument"); return; }
    H2ERI_build_Coulomb_multi(h, 1, den_mat, J_mat);
}

void H2ERI_build_Coulomb_dev(H2ERI_p h, const double *den_dev, double *J_dev, void *stream) {
    if (!h || !den_dev || !J_dev) { set_error(ERR_ARG, "H2ERI_build_Coulomb_dev: null argument"); return; }
    coulomb_dev_impl(h, 1, den_dev, J_dev, (hipStream_t)stream);
}

void H2ERI_build_Coulomb_multi_dev(H2ERI_p h, const int ndens, const double *den_dev, double *J_dev, void *stream) {
    if (!h || ndens < 0 || (ndens > 0 && (!den_dev || !J_dev))) { set_error(ERR_ARG, "H2ERI_build_Coulomb_multi_dev: bad argument"); return; }
    if (ndens > 0) coulomb_dev_impl(h, ndens, den_dev, J_dev, (hipStream_t)stream);
}

// Captures the device hot path for fixed device buffers into a HIP graph held
// by the handle (replacing a previous one): kind 0 is the batched matvec
// (H2ERI_matvec_multi_dev with n vectors), kind 1 the Coulomb build
// (H2ERI_build_Coulomb_multi_dev with n densities), kind 2 the exchange build
// (H2ERI_build_exchange_dev, n = 1; the density screening included).  Nothing runs during the
// capture; H2ERI_graph_replay enqueues the captured sequence.  The graph
// refers to in_dev, out_dev and the handle's buffers: changing the precision,
// tile size or schedule, or destroying the handle, releases it.
int H2ERI_graph_capture(H2ERI_p h, const int kind, const int n, const double *in_dev, double *out_dev) {
    if (!h || kind < 0 || kind > 2 || n < 1 || (kind == 2 && n != 1) || !in_dev || !out_dev) {
        set_error(ERR_ARG, "H2ERI_graph_capture: bad argument");
        return -1;
    }
    // build every plan and buffer first: allocation is not allowed while capturing
    DevState *d = dstate_get(h);
    if (kind == 2) {
        if (!exchange_prepare(h)) return -1;
    } else {
        if (!upload_storage(h)) return -1;
        for (int done = 0; done < n;) {
            const int rem = n - done;
            const int nv = nv_pass(rem);
            if (!get_plan(h, nv)) return -1;
            done += nv;
        }
    }
    if (!d->side && hipStreamCreateWithFlags(&d->side, hipStreamNonBlocking) != hipSuccess) { set_error(ERR_DEVICE, "stream"); return -1; }
    if (!d->ev_fork && hipEventCreateWithFlags(&d->ev_fork, hipEventDisableTiming) != hipSuccess) { set_error(ERR_DEVICE, "event"); return -1; }
    if (!d->ev_join && hipEventCreateWithFlags(&d->ev_join, hipEventDisableTiming) != hipSuccess) { set_error(ERR_DEVICE, "event"); return -1; }
    release_graph(d);   // also frees the previous graph's buffers
    if (kind == 1) {
        // the graph's own x, y and half-matrix buffers (nothing else resizes them)
        if (!d->d_gidx) { set_error(ERR_ARG, "Coulomb build needs shell-pair metadata"); return -1; }
        const size_t nx = (size_t)h->krnl_mat_size * n, nn = (size_t)h->num_bf * h->num_bf;
        if (hipMalloc(&d->g_x, sizeof(double) * nx) != hipSuccess || hipMalloc(&d->g_y, sizeof(double) * nx) != hipSuccess ||
            hipMalloc(&d->g_Jh, sizeof(double) * nn) != hipSuccess) {
            release_graph(d);
            set_error(ERR_DEVICE, "H2ERI_graph_capture: allocation failed");
            return -1;
        }
    }
    hipStream_t cs;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) { set_error(ERR_DEVICE, "capture stream"); return -1; }
    hipGraph_t g = nullptr;
    bool ok = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (ok) {
        d->capturing = true;
        ok = kind == 0   ? matvec_multi_dev_impl(h, n, in_dev, out_dev, cs)
             : kind == 1 ? coulomb_run(h, n, in_dev, out_dev, cs, d->g_x, d->g_y, d->g_Jh)
                         : exchange_enqueue(h, in_dev, out_dev, cs);
        d->capturing = false;
        const bool end_ok = hipStreamEndCapture(cs, &g) == hipSuccess;
        ok = ok && end_ok && g;
    }
    if (ok && hipGraphInstantiate(&d->graph, g, nullptr, nullptr, 0) != hipSuccess) {
        d->graph = nullptr;
        ok = false;
    }
    if (g) (void)hipGraphDestroy(g);
    (void)hipStreamDestroy(cs);
    if (!ok) {
        release_graph(d);
        if (!has_error()) set_error(ERR_DEVICE, "H2ERI_graph_capture: capture failed");
        return -1;
    }
    d->graph_kind = kind;
    return 0;
}

// Enqueues the captured graph on `stream` (a hipStream_t, NULL = default).
int H2ERI_graph_replay(H2ERI_p h, void *stream) {
    DevState *d = h ? dstate(h) : nullptr;
    if (!d || !d->graph) { set_error(ERR_ARG, "H2ERI_graph_replay: nothing captured"); return -1; }
    if (hipGraphLaunch(d->graph, (hipStream_t)stream) != hipSuccess) { set_error(ERR_DEVICE, "hipGraphLaunch failed"); return -1; }
    return 0;
}

void H2ERI_set_precision(H2ERI_p h, const int ub_bits) {
    if (!h || (ub_bits != 32 && ub_bits != 64)) { set_error(ERR_ARG, "H2ERI_set_precision: 32 or 64"); return; }
    DevState *d = dstate_get(h);
    if (d->ub_bits != ub_bits) { d->ub_bits = ub_bits; release_buffers(d); }
}

size_t H2ERI_upload(H2ERI_p h) {
    if (h && h->BD_JIT) return upload_storage(h) ? dstate(h)->resident_bytes : 0;
    if (!h || !get_plan(h, 1)) return 0;
    DevState *d = dstate(h);
    // everything a single-vector pass needs besides its vectors, so that the
    // allocation-free entry point finds it in place
    if (!d->side && hipStreamCreateWithFlags(&d->side, hipStreamNonBlocking) != hipSuccess) { set_error(ERR_DEVICE, "stream"); return 0; }
    if (!d->ev_fork && hipEventCreateWithFlags(&d->ev_fork, hipEventDisableTiming) != hipSuccess) { set_error(ERR_DEVICE, "event"); return 0; }
    if (!d->ev_join && hipEventCreateWithFlags(&d->ev_join, hipEventDisableTiming) != hipSuccess) { set_error(ERR_DEVICE, "event"); return 0; }
    if (!noalloc_scratch(d->dev)) return 0;
    return d->resident_bytes;
}

void H2ERI_set_phase_timing(H2ERI_p h, const int on) {
    if (!h) return;
    dstate_get(h)->phase_timing = on != 0;
}

void H2ERI_get_phase_ms(H2ERI_p h, double *ms7) {
    if (!h || !ms7) return;
    DevState *d = dstate_get(h);
    collect_phase_ms(h, d);
    for (int i = 0; i < 7; i++) ms7[i] = d->phase_ms[i];
}

int H2ERI_get_launch_stats(H2ERI_p h, const int max_launch, double *out5) {
    if (!h) return 0;
    DevState *d = dstate_get(h);
    collect_phase_ms(h, d);
    const int n = (int)d->launches.size();
    for (int i = 0; i < n && i < max_launch && out5; i++) {
        const LaunchRec &r = d->launches[i];
        out5[5 * i + 0] = r.phase;
        out5[5 * i + 1] = r.start;
        out5[5 * i + 2] = r.blocks;
        out5[5 * i + 3] = r.mat_bytes;
        out5[5 * i + 4] = r.ms;
    }
    return n;
}

void H2ERI_matvec_bytes(H2ERI_p h, double *out6) {
    if (!h || !out6) return;
    DevPlan *dp = get_plan(h, 1);
    if (!dp) return;
    const Storage &S = dstate(h)->st;
    const double alg = S.bytes_U + S.bytes_B + S.bytes_D + dp->P.bytes_vec;
    const size_t *ms = h->mat_size;
    const double ref = 8.0 * (double)(ms[MV_FWD_SIZE_IDX] + ms[MV_MID_SIZE_IDX] + ms[MV_BWD_SIZE_IDX] + ms[MV_DEN_SIZE_IDX]);
    out6[0] = alg;
    out6[1] = ref;
    out6[2] = S.bytes_U;
    out6[3] = S.bytes_B;
    out6[4] = S.bytes_D;
    out6[5] = dp->P.bytes_vec;
}

// Near-field schedule (tuning knob; re-plans): 0 balances the near-field
// tiles over the sweep launches, 1 runs them in three launches of their own
// ahead of the sweeps, 2 + g (default 3) runs them on a second stream,
// concurrently with the sweeps and joined before the gather, as g + 1
// byte-balanced groups of blocks pipelined over g + 3 launches.
void H2ERI_set_schedule(H2ERI_p h, const int mode) {
    if (!h || mode < 0 || mode > 257) { set_error(ERR_ARG, "H2ERI_set_schedule: 0 to 257"); return; }
    DevState *d = dstate_get(h);
    for (int i = 0; i < NWIDTH; i++)
        if (d->cfg[i].nf_schedule != mode) { d->cfg[i].nf_schedule = mode; release_plan(d, i); }
}

// Executor of the matvec (re-plans): 0 launches every stage separately (the
// near field on a second stream for schedules >= 2); 1 runs the whole pass as
// one persistent launch whose workgroups take the plan's tasks in a
// dependency-checked order (plus a small launch that clears its counters).
// The staged plan runs matrix parts up to 32 KB as wave tiles, whose column
// sums are added in another order, so the two agree to rounding (bitwise
// when the plan has no such part).
void H2ERI_set_executor(H2ERI_p h, const int mode) {
    if (!h || (mode != 0 && mode != 1)) { set_error(ERR_ARG, "H2ERI_set_executor: 0 or 1"); return; }
    DevState *d = dstate_get(h);
    for (int i = 0; i < NWIDTH; i++)
        if (d->cfg[i].executor != mode) { d->cfg[i].executor = mode; release_plan(d, i); }
}

int H2ERI_get_executor(H2ERI_p h) { return h && h->dev ? dstate(h)->cfg[0].executor : 0; }

// Waits for the device and reports whether a wait of the persistent executor
// exceeded its bound since the last call (which would mean a planning fault:
// the results of that pass are not valid).  Returns 0 when clean, 1 with the
// error set, -1 on a device error.  Clears the flag.
int H2ERI_device_status(H2ERI_p h) {
    if (!h) return -1;
    DevState *d = dstate_get(h);
    if (hipDeviceSynchronize() != hipSuccess) { set_error(ERR_DEVICE, "H2ERI_device_status: device error"); return -1; }
    return queue_flags(d);
}

// Bound of every dependency wait of the persistent executor, in polls
// (tuning and tests; no re-plan).
void H2ERI_set_spin_limit(H2ERI_p h, const int polls) {
    if (!h || polls < 0) { set_error(ERR_ARG, "H2ERI_set_spin_limit: polls >= 0"); return; }
    dstate_get(h)->spin_limit = polls;
}

// Queue statistics of the nv-vector plan (building it when needed):
// out4 = {tasks, tasks started without a ticket, workgroups, list-scheduling
// estimate of the pass in microseconds}.
int H2ERI_queue_stats(H2ERI_p h, const int nv, double *out4) {
    if (!h || !out4 || !nv_valid(nv)) { set_error(ERR_ARG, "H2ERI_queue_stats: bad argument"); return -1; }
    DevPlan *dp = get_plan(h, nv);
    if (!dp || !ensure_queue(dstate(h), *dp)) return -1;
    out4[0] = dp->q_n;
    out4[1] = dp->q_static;
    out4[2] = dp->q_grid;
    out4[3] = dp->q_est_us;
    return 0;
}

// Matrix parts up to `bytes` run as wave tiles in the staged executor (one
// wave each, four per workgroup); 0 disables them (tuning knob; re-plans).
void H2ERI_set_wave_tile_bytes(H2ERI_p h, const double bytes) {
    if (!h || bytes < 0) { set_error(ERR_ARG, "H2ERI_set_wave_tile_bytes: bytes >= 0"); return; }
    DevState *d = dstate_get(h);
    for (int i = 0; i < NWIDTH; i++)
        if (d->cfg[i].wave_bytes != bytes) { d->cfg[i].wave_bytes = bytes; release_plan(d, i); }
}

// Workgroups of every near-field launch of the staged executor on the second
// stream (tuning knob; no re-plan, releases a captured graph): 0 gives each
// task a workgroup of its own; n > 0 at most n, each taking tasks
// blockIdx.x, blockIdx.x + n, ..., so the sweeps' launches beside them keep
// CUs of their own.
void H2ERI_set_near_field_grid(H2ERI_p h, const int n) {
    if (!h || n < 0) { set_error(ERR_ARG, "H2ERI_set_near_field_grid: n >= 0"); return; }
    DevState *d = dstate_get(h);
    release_graph(d);
    for (int i = 0; i < NWIDTH; i++) d->cfg[i].nf_grid = n;
}

// Tile size of the bi-GEMV tasks in bytes (tuning knob; re-plans).
void H2ERI_set_tile_bytes(H2ERI_p h, const double bytes) {
    if (!h || bytes < 1024) return;
    DevState *d = dstate_get(h);
    for (int i = 0; i < NWIDTH; i++) d->cfg[i].tile_bytes = bytes;
    release_plans(d);
}

// Autotuner: times the device matvec of nv vectors for every combination of
// executor, tile size, near-field schedule and (staged executor) wave-tile
// limit in the candidate lists (HIP events around `reps` back-to-back calls
// after one warm-up) and keeps the fastest.  out5 = {best ms per call, tile
// bytes, schedule, executor, wave-tile bytes}.  Returns 0, or -1 with the
// error set.
int H2ERI_autotune(H2ERI_p h, const int nv, const int reps, double *out6) {
    if (!h || !nv_valid(nv) || reps < 1) { set_error(ERR_ARG, "H2ERI_autotune: bad argument"); return -1; }
    if (h->BD_JIT) { set_error(ERR_UNSUPPORTED, "H2ERI_autotune: a BD_JIT handle has no stored-block plans"); return -1; }
    if (!upload_storage(h)) return -1;
    DevState *d = dstate(h);
    const size_t N = (size_t)h->krnl_mat_size;
    if (!ensure_io(d, N * nv)) return -1;
    // inputs: a fixed pattern (the time does not depend on the values)
    hipLaunchKernelGGL(zero_kernel, dim3(1024), dim3(256), 0, 0, d->d_x, N * nv);
    if (hipGetLastError() != hipSuccess) { set_error(ERR_DEVICE, "autotune: zero"); return -1; }
    const double tiles[] = {65536.0, 131072.0, 262144.0, 524288.0};
    // schedules: near field in the sweep launches (0), or on a second stream in
    // 1, 2, 3, 5 or 9 pipelined groups (more groups bring each low-rank block's
    // second read of its smaller half closer to the first, at the cost of more
    // launches)
    const int scheds[] = {0, 2, 3, 4, 6, 10};
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) { set_error(ERR_DEVICE, "autotune: events"); return -1; }
    const int vi = nv_index(nv);
    Tune &c = d->cfg[vi];
    Tune best_cfg = c;
    double best = 1e300;
    const double waves[] = {0.0, 32768.0};
    const bool was_timing = d->phase_timing;
    d->phase_timing = false;
    bool ok = true;
    // times the current configuration c and keeps it when it is the fastest
    auto time_cfg = [&]() -> bool {
        if (!matvec_multi_dev_impl(h, nv, d->d_x, d->d_y, nullptr)) return false;   // plans, uploads, warms up
        if (hipEventRecord(e0, nullptr) != hipSuccess) return false;
        for (int r = 0; r < reps; r++)
            if (!matvec_multi_dev_impl(h, nv, d->d_x, d->d_y, nullptr)) return false;
        if (hipEventRecord(e1, nullptr) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return false;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return false;
        if (ms / reps < best) { best = ms / reps; best_cfg = c; }
        if (h->print_dbginfo) {
            fprintf(stderr, "[H2ERI_autotune] nv %d: executor %d, tile %.0f KB, schedule %d, wave tiles %.0f KB, "
                    "near-field grid %d: %.4f ms\n", nv, c.executor, c.tile_bytes / 1024.0, c.nf_schedule,
                    c.wave_bytes / 1024.0, c.nf_grid, ms / reps);
            fflush(stderr);
        }
        return true;
    };
    for (int ex = 0; ok && ex < 2; ex++)
    for (int wi = 0; ok && wi < (ex == 0 && nv != 16 ? 2 : 1); wi++)   // no wave tiles at 16 vectors per pass
    for (double tb : tiles) {
        for (int sc : scheds) {
            release_plan(d, vi);
            c.executor = ex;
            c.wave_bytes = ex == 0 ? waves[wi] : 0.0;
            c.tile_bytes = tb;
            c.nf_schedule = sc;
            c.nf_grid = 0;
            ok = time_cfg();
            if (!ok) break;
        }
        if (!ok) break;
    }
    // the best staged configuration with a second stream, its near-field
    // launches capped at 3/4 and 1/2 of the workgroups the GPU holds at once
    if (ok && best_cfg.executor == 0 && best_cfg.nf_schedule >= 2) {
        int ncu = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, d->dev) != hipSuccess || ncu <= 0) ncu = 256;
        const int resident = ncu * (nv == 16 ? waves_per_simd<16>() : waves_per_simd<1>());   // 4 waves each: one per SIMD
        const Tune base = best_cfg;
        for (int frac : {3, 2}) {
            if (c.executor != base.executor || c.tile_bytes != base.tile_bytes || c.nf_schedule != base.nf_schedule ||
                c.wave_bytes != base.wave_bytes)
                release_plan(d, vi);
            c = base;
            c.nf_grid = resident * frac / 4;
            ok = time_cfg();
            if (!ok) break;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    d->phase_timing = was_timing;
    release_plan(d, vi);
    c = best_cfg;
    if (!ok) {
        if (!has_error()) set_error(ERR_DEVICE, "autotune: timing failed");
        return -1;
    }
    if (out6) {
        out6[0] = best; out6[1] = c.tile_bytes; out6[2] = c.nf_schedule; out6[3] = c.executor; out6[4] = c.wave_bytes;
        out6[5] = c.nf_grid;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Molecule groups: one persistent launch over the tasks of several representations
// ---------------------------------------------------------------------------
struct H2EGroup {
    std::vector<H2ERI_p> hs;
    std::vector<const double *> xs;
    std::vector<double *> ys;
    int mode = 0;                 // 0: one grouped persistent launch, 1: each molecule's own pass in turn
    std::vector<DevPlan> plans;   // per molecule: workspace and task arrays (one vector)
    MolPtrs *d_mols = nullptr;
    QTask *d_q = nullptr;
    int *d_need = nullptr, *d_ctr = nullptr;
    int q_n = 0, q_static = 0, q_nlist = 0, q_grid = 0;
    double q_est_us = 0;
};

static void group_free(H2EGroup *g) {
    for (DevPlan &p : g->plans) {
        void **ps[] = {(void **)&p.Wv, (void **)&p.d_t64, (void **)&p.d_t32, (void **)&p.d_rt, (void **)&p.d_rsrc,
                       (void **)&p.d_vsrc};
        for (void **q : ps) free_dev(*q);
    }
    void **ps[] = {(void **)&g->d_mols, (void **)&g->d_q, (void **)&g->d_need, (void **)&g->d_ctr};
    for (void **q : ps) free_dev(*q);
    delete g;
}

// Groups nh representations (single vector each) for H2ERI_group_matvec: y_k
// := A_k x_k for k < nh with fixed device buffers X_dev[k], Y_dev[k].  The
// tasks of all the molecules' plans form one persistent queue: list
// scheduling orders them jointly on the resident workgroups, longest
// remaining path first, so the small molecules' work fills the gaps of the
// big ones' critical paths and one launch (plus the counter reset) runs the
// whole group.  The handles must outlive the group and keep their precision;
// returns NULL with the error set on failure.
void *H2ERI_group_create(const int nh, H2ERI_p *handles, const double *const *X_dev, double *const *Y_dev) {
    if (nh < 1 || !handles || !X_dev || !Y_dev) { set_error(ERR_ARG, "H2ERI_group_create: bad argument"); return nullptr; }
    H2EGroup *g = new H2EGroup();
    g->plans.resize(nh);
    std::vector<MolPtrs> mols(nh);
    int dev = 0, ncu = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, queue_kernel<1, true>, WG, 0) != hipSuccess) {
        set_error(ERR_DEVICE, "H2ERI_group_create: occupancy query failed");
        delete g;
        return nullptr;
    }
    const int G = ncu * std::max(per, 1);
    std::vector<QGraph> graphs(nh);
    for (int k = 0; k < nh; k++) {
        H2ERI_p h = handles[k];
        if (!h || !X_dev[k] || !Y_dev[k]) { set_error(ERR_ARG, "H2ERI_group_create: null handle or buffer %d", k); group_free(g); return nullptr; }
        if (h->BD_JIT) { set_error(ERR_UNSUPPORTED, "H2ERI_group_create: handle %d is BD_JIT", k); group_free(g); return nullptr; }
        if (!upload_storage(h)) { group_free(g); return nullptr; }
        DevState *d = dstate(h);
        DevPlan &dp = g->plans[k];
        if (!build_plan(h, d->st, dp.P, 1, d->cfg[0].tile_bytes, d->cfg[0].nf_schedule, 1, 0.0)) { group_free(g); return nullptr; }
        if (!queue_graph(dp.P, 1, G, graphs[k])) { group_free(g); return nullptr; }
        const Plan &P = dp.P;
        auto up = [&](void **dst, const void *src, size_t bytes) -> bool {
            if (bytes == 0) bytes = 64;
            hipError_t e = hipMalloc(dst, bytes);
            if (e == hipSuccess && src) e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
            if (e != hipSuccess) { set_error(ERR_DEVICE, "group upload: %s", hipGetErrorString(e)); return false; }
            return true;
        };
        bool ok = up((void **)&dp.Wv, nullptr, (size_t)P.nW * 8) &&
                  up((void **)&dp.d_t64, P.t64.empty() ? nullptr : P.t64.data(), P.t64.size() * sizeof(BTask)) &&
                  up((void **)&dp.d_t32, P.t32.empty() ? nullptr : P.t32.data(), P.t32.size() * sizeof(BTask)) &&
                  up((void **)&dp.d_rt, P.rt.empty() ? nullptr : P.rt.data(), P.rt.size() * sizeof(RTask)) &&
                  up((void **)&dp.d_rsrc, P.rsrc.empty() ? nullptr : P.rsrc.data(), P.rsrc.size() * sizeof(int)) &&
                  up((void **)&dp.d_vsrc, P.vsrc.empty() ? nullptr : P.vsrc.data(), P.vsrc.size() * sizeof(int));
        if (!ok) { group_free(g); return nullptr; }
        mols[k] = MolPtrs{d->S64, d->S32, dp.Wv, dp.d_t64, dp.d_t32, dp.d_rt, dp.d_vsrc, dp.d_rsrc, X_dev[k], Y_dev[k],
                          (long long)h->krnl_mat_size};
        g->hs.push_back(h);
        g->xs.push_back(X_dev[k]);
        g->ys.push_back(Y_dev[k]);
    }
    QPlan Q;
    if (!queue_schedule(graphs, G, Q)) { group_free(g); return nullptr; }
    const size_t nq = std::max<size_t>(Q.tasks.size(), 1), nl = Q.need.size();
    bool ok = hipMalloc(&g->d_mols, sizeof(MolPtrs) * nh) == hipSuccess &&
              hipMemcpy(g->d_mols, mols.data(), sizeof(MolPtrs) * nh, hipMemcpyHostToDevice) == hipSuccess &&
              hipMalloc(&g->d_q, nq * sizeof(QTask)) == hipSuccess &&
              (Q.tasks.empty() || hipMemcpy(g->d_q, Q.tasks.data(), Q.tasks.size() * sizeof(QTask), hipMemcpyHostToDevice) == hipSuccess) &&
              hipMalloc(&g->d_need, std::max<size_t>(nl, 1) * sizeof(int)) == hipSuccess &&
              (nl == 0 || hipMemcpy(g->d_need, Q.need.data(), nl * sizeof(int), hipMemcpyHostToDevice) == hipSuccess) &&
              hipMalloc(&g->d_ctr, (QC_LISTS + nl) * sizeof(int)) == hipSuccess &&
              hipMemset(g->d_ctr, 0, (QC_LISTS + nl) * sizeof(int)) == hipSuccess && hipDeviceSynchronize() == hipSuccess;
    if (!ok) {
        set_error(ERR_DEVICE, "H2ERI_group_create: upload failed");
        group_free(g);
        return nullptr;
    }
    g->q_n = (int)Q.tasks.size();
    g->q_static = Q.nstatic;
    g->q_nlist = (int)nl;
    g->q_grid = std::min(G, std::max(g->q_n, 1));
    g->q_est_us = Q.est_us;
    return g;
}

// Enqueues one matvec of every molecule of the group on `stream`: one grouped
// persistent launch (mode 0), or each molecule's own pass in turn with its
// handle's configuration (mode 1).
int H2ERI_group_matvec(void *group, void *stream) {
    H2EGroup *g = (H2EGroup *)group;
    if (!g) { set_error(ERR_ARG, "H2ERI_group_matvec: null group"); return -1; }
    hipStream_t s = (hipStream_t)stream;
    if (g->mode == 1) {
        for (size_t k = 0; k < g->hs.size(); k++)
            if (!matvec_multi_dev_impl(g->hs[k], 1, g->xs[k], g->ys[k], s)) return -1;
        return 0;
    }
    hipLaunchKernelGGL(queue_reset_kernel, dim3((std::max(g->q_nlist, QC_LISTS) + 255) / 256), dim3(256), 0, s, g->d_ctr,
                       g->q_nlist);
    hipLaunchKernelGGL((queue_kernel<1, true>), dim3(g->q_grid), dim3(WG), 0, s, g->d_q, g->q_n, g->d_need,
                       g->d_ctr, QUEUE_SPIN_LIMIT, g->d_mols, (const BTask *)nullptr, (const BTask *)nullptr,
                       (const RTask *)nullptr, (const int *)nullptr, (const int *)nullptr, (const double *)nullptr,
                       (const float *)nullptr, (double *)nullptr, (const double *)nullptr, (double *)nullptr, 0LL);
    if (hipGetLastError() != hipSuccess) { set_error(ERR_DEVICE, "H2ERI_group_matvec: launch failed"); return -1; }
    return 0;
}

// Synchronises and reports whether a dependency wait of the group's
// persistent launch exceeded its bound since the last call (0 clean, 1 with
// the error set, -1 on a device error); clears the flag.
int H2ERI_group_status(void *group) {
    H2EGroup *g = (H2EGroup *)group;
    if (!g) { set_error(ERR_ARG, "H2ERI_group_status: null group"); return -1; }
    int e = 0;
    if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(&e, g->d_ctr + QC_ERROR, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
        set_error(ERR_DEVICE, "H2ERI_group_status: device error");
        return -1;
    }
    for (H2ERI_p h : g->hs) {   // the molecules' own passes (mode 1)
        const int r = queue_flags(dstate(h));
        if (r < 0) return -1;
        if (r > 0) e = 1;
    }
    if (!e) return 0;
    const int zero = 0;
    if (hipMemcpy(g->d_ctr + QC_ERROR, &zero, sizeof(int), hipMemcpyHostToDevice) != hipSuccess) return -1;
    set_error(ERR_DEVICE, "group: a dependency wait exceeded its bound");
    return 1;
}

void H2ERI_group_destroy(void *group) {
    if (group) group_free((H2EGroup *)group);
}

// Mode of H2ERI_group_matvec: 0 one grouped persistent launch, 1 each
// molecule's own pass in turn.  Returns 0, -1 on a bad argument.
int H2ERI_group_set_mode(void *group, const int mode) {
    H2EGroup *g = (H2EGroup *)group;
    if (!g || mode < 0 || mode > 1) { set_error(ERR_ARG, "H2ERI_group_set_mode: bad argument"); return -1; }
    g->mode = mode;
    return 0;
}

// Times both modes (reps group matvecs after a warm-up each, HIP events on
// the null stream) and keeps the faster: the grouped launch wins when the
// molecules are small and leave the GPU idle alone, the molecules' own passes
// when each fills the GPU.  out3 = {ms of mode 0, ms of mode 1, mode kept}.
// Returns 0, -1 on error.
int H2ERI_group_autotune(void *group, const int reps, double *out3) {
    H2EGroup *g = (H2EGroup *)group;
    if (!g || reps < 1) { set_error(ERR_ARG, "H2ERI_group_autotune: bad argument"); return -1; }
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) { set_error(ERR_DEVICE, "group autotune: events"); return -1; }
    double ms[2] = {0.0, 0.0};
    bool ok = true;
    for (int m = 0; ok && m < 2; m++) {
        g->mode = m;
        ok = H2ERI_group_matvec(g, nullptr) == 0 && hipEventRecord(e0, nullptr) == hipSuccess;
        for (int r = 0; ok && r < reps; r++) ok = H2ERI_group_matvec(g, nullptr) == 0;
        float t = 0.f;
        ok = ok && hipEventRecord(e1, nullptr) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
             hipEventElapsedTime(&t, e0, e1) == hipSuccess;
        ms[m] = t / reps;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!ok) {
        if (!has_error()) set_error(ERR_DEVICE, "group autotune: timing failed");
        return -1;
    }
    if (H2ERI_group_status(g) != 0) return -1;
    g->mode = ms[1] < ms[0] ? 1 : 0;
    if (out3) { out3[0] = ms[0]; out3[1] = ms[1]; out3[2] = g->mode; }
    return 0;
}

size_t h2eri_struct_size(void) { return sizeof(struct H2ERI); }


// Verification of a plan on the host (no HIP call): executes every tile and
// reduce task with the kernels' semantics in launch order for nv vectors
// (X, Y vector-major, nv x N), and checks that within one launch no task
// reads a workspace element another task of that launch writes and no
// element is stored twice.  For tests of the planner; the device entry points
// never call it.  Returns the number of conflicts, -1 on failure.
int H2ERI_plan_emulate(H2ERI_p h, const int ub_bits, const int nv, const double *X, double *Y) {
    if (!h || !h->U || !h->c_B_blks || !h->c_D_blks || !X || !Y || !nv_valid(nv)) {
        set_error(ERR_ARG, "plan_emulate: bad args");
        return -1;
    }
    if (h->BD_JIT) {   // no stored blocks to plan (H2ERI_jit_emulate runs that form)
        set_error(ERR_UNSUPPORTED, "plan_emulate: a BD_JIT handle has no stored blocks");
        return -1;
    }
    Storage S;
    S.build(h, ub_bits == 32);
    Plan P;
    DevState *ds = dstate_get(h);
    const Tune &c = ds->cfg[nv_index(nv)];
    if (!build_plan(h, S, P, nv, c.tile_bytes, c.nf_schedule, 0, c.wave_bytes)) return -1;
    const long long N = h->krnl_mat_size;
    std::vector<double> Wv((size_t)P.nW * nv, std::nan(""));
    std::vector<int> wmark((size_t)P.nW, -1), rmark((size_t)P.nW, -1);
    std::vector<int> ymark((size_t)N, -1);
    // elements written on the near-field stream: the main stream may touch
    // them only in the gather (after the join)
    std::vector<char> side_w((size_t)P.nW, 0);
    bool in_side = false, before_join = true;
    int conflicts = 0;
    auto rdw = [&](long long e, int task) {
        if (wmark[e] >= 0 && wmark[e] != task) conflicts++;
        if (!in_side && before_join && side_w[e]) conflicts++;
        rmark[e] = task;
    };
    auto wr = [&](long long e, int task) {
        if ((wmark[e] >= 0 && wmark[e] != task) || (rmark[e] >= 0 && rmark[e] != task)) conflicts++;
        if (in_side) side_w[e] = 1;
        else if (side_w[e]) conflicts++;
        wmark[e] = task;
    };
    // launch order of the emulation: near-field stream first, then the main stream
    std::vector<Stage> order;
    for (const Stage &st : P.stages) if (st.stream == 1) order.push_back(st);
    for (const Stage &st : P.stages) if (st.stream == 0) order.push_back(st);
    auto vin = [&](int b, int e, long long idx, int v, int task) -> double {
        double s = 0.0;
        for (int j = b; j < e; j++) {
            const int src = P.vsrc[j];
            if (src < 0) return X[(long long)v * N + (-src - 1) + idx];
            rdw(src + idx, task);
            s += Wv[(size_t)(src + idx) * nv + v];
        }
        return s;
    };
    for (const Stage &st : order) {
        std::fill(wmark.begin(), wmark.end(), -1);
        std::fill(rmark.begin(), rmark.end(), -1);
        in_side = st.stream == 1;
        before_join = st.phase != PH_GATHER;
        const int ntile = st.t64_cnt + st.t32_cnt + st.w64_cnt + st.w32_cnt;
        for (int q = 0; q < ntile; q++) {
            int k = q;
            const BTask *tp;
            bool f32;
            if (k < st.t64_cnt) { tp = &P.t64[st.t64_beg + k]; f32 = false; }
            else if ((k -= st.t64_cnt) < st.t32_cnt) { tp = &P.t32[st.t32_beg + k]; f32 = true; }
            else if ((k -= st.t32_cnt) < st.w64_cnt) { tp = &P.w64[st.w64_beg + k]; f32 = false; }
            else { k -= st.w64_cnt; tp = &P.w32[st.w32_beg + k]; f32 = true; }
            const BTask &t = *tp;
            auto M = [&](int r, int c) -> double {
                const long long e = t.mat + (long long)r * t.ld + c;
                return f32 ? (double)S.h32[e] : S.h64[e];
            };
            for (int v = 0; v < nv; v++) {
                std::vector<double> x0(t.ncol, 0.0), y1(t.nrow, 0.0);
                if (t.out0 >= 0)
                    for (int c = 0; c < t.ncol; c++) x0[c] = vin(t.v0b, t.v0e, t.v0off + c, v, q);
                if (t.out1 >= 0)
                    for (int r = 0; r < t.nrow; r++) y1[r] = vin(t.v1b, t.v1e, t.v1off + r, v, q);
                if (t.out0 >= 0)
                    for (int r = 0; r < t.nrow; r++) {
                        double d = 0.0;
                        for (int c = 0; c < t.ncol; c++) d += M(r, c) * x0[c];
                        wr(t.out0 + r, q);
                        Wv[(size_t)(t.out0 + r) * nv + v] = d;
                    }
                if (t.out1 >= 0)
                    for (int c = 0; c < t.ncol; c++) {
                        double s = 0.0;
                        for (int r = 0; r < t.nrow; r++) s += M(r, c) * y1[r];
                        wr(t.out1 + c, q);
                        Wv[(size_t)(t.out1 + c) * nv + v] = s;
                    }
            }
        }
        for (int q = 0; q < st.r_cnt; q++) {
            const RTask &r = P.rt[st.r_beg + q];
            const int task = ntile + q;
            for (int i = 0; i < r.len; i++) {
                std::vector<double> s(nv, 0.0);
                for (int j = r.src_beg; j < r.src_end; j++) {
                    const long long e = P.rsrc[j] + r.soff + i;
                    rdw(e, task);
                    for (int v = 0; v < nv; v++) s[v] += Wv[(size_t)e * nv + v];
                }
                if (r.flags & 1) {
                    if (ymark[r.dst + i] >= 0) conflicts++;
                    ymark[r.dst + i] = task;
                    for (int v = 0; v < nv; v++) Y[(size_t)v * N + r.dst + i] = s[v];
                } else {
                    wr(r.dst + i, task);
                    for (int v = 0; v < nv; v++) Wv[(size_t)(r.dst + i) * nv + v] = s[v];
                }
            }
        }
    }
    for (long long i = 0; i < N; i++)
        if (ymark[i] < 0) conflicts++;   // every output element written exactly once
    return conflicts;
}

// Verification of the persistent executor's queue on the host (no HIP call):
// builds the plan and its queue for G = `workgroups`, executes the tasks in
// ticket order with the kernels' semantics, and counts every read of a
// workspace element whose writer is not a producer of the list the read goes
// through (a dependency the counters would not wait for), every read of an
// unwritten element, and every output element not written exactly once.
// Returns that count, or -1 on failure.
int H2ERI_queue_emulate(H2ERI_p h, const int ub_bits, const int nv, const int workgroups, const double *X, double *Y) {
    if (!h || !h->U || !h->c_B_blks || !h->c_D_blks || !X || !Y || !nv_valid(nv) ||
        workgroups < 1 || h->BD_JIT) {
        set_error(ERR_ARG, "queue_emulate: bad args (or a BD_JIT handle, which has no stored blocks)");
        return -1;
    }
    Storage S;
    S.build(h, ub_bits == 32);
    Plan P;
    DevState *ds = dstate_get(h);
    const Tune &c = ds->cfg[nv_index(nv)];
    if (!build_plan(h, S, P, nv, c.tile_bytes, c.nf_schedule, 1, 0.0)) return -1;
    QPlan Q;
    if (!build_queue(P, nv, workgroups, Q)) return -1;
    const long long N = h->krnl_mat_size;
    std::vector<double> Wv((size_t)P.nW * nv, 0.0);
    std::vector<int> writer((size_t)P.nW, -1), ymark((size_t)N, -1);
    int conflicts = 0;
    for (size_t k = 0; k < Q.tasks.size(); k++) {
        const QTask &q = Q.tasks[k];
        auto rd = [&](long long e, int list) -> double {
            const int w = e >= 0 && e < P.nW ? writer[e] : -2;
            if (w < 0) {
                conflicts++;
                return 0.0;
            }
            const QTask &wq = Q.tasks[w];
            if (list < 0 || (wq.out0 != list && wq.out1 != list)) conflicts++;
            return 1.0;
        };
        auto vin = [&](int b, int e, long long idx, int v, int list) -> double {
            double s = 0.0;
            for (int j = b; j < e; j++) {
                const int src = P.vsrc[j];
                if (src < 0) return X[(long long)v * N + (-src - 1) + idx];
                rd(src + idx, list);
                s += Wv[(size_t)(src + idx) * nv + v];
            }
            return s;
        };
        auto wr = [&](long long e) {
            if (writer[e] >= 0) conflicts++;
            writer[e] = (int)k;
        };
        if (q.kind < 2) {
            const bool f32 = q.kind == 1;
            const BTask &t = (f32 ? P.t32 : P.t64)[q.idx];
            auto M = [&](int r, int c) -> double {
                const long long e = t.mat + (long long)r * t.ld + c;
                return f32 ? (double)S.h32[e] : S.h64[e];
            };
            std::vector<std::vector<double>> o0(nv), o1(nv);
            for (int v = 0; v < nv; v++) {
                std::vector<double> x0(t.ncol, 0.0), y1(t.nrow, 0.0);
                if (t.out0 >= 0)
                    for (int c = 0; c < t.ncol; c++) x0[c] = vin(t.v0b, t.v0e, t.v0off + c, v, q.in0);
                if (t.out1 >= 0)
                    for (int r = 0; r < t.nrow; r++) y1[r] = vin(t.v1b, t.v1e, t.v1off + r, v, q.in1 >= 0 ? q.in1 : q.in0);
                if (t.out0 >= 0)
                    for (int r = 0; r < t.nrow; r++) {
                        double d = 0.0;
                        for (int c = 0; c < t.ncol; c++) d += M(r, c) * x0[c];
                        o0[v].push_back(d);
                    }
                if (t.out1 >= 0)
                    for (int c = 0; c < t.ncol; c++) {
                        double s = 0.0;
                        for (int r = 0; r < t.nrow; r++) s += M(r, c) * y1[r];
                        o1[v].push_back(s);
                    }
            }
            if (t.out0 >= 0)
                for (int r = 0; r < t.nrow; r++) {
                    wr(t.out0 + r);
                    for (int v = 0; v < nv; v++) Wv[(size_t)(t.out0 + r) * nv + v] = o0[v][r];
                }
            if (t.out1 >= 0)
                for (int c = 0; c < t.ncol; c++) {
                    wr(t.out1 + c);
                    for (int v = 0; v < nv; v++) Wv[(size_t)(t.out1 + c) * nv + v] = o1[v][c];
                }
        } else {
            const RTask &r = P.rt[q.idx];
            for (int i = 0; i < r.len; i++) {
                std::vector<double> s(nv, 0.0);
                for (int j = r.src_beg; j < r.src_end; j++) {
                    const long long e = P.rsrc[j] + r.soff + i;
                    rd(e, q.in0);
                    for (int v = 0; v < nv; v++) s[v] += Wv[(size_t)e * nv + v];
                }
                if (r.flags & 1) {
                    if (ymark[r.dst + i] >= 0) conflicts++;
                    ymark[r.dst + i] = (int)k;
                    for (int v = 0; v < nv; v++) Y[(size_t)v * N + r.dst + i] = s[v];
                } else {
                    wr(r.dst + i);
                    for (int v = 0; v < nv; v++) Wv[(size_t)(r.dst + i) * nv + v] = s[v];
                }
            }
        }
    }
    for (long long i = 0; i < N; i++)
        if (ymark[i] < 0) conflicts++;
    return conflicts;
}

// Builds the device plans of every batch width on the host only and runs
// their bounds checks; no HIP call is made.  Returns the number of tile
// tasks of the single-vector plan, or -1 on failure.
// Work figures of the nv-vector plan with the current configuration, built
// on the host only (no HIP call): out8 = {tile tasks, reduce tasks, stages,
// workspace elements per vector, matrix bytes streamed per pass, slot
// elements written by tiles per vector, slot elements read by tiles and
// reduce tasks per vector, external output elements}.  0, -1 on failure.
int H2ERI_plan_stats(H2ERI_p h, const int ub_bits, const int nv, double *out8) {
    if (!h || !h->U || !h->c_B_blks || !h->c_D_blks || !out8 || !nv_valid(nv)) { set_error(ERR_ARG, "plan_stats: bad argument"); return -1; }
    if (h->BD_JIT) { set_error(ERR_UNSUPPORTED, "plan_stats: a BD_JIT handle has no stored blocks (H2ERI_jit_stats)"); return -1; }
    Storage S;
    S.build(h, ub_bits == 32);
    DevState *ds = dstate_get(h);
    const Tune &c = ds->cfg[nv_index(nv)];
    Plan P;
    if (!build_plan(h, S, P, nv, c.tile_bytes, c.nf_schedule, c.executor, c.wave_bytes)) return -1;
    double mat = 0, wr = 0, rd = 0, rd_r = 0, wr_r = 0;
    for (const Stage &st : P.stages) mat += st.mat_bytes;
    for (const auto *L : {&P.t64, &P.t32, &P.w64, &P.w32})
        for (const BTask &t : *L) {
            if (t.out0 >= 0) { wr += t.nrow; if (P.vsrc[t.v0b] >= 0) rd += (double)t.ncol * (t.v0e - t.v0b); }
            if (t.out1 >= 0) { wr += t.ncol; if (P.vsrc[t.v1b] >= 0) rd += (double)t.nrow * (t.v1e - t.v1b); }
        }
    double ext = 0;
    for (const RTask &r : P.rt) {
        rd_r += (double)r.len * (r.src_end - r.src_beg);
        if (r.flags & 1) ext += r.len;
        else wr_r += r.len;
    }
    if (getenv("H2ERI_PLAN_STATS_DETAIL")) {
        fprintf(stderr, "plan nv %d: tiles write %.3g read %.3g, reduce tasks write %.3g read %.3g elements per vector\n",
                nv, wr, rd, wr_r, rd_r);
        // per phase: tile writes / reads, reduce writes / reads
        double pw[PH_N + 1] = {0}, pr[PH_N + 1] = {0}, qw[PH_N + 1] = {0}, qr[PH_N + 1] = {0};
        for (const Stage &st : P.stages) {
            const int ph = std::min(std::max(st.phase, 0), (int)PH_N);
            for (int pass = 0; pass < 2; pass++) {
                const std::vector<BTask> &L = pass == 0 ? P.t64 : P.t32;
                const int b = pass == 0 ? st.t64_beg : st.t32_beg, n = pass == 0 ? st.t64_cnt : st.t32_cnt;
                for (int i = b; i < b + n; i++) {
                    const BTask &t = L[i];
                    if (t.out0 >= 0) { pw[ph] += t.nrow; if (P.vsrc[t.v0b] >= 0) pr[ph] += (double)t.ncol * (t.v0e - t.v0b); }
                    if (t.out1 >= 0) { pw[ph] += t.ncol; if (P.vsrc[t.v1b] >= 0) pr[ph] += (double)t.nrow * (t.v1e - t.v1b); }
                }
            }
            for (int i = st.r_beg; i < st.r_beg + st.r_cnt; i++) {
                const RTask &r = P.rt[i];
                qr[ph] += (double)r.len * (r.src_end - r.src_beg);
                if (!(r.flags & 1)) qw[ph] += r.len;
            }
        }
        static const char *nm[] = {"up", "B", "y1", "down", "D", "gather", "?"};
        for (int ph = 0; ph <= PH_N; ph++)
            if (pw[ph] + pr[ph] + qw[ph] + qr[ph] > 0)
                fprintf(stderr, "  phase %s: tiles write %.3g read %.3g, reduces write %.3g read %.3g\n", nm[std::min(ph, 6)],
                        pw[ph], pr[ph], qw[ph], qr[ph]);
    }
    wr += wr_r;
    rd += rd_r;
    out8[0] = (double)(P.t64.size() + P.t32.size() + P.w64.size() + P.w32.size());
    out8[1] = (double)P.rt.size();
    out8[2] = (double)P.stages.size();
    out8[3] = (double)P.nW;
    out8[4] = mat;
    out8[5] = wr;
    out8[6] = rd;
    out8[7] = ext;
    return 0;
}

int H2ERI_plan_check(H2ERI_p h, const int ub_bits) {
    if (!h || !h->U || !h->c_B_blks || !h->c_D_blks) { set_error(ERR_ARG, "plan_check: nothing built"); return -1; }
    if (h->BD_JIT) { set_error(ERR_UNSUPPORTED, "plan_check: a BD_JIT handle has no stored blocks"); return -1; }
    Storage S;
    S.build(h, ub_bits == 32);
    DevState *ds = dstate_get(h);
    int tasks = 0;
    for (int nv : NV_VARIANTS)
        for (int ex = 0; ex < 2; ex++) {
            Plan P;
            const Tune &c = ds->cfg[nv_index(nv)];
            if (!build_plan(h, S, P, nv, c.tile_bytes, c.nf_schedule, ex, c.wave_bytes)) return -1;
            if (nv == 1 && ex == 0) tasks = (int)(P.t64.size() + P.t32.size() + P.w64.size() + P.w32.size());
        }
    return tasks;
}


}  // extern "C"
