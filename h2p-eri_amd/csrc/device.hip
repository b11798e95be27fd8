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
//     one flat storage array with its row stride padded to the vector width;
//   * each phase is a list of tiles, "bi-GEMV tasks": tile x in0 -> out0
//     (row dots), tile^T x in1 -> out1 (column sums), executed by one
//     workgroup each in one launch;
//   * every tile output goes to a private slot; a segmented-reduce launch
//     sums the slots of each destination in a fixed order, so results are
//     bit-reproducible and no atomics are needed;
//   * U and B may be stored in FP32 (accumulation stays FP64), D stays FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <numeric>

#include "h2eri_internal.h"

namespace h2e {

// ---------------------------------------------------------------------------
// Task descriptors
// ---------------------------------------------------------------------------
struct BTask {
    long long mat;   // element offset into the storage array of its precision
    int nrow, ncol, ld;
    int lpr;         // lanes per row: 4, 8, 16, 32 or 64
    int in0, in1;    // workspace offsets (-1: unused)
    int out0, out1;  // workspace offsets (-1: unused)
    int flags;       // bit 0: out0 accumulates (+=) instead of storing
    int pad_;
};
static_assert(sizeof(BTask) == 48, "BTask layout");

struct RTask {
    long long dst;   // offset into workspace, or into the external output
    int len;         // <= 64 elements (one chunk of a destination segment)
    int src_beg, src_end;
    int flags;       // bit 0: destination is the external output vector;
                     // bit 1: accumulate (dst += sum) instead of storing
    int soff;        // element offset of this chunk inside every source slot
    int pad_;
};
static_assert(sizeof(RTask) == 32, "RTask layout");
static const int RCHUNK = 64;       // destination elements per reduce workgroup

static const int WG = 256;          // threads per workgroup
static const int COLS_PER_LANE = 8; // columns a lane holds in registers
static const int MAX_TILE_COLS = 64 * COLS_PER_LANE;

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
template <typename T> struct VecT;
template <> struct VecT<double> { typedef double2 type; static const int W = 2; };
template <> struct VecT<float> { typedef float4 type; static const int W = 4; };

__device__ __forceinline__ double vget(const double2 &v, int i) { return i == 0 ? v.x : v.y; }
__device__ __forceinline__ double vget(const float4 &v, int i) {
    return (double)(i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w)));
}

template <typename T>
__device__ void bgemv_tile(const BTask &t, const T *__restrict__ S, double *__restrict__ Wv, double *red) {
    typedef typename VecT<T>::type V;
    constexpr int W = VecT<T>::W;
    constexpr int KP = COLS_PER_LANE / W;
    const int lpr = t.lpr;
    const int tid = threadIdx.x;
    const int lane = tid & (lpr - 1);
    const int grp = tid / lpr;
    const int G = WG / lpr;
    const int R = t.nrow, C = t.ncol;
    const bool do0 = t.out0 >= 0, do1 = t.out1 >= 0;

    double x0[KP][W];
    double a1[KP][W];
    #pragma unroll
    for (int k = 0; k < KP; k++)
        #pragma unroll
        for (int w = 0; w < W; w++) {
            const int c = (k * lpr + lane) * W + w;
            x0[k][w] = (do0 && c < C) ? Wv[t.in0 + c] : 0.0;
            a1[k][w] = 0.0;
        }
    const T *base = S + t.mat;
    // RU rows per iteration: all RU*KP vector loads are issued before any is
    // consumed, so every lane keeps RU*KP requests in flight.
    constexpr int RU = 4;
    for (int r0 = grp; r0 < R; r0 += RU * G) {
        V v[RU][KP];
        #pragma unroll
        for (int u = 0; u < RU; u++) {
            const int r = r0 + u * G;
            const T *row = base + (long long)r * t.ld;
            #pragma unroll
            for (int k = 0; k < KP; k++) {
                const int c0 = (k * lpr + lane) * W;
                if (r < R && c0 < C) v[u][k] = *reinterpret_cast<const V *>(row + c0);
                else v[u][k] = V{};
            }
        }
        if (do1) {
            #pragma unroll
            for (int u = 0; u < RU; u++) {
                const int r = r0 + u * G;
                const double y = (r < R) ? Wv[t.in1 + r] : 0.0;
                #pragma unroll
                for (int k = 0; k < KP; k++)
                    #pragma unroll
                    for (int w = 0; w < W; w++) a1[k][w] = fma(vget(v[u][k], w), y, a1[k][w]);
            }
        }
        if (do0) {
            double d[RU];
            #pragma unroll
            for (int u = 0; u < RU; u++) {
                d[u] = 0.0;
                #pragma unroll
                for (int k = 0; k < KP; k++)
                    #pragma unroll
                    for (int w = 0; w < W; w++) d[u] = fma(vget(v[u][k], w), x0[k][w], d[u]);
            }
            for (int m = lpr >> 1; m > 0; m >>= 1) {
                #pragma unroll
                for (int u = 0; u < RU; u++) d[u] += __shfl_xor(d[u], m, lpr);
            }
            if (lane == 0) {
                #pragma unroll
                for (int u = 0; u < RU; u++) {
                    const int r = r0 + u * G;
                    if (r < R) {
                        if (t.flags & 1) Wv[t.out0 + r] += d[u];
                        else Wv[t.out0 + r] = d[u];
                    }
                }
            }
        }
    }
    if (do1) {
        // reduce across row groups: first inside the wave, then across waves
        #pragma unroll
        for (int k = 0; k < KP; k++)
            #pragma unroll
            for (int w = 0; w < W; w++)
                for (int m = lpr; m < 64; m <<= 1) a1[k][w] += __shfl_xor(a1[k][w], m, 64);
        const int wave = tid >> 6;
        const int lw = tid & 63;
        if (lw < lpr) {
            #pragma unroll
            for (int k = 0; k < KP; k++)
                #pragma unroll
                for (int w = 0; w < W; w++) red[wave * MAX_TILE_COLS + (k * lpr + lw) * W + w] = a1[k][w];
        }
        __syncthreads();
        for (int c = tid; c < C; c += WG) {
            double s = red[c] + red[MAX_TILE_COLS + c] + red[2 * MAX_TILE_COLS + c] + red[3 * MAX_TILE_COLS + c];
            Wv[t.out1 + c] = s;
        }
    }
}

// One launch runs n64 FP64 tasks followed by n32 FP32 tasks.
__global__ __launch_bounds__(WG) void bgemv_kernel(const BTask *__restrict__ t64, int n64, const BTask *__restrict__ t32,
                                                   int n32, const double *__restrict__ S64,
                                                   const float *__restrict__ S32, double *__restrict__ Wv) {
    __shared__ double red[4 * MAX_TILE_COLS];
    const int b = blockIdx.x;
    if (b < n64) {
        const BTask t = t64[b];
        bgemv_tile<double>(t, S64, Wv, red);
    } else if (b - n64 < n32) {
        const BTask t = t32[b - n64];
        bgemv_tile<float>(t, S32, Wv, red);
    }
}

// Segmented fixed-order sum of partial slots into each destination chunk.
// Lane i owns element i of the chunk; wave w sums the sources w, w+4, ...
// (four loads in flight per lane), then the four wave sums are added in
// wave order, so the result is deterministic.
__global__ __launch_bounds__(WG) void reduce_kernel(const RTask *__restrict__ rt, const int *__restrict__ src,
                                                    double *__restrict__ Wv, double *__restrict__ ext) {
    __shared__ double red[4][RCHUNK];
    const RTask t = rt[blockIdx.x];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double s = 0.0;
    if (lane < t.len) {
        const double *base = Wv + t.soff + lane;
        int j = t.src_beg + w;
        for (; j + 12 < t.src_end; j += 16) {
            const double a0 = base[src[j]], a1 = base[src[j + 4]], a2 = base[src[j + 8]], a3 = base[src[j + 12]];
            s += a0;
            s += a1;
            s += a2;
            s += a3;
        }
        for (; j < t.src_end; j += 4) s += base[src[j]];
    }
    red[w][lane] = s;
    __syncthreads();
    if (w == 0 && lane < t.len) {
        const double v = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
        double *dst = ((t.flags & 1) ? ext : Wv) + t.dst + lane;
        if (t.flags & 2) *dst += v;
        else *dst = v;
    }
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

__global__ void zero_kernel(double *__restrict__ p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0.0;
}

// ---------------------------------------------------------------------------
// Plan
// ---------------------------------------------------------------------------
enum Phase { PH_UP = 0, PH_B = 1, PH_Y1 = 2, PH_DOWN = 3, PH_D = 4, PH_GATHER = 5, PH_N = 6 };

struct Stage {
    int phase;
    int t64_beg, t64_cnt, t32_beg, t32_cnt;
    int r_beg, r_cnt;
    double mat_bytes;   // matrix bytes streamed by this stage's tile launch
};

// One timed launch of the last matvec (phase timing mode).
struct LaunchRec {
    int phase, kind;    // kind 0: bi-GEMV tiles, 1: segmented reduce
    int blocks;
    double mat_bytes;
    float ms;
};

struct Segment {
    long long off;   // workspace offset, or output offset when ext
    int len;
    bool ext;
    std::vector<int> srcs;
    bool acc = false;   // reduce adds to the destination
};

struct Plan {
    std::vector<BTask> t64, t32;
    std::vector<RTask> rt;
    std::vector<int> rsrc;
    std::vector<Stage> stages;
    long long nW = 0;
    long long n64 = 0, n32 = 0;
    double bytes_U = 0, bytes_B = 0, bytes_D = 0, bytes_vec = 0;
};

struct DevState {
    int dev = -1;
    int ub_bits = 64;
    bool ready = false;
    double tile_bytes = 128.0 * 1024.0;
    int nf_schedule = 0;      // 0: near field balanced over the sweep stages, 1: own stages first
    Plan plan;
    double *S64 = nullptr;
    float *S32 = nullptr;
    double *Wv = nullptr;
    BTask *d_t64 = nullptr, *d_t32 = nullptr;
    RTask *d_rt = nullptr;
    int *d_rsrc = nullptr;
    // Coulomb
    int *d_gidx = nullptr;
    double *d_D = nullptr, *d_J = nullptr, *d_Jh = nullptr, *d_y = nullptr;
    double *d_xio = nullptr;    // host-pointer path staging
    size_t xio_len = 0;
    bool phase_timing = false;
    std::vector<hipEvent_t> evs;
    std::vector<LaunchRec> launches;
    bool collected = true;   // launch times of the last timed call already read
    double phase_ms[7] = {0, 0, 0, 0, 0, 0, 0};
    size_t resident_bytes = 0;
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

static void release_buffers(DevState *d) {
    void **ps[] = {(void **)&d->S64, (void **)&d->S32, (void **)&d->Wv, (void **)&d->d_t64, (void **)&d->d_t32,
                   (void **)&d->d_rt, (void **)&d->d_rsrc, (void **)&d->d_gidx, (void **)&d->d_D,
                   (void **)&d->d_J, (void **)&d->d_Jh, (void **)&d->d_y, (void **)&d->d_xio};
    for (void **p : ps) free_dev(*p);
    d->xio_len = 0;
    d->ready = false;
    d->resident_bytes = 0;
}

void device_release(H2ERI_p h) {
    DevState *d = dstate(h);
    if (!d) return;
    release_buffers(d);
    for (hipEvent_t e : d->evs) (void)hipEventDestroy(e);
    delete d;
    h->dev = nullptr;
}

void device_invalidate(H2ERI_p h) {
    DevState *d = dstate(h);
    if (d) release_buffers(d);
}

static DevState *dstate_get(H2ERI_p h) {
    if (!h->dev) h->dev = new DevState();
    return dstate(h);
}

namespace {

struct Builder {
    H2ERI_p h;
    Plan &P;
    int ub_bits;
    double tile_bytes;
    std::vector<double> host64;
    std::vector<float> host32;
    std::vector<Segment> segs;

    int nf_schedule = 0;
    Builder(H2ERI_p h_, Plan &p_, int bits, double tb, int nfs = 0)
        : h(h_), P(p_), ub_bits(bits), tile_bytes(tb), nf_schedule(nfs) {}

    long long walloc(long long n) {
        long long o = P.nW;
        P.nW += (n + 7) & ~7LL;   // keep 64-byte alignment of every region
        return o;
    }

    // store a row-major nrow x ncol matrix (source ld = ncol) with padded ld
    long long store(const double *src, int nrow, int ncol, bool fp32, int &ld_out) {
        const int W = fp32 ? 4 : 2;
        int ld = (ncol + W - 1) / W * W;
        ld_out = ld;
        if (fp32) {
            long long off = (long long)host32.size();
            size_t n = (size_t)nrow * ld;
            host32.resize(host32.size() + ((n + 63) & ~(size_t)63), 0.0f);
            for (int r = 0; r < nrow; r++)
                for (int c = 0; c < ncol; c++) host32[off + (size_t)r * ld + c] = (float)src[(size_t)r * ncol + c];
            return off;
        }
        long long off = (long long)host64.size();
        size_t n = (size_t)nrow * ld;
        host64.resize(host64.size() + ((n + 31) & ~(size_t)31), 0.0);
        for (int r = 0; r < nrow; r++) memcpy(&host64[off + (size_t)r * ld], src + (size_t)r * ncol, sizeof(double) * ncol);
        return off;
    }

    int new_seg(long long off, int len, bool ext) {
        segs.push_back(Segment{off, len, ext, {}, false});
        return (int)segs.size() - 1;
    }

    // Tile an nrow x ncol matrix into tasks.  dst0/dst1 are segment ids (-1 none).
    // out0 of the matrix goes to dst0, out1 to dst1.  When acc0_off >= 0 the
    // row dots are accumulated straight into the workspace at acc0_off (one
    // column tile only; unique writer).
    void add_matrix(std::vector<BTask> &tl, long long mat, int nrow, int ncol, int ld, bool fp32, int in0, int in1,
                    int dst0, int dst1, long long acc0_off = -1) {
        if (nrow <= 0 || ncol <= 0) return;
        const double esz = fp32 ? 4.0 : 8.0;
        const int nct = (ncol + MAX_TILE_COLS - 1) / MAX_TILE_COLS;
        // balanced column tile width, a multiple of 8 so every tile starts on a
        // vector-aligned column (<= MAX_TILE_COLS since that is a multiple of 8)
        const int ctw = ((ncol + nct - 1) / nct + 7) / 8 * 8;
        int rows = (int)std::max(1.0, tile_bytes / (esz * ctw));
        const int nrt = (nrow + rows - 1) / rows;
        rows = (nrow + nrt - 1) / nrt;                 // balanced row tiles
        // partial slots: out0 per column tile (full length), out1 per row tile
        std::vector<long long> s0(nct, -1), s1(nrt, -1);
        if (dst0 >= 0 && acc0_off < 0)
            for (int c = 0; c < nct; c++) { s0[c] = walloc(nrow); segs[dst0].srcs.push_back((int)s0[c]); }
        if (dst1 >= 0)
            for (int r = 0; r < nrt; r++) { s1[r] = walloc(ncol); segs[dst1].srcs.push_back((int)s1[r]); }
        for (int rt = 0; rt < nrt; rt++) {
            const int r0 = rt * rows, nr = std::min(rows, nrow - r0);
            for (int ct = 0; ct < nct; ct++) {
                const int c0 = ct * ctw, nc = std::min(ctw, ncol - c0);
                BTask t;
                t.mat = mat + (long long)r0 * ld + c0;
                t.nrow = nr;
                t.ncol = nc;
                t.ld = ld;
                int lpr = 4;
                while (lpr * COLS_PER_LANE < nc) lpr <<= 1;
                t.lpr = lpr;
                t.in0 = (dst0 >= 0 || acc0_off >= 0) ? in0 + c0 : -1;
                t.in1 = dst1 >= 0 ? in1 + r0 : -1;
                if (acc0_off >= 0) { t.out0 = (int)(acc0_off + r0); t.flags = 1; }
                else { t.out0 = dst0 >= 0 ? (int)(s0[ct] + r0) : -1; t.flags = 0; }
                t.out1 = dst1 >= 0 ? (int)(s1[rt] + c0) : -1;
                t.pad_ = 0;
                if (t.out0 < 0 && t.out1 < 0) continue;
                tl.push_back(t);
            }
        }
    }

    // ---- stages: tiles of stage s run in one launch, then the reduces of s ----
    std::vector<std::vector<BTask>> S64, S32;
    std::vector<std::vector<int>> SR;
    std::vector<int> sphase;
    std::vector<double> sload;

    std::vector<BTask> &T(int st, bool fp32) { return fp32 ? S32[st] : S64[st]; }

    static double tasks_bytes(const std::vector<BTask> &tl, size_t from, double esz) {
        double b = 0;
        for (size_t i = from; i < tl.size(); i++) b += esz * tl[i].nrow * tl[i].ncol;
        return b;
    }

    // add_matrix into stage st, accounting its bytes to the stage load
    void add_at(int st, long long mat, int nrow, int ncol, int ld, bool fp32, long long in0, long long in1, int dst0,
                int dst1, long long acc0 = -1) {
        std::vector<BTask> &tl = T(st, fp32);
        size_t n0 = tl.size();
        add_matrix(tl, mat, nrow, ncol, ld, fp32, (int)in0, (int)in1, dst0, dst1, acc0);
        sload[st] += tasks_bytes(tl, n0, fp32 ? 4.0 : 8.0);
    }

    // Dense block: out0 = M in0 -> dst0, out1 = M^T in1 -> dst1, one pass.
    void add_dense(H2E_dense_mat_p blk, bool fp32, long long in0, long long in1, int dst0, int dst1, int st,
                   double &bytes) {
        int ld;
        long long off = store(blk->data, blk->nrow, blk->ncol, fp32, ld);
        add_at(st, off, blk->nrow, blk->ncol, ld, fp32, in0, in1, dst0, dst1);
        bytes += (fp32 ? 4.0 : 8.0) * blk->nrow * blk->ncol;
    }

    // Low-rank block M = U V (U: nrow x k, V^T: ncol x k), out0 = U (V in0),
    // out1 = V^T (U^T in1), in stages a < b < c.  With both outputs the
    // smaller half is streamed in a and c and the other half once in b with a
    // fused bi-GEMV, so 1.5 passes instead of 2; with out0 only, V^T then U.
    void add_lowrank(H2E_dense_mat_p blk, bool fp32, long long in0, long long in1, int dst0, int dst1, int a, int b,
                     int c, double &bytes) {
        const int nr = blk->nrow, nc = blk->ncol, k = -blk->ld;
        int Uld, Vld;
        const long long Uo = store(blk->data, nr, k, fp32, Uld);
        const long long Vo = store(blk->data + (size_t)nr * k, nc, k, fp32, Vld);
        bytes += (fp32 ? 4.0 : 8.0) * (double)(nr + nc) * k;
        const long long t_off = walloc(k), s_off = walloc(k);
        const int st_t = new_seg(t_off, k, false), st_s = new_seg(s_off, k, false);
        if (dst1 < 0) {
            add_at(a, Vo, nc, k, Vld, fp32, -1, in0, -1, st_t);          // t = V in0
            SR[a].push_back(st_t);
            add_at(b, Uo, nr, k, Uld, fp32, t_off, -1, dst0, -1);        // out0 = U t
            return;
        }
        if (nr <= nc) {   // U twice
            add_at(a, Uo, nr, k, Uld, fp32, -1, in1, -1, st_s);          // s = U^T in1
            SR[a].push_back(st_s);
            add_at(b, Vo, nc, k, Vld, fp32, s_off, in0, dst1, st_t);     // out1 = V^T s, t = V in0
            SR[b].push_back(st_t);
            add_at(c, Uo, nr, k, Uld, fp32, t_off, -1, dst0, -1);        // out0 = U t
        } else {          // V^T twice
            add_at(a, Vo, nc, k, Vld, fp32, -1, in0, -1, st_t);          // t = V in0
            SR[a].push_back(st_t);
            add_at(b, Uo, nr, k, Uld, fp32, t_off, in1, dst0, st_s);     // out0 = U t, s = U^T in1
            SR[b].push_back(st_s);
            add_at(c, Vo, nc, k, Vld, fp32, s_off, -1, dst1, -1);        // out1 = V^T s
        }
    }

    void finish_stages() {
        auto key = [](const BTask &t) { return (long long)t.nrow * t.ncol; };
        for (size_t st = 0; st < S64.size(); st++) {
            std::stable_sort(S64[st].begin(), S64[st].end(), [&](const BTask &x, const BTask &y) { return key(x) > key(y); });
            std::stable_sort(S32[st].begin(), S32[st].end(), [&](const BTask &x, const BTask &y) { return key(x) > key(y); });
            Stage s;
            s.phase = sphase[st];
            s.mat_bytes = tasks_bytes(S64[st], 0, 8.0) + tasks_bytes(S32[st], 0, 4.0);
            s.t64_beg = (int)P.t64.size();
            s.t64_cnt = (int)S64[st].size();
            s.t32_beg = (int)P.t32.size();
            s.t32_cnt = (int)S32[st].size();
            P.t64.insert(P.t64.end(), S64[st].begin(), S64[st].end());
            P.t32.insert(P.t32.end(), S32[st].begin(), S32[st].end());
            s.r_beg = (int)P.rt.size();
            for (int sid : SR[st]) {
                Segment &g = segs[sid];
                if (g.len <= 0) continue;
                const int sb = (int)P.rsrc.size();
                P.rsrc.insert(P.rsrc.end(), g.srcs.begin(), g.srcs.end());
                for (int e0 = 0; e0 < g.len; e0 += RCHUNK) {
                    RTask r;
                    r.dst = g.off + e0;
                    r.len = std::min(RCHUNK, g.len - e0);
                    r.src_beg = sb;
                    r.src_end = (int)P.rsrc.size();
                    r.flags = (g.ext ? 1 : 0) | (g.acc ? 2 : 0);
                    r.soff = e0;
                    r.pad_ = 0;
                    P.rt.push_back(r);
                }
            }
            s.r_cnt = (int)P.rt.size() - s.r_beg;
            if (s.t64_cnt + s.t32_cnt + s.r_cnt > 0) P.stages.push_back(s);
        }
    }

    bool build() {
        const int n_node = h->n_node, nleaf = h->n_leaf_node, N = h->krnl_mat_size;
        const bool ub32 = ub_bits == 32;
        P = Plan();
        const long long x_off = walloc(N);
        // y0 / y1 with each node's children contiguous (children order)
        std::vector<long long> y0o(n_node, -1), y1o(n_node, -1);
        std::vector<int> rank(n_node, 0);
        for (int i = 0; i < n_node; i++) rank[i] = h->U[i] ? h->U[i]->ncol : 0;
        long long tot_r = 0;
        for (int i = 0; i < n_node; i++) tot_r += rank[i];
        const long long y0base = walloc(tot_r), y1base = walloc(tot_r);
        {
            long long pos = 0;
            y0o[h->root_idx] = pos;
            pos += rank[h->root_idx];
            for (int lv = 0; lv <= h->max_level; lv++)
                for (int j = 0; j < h->level_n_node[lv]; j++) {
                    int node = h->level_nodes[(size_t)lv * nleaf + j];
                    for (int k = 0; k < h->n_child[node]; k++) {
                        int ch = h->children[(size_t)node * 8 + k];
                        y0o[ch] = pos;
                        pos += rank[ch];
                    }
                }
            for (int i = 0; i < n_node; i++) { y1o[i] = y1base + y0o[i]; y0o[i] = y0base + y0o[i]; }
        }
        auto xseg = [&](int node) { return x_off + h->mat_cluster[2 * node]; };
        auto nlen = [&](int node) { return h->mat_cluster[2 * node + 1] - h->mat_cluster[2 * node] + 1; };
        std::vector<int> seg_y0(n_node, -1), seg_y1(n_node, -1), seg_y(n_node, -1);
        for (int i = 0; i < n_node; i++) {
            if (h->node_level[i] >= h->min_adm_level) {
                seg_y0[i] = new_seg(y0o[i], rank[i], false);
                seg_y1[i] = new_seg(y1o[i], rank[i], false);
            }
            if (h->n_child[i] == 0) seg_y[i] = new_seg(h->mat_cluster[2 * i], nlen(i), true);
        }
        // ---- stage layout along the critical path
        const int minL = h->min_adm_level, maxL = h->max_level;
        const int nU = (minL <= maxL) ? (maxL - minL + 1) : 0;
        const int s0 = nf_schedule == 1 ? 3 : 0;   // near field first: three stages of its own
        const int sBA = s0 + nU, sBB = sBA + 1, sBC = sBA + 2;
        const int sD0 = sBA + 3;
        const int sG = sD0 + nU;
        const int nst = sG + 1;
        S64.assign(nst, {});
        S32.assign(nst, {});
        SR.assign(nst, {});
        sload.assign(nst, 0.0);
        sphase.assign(nst, PH_GATHER);
        for (int q = 0; q < s0; q++) sphase[q] = PH_D;
        for (int s2 = 0; s2 < nU; s2++) { sphase[s0 + s2] = PH_UP; sphase[sD0 + s2] = PH_DOWN; }
        sphase[sBA] = PH_B; sphase[sBB] = PH_B; sphase[sBC] = PH_Y1;
        // storage of U
        std::vector<long long> Uoff(n_node, -1);
        std::vector<int> Uld(n_node, 0);
        for (int i = 0; i < n_node; i++) {
            H2E_dense_mat_p U = h->U[i];
            if (!U || U->nrow == 0 || U->ncol == 0) continue;
            Uoff[i] = store(U->data, U->nrow, U->ncol, ub32, Uld[i]);
            P.bytes_U += 2.0 * (ub32 ? 4.0 : 8.0) * U->nrow * U->ncol;
        }
        // ---- upward sweep: level maxL - s at stage s
        for (int lv = maxL; lv >= minL && nU > 0; lv--) {
            const int st = s0 + maxL - lv;
            for (int j = 0; j < h->level_n_node[lv]; j++) {
                int node = h->level_nodes[(size_t)lv * nleaf + j];
                SR[st].push_back(seg_y0[node]);
                if (Uoff[node] < 0) continue;
                H2E_dense_mat_p U = h->U[node];
                long long in1 = (h->n_child[node] == 0) ? xseg(node) : y0o[h->children[(size_t)node * 8 + 0]];
                add_at(st, Uoff[node], U->nrow, U->ncol, Uld[node], ub32, -1, in1, -1, seg_y0[node]);
            }
        }
        // ---- coupling blocks in stages sBA..sBC, y1 reduced after sBC
        for (int i = 0; i < h->n_B; i++) {
            H2E_dense_mat_p blk = h->c_B_blks[i];
            if (blk->nrow == 0 || blk->ncol == 0) continue;
            int n0 = h->r_adm_pairs[2 * i], n1 = h->r_adm_pairs[2 * i + 1];
            int l0 = h->node_level[n0], l1 = h->node_level[n1];
            long long in0, in1;
            int d0, d1;
            if (l0 == l1) { in0 = y0o[n1]; in1 = y0o[n0]; d0 = seg_y1[n0]; d1 = seg_y1[n1]; }
            else if (l0 > l1) { in0 = xseg(n1); in1 = y0o[n0]; d0 = seg_y1[n0]; d1 = seg_y[n1]; }
            else { in0 = y0o[n1]; in1 = xseg(n0); d0 = seg_y[n0]; d1 = seg_y1[n1]; }
            if (blk->ld >= 0) add_dense(blk, ub32, in0, in1, d0, d1, sBA, P.bytes_B);
            else add_lowrank(blk, ub32, in0, in1, d0, d1, sBA, sBB, sBC, P.bytes_B);
        }
        for (int i = 0; i < n_node; i++)
            if (seg_y1[i] >= 0) SR[sBC].push_back(seg_y1[i]);
        // ---- downward sweep: level minL + s at stage sD0 + s
        for (int lv = minL; lv <= maxL && nU > 0; lv++) {
            const int st = sD0 + (lv - minL);
            for (int j = 0; j < h->level_n_node[lv]; j++) {
                int node = h->level_nodes[(size_t)lv * nleaf + j];
                if (Uoff[node] < 0) continue;
                H2E_dense_mat_p U = h->U[node];
                if (h->n_child[node] == 0) {
                    add_at(st, Uoff[node], U->nrow, U->ncol, Uld[node], ub32, y1o[node], -1, seg_y[node], -1);
                } else if (U->ncol <= MAX_TILE_COLS) {
                    // one column tile: each row has a unique writer, add in place
                    long long acc = y1o[h->children[(size_t)node * 8 + 0]];
                    add_at(st, Uoff[node], U->nrow, U->ncol, Uld[node], ub32, y1o[node], -1, -1, -1, acc);
                } else {
                    // several column tiles: partial slots, reduce adds into y1 of the children
                    int sg = new_seg(y1o[h->children[(size_t)node * 8 + 0]], U->nrow, false);
                    segs[sg].acc = true;
                    add_at(st, Uoff[node], U->nrow, U->ncol, Uld[node], ub32, y1o[node], -1, sg, -1);
                    SR[st].push_back(sg);
                }
            }
        }
        // ---- near field: depends on x only, balanced over all stages
        std::vector<int> order(h->n_D);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return blk_elems(h->c_D_blks[a]) > blk_elems(h->c_D_blks[b]); });
        for (int i : order) {
            H2E_dense_mat_p blk = h->c_D_blks[i];
            if (blk->nrow == 0 || blk->ncol == 0) continue;
            long long in0, in1;
            int d0, d1;
            if (i < nleaf) {
                int node = h->height_nodes[i];
                in0 = xseg(node); in1 = -1; d0 = seg_y[node]; d1 = -1;
            } else {
                int n0 = h->r_inadm_pairs[2 * (i - nleaf)], n1 = h->r_inadm_pairs[2 * (i - nleaf) + 1];
                in0 = xseg(n1); in1 = xseg(n0); d0 = seg_y[n0]; d1 = seg_y[n1];
            }
            const double bytes = 8.0 * (double)blk_elems(blk);
            if (nf_schedule == 1) {
                if (blk->ld >= 0) add_dense(blk, false, in0, in1, d0, d1, 0, P.bytes_D);
                else add_lowrank(blk, false, in0, in1, d0, d1, 0, 1, 2, P.bytes_D);
            } else if (blk->ld >= 0) {
                int best = 0;
                for (int st = 1; st < nst; st++) if (sload[st] < sload[best]) best = st;
                add_dense(blk, false, in0, in1, d0, d1, best, P.bytes_D);
            } else if (-blk->ld > 0) {
                const int span = (d1 < 0) ? 2 : 3;
                int best = 0;
                double bestv = 1e300;
                for (int a = 0; a + span - 1 < nst; a++) {
                    double v = 0;
                    for (int q = 0; q < span; q++) v = std::max(v, sload[a + q] + bytes / span);
                    if (v < bestv) { bestv = v; best = a; }
                }
                add_lowrank(blk, false, in0, in1, d0, d1, best, best + 1, best + 2, P.bytes_D);
            }
        }
        for (int i = 0; i < n_node; i++)
            if (seg_y[i] >= 0) SR[sG].push_back(seg_y[i]);
        finish_stages();
        P.n64 = (long long)host64.size();
        P.n32 = (long long)host32.size();
        P.bytes_vec = 8.0 * (2.0 * N + 4.0 * (double)tot_r);
        if (P.nW >= (1LL << 31)) { set_error(ERR_UNSUPPORTED, "workspace exceeds 2^31 doubles"); return false; }
        return validate();
    }

    // Every access of every launch checked on the host before any launch:
    // tile reads inside storage and vector-aligned, vector reads/writes inside
    // the workspace, reduce sources/destinations inside their arrays.
    bool validate() const {
        const long long N = h->krnl_mat_size;
        auto chk_tasks = [&](const std::vector<BTask> &tl, long long nstore, int W) -> bool {
            for (const BTask &t : tl) {
                bool ok = t.nrow > 0 && t.ncol > 0 && t.ncol <= MAX_TILE_COLS && t.ld >= t.ncol && t.ld % W == 0 &&
                          t.mat % W == 0 && t.mat >= 0 && (t.lpr & (t.lpr - 1)) == 0 && t.lpr >= 4 && t.lpr <= 64 &&
                          t.lpr * COLS_PER_LANE >= t.ncol;
                // last vector of the last row
                long long last = t.mat + (long long)(t.nrow - 1) * t.ld + ((t.ncol - 1) / W) * W + W;
                ok = ok && last <= nstore;
                if (t.out0 >= 0) ok = ok && t.in0 >= 0 && t.in0 + t.ncol <= P.nW && t.out0 + t.nrow <= P.nW;
                if (t.out1 >= 0) ok = ok && t.in1 >= 0 && t.in1 + t.nrow <= P.nW && t.out1 + t.ncol <= P.nW;
                ok = ok && (t.out0 >= 0 || t.out1 >= 0);
                if (!ok) {
                    set_error(ERR_ARG, "device plan check failed: task mat=%lld %dx%d ld=%d lpr=%d", t.mat, t.nrow,
                              t.ncol, t.ld, t.lpr);
                    return false;
                }
            }
            return true;
        };
        if (!chk_tasks(P.t64, (long long)host64.size(), 2) || !chk_tasks(P.t32, (long long)host32.size(), 4))
            return false;
        for (const RTask &r : P.rt) {
            bool ok = r.len > 0 && r.len <= RCHUNK && r.dst >= 0 && r.soff >= 0 && r.src_beg >= 0 &&
                      r.src_end <= (int)P.rsrc.size() && r.src_beg <= r.src_end &&
                      r.dst + r.len <= ((r.flags & 1) ? N : P.nW);
            for (int j = r.src_beg; ok && j < r.src_end; j++)
                ok = P.rsrc[j] >= 0 && P.rsrc[j] + r.soff + r.len <= P.nW;
            if (!ok) { set_error(ERR_ARG, "device plan check failed: reduce dst=%lld len=%d", r.dst, r.len); return false; }
        }
        for (const Stage &s : P.stages) {
            if (s.t64_beg + s.t64_cnt > (int)P.t64.size() || s.t32_beg + s.t32_cnt > (int)P.t32.size() ||
                s.r_beg + s.r_cnt > (int)P.rt.size()) {
                set_error(ERR_ARG, "device plan check failed: stage ranges");
                return false;
            }
        }
        return true;
    }
};

}  // namespace

static bool ensure_events(DevState *d, size_t n) {
    while (d->evs.size() < n) {
        hipEvent_t e;
        HIP_OK(hipEventCreate(&e));
        d->evs.push_back(e);
    }
    return true;
}

static bool upload(H2ERI_p h) {
    DevState *d = dstate_get(h);
    if (d->ready) return true;
    if (!h->U || !h->c_B_blks || !h->c_D_blks) { set_error(ERR_ARG, "no H2 representation built or loaded"); return false; }
    if (h->BD_JIT) { set_error(ERR_UNSUPPORTED, "BD_JIT representations are not supported"); return false; }
    HIP_OK(hipGetDevice(&d->dev));
    release_buffers(d);
    Builder b(h, d->plan, d->ub_bits, d->tile_bytes, d->nf_schedule);
    if (!b.build()) return false;
    Plan &P = d->plan;
    auto alloc_copy = [&](void **dst, const void *src, size_t bytes) -> bool {
        if (bytes == 0) bytes = 64;
        HIP_OK(hipMalloc(dst, bytes));
        if (src) HIP_OK(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
        d->resident_bytes += bytes;
        return true;
    };
    if (!alloc_copy((void **)&d->S64, b.host64.empty() ? nullptr : b.host64.data(), b.host64.size() * 8)) return false;
    if (!alloc_copy((void **)&d->S32, b.host32.empty() ? nullptr : b.host32.data(), b.host32.size() * 4)) return false;
    if (!alloc_copy((void **)&d->Wv, nullptr, (size_t)P.nW * 8)) return false;
    HIP_OK(hipMemset(d->Wv, 0, (size_t)P.nW * 8));
    if (!alloc_copy((void **)&d->d_t64, P.t64.empty() ? nullptr : P.t64.data(), P.t64.size() * sizeof(BTask))) return false;
    if (!alloc_copy((void **)&d->d_t32, P.t32.empty() ? nullptr : P.t32.data(), P.t32.size() * sizeof(BTask))) return false;
    if (!alloc_copy((void **)&d->d_rt, P.rt.empty() ? nullptr : P.rt.data(), P.rt.size() * sizeof(RTask))) return false;
    if (!alloc_copy((void **)&d->d_rsrc, P.rsrc.empty() ? nullptr : P.rsrc.data(), P.rsrc.size() * sizeof(int))) return false;
    // Coulomb index map: x index -> (D element, 2x flag)
    if (h->num_bf > 0 && h->sp_shell_idx && h->shell_bf_sidx) {
        const int np = h->num_sp, nbf = h->num_bf, N = h->krnl_mat_size;
        std::vector<int> g(N);
        for (int i = 0; i < np; i++) {
            int s0 = h->sp_shell_idx[i], s1 = h->sp_shell_idx[i + np];
            int srow = h->shell_bf_sidx[s0], nrow = h->shell_bf_sidx[s0 + 1] - srow;
            int scol = h->shell_bf_sidx[s1], ncol = h->shell_bf_sidx[s1 + 1] - scol;
            int two = (s0 != s1) ? 1 : 0;
            int xo = h->sp_bfp_sidx[i];
            for (int j = 0; j < ncol; j++)
                for (int k = 0; k < nrow; k++) g[xo + j * nrow + k] = (((scol + j) * nbf + srow + k) << 1) | two;
        }
        if (!alloc_copy((void **)&d->d_gidx, g.data(), (size_t)N * sizeof(int))) return false;
    }
    d->ready = true;
    return true;
}

static void launch_bgemv(int nb, const Stage &st, DevState *d, hipStream_t s) {
    hipLaunchKernelGGL(bgemv_kernel, dim3(nb), dim3(WG), 0, s, d->d_t64 + st.t64_beg, st.t64_cnt,
                       d->d_t32 + st.t32_beg, st.t32_cnt, d->S64, d->S32, d->Wv);
}

// Enqueue one matvec: x already at Wv[0..N), result written to y_dev.  In
// phase-timing mode an event is recorded around every launch.
static bool run_matvec(H2ERI_p h, DevState *d, double *y_dev, hipStream_t s) {
    (void)h;
    const Plan &P = d->plan;
    const bool tim = d->phase_timing;
    if (tim) {
        if (!ensure_events(d, 2 * P.stages.size() + 1)) return false;
        d->launches.clear();
        d->collected = false;
        HIP_OK(hipEventRecord(d->evs[0], s));
    }
    for (size_t si = 0; si < P.stages.size(); si++) {
        const Stage &st = P.stages[si];
        const int nb = st.t64_cnt + st.t32_cnt;
        if (nb > 0) {
            launch_bgemv(nb, st, d, s);
            HIP_OK(hipGetLastError());
            if (tim) {
                d->launches.push_back(LaunchRec{st.phase, 0, nb, st.mat_bytes, 0.f});
                HIP_OK(hipEventRecord(d->evs[d->launches.size()], s));
            }
        }
        if (st.r_cnt > 0) {
            hipLaunchKernelGGL(reduce_kernel, dim3(st.r_cnt), dim3(WG), 0, s, d->d_rt + st.r_beg, d->d_rsrc, d->Wv, y_dev);
            HIP_OK(hipGetLastError());
            if (tim) {
                d->launches.push_back(LaunchRec{st.phase, 1, st.r_cnt, 0.0, 0.f});
                HIP_OK(hipEventRecord(d->evs[d->launches.size()], s));
            }
        }
    }
    return true;
}

// Per-launch and per-phase milliseconds of the last timed call (the stream
// must have been synchronised).
static void collect_phase_ms(H2ERI_p h, DevState *d) {
    if (!d->phase_timing || d->launches.empty() || d->collected) return;
    d->collected = true;
    for (int p = 0; p < 7; p++) d->phase_ms[p] = 0.0;
    for (size_t i = 0; i < d->launches.size(); i++) {
        float ms = 0.f;
        if (hipEventSynchronize(d->evs[i + 1]) != hipSuccess ||
            hipEventElapsedTime(&ms, d->evs[i], d->evs[i + 1]) != hipSuccess)
            ms = 0.f;
        d->launches[i].ms = ms;
        d->phase_ms[d->launches[i].phase] += ms;
        d->phase_ms[6] += ms;
    }
    // reference timer slots (seconds): forward, intermediate, backward, dense, vector ops
    h->timers[MV_FWD_TIMER_IDX] += d->phase_ms[PH_UP] * 1e-3;
    h->timers[MV_MID_TIMER_IDX] += (d->phase_ms[PH_B] + d->phase_ms[PH_Y1]) * 1e-3;
    h->timers[MV_BWD_TIMER_IDX] += d->phase_ms[PH_DOWN] * 1e-3;
    h->timers[MV_DEN_TIMER_IDX] += d->phase_ms[PH_D] * 1e-3;
    h->timers[MV_VOP_TIMER_IDX] += d->phase_ms[PH_GATHER] * 1e-3;
}

static bool matvec_dev_impl(H2ERI_p h, const double *x_dev, double *y_dev, hipStream_t s) {
    if (!upload(h)) return false;
    DevState *d = dstate(h);
    const int N = h->krnl_mat_size;
    HIP_OK(hipMemcpyAsync(d->Wv, x_dev, sizeof(double) * (size_t)N, hipMemcpyDeviceToDevice, s));
    return run_matvec(h, d, y_dev, s);
}

static bool ensure_xio(DevState *d, size_t n) {
    if (d->xio_len >= n) return true;
    free_dev((void *&)d->d_xio);
    HIP_OK(hipMalloc(&d->d_xio, sizeof(double) * n));
    d->xio_len = n;
    return true;
}

static bool matvec_host_impl(H2ERI_p h, const double *x, double *y) {
    if (!upload(h)) return false;
    DevState *d = dstate(h);
    const size_t N = (size_t)h->krnl_mat_size;
    if (!ensure_xio(d, N)) return false;
    HIP_OK(hipMemcpy(d->Wv, x, sizeof(double) * N, hipMemcpyHostToDevice));
    if (!run_matvec(h, d, d->d_xio, nullptr)) return false;
    HIP_OK(hipMemcpy(y, d->d_xio, sizeof(double) * N, hipMemcpyDeviceToHost));
    collect_phase_ms(h, d);
    h->n_matvec++;
    return true;
}

static bool coulomb_dev_impl(H2ERI_p h, const double *D_dev, double *J_dev, hipStream_t s) {
    if (!upload(h)) return false;
    DevState *d = dstate(h);
    if (!d->d_gidx) { set_error(ERR_ARG, "Coulomb build needs shell-pair metadata"); return false; }
    const int N = h->krnl_mat_size, nbf = h->num_bf;
    if (!d->d_Jh) HIP_OK(hipMalloc(&d->d_Jh, sizeof(double) * (size_t)nbf * nbf));
    if (!d->d_y) HIP_OK(hipMalloc(&d->d_y, sizeof(double) * (size_t)N));
    hipLaunchKernelGGL(uncontract_kernel, dim3((N + 255) / 256), dim3(256), 0, s, d->d_gidx, D_dev, d->Wv, N);
    HIP_OK(hipGetLastError());
    if (!run_matvec(h, d, d->d_y, s)) return false;
    size_t nn = (size_t)nbf * nbf;
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)std::min<size_t>((nn + 255) / 256, 4096)), dim3(256), 0, s, d->d_Jh, nn);
    HIP_OK(hipGetLastError());
    hipLaunchKernelGGL(contract_scatter_kernel, dim3((N + 255) / 256), dim3(256), 0, s, d->d_gidx, d->d_y, d->d_Jh, N);
    HIP_OK(hipGetLastError());
    dim3 grid((nbf + 31) / 32, (nbf + 31) / 32);
    hipLaunchKernelGGL(symmetrize_kernel, grid, dim3(256), 0, s, d->d_Jh, J_dev, nbf);
    HIP_OK(hipGetLastError());
    return true;
}

}  // namespace h2e

using namespace h2e;

extern "C" {

void H2ERI_matvec(H2ERI_p h, const double *x, double *y) {
    if (!h || !x || !y) { set_error(ERR_ARG, "H2ERI_matvec: null argument"); return; }
    matvec_host_impl(h, x, y);
}

void H2ERI_matvec_dev(H2ERI_p h, const double *x_dev, double *y_dev, void *stream) {
    if (!h || !x_dev || !y_dev) { set_error(ERR_ARG, "H2ERI_matvec_dev: null argument"); return; }
    matvec_dev_impl(h, x_dev, y_dev, (hipStream_t)stream);
}

void H2ERI_matvec_multi(H2ERI_p h, const int nvec, const double *X, double *Y) {
    if (!h || nvec < 0 || (nvec > 0 && (!X || !Y))) { set_error(ERR_ARG, "H2ERI_matvec_multi: bad argument"); return; }
    const size_t N = (size_t)h->krnl_mat_size;
    for (int k = 0; k < nvec; k++) {
        if (!matvec_host_impl(h, X + k * N, Y + k * N)) return;
    }
}

void H2ERI_matvec_multi_dev(H2ERI_p h, const int nvec, const double *X_dev, double *Y_dev, void *stream) {
    if (!h || nvec < 0) { set_error(ERR_ARG, "H2ERI_matvec_multi_dev: bad argument"); return; }
    const size_t N = (size_t)h->krnl_mat_size;
    for (int k = 0; k < nvec; k++)
        if (!matvec_dev_impl(h, X_dev + k * N, Y_dev + k * N, (hipStream_t)stream)) return;
}

void H2ERI_build_Coulomb(H2ERI_p h, const double *den_mat, double *J_mat) {
    if (!h || !den_mat || !J_mat) { set_error(ERR_ARG, "H2ERI_build_Coulomb: null argument"); return; }
    if (!upload(h)) return;
    DevState *d = dstate(h);
    const size_t nn = (size_t)h->num_bf * h->num_bf;
    if (!d->d_D) { if (hipMalloc(&d->d_D, nn * 8) != hipSuccess) { set_error(ERR_DEVICE, "alloc D"); return; } }
    if (!d->d_J) { if (hipMalloc(&d->d_J, nn * 8) != hipSuccess) { set_error(ERR_DEVICE, "alloc J"); return; } }
    if (hipMemcpy(d->d_D, den_mat, nn * 8, hipMemcpyHostToDevice) != hipSuccess) { set_error(ERR_DEVICE, "copy D"); return; }
    if (!coulomb_dev_impl(h, d->d_D, d->d_J, nullptr)) return;
    if (hipMemcpy(J_mat, d->d_J, nn * 8, hipMemcpyDeviceToHost) != hipSuccess) { set_error(ERR_DEVICE, "copy J"); return; }
    collect_phase_ms(h, d);
    h->n_matvec++;
}

void H2ERI_build_Coulomb_multi(H2ERI_p h, const int ndens, const double *den_mats, double *J_mats) {
    const size_t nn = h ? (size_t)h->num_bf * h->num_bf : 0;
    for (int k = 0; k < ndens; k++) {
        H2ERI_build_Coulomb(h, den_mats + k * nn, J_mats + k * nn);
        if (h2eri_last_error()) return;
    }
}

void H2ERI_build_Coulomb_dev(H2ERI_p h, const double *den_dev, double *J_dev, void *stream) {
    if (!h || !den_dev || !J_dev) { set_error(ERR_ARG, "H2ERI_build_Coulomb_dev: null argument"); return; }
    coulomb_dev_impl(h, den_dev, J_dev, (hipStream_t)stream);
}

void H2ERI_set_precision(H2ERI_p h, const int ub_bits) {
    if (!h || (ub_bits != 32 && ub_bits != 64)) { set_error(ERR_ARG, "H2ERI_set_precision: 32 or 64"); return; }
    DevState *d = dstate_get(h);
    if (d->ub_bits != ub_bits) { d->ub_bits = ub_bits; release_buffers(d); }
}

size_t H2ERI_upload(H2ERI_p h) {
    if (!h || !upload(h)) return 0;
    return dstate(h)->resident_bytes;
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
        out5[5 * i + 1] = r.kind;
        out5[5 * i + 2] = r.blocks;
        out5[5 * i + 3] = r.mat_bytes;
        out5[5 * i + 4] = r.ms;
    }
    return n;
}

void H2ERI_matvec_bytes(H2ERI_p h, double *out6) {
    if (!h || !out6) return;
    if (!upload(h)) return;
    const Plan &P = dstate(h)->plan;
    double alg = P.bytes_U + P.bytes_B + P.bytes_D + P.bytes_vec;
    const size_t *ms = h->mat_size;
    double ref = 8.0 * (double)(ms[MV_FWD_SIZE_IDX] + ms[MV_MID_SIZE_IDX] + ms[MV_BWD_SIZE_IDX] + ms[MV_DEN_SIZE_IDX]);
    out6[0] = alg;
    out6[1] = ref;
    out6[2] = P.bytes_U;
    out6[3] = P.bytes_B;
    out6[4] = P.bytes_D;
    out6[5] = P.bytes_vec;
}

// Near-field schedule (tuning knob; re-plans): 0 balances the near-field
// tiles over the sweep launches, 1 runs them in three launches of their own.
void H2ERI_set_schedule(H2ERI_p h, const int mode) {
    if (!h || (mode != 0 && mode != 1)) { set_error(ERR_ARG, "H2ERI_set_schedule: 0 or 1"); return; }
    DevState *d = dstate_get(h);
    if (d->nf_schedule != mode) { d->nf_schedule = mode; release_buffers(d); }
}

// Tile size of the bi-GEMV tasks in bytes (tuning knob; re-plans).
void H2ERI_set_tile_bytes(H2ERI_p h, const double bytes) {
    if (!h || bytes < 1024) return;
    DevState *d = dstate_get(h);
    d->tile_bytes = bytes;
    release_buffers(d);
}

size_t h2eri_struct_size(void) { return sizeof(struct H2ERI); }

void H2ERI_build_exchange(H2ERI_p h, const double *den_mat, double *K_mat) {
    (void)h; (void)den_mat; (void)K_mat;
    set_error(ERR_UNSUPPORTED, "H2ERI_build_exchange: not available in this build yet");
}

// Verification of a plan on the host (no HIP call): executes every tile and
// reduce task with the kernels' semantics in launch order, and checks that
// within one launch no task reads an element another task of that launch
// writes and no element is stored by two tasks.  For tests of the planner;
// the device entry points never call it.
int H2ERI_plan_emulate(H2ERI_p h, const int ub_bits, const double *x, double *y) {
    if (!h || !h->U || !h->c_B_blks || !h->c_D_blks || !x || !y) { set_error(ERR_ARG, "plan_emulate: bad args"); return -1; }
    Plan P;
    Builder b(h, P, ub_bits == 32 ? 32 : 64, dstate_get(h)->tile_bytes, dstate_get(h)->nf_schedule);
    if (!b.build()) return -1;
    const long long N = h->krnl_mat_size;
    std::vector<double> Wv((size_t)P.nW, 0.0);
    std::memcpy(Wv.data(), x, sizeof(double) * (size_t)N);
    std::vector<int> wmark((size_t)P.nW, -1), rmark((size_t)P.nW, -1), ymark((size_t)N, -1);
    int races = 0;
    auto write_el = [&](long long e, int task, bool acc) {
        if (wmark[e] >= 0 && wmark[e] != task && !acc) races++;
        if (rmark[e] >= 0 && rmark[e] != task) races++;
        wmark[e] = task;
    };
    auto read_el = [&](long long e, int task) {
        if (wmark[e] >= 0 && wmark[e] != task) races++;
        rmark[e] = task;
    };
    int launch = 0;
    for (const Stage &st : P.stages) {
        // tile launch
        std::fill(wmark.begin(), wmark.end(), -1);
        std::fill(rmark.begin(), rmark.end(), -1);
        for (int q = 0; q < st.t64_cnt + st.t32_cnt; q++) {
            const bool f32 = q >= st.t64_cnt;
            const BTask &t = f32 ? P.t32[st.t32_beg + q - st.t64_cnt] : P.t64[st.t64_beg + q];
            const int id = q;
            auto M = [&](int r, int c) -> double {
                long long e = t.mat + (long long)r * t.ld + c;
                return f32 ? (double)b.host32[e] : b.host64[e];
            };
            if (t.out0 >= 0) {
                for (int c = 0; c < t.ncol; c++) read_el(t.in0 + c, id);
                for (int r = 0; r < t.nrow; r++) {
                    double d = 0.0;
                    for (int c = 0; c < t.ncol; c++) d += M(r, c) * Wv[t.in0 + c];
                    write_el(t.out0 + r, id, (t.flags & 1) != 0);
                    if (t.flags & 1) Wv[t.out0 + r] += d;
                    else Wv[t.out0 + r] = d;
                }
            }
            if (t.out1 >= 0) {
                for (int r = 0; r < t.nrow; r++) read_el(t.in1 + r, id);
                for (int c = 0; c < t.ncol; c++) {
                    double s = 0.0;
                    for (int r = 0; r < t.nrow; r++) s += M(r, c) * Wv[t.in1 + r];
                    write_el(t.out1 + c, id, false);
                    Wv[t.out1 + c] = s;
                }
            }
        }
        launch++;
        // reduce launch
        std::fill(wmark.begin(), wmark.end(), -1);
        std::fill(rmark.begin(), rmark.end(), -1);
        for (int q = 0; q < st.r_cnt; q++) {
            const RTask &r = P.rt[st.r_beg + q];
            for (int i = 0; i < r.len; i++) {
                double s = 0.0;
                for (int j = r.src_beg; j < r.src_end; j++) {
                    read_el(P.rsrc[j] + r.soff + i, q);
                    s += Wv[P.rsrc[j] + r.soff + i];
                }
                if (r.flags & 1) {
                    long long e = r.dst + i;
                    if (ymark[e] == launch) races++;
                    ymark[e] = launch;
                    y[e] = (r.flags & 2) ? y[e] + s : s;
                } else {
                    write_el(r.dst + i, q, false);
                    Wv[r.dst + i] = (r.flags & 2) ? Wv[r.dst + i] + s : s;
                }
            }
        }
        launch++;
    }
    return races;
}

// Builds the device plan on the host only and runs its bounds check; no HIP
// call is made.  Returns the number of tile tasks, or -1 on failure.
int H2ERI_plan_check(H2ERI_p h, const int ub_bits) {
    if (!h || !h->U || !h->c_B_blks || !h->c_D_blks) { set_error(ERR_ARG, "plan_check: nothing built"); return -1; }
    Plan P;
    double tb = dstate_get(h)->tile_bytes;
    Builder b(h, P, ub_bits == 32 ? 32 : 64, tb, dstate_get(h)->nf_schedule);
    if (!b.build()) return -1;
    return (int)(P.t64.size() + P.t32.size());
}

void H2ERI_exchange_workbuf_init(H2ERI_p) {}
void H2ERI_exchange_workbuf_free(H2ERI_p) {}

}  // extern "C"
