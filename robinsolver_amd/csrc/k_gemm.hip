// k_gemm.hip -- batched complex GEMM on the gfx950 matrix cores.
//
// Replaces the zgemm/zgemv call sites of the reference merge and solve passes
// (robin_tree_module.f90:550-605, 898-959, 1094-1109) and the zgetrs-based
// block products of the merge (robin_tree_module.f90:482-534), which this
// build expresses as products with explicit interface inverses.
//
// Complex products run as three real MFMA products (Gauss's 3M) on interleaved
// storage:
//   P1 += Ar*Br,  P2 += Ai*Bi,  P3 += (Ar+Ai)*(Br+Bi),
//   Cr = P1 - P2,  Ci = P3 - P1 - P2,
// with v_mfma_f64_16x16x4f64 (complex128) or v_mfma_f32_16x16x4f32
// (complex64, FP32 accumulation); the sums Ar+Ai and Br+Bi are formed from the
// operands as they are read from LDS.  The error bound of Ci carries
// (|Ar|+|Ai|)(|Br|+|Bi|) in place of |A||B|, at most twice as large.  A 256-thread workgroup computes a 64 x BN
// output tile from 64 x 16 and 16 x BN operand tiles staged in LDS; BN = 64 for
// wide products (merge), BN = 16 for the few-RHS solve passes, which then stay
// HBM-bound rather than MFMA-bound; gemm_dma_kernel does the same with its
// operands copied from memory straight into LDS stages (up to 128 x 128 tiles,
// 256 or 512 threads).  The K reduction of every output column is a fixed
// sequence that does not depend on N, on the other columns or on the tile, so
// a right-hand side gets the same bits alone or batched with others.
//
// Each problem of a batch may scatter its output rows/columns through index
// maps (the face permutation between a child's and its parent's boundary
// ordering), read B from two row ranges (the two children's interface
// vectors) or gather its rows through a row map (a node's incoming data, held
// where the levels above wrote it), and add beta*Din in the epilogue.
#include <algorithm>
#include <type_traits>

#include "rs_common.hpp"
#include "rs_kernels.hpp"

namespace rs {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename R> struct acc_t;
template <> struct acc_t<double> { using T = d4; };
template <> struct acc_t<float> { using T = f4; };

__device__ inline d4 mfma(double a, double b, d4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
__device__ inline f4 mfma(float a, float b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

// Row of accumulator element v held by `lane` in a 16x16 output block.  Both
// instructions read A[i][k] from lane i + 16k and B[k][j] from lane j + 16k and
// put column j = lane % 16 in the lane, but they order the rows differently:
// the f32 form keeps four consecutive rows per lane group, the f64 form
// interleaves the four lane groups.
template <typename R> __device__ inline int out_row(int lane, int v);
template <> __device__ inline int out_row<float>(int lane, int v) { return 4 * (lane >> 4) + v; }
template <> __device__ inline int out_row<double>(int lane, int v) { return (lane >> 4) + 4 * v; }

constexpr int BMX = 64;   // row-tile height assumed by the K-split rule (whatever the tile)
constexpr int KALIGN = 32; // K-split slices start at multiples of this (whatever the chunk depth)

// Epilogue through LDS: the accumulators (MFMA layout: 16 columns per wave
// instruction) go to a column-major tile Ct [BN][BM + 1], and each thread then
// handles consecutive rows of one column, so stores (and the Din reads) are
// contiguous runs of a column.
template <typename R, int BM, int TM, int TN>
__device__ __forceinline__ void gemm_acc_to_lds(c_t<R>* Ct, const typename acc_t<R>::T (&p1)[TM][TN],
                                       const typename acc_t<R>::T (&p2)[TM][TN],
                                       const typename acc_t<R>::T (&p3)[TM][TN], int wm, int wn, int lane) {
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = wm + i * 16 + out_row<R>(lane, v);
                const int col = wn + j * 16 + (lane & 15);
                Ct[col * (BM + 1) + row] = mk<R>(p1[i][j][v] - p2[i][j][v], (p3[i][j][v] - p1[i][j][v]) - p2[i][j][v]);
            }
}

// The tile from Ct to the K-slice workspace, or through the problem's maps and
// alpha / beta * Din to its output.
template <typename S, typename R, int BM, int BN, int NT = 256>
__device__ __forceinline__ void gemm_tile_out(const c_t<R>* Ct, const GemmDesc& D, c_t<R>* ws, int p, int ks, int kslices,
                                     int M, int N, int m0, int n0, int tid) {
    using V = c_t<R>;
    using VS = c_t<S>;
    constexpr int NE = BM * BN / NT;
    if (kslices > 1) {
        V* W = ws + ((size_t)p * kslices + ks) * (size_t)M * N;
#pragma unroll
        for (int q = 0; q < NE; ++q) {
            const int e = tid + NT * q;
            const int r = e % BM, c = e / BM;
            const int row = m0 + r, col = n0 + c;
            if (row < M && col < N) W[(size_t)row + (size_t)col * M] = Ct[c * (BM + 1) + r];
        }
        return;
    }
    VS* const Out = (VS*)D.Out;
    const VS* const Din = (const VS*)D.Din;
    const int* const rmap = D.rmap;
    const int* const cmap = D.cmap;
    const int ldd = D.ldd, ldo = D.ldo;
    const VS alpha = mk<S>((S)D.ar, (S)D.ai), beta = mk<S>((S)D.br, (S)D.bi);
    // In groups of up to eight elements: all index-map and Din loads of a group
    // are issued back to back (branches only around whole loops), then the
    // results are formed and stored.  (Whole-tile groups would hold too many
    // registers for the largest tiles.)
    constexpr int NG = NE < 8 ? NE : 8;
    static_assert(NE % NG == 0, "epilogue group size");
    const bool dmap = D.dmap != 0;
#pragma unroll
    for (int g0 = 0; g0 < NE; g0 += NG) {
        int rowc[NG], colc[NG], orow[NG], ocol[NG];
        bool ok[NG];
#pragma unroll
        for (int q = 0; q < NG; ++q) {
            const int e = tid + NT * (g0 + q);
            const int row = m0 + e % BM, col = n0 + e / BM;
            ok[q] = row < M && col < N;
            rowc[q] = row < M ? row : M - 1;
            colc[q] = col < N ? col : N - 1;
        }
        if (rmap) {
#pragma unroll
            for (int q = 0; q < NG; ++q) orow[q] = gldi(rmap + rowc[q]);
        } else {
#pragma unroll
            for (int q = 0; q < NG; ++q) orow[q] = rowc[q];
        }
        if (cmap) {
#pragma unroll
            for (int q = 0; q < NG; ++q) ocol[q] = gldi(cmap + colc[q]);
        } else {
#pragma unroll
            for (int q = 0; q < NG; ++q) ocol[q] = colc[q];
        }
        VS val[NG];
#pragma unroll
        for (int q = 0; q < NG; ++q) {
            const int e = tid + NT * (g0 + q);
            const V c = Ct[(e / BM) * (BM + 1) + e % BM];
            val[q] = cmul(alpha, mk<S>((S)c.x, (S)c.y));
        }
        if (Din) {
            VS din[NG];
            if (dmap) {
#pragma unroll
                for (int q = 0; q < NG; ++q) din[q] = gld(Din + ((size_t)orow[q] + (size_t)ocol[q] * ldd));
            } else {
#pragma unroll
                for (int q = 0; q < NG; ++q) din[q] = gld(Din + ((size_t)rowc[q] + (size_t)colc[q] * ldd));
            }
#pragma unroll
            for (int q = 0; q < NG; ++q) val[q] = cfma(beta, din[q], val[q]);
        }
#pragma unroll
        for (int q = 0; q < NG; ++q)
            if (ok[q]) gst(Out + ((size_t)orow[q] + (size_t)ocol[q] * ldo), val[q]);
    }
}

// One workgroup computes a 64 x BN tile over the K range of its slice.  The
// next 64 x 32 / 32 x BN operand tiles are fetched into registers while the
// current ones, staged in LDS, feed the matrix cores.  With kslices > 1 the raw
// partial products go to a workspace and gemm_reduce_kernel sums the slices in
// a fixed order.
//
// S is the precision of A, B, Din and Out, R the precision of the products:
// S = double, R = float is the complex64 merge of the mixed-precision
// factorization (operands rounded to complex64 as they are staged, FP32
// accumulation, the epilogue and the stored result in FP64).
//
// BK: K depth of one staged chunk; DB: two LDS chunk buffers (one barrier per
// chunk) instead of one (two).  Neither changes the order in which a column's
// products are accumulated.
// (Measured on the solve-pass shapes: at N = 256, BK = 16 with two buffers beat
// BK = 32 or 64 with one by 10-25%; the 16-column tiles, which carry the
// few-RHS matrix streams, do best with BK = 32 and one buffer, 12% ahead of
// BK = 16 with two on the 3D top levels at N = 16.)  The MFMA steps of a chunk
// are always fully unrolled.
//
// One launch may hold several groups of problems (GemmLaunch): each group has
// its own shape, descriptors, K split and workspace, and owns a consecutive
// range of workgroups.  A problem runs the same instruction sequence whether
// its group is launched alone or with others.
template <typename S, typename R, int BM, int BN, int BK = (BN == 16 ? 32 : 16), bool DB = (BN != 16)>
__global__ void __launch_bounds__(256) gemm_kernel(GemmLaunch GL) {
    using V = c_t<R>;
    using VS = c_t<S>;
    using A4 = typename acc_t<R>::T;
    // four waves: 2x2 for 128x64, 64x128, 64x64, 32x64, 32x32 and 64x32 tiles,
    // 4x1 for 64x16 and 128x16, 1x4 for 16x64
    constexpr int WAVES_N = (BN == 16) ? 1 : (BM == 16 ? 4 : 2);
    constexpr int WAVES_M = 4 / WAVES_N;
    constexpr int WM = BM / WAVES_M, WN = BN / WAVES_N;
    constexpr int TM = WM / 16, TN = WN / 16;
    constexpr int NA = BM * BK / 256, NB = BK * BN / 256;
    // operand tiles; after the K loop the same memory holds the output tile
    // (column stride BM + 1, so the 16 columns of one MFMA row group fall in
    // different banks)
    // (operand rows padded by one element against bank conflicts)
    constexpr int CHUNK = BK * (BM + BN + 2);
    constexpr int SMEM = ((DB ? 2 : 1) * CHUNK > BN * (BM + 1)) ? (DB ? 2 : 1) * CHUNK : BN * (BM + 1);
    __shared__ V smem[SMEM];

    // this workgroup's group (static indices only: the groups stay in kernel arguments)
    const GemmDesc* descs = GL.g[0].descs;
    V* ws = (V*)GL.g[0].ws;
    int M = GL.g[0].M, N = GL.g[0].N, K = GL.g[0].K, tilesM = GL.g[0].tilesM, tilesN = GL.g[0].tilesN,
        kslices = GL.g[0].ks, kchunk = GL.g[0].kchunk, wg = blockIdx.x;
#pragma unroll
    for (int q = 1; q < GEMM_MAX_GROUPS; ++q)
        if (q < GL.n && (int)blockIdx.x >= GL.g[q].wg0) {
            descs = GL.g[q].descs; ws = (V*)GL.g[q].ws;
            M = GL.g[q].M; N = GL.g[q].N; K = GL.g[q].K; tilesM = GL.g[q].tilesM; tilesN = GL.g[q].tilesN;
            kslices = GL.g[q].ks; kchunk = GL.g[q].kchunk; wg = blockIdx.x - GL.g[q].wg0;
        }
    const int tiles = tilesM * tilesN;
    const int per = tiles * kslices;
    const int p = wg / per;
    const int rem = wg - p * per;
    const int ks = rem / tiles;
    const int t = rem - ks * tiles;
    const int tm = t % tilesM, tn = t / tilesM;
    const GemmDesc& D = descs[p];
    // descriptor fields into registers once (the loop's stores could alias them)
    const VS* const A = (const VS*)D.A;
    const VS* const B = (const VS*)D.B;
    const VS* const B2 = (const VS*)D.B2;
    const int* const bmap = D.bmap;
    const int lda = D.lda, ldb = D.ldb, ksplit = D.ksplit;
    const int m0 = tm * BM, n0 = tn * BN;
    const int kbeg = ks * kchunk, kend = min(K, kbeg + kchunk);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave % WAVES_M) * WM, wn = (wave / WAVES_M) * WN;

    A4 p1[TM][TN], p2[TM][TN], p3[TM][TN];
    for (int i = 0; i < TM; ++i)
        for (int j = 0; j < TN; ++j)
            for (int v = 0; v < 4; ++v) { p1[i][j][v] = R(0); p2[i][j][v] = R(0); p3[i][j][v] = R(0); }

    // Operand prefetch: every load is issued unconditionally from an index
    // clamped into range (a load under a branch makes the compiler wait for all
    // outstanding loads where the branches join); elements outside the matrix
    // or the K range are replaced by zero when staged.
    VS ra[NA], rb[NB];
    bool oka[NA], okb[NB];
    // a row-mapped B: this thread's B row is the same in every column (k =
    // tid % BK); its index is loaded one chunk ahead, so that the data loads do
    // not wait for it
    int bix = 0;
    if (bmap && kbeg < kend) bix = gldi(bmap + min(kbeg + (int)(tid % BK), kend - 1));
    auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < NA; ++q) {
            const int e = tid + 256 * q;
            const int r = e % BM, k = e / BM;
            const int gr = m0 + r, gk = k0 + k;
            oka[q] = gr < M && gk < kend;
            ra[q] = gld(A + ((size_t)min(gr, M - 1) + (size_t)min(gk, kend - 1) * lda));
        }
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            const int e = tid + 256 * q;
            const int k = e % BK, c = e / BK;
            const int gk = min(k0 + k, kend - 1), gc = min(n0 + c, N - 1);
            okb[q] = k0 + k < kend && n0 + c < N;
            const VS* src = bmap ? B + ((size_t)bix + (size_t)gc * ldb)
                                 : (gk < ksplit) ? B + ((size_t)gk + (size_t)gc * ldb)
                                                 : B2 + ((size_t)(gk - ksplit) + (size_t)gc * ldb);
            rb[q] = gld(src);
        }
        if (bmap) bix = gldi(bmap + min(k0 + BK + (int)(tid % BK), kend - 1));
    };
    // chunk buffer b: As [BK][BM + 1], then Bs [BK][BN + 1]
    auto stash = [&](int b) {
        V (*As)[BM + 1] = (V (*)[BM + 1])(smem + b * CHUNK);
        V (*Bs)[BN + 1] = (V (*)[BN + 1])(smem + b * CHUNK + BK * (BM + 1));
#pragma unroll
        for (int q = 0; q < NA; ++q) {
            const int e = tid + 256 * q;
            As[e / BM][e % BM] = oka[q] ? mk<R>((R)ra[q].x, (R)ra[q].y) : cz<R>();
        }
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            const int e = tid + 256 * q;
            Bs[e % BK][e / BK] = okb[q] ? mk<R>((R)rb[q].x, (R)rb[q].y) : cz<R>();
        }
    };

    if (kbeg < kend) {
        fetch(kbeg);
        stash(0);
        __syncthreads();
    }
    int cur = 0;
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
        const bool more = k0 + BK < kend;
        if (more) fetch(k0 + BK);
        V (*As)[BM + 1] = (V (*)[BM + 1])(smem + cur * CHUNK);
        V (*Bs)[BN + 1] = (V (*)[BN + 1])(smem + cur * CHUNK + BK * (BM + 1));
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            V af[TM], bf[TN];
            R as[TM], bs[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) {
                af[i] = As[kk + (lane >> 4)][wm + i * 16 + (lane & 15)];
                as[i] = af[i].x + af[i].y;
            }
#pragma unroll
            for (int j = 0; j < TN; ++j) {
                bf[j] = Bs[kk + (lane >> 4)][wn + j * 16 + (lane & 15)];
                bs[j] = bf[j].x + bf[j].y;
            }
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j) {
                    p1[i][j] = mfma(af[i].x, bf[j].x, p1[i][j]);
                    p2[i][j] = mfma(af[i].y, bf[j].y, p2[i][j]);
                    p3[i][j] = mfma(as[i], bs[j], p3[i][j]);
                }
        }
        if constexpr (DB) {
            // the other buffer was last read before the previous barrier
            if (more) stash(cur ^ 1);
            __syncthreads();
            cur ^= 1;
        } else {
            __syncthreads();
            if (more) {
                stash(0);
                __syncthreads();
            }
        }
    }

    // the K loop ended with a barrier, so the operand tiles are free
    gemm_acc_to_lds<R, BM, TM, TN>(smem, p1, p2, p3, wm, wn, lane);
    __syncthreads();
    gemm_tile_out<S, R, BM, BN>(smem, D, ws, p, ks, kslices, M, N, m0, n0, tid);
}

// 16 zero bytes: the source of staged operand elements beyond the K range
__device__ __align__(16) float4 g_zero16 = {0.f, 0.f, 0.f, 0.f};

// s_waitcnt vmcnt(n) alone (expcnt and lgkmcnt not waited for)
template <int n> __device__ inline void wait_vm() {
    static_assert(n >= 0 && n < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((n & 15) | ((n >> 4) << 14) | (7 << 4) | (15 << 8));
}

// Copies from global memory straight into LDS: each lane's 16 (4) bytes at g
// land at LDS byte address lds + 16 (4) x lane; lds is the same in every lane.
// Written as asm (with the LDS base in M0, one wait state between setting M0
// and the copy, as the compiler leaves for its own LDS copies), so that the
// compiler keeps no count of them: the kernel waits for them itself (wait_vm)
// before the barrier after which they are read.  No other instruction of
// these kernels uses M0.
__device__ inline void lds_copy16(const void* g, unsigned lds) {
    asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(g), "s"(lds)
                 : "memory", "m0");
}
__device__ inline void lds_copy4(const void* g, unsigned lds) {
    asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dword %0, off" ::"v"(g), "s"(lds)
                 : "memory", "m0");
}
// workgroup barrier after this wave's LDS reads (lgkmcnt(0)), without the
// release fence's wait for outstanding global loads
__device__ inline void bare_barrier() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(15 | (3 << 14) | (7 << 4));
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// gemm_kernel with the operands copied from memory straight into LDS
// (global_load_lds_dwordx4: each lane's 16 bytes land at the wave's LDS base
// plus 16 x lane, no registers, no conversion) through NS chunk stages, so
// that the loads of NS - 1 chunks are in flight while one is multiplied; one
// barrier per chunk.  The products, their order over K (16-deep chunks from
// the slice start, k ascending in steps of four, Gauss's three real
// products) and the epilogue are those of gemm_kernel, so the results are the
// same bits.  The operands are staged as stored (precision S) and rounded to
// the precision of the products R as they are read from LDS.
//
// Stage layout in 16-byte units (EPU complex elements each): A as [k][BM/EPU]
// with unit index i/EPU xor swzA(k), B as [n][BK/EPU] (k contiguous, as B is
// stored) with unit index (k/EPU) xor swzB(n); the MFMA operand reads (16 lanes
// on 16 consecutive rows or columns at one k) then fall in distinct banks.
// Each lane loads the unit whose position in the stage is its own, so the
// swizzle is applied to the global addresses.  Elements beyond the K range
// come from g_zero16; rows beyond M and columns beyond N from in-range data
// (their products land in output rows and columns that are not stored).
// A row-mapped B (the down pass's incoming rows) moves in 16-byte units too when
// the map sends every row pair to two contiguous rows starting at an even one
// (GemmDesc::bpair; complex128 units are single rows): the map over the slice's
// K range is first copied into LDS.  Operands that cannot move in 16-byte units
// (other row maps, odd leading dimensions or offsets) move in 4-byte pieces
// into the same layout.
template <typename S, typename R, int BM, int BN, int NS, int NT = 256, int WAVES_N = 2>
__global__ void __launch_bounds__(NT) gemm_dma_kernel(GemmLaunch GL) {
    using V = c_t<S>;    // operands as stored and staged
    using VR = c_t<R>;   // operands of the products
    using A4 = typename acc_t<R>::T;
    constexpr int BK = 16;   // (32-deep chunks and B in padded columns, moved in 4-byte pieces, measured slower)
    constexpr int EPU = 16 / (int)sizeof(V);
    constexpr int NW = NT / 64;
    constexpr int WAVES_M = NW / WAVES_N;
    constexpr int WM = BM / WAVES_M, WN = BN / WAVES_N;
    constexpr int TM = WM / 16, TN = WN / 16;
    constexpr int AROW = BM / EPU, BCOL = BK / EPU;   // units per A k-row, per B column
    constexpr int AU = BK * AROW, BU = BN * BCOL;      // units per stage
    constexpr int NUA = AU / NT, NUB = BU / NT;        // 16-byte loads per thread and chunk
    static_assert(AU % NT == 0 && BU % NT == 0 && AROW >= 16 && TM >= 1 && TN >= 1, "tile shape");
    constexpr int STAGE = (AU + BU) * EPU;             // elements per stage
    constexpr int CT = (BN * (BM + 1) * (int)sizeof(VR) + (int)sizeof(V) - 1) / (int)sizeof(V);   // output tile
    constexpr int SMEM = NS * STAGE > CT ? NS * STAGE : CT;
    __shared__ __align__(16) V smem[SMEM];
    constexpr int PBL = 512;   // map entries of a row-mapped B (one per 16-byte unit) held in LDS
    __shared__ int pbl[PBL];

    const GemmDesc* descs = GL.g[0].descs;
    VR* ws = (VR*)GL.g[0].ws;
    int M = GL.g[0].M, N = GL.g[0].N, K = GL.g[0].K, tilesM = GL.g[0].tilesM, tilesN = GL.g[0].tilesN,
        kslices = GL.g[0].ks, kchunk = GL.g[0].kchunk, wg = blockIdx.x;
#pragma unroll
    for (int q = 1; q < GEMM_MAX_GROUPS; ++q)
        if (q < GL.n && (int)blockIdx.x >= GL.g[q].wg0) {
            descs = GL.g[q].descs; ws = (VR*)GL.g[q].ws;
            M = GL.g[q].M; N = GL.g[q].N; K = GL.g[q].K; tilesM = GL.g[q].tilesM; tilesN = GL.g[q].tilesN;
            kslices = GL.g[q].ks; kchunk = GL.g[q].kchunk; wg = blockIdx.x - GL.g[q].wg0;
        }
    const int tiles = tilesM * tilesN;
    const int per = tiles * kslices;
    const int p = wg / per;
    const int rem = wg - p * per;
    const int ks = rem / tiles;
    const int t = rem - ks * tiles;
    const int tm = t % tilesM, tn = t / tilesM;
    const GemmDesc& D = descs[p];
    const V* const A = (const V*)D.A;
    const V* const B = (const V*)D.B;
    const V* const B2 = (const V*)D.B2;
    const int* const bmap = D.bmap;
    const int lda = D.lda, ldb = D.ldb, ksplit = D.ksplit;
    const int m0 = tm * BM, n0 = tn * BN;
    const int kbeg = ks * kchunk, kend = min(K, kbeg + kchunk);
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = (wave % WAVES_M) * WM, wn = (wave / WAVES_M) * WN;
    const unsigned lds0 = (unsigned)(size_t)(__attribute__((address_space(3))) V*)smem;

    A4 p1[TM][TN], p2[TM][TN], p3[TM][TN];
    for (int i = 0; i < TM; ++i)
        for (int j = 0; j < TN; ++j)
            for (int v = 0; v < 4; ++v) { p1[i][j][v] = R(0); p2[i][j][v] = R(0); p3[i][j][v] = R(0); }

    auto swzA = [](int k) { return EPU == 2 ? (k & 1) << 3 : 0; };
    auto swzB = [](int n) { return EPU == 2 ? (n >> 1) & 7 : n & 15; };
    // 16-byte units need both operands' elements 16-byte aligned in pairs; a
    // row-mapped B in 16-byte units needs its rows in contiguous even pairs
    // (D.bpair) and its map over the slice's K range in LDS
    const bool mapped = bmap && (EPU == 1 || D.bpair) && (kend - kbeg) <= EPU * PBL;
    const bool fast = (!bmap || mapped) && ((((size_t)A | (size_t)B | (size_t)(B2 ? B2 : B)) & 15) == 0) &&
                      (EPU == 1 || (((lda | ldb | (bmap ? 0 : ksplit) | M | K | kbeg) & 1) == 0));
    if (fast && mapped) {
        for (int j = tid; j < (kend - kbeg + EPU - 1) / EPU; j += NT) pbl[j] = gldi(bmap + kbeg + EPU * j);
        __syncthreads();
    }
    const char* zero = (const char*)&g_zero16;
    // 16-byte copies: each thread's source addresses advance by one chunk per
    // issue (chunks are issued in order); an element beyond the K range comes
    // from g_zero16
    const char* pa[NUA];
    int ka[NUA];
#pragma unroll
    for (int q = 0; q < NUA; ++q) {
        const int u = (q * NW + wave) * 64 + lane;
        const int k = u / AROW, ip = (u % AROW) ^ swzA(k);
        const int row = min(m0 + ip * EPU, M - EPU);
        ka[q] = kbeg + k;
        pa[q] = (const char*)(A + (size_t)row + (size_t)ka[q] * lda);
    }
    const size_t astep = (size_t)BK * lda * sizeof(V);
    const char* pb1[NUB];
    const char* pb2[NUB];
    int kb[NUB];
#pragma unroll
    for (int q = 0; q < NUB; ++q) {
        const int u = (q * NW + wave) * 64 + lane;
        const int n = u / BCOL, kp = (u % BCOL) ^ swzB(n);
        const size_t col = (size_t)min(n0 + n, N - 1) * ldb;
        kb[q] = kbeg + kp * EPU;
        pb1[q] = (const char*)(B + col);
        pb2[q] = (const char*)((B2 ? B2 : B) + col) - (size_t)ksplit * sizeof(V);
    }
    auto issue16 = [&](int k0, int st) {
        const unsigned sA = lds0 + st * STAGE * (int)sizeof(V);
        const unsigned sB = sA + AU * 16;
#pragma unroll
        for (int q = 0; q < NUA; ++q) {
            lds_copy16(ka[q] < kend ? pa[q] : zero, sA + (q * NW + wave) * 1024);
            ka[q] += BK;
            pa[q] += astep;
        }
        if (mapped) {
#pragma unroll
            for (int q = 0; q < NUB; ++q) {
                const int gk = kb[q];
                const int row = pbl[min((gk - kbeg) / EPU, PBL - 1)];
                lds_copy16(gk < kend ? pb1[q] + (size_t)row * sizeof(V) : zero, sB + (q * NW + wave) * 1024);
                kb[q] += BK;
            }
        } else {
#pragma unroll
            for (int q = 0; q < NUB; ++q) {
                const int gk = kb[q];
                const char* b = (gk < ksplit ? pb1[q] : pb2[q]) + (size_t)gk * sizeof(V);
                lds_copy16(gk < kend ? b : zero, sB + (q * NW + wave) * 1024);
                kb[q] += BK;
            }
        }
    };
    // the same layout in 4-byte pieces (4 per unit), element by element
    auto issue4 = [&](int k0, int st) {
        const unsigned sA = lds0 + st * STAGE * (int)sizeof(V);
        const unsigned sB = sA + AU * 16;
#pragma unroll
        for (int q = 0; q < 4 * NUA; ++q) {
            const int d = (q * NW + wave) * 64 + lane;
            const int u = d >> 2, w = d & 3;
            const int k = u / AROW, ip = (u % AROW) ^ swzA(k);
            const int e = w / (4 / EPU), c = w % (4 / EPU);
            const int row = min(m0 + ip * EPU + e, M - 1), gk = k0 + k;
            const char* src = gk < kend ? (const char*)(A + (size_t)row + (size_t)gk * lda) + 4 * c : zero;
            lds_copy4(src, sA + (q * NW + wave) * 256);
        }
#pragma unroll
        for (int q = 0; q < 4 * NUB; ++q) {
            const int d = (q * NW + wave) * 64 + lane;
            const int u = d >> 2, w = d & 3;
            const int n = u / BCOL, kp = (u % BCOL) ^ swzB(n);
            const int e = w / (4 / EPU), c = w % (4 / EPU);
            const int col = min(n0 + n, N - 1), gk = k0 + kp * EPU + e;
            const char* src = zero;
            if (gk < kend) {
                const V* b = bmap ? B + gldi(bmap + gk) : (gk < ksplit ? B + (size_t)gk : B2 + (size_t)(gk - ksplit));
                src = (const char*)(b + (size_t)col * ldb) + 4 * c;
            }
            lds_copy4(src, sB + (q * NW + wave) * 256);
        }
    };
    auto compute = [&](int st) {
        const V* sA = smem + st * STAGE;
        const V* sB = sA + AU * EPU;
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            const int k = kk + (lane >> 4);
            VR af[TM], bf[TN];
            R as[TM], bs[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) {
                const int r = wm + i * 16 + (lane & 15);
                const V x = sA[(k * AROW + ((r / EPU) ^ swzA(k))) * EPU + r % EPU];
                af[i] = mk<R>((R)x.x, (R)x.y);
                as[i] = af[i].x + af[i].y;
            }
#pragma unroll
            for (int j = 0; j < TN; ++j) {
                const int n = wn + j * 16 + (lane & 15);
                const V x = sB[(n * BCOL + ((k / EPU) ^ swzB(n))) * EPU + k % EPU];
                bf[j] = mk<R>((R)x.x, (R)x.y);
                bs[j] = bf[j].x + bf[j].y;
            }
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j) {
                    p1[i][j] = mfma(af[i].x, bf[j].x, p1[i][j]);
                    p2[i][j] = mfma(af[i].y, bf[j].y, p2[i][j]);
                    p3[i][j] = mfma(as[i], bs[j], p3[i][j]);
                }
        }
    };
    const int nch = kbeg < kend ? (kend - kbeg + BK - 1) / BK : 0;
    // chunk c is multiplied once this thread's loads of it are done (those of
    // the chunks after it may still be in flight) and every wave has passed
    // the barrier; the stage refilled after that barrier is the one multiplied
    // in the previous iteration
    auto run = [&](auto issue, auto per_chunk) {
        constexpr int PC = decltype(per_chunk)::value;
#pragma unroll
        for (int c = 0; c < NS - 1; ++c)
            if (c < nch) issue(kbeg + c * BK, c);
        for (int c = 0; c < nch; ++c) {
            const int ahead = min(NS - 2, nch - 1 - c);
            if (ahead >= 2) wait_vm<(2 * PC < 63 ? 2 * PC : 63)>();
            else if (ahead == 1) wait_vm<PC>();
            else wait_vm<0>();
            bare_barrier();
            if (c + NS - 1 < nch) issue(kbeg + (c + NS - 1) * BK, (c + NS - 1) % NS);
            compute(c % NS);
        }
    };
    static_assert(NS >= 2 && NS <= 4, "stages");
    if (fast) run(issue16, std::integral_constant<int, NUA + NUB>());
    else run(issue4, std::integral_constant<int, 4 * (NUA + NUB)>());
    wait_vm<0>();
    __syncthreads();
    gemm_acc_to_lds<R, BM, TM, TN>((VR*)smem, p1, p2, p3, wm, wn, lane);
    __syncthreads();
    gemm_tile_out<S, R, BM, BN, NT>((const VR*)smem, D, ws, p, ks, kslices, M, N, m0, n0, tid);
}

// Products of at most 16 columns per tile (one-RHS and few-RHS solve passes),
// which stream A once: each wave computes a 16-row block of a 64-row tile and
// reads its A operand straight from memory in the MFMA register layout (lane l
// holds A[row l % 16][k l / 16], 16 consecutive rows of one column per 16
// lanes), with SK k-steps of operands in flight in each of two register
// stages; no LDS, no barrier.  The MFMA sequence of every 16 x 16 output block
// (k ascending in steps of four, three real products) is that of gemm_kernel,
// so the results are the same bits as with any other tile; the K split, the
// workspace and the epilogue are the same too.
template <typename R, int SK>
__global__ void __launch_bounds__(256) gemm_skinny_kernel(GemmLaunch GL) {
    using V = c_t<R>;
    using A4 = typename acc_t<R>::T;
    const GemmDesc* descs = GL.g[0].descs;
    V* ws = (V*)GL.g[0].ws;
    int M = GL.g[0].M, N = GL.g[0].N, K = GL.g[0].K, tilesM = GL.g[0].tilesM, tilesN = GL.g[0].tilesN,
        kslices = GL.g[0].ks, kchunk = GL.g[0].kchunk, wg = blockIdx.x;
#pragma unroll
    for (int q = 1; q < GEMM_MAX_GROUPS; ++q)
        if (q < GL.n && (int)blockIdx.x >= GL.g[q].wg0) {
            descs = GL.g[q].descs; ws = (V*)GL.g[q].ws;
            M = GL.g[q].M; N = GL.g[q].N; K = GL.g[q].K; tilesM = GL.g[q].tilesM; tilesN = GL.g[q].tilesN;
            kslices = GL.g[q].ks; kchunk = GL.g[q].kchunk; wg = blockIdx.x - GL.g[q].wg0;
        }
    const int tiles = tilesM * tilesN;
    const int per = tiles * kslices;
    const int p = wg / per;
    const int rem = wg - p * per;
    const int ks = rem / tiles;
    const int t = rem - ks * tiles;
    const int tm = t % tilesM, tn = t / tilesM;
    const GemmDesc& D = descs[p];
    const V* const A = (const V*)D.A;
    const V* const B = (const V*)D.B;
    const V* const B2 = (const V*)D.B2;
    const int* const bmap = D.bmap;
    const int lda = D.lda, ldb = D.ldb, ksplit = D.ksplit;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, kq = lane >> 4;
    const int m0 = tm * 64 + wave * 16, n0 = tn * 16;
    const int kbeg = ks * kchunk, kend = min(K, kbeg + kchunk);
    const int arow = m0 + li, bcol = n0 + li;
    const bool rowok = arow < M, colok = bcol < N;
    const V* const Ar = A + (size_t)(rowok ? arow : M - 1);
    const size_t boff = (size_t)(colok ? bcol : N - 1) * ldb;

    A4 p1, p2, p3;
    for (int v = 0; v < 4; ++v) { p1[v] = R(0); p2[v] = R(0); p3[v] = R(0); }
    // two stages of SK k-steps (4 k each); every load is issued unconditionally
    // from a clamped index and zeroed outside the matrix or the K range
    V a0[SK], b0[SK], a1[SK], b1[SK];
    // a row-mapped B: the indices of a stage are loaded one stage ahead (the
    // stages are fetched in k order)
    int bix[SK];
    auto bidx = [&](int k0) {
#pragma unroll
        for (int u = 0; u < SK; ++u) {
            const int k = k0 + 4 * u + kq;
            bix[u] = gldi(bmap + (k < kend ? k : kend - 1));
        }
    };
    if (bmap && kbeg < kend) bidx(kbeg);
    auto fetch = [&](int k0, V (&a)[SK], V (&b)[SK]) {
#pragma unroll
        for (int u = 0; u < SK; ++u) {
            const int k = k0 + 4 * u + kq;
            const int kc = k < kend ? k : kend - 1;
            a[u] = gld(Ar + (size_t)kc * lda);
            b[u] = gld(bmap ? B + boff + bix[u] : (kc < ksplit ? B + boff + kc : B2 + boff + (kc - ksplit)));
        }
        if (bmap) bidx(k0 + 4 * SK);
    };
    auto step = [&](int k0, V (&a)[SK], V (&b)[SK]) {
#pragma unroll
        for (int u = 0; u < SK; ++u) {
            const int k = k0 + 4 * u;
            if (k >= kend) break;
            const bool ok = k + kq < kend;
            const V av = (ok && rowok) ? a[u] : cz<R>();
            const V bv = (ok && colok) ? b[u] : cz<R>();
            p1 = mfma(av.x, bv.x, p1);
            p2 = mfma(av.y, bv.y, p2);
            p3 = mfma(av.x + av.y, bv.x + bv.y, p3);
        }
    };
    constexpr int KS = 4 * SK;
    if (kbeg < kend) fetch(kbeg, a0, b0);
    for (int k0 = kbeg; k0 < kend; k0 += 2 * KS) {
        if (k0 + KS < kend) fetch(k0 + KS, a1, b1);
        step(k0, a0, b0);
        if (k0 + KS >= kend) break;
        if (k0 + 2 * KS < kend) fetch(k0 + 2 * KS, a0, b0);
        step(k0 + KS, a1, b1);
    }

    // epilogue: lane holds rows out_row(lane, v) of column li of the wave's block
    if (kslices > 1) {
        V* W = ws + ((size_t)p * kslices + ks) * (size_t)M * N;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = m0 + out_row<R>(lane, v), col = bcol;
            if (row < M && colok) W[(size_t)row + (size_t)col * M] = mk<R>(p1[v] - p2[v], (p3[v] - p1[v]) - p2[v]);
        }
        return;
    }
    V* const Out = (V*)D.Out;
    const V* const Din = (const V*)D.Din;
    const V alpha = mk<R>((R)D.ar, (R)D.ai), beta = mk<R>((R)D.br, (R)D.bi);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int row = m0 + out_row<R>(lane, v);
        const bool ok = row < M && colok;
        const int rowc = row < M ? row : M - 1, colc = colok ? bcol : N - 1;
        const int orow = D.rmap ? gldi(D.rmap + rowc) : rowc;
        const int ocol = D.cmap ? gldi(D.cmap + colc) : colc;
        V val = cmul(alpha, mk<R>(p1[v] - p2[v], (p3[v] - p1[v]) - p2[v]));
        if (Din) {
            const size_t di = D.dmap ? (size_t)orow + (size_t)ocol * D.ldd : (size_t)rowc + (size_t)colc * D.ldd;
            val = cfma(beta, gld(Din + di), val);
        }
        if (ok) gst(Out + ((size_t)orow + (size_t)ocol * D.ldo), val);
    }
}

// Products of at most NC columns without the matrix cores: one lane per
// (problem, K slice, output row), lanes of a wave on consecutive rows, so each
// load of A is one coalesced column segment and nothing is computed for the
// 15 of 16 MFMA columns a one-column product leaves empty.  A
// v_mfma_*_16x16x4 accumulates its four k terms as four fused multiply-adds
// in k order, so the chains below -- k ascending from the slice start in
// groups of four, the last group padded with zero terms, the three real
// products of Gauss's form -- give the bits of every other tile; the K split,
// the workspace and the epilogue are those of gemm_skinny_kernel.
template <typename R, int NC, int U>
__global__ void __launch_bounds__(256) gemv_kernel(GemmLaunch GL) {
    using V = c_t<R>;
    static_assert(U % 4 == 0, "k steps come in groups of four");
    const GemmDesc* descs = GL.g[0].descs;
    V* ws = (V*)GL.g[0].ws;
    int M = GL.g[0].M, N = GL.g[0].N, K = GL.g[0].K, nprob = GL.g[0].nprob, kslices = GL.g[0].ks,
        kchunk = GL.g[0].kchunk, wg = blockIdx.x;
#pragma unroll
    for (int q = 1; q < GEMM_MAX_GROUPS; ++q)
        if (q < GL.n && (int)blockIdx.x >= GL.g[q].wg0) {
            descs = GL.g[q].descs; ws = (V*)GL.g[q].ws;
            M = GL.g[q].M; N = GL.g[q].N; K = GL.g[q].K; nprob = GL.g[q].nprob;
            kslices = GL.g[q].ks; kchunk = GL.g[q].kchunk; wg = blockIdx.x - GL.g[q].wg0;
        }
    const long long chain = (long long)wg * 256 + threadIdx.x;
    if (chain >= (long long)nprob * kslices * M) return;
    const int row = (int)(chain % M);
    const long long ps = chain / M;
    const int ks = (int)(ps % kslices), p = (int)(ps / kslices);
    const GemmDesc& D = descs[p];
    const V* const Ar = (const V*)D.A + row;
    const V* const B = (const V*)D.B;
    const V* const B2 = (const V*)D.B2;
    const int* const bmap = D.bmap;
    const int lda = D.lda, ldb = D.ldb, ksplit = D.ksplit;
    const int kbeg = ks * kchunk, kend = min(K, kbeg + kchunk);
    const int kend4 = kbeg + ((kend - kbeg + 3) & ~3);   // end of the last group of four

    R p1[NC], p2[NC], p3[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) { p1[c] = R(0); p2[c] = R(0); p3[c] = R(0); }
    // two register stages of U k-steps: the loads of the next stage are in
    // flight while the current one is summed
    V a0[U], b0[U][NC], a1[U], b1[U][NC];
    auto fetch = [&](int k0, V (&a)[U], V (&b)[U][NC]) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = k0 + u;
            const bool ok = k < kend;
            const int kc = ok ? k : kend - 1;
            a[u] = gld(Ar + (size_t)kc * lda);
            const int bo = bmap ? gldi(bmap + kc) : 0;
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const size_t boff = (size_t)(c < N ? c : N - 1) * ldb;
                b[u][c] = gld(bmap ? B + boff + bo : (kc < ksplit ? B + boff + kc : B2 + boff + (kc - ksplit)));
            }
            if (!ok) {
                a[u] = cz<R>();
#pragma unroll
                for (int c = 0; c < NC; ++c) b[u][c] = cz<R>();
            }
        }
    };
    auto step = [&](int k0, V (&a)[U], V (&b)[U][NC]) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (k0 + u >= kend4) break;
            const R as = a[u].x + a[u].y;
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                p1[c] = fma(a[u].x, b[u][c].x, p1[c]);
                p2[c] = fma(a[u].y, b[u][c].y, p2[c]);
                p3[c] = fma(as, b[u][c].x + b[u][c].y, p3[c]);
            }
        }
    };
    if (kbeg < kend4) fetch(kbeg, a0, b0);
    for (int k0 = kbeg; k0 < kend4; k0 += 2 * U) {
        if (k0 + U < kend4) fetch(k0 + U, a1, b1);
        step(k0, a0, b0);
        if (k0 + U >= kend4) break;
        if (k0 + 2 * U < kend4) fetch(k0 + 2 * U, a0, b0);
        step(k0 + U, a1, b1);
    }

    if (kslices > 1) {
        V* W = ws + ((size_t)p * kslices + ks) * (size_t)M * N;
#pragma unroll
        for (int c = 0; c < NC; ++c)
            if (c < N) W[(size_t)row + (size_t)c * M] = mk<R>(p1[c] - p2[c], (p3[c] - p1[c]) - p2[c]);
        return;
    }
    V* const Out = (V*)D.Out;
    const V* const Din = (const V*)D.Din;
    const V alpha = mk<R>((R)D.ar, (R)D.ai), beta = mk<R>((R)D.br, (R)D.bi);
    const int orow = D.rmap ? gldi(D.rmap + row) : row;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        if (c >= N) break;
        const int ocol = D.cmap ? gldi(D.cmap + c) : c;
        V val = cmul(alpha, mk<R>(p1[c] - p2[c], (p3[c] - p1[c]) - p2[c]));
        if (Din) {
            const size_t di = D.dmap ? (size_t)orow + (size_t)ocol * D.ldd : (size_t)row + (size_t)c * D.ldd;
            val = cfma(beta, gld(Din + di), val);
        }
        gst(Out + ((size_t)orow + (size_t)ocol * D.ldo), val);
    }
}

// Sum of the K slices in slice order, then the epilogue of gemm_kernel.
// (The slices are summed in precision R, the epilogue runs in S.)
template <typename S, typename R>
__global__ void gemm_reduce_kernel(const GemmDesc* __restrict__ descs, int M, int N, int kslices,
                                   const c_t<R>* __restrict__ ws) {
    using V = c_t<R>;
    using VS = c_t<S>;
    const size_t mn = (size_t)M * N;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int p = blockIdx.y;
    if (e >= mn) return;
    const int row = (int)(e % M), col = (int)(e / M);
    const GemmDesc& D = descs[p];
    const V* W = ws + (size_t)p * kslices * mn + e;
    V sum = W[0];
    for (int k = 1; k < kslices; ++k) sum = cadd(sum, W[(size_t)k * mn]);
    VS val = cmul(mk<S>((S)D.ar, (S)D.ai), mk<S>((S)sum.x, (S)sum.y));
    const int orow = D.rmap ? gldi(D.rmap + row) : row;
    const int ocol = D.cmap ? gldi(D.cmap + col) : col;
    const VS* Din = (const VS*)D.Din;
    if (Din) {
        size_t di = D.dmap ? (size_t)orow + (size_t)ocol * D.ldd : (size_t)row + (size_t)col * D.ldd;
        val = cfma(mk<S>((S)D.br, (S)D.bi), gld(Din + di), val);
    }
    gst((VS*)D.Out + ((size_t)orow + (size_t)ocol * D.ldo), val);
}

}  // namespace

int gemm_kslices(int nprob, int M, int K, int ks_target) {
    // split K only when the row tiles of all problems cannot fill the chip;
    // the choice never depends on N
    const long long rowtiles = (long long)nprob * ((M + BMX - 1) / BMX);
    if (rowtiles >= ks_target / 2 || K < 256) return 1;
    long long ks = (ks_target + rowtiles - 1) / rowtiles;
    if (ks > K / 128) ks = K / 128;
    if (ks > 16) ks = 16;
    return ks < 1 ? 1 : (int)ks;
}

size_t gemm_ws_elems(int nprob, int M, int N, int K, int ks_target) {
    const int ks = gemm_kslices(nprob, M, K, ks_target);
    return ks > 1 ? (size_t)nprob * ks * M * N : 0;
}

// (few columns: the skinny tiles stream the tall products of the top levels
// fastest, the LDS-staged 64 x 16 tiles the many small ones below)
int gemm_tile_rule(int M, int N) {
    if (N <= 16) return M >= 512 ? GEMM_TILE_SKINNY : GEMM_TILE_64x16;
    return M <= 16 ? GEMM_TILE_16x64 : (M <= 32 ? GEMM_TILE_32x64 : GEMM_TILE_64x64);
}

namespace {

// One launch over up to GEMM_MAX_GROUPS groups of problems, all with one tile
// shape; then the K-slice sums of the groups that were split.
template <typename S, typename R>
void launch_gemm_groups(const GemmGroup* grp, int ngroups, hipStream_t st, int tile) {
    GemmLaunch GL = {};
    int tile_M = 0, tile_N = 0;
    for (int q = 0; q < ngroups; ++q)
        if (grp[q].nprob > 0 && grp[q].M > 0 && grp[q].N > 0) {
            tile_M = std::max(tile_M, grp[q].M);
            tile_N = std::max(tile_N, grp[q].N);
        }
    if (!tile_M) return;
    // Every 16x16 output block runs the same MFMA sequence over K whatever the
    // tile, so the tile shape (by rule from M and N, or chosen by the caller's
    // autotuning) never changes the results.
    if (tile < 0 || tile >= GEMM_TILES) tile = gemm_tile_rule(tile_M, tile_N);
    // the skinny kernel runs the products of the mixed-precision merge (S != R) as
    // 64x16, and so does the matrix-core-free kernel those and products of more
    // than GEMM_GEMV_MAX_N columns
    if (tile == GEMM_TILE_SKINNY && !std::is_same<S, R>::value) tile = GEMM_TILE_64x16;
    if (tile == GEMM_TILE_GEMV && (!std::is_same<S, R>::value || tile_N > GEMM_GEMV_MAX_N)) tile = GEMM_TILE_64x16;
    if ((tile == GEMM_TILE_128x64 || tile == GEMM_TILE_64x128) && !std::is_same<R, float>::value) tile = GEMM_TILE_64x64;
    // the direct-to-LDS tiles: complex128 operands (complex128 products or the
    // mixed-precision merge) without the 64 x 128 and 128 x 128 tiles (64 x 64
    // under these choices)
    if ((tile == GEMM_TILE_D128x128 || tile == GEMM_TILE_D64x128) && !std::is_same<S, float>::value)
        tile = GEMM_TILE_D64x64;
    static const int TBM[GEMM_TILES] = {64, 32, 16, 64, 32, 64, 128, 64, 1, 128, 64, 64, 128, 64, 64, 128},
                     TBN[GEMM_TILES] = {64, 64, 64, 16, 32, 32, 16, 16, GEMM_GEMV_MAX_N, 64, 128, 64, 64, 128, 32, 128};
    const int bm = TBM[tile], bn = TBN[tile];
    int wg = 0;
    for (int q = 0; q < ngroups; ++q) {
        const GemmGroup& g = grp[q];
        if (g.nprob <= 0 || g.M <= 0 || g.N <= 0) continue;
        int ks = gemm_kslices(g.nprob, g.M, g.K, g.ks_target);
        if (!g.ws) ks = 1;
        const int kchunk = ((g.K + ks - 1) / ks + KALIGN - 1) / KALIGN * KALIGN;
        ks = (g.K + kchunk - 1) / kchunk;
        GemmLaunch::Grp& e = GL.g[GL.n++];
        e.descs = g.descs; e.ws = g.ws;
        e.M = g.M; e.N = g.N; e.K = g.K;
        e.tilesM = (g.M + bm - 1) / bm; e.tilesN = (g.N + bn - 1) / bn;
        e.ks = ks; e.kchunk = kchunk; e.wg0 = wg; e.nprob = g.nprob;
        if (tile == GEMM_TILE_GEMV)   // one lane per (problem, slice, row)
            wg += (int)(((long long)g.nprob * ks * g.M + 255) / 256);
        else
            wg += g.nprob * e.tilesM * e.tilesN * ks;
    }
    auto run = [&](auto kern, int nt = 256) { hipLaunchKernelGGL(kern, dim3(wg), dim3(nt), 0, st, GL); };
    switch (tile) {
        case GEMM_TILE_SKINNY:
            if constexpr (std::is_same<S, R>::value) run(gemm_skinny_kernel<R, 8>);
            break;
        case GEMM_TILE_GEMV:
            if constexpr (std::is_same<S, R>::value) {
                if (tile_N == 1) run(gemv_kernel<R, 1, 8>);
                else run(gemv_kernel<R, GEMM_GEMV_MAX_N, 4>);
            }
            break;
        case GEMM_TILE_64x16: run(gemm_kernel<S, R, 64, 16>); break;
        case GEMM_TILE_16x64: run(gemm_kernel<S, R, 16, 64>); break;
        case GEMM_TILE_32x64: run(gemm_kernel<S, R, 32, 64>); break;
        case GEMM_TILE_32x32: run(gemm_kernel<S, R, 32, 32>); break;
        case GEMM_TILE_64x32: run(gemm_kernel<S, R, 64, 32>); break;
        case GEMM_TILE_128x16: run(gemm_kernel<S, R, 128, 16>); break;
        case GEMM_TILE_128x64:
            if constexpr (std::is_same<R, float>::value) run(gemm_kernel<S, R, 128, 64>);
            break;
        case GEMM_TILE_64x128:
            if constexpr (std::is_same<R, float>::value) run(gemm_kernel<S, R, 64, 128>);
            break;
        // (stage counts and wave layouts from timings of the solve-pass shapes)
        case GEMM_TILE_D64x64:
            run(gemm_dma_kernel<S, R, 64, 64, std::is_same<S, float>::value ? 3 : 2>);
            break;
        case GEMM_TILE_D64x32:
            run(gemm_dma_kernel<S, R, 64, 32, std::is_same<S, float>::value ? 4 : 2>);
            break;
        case GEMM_TILE_D128x64:   // eight waves of 32 x 32
            run(gemm_dma_kernel<S, R, 128, 64, std::is_same<S, float>::value ? 3 : 2, 512, 2>, 512);
            break;
        case GEMM_TILE_D64x128:   // eight waves of 32 x 32
            if constexpr (std::is_same<S, float>::value) run(gemm_dma_kernel<S, R, 64, 128, 3, 512, 4>, 512);
            break;
        case GEMM_TILE_D128x128:   // eight waves of 64 x 64
            if constexpr (std::is_same<S, float>::value) run(gemm_dma_kernel<S, R, 128, 128, 2, 512, 2>, 512);
            break;
        default: run(gemm_kernel<S, R, 64, 64>); break;
    }
    for (int q = 0; q < GL.n; ++q) {
        const GemmLaunch::Grp& e = GL.g[q];
        if (e.ks <= 1) continue;
        const size_t mn = (size_t)e.M * e.N;
        const int nprob = e.nprob;
        hipLaunchKernelGGL((gemm_reduce_kernel<S, R>), dim3((unsigned)((mn + 255) / 256), nprob), dim3(256), 0, st,
                           e.descs, e.M, e.N, e.ks, (const c_t<R>*)e.ws);
    }
}

template <typename S, typename R>
void launch_gemm_sr(const GemmDesc* d_descs, int nprob, int M, int N, int K, hipStream_t st, void* ws, int tile,
                    int ks_target) {
    GemmGroup g;
    g.descs = d_descs; g.nprob = nprob; g.M = M; g.N = N; g.K = K; g.ws = ws; g.ks_target = ks_target;
    launch_gemm_groups<S, R>(&g, 1, st, tile);
}

}  // namespace

template <typename R>
void launch_gemm_grouped(const GemmGroup* groups, int ngroups, hipStream_t st, int tile) {
    for (int q0 = 0; q0 < ngroups; q0 += GEMM_MAX_GROUPS)
        launch_gemm_groups<R, R>(groups + q0, std::min(GEMM_MAX_GROUPS, ngroups - q0), st, tile);
}
template void launch_gemm_grouped<double>(const GemmGroup*, int, hipStream_t, int);
template void launch_gemm_grouped<float>(const GemmGroup*, int, hipStream_t, int);

template <typename R>
void launch_gemm(const GemmDesc* d_descs, int nprob, int M, int N, int K, hipStream_t st, void* ws, int tile,
                 int ks_target) {
    launch_gemm_sr<R, R>(d_descs, nprob, M, N, K, st, ws, tile, ks_target);
}

void launch_gemm_c64(const GemmDesc* d_descs, int nprob, int M, int N, int K, hipStream_t st, void* ws, int tile,
                     int ks_target) {
    launch_gemm_sr<double, float>(d_descs, nprob, M, N, K, st, ws, tile, ks_target);
}

template void launch_gemm<double>(const GemmDesc*, int, int, int, int, hipStream_t, void*, int, int);
template void launch_gemm<float>(const GemmDesc*, int, int, int, int, hipStream_t, void*, int, int);

}  // namespace rs
