// Split-precision (hi+lo) GEMMs for the dot-product GP kernel on gfx950 MFMA.
//
// A FP64 value x is carried as hi = round(x), lo = round(x - hi) in BF16 (or
// FP16); products use three MFMA passes hi*hi + hi*lo + lo*hi with FP32
// accumulation (v_mfma_f32_16x16x32_bf16 / _f16).  The descriptor
// normalisation, the K^zeta epilogue, the energy sum and every force/virial
// reduction stay FP64.
//
// Operand layout of the 16x16x32 instructions (wave64): lane l holds 8
// consecutive k of row l%16 of A (k = 8*(l/16) .. +7) and 8 consecutive k of
// column l%16 of B; C lane l holds rows 4*(l/16)+v, column l%16.  Both
// operands are therefore read as 16-byte vectors from k-contiguous storage,
// straight from global memory (L2-resident), without LDS staging.
#include "gapx_kernels.h"

#include <algorithm>
#include <atomic>

namespace gapx {

typedef float float4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef short short8_t __attribute__((ext_vector_type(8)));

static inline int cdivm(long a, long b) { return (int)((a + b - 1) / b); }

template <int PREC>
struct Half;
template <>
struct Half<1> {   // bf16
    typedef __bf16 T;
    typedef bf16x8_t V;
    static __device__ __forceinline__ T from(double x) { return (T)(float)x; }
    static __device__ __forceinline__ double to(T x) { return (double)(float)x; }
    static __device__ __forceinline__ float4_t mfma(V a, V b, float4_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};
template <>
struct Half<2> {   // fp16
    typedef _Float16 T;
    typedef f16x8_t V;
    // through float: the hardware converts float to half; the lo part keeps the residual
    static __device__ __forceinline__ T from(double x) { return (T)(float)x; }
    static __device__ __forceinline__ double to(T x) { return (double)(float)x; }
    static __device__ __forceinline__ float4_t mfma(V a, V b, float4_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};

// split rows x cols FP64 (row-major, ld src_ld) into hi/lo with leading dimension
// ld_out; transpose=true writes the transpose (cols x rows).  Padding is zero.
template <int PREC>
__global__ void k_split(int rows, int cols, int ld_out, int rows_pad, const double *src, int src_ld, bool transpose,
                        typename Half<PREC>::T *hi, typename Half<PREC>::T *lo, double scale) {
    typedef Half<PREC> H;
    const size_t tot = transpose ? (size_t)cols * ld_out : (size_t)rows_pad * ld_out;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < tot; idx += (size_t)gridDim.x * blockDim.x) {
        int r, c;
        if (!transpose) { r = (int)(idx / ld_out); c = (int)(idx % ld_out); }
        else { c = (int)(idx / ld_out); r = (int)(idx % ld_out); }
        double x = (r < rows && c < cols) ? src[(size_t)r * src_ld + c] * scale : 0.0;
        typename H::T h = H::from(x);
        typename H::T l = H::from(x - H::to(h));
        hi[idx] = h;
        lo[idx] = l;
    }
}

void launch_split_matrix(int prec, int rows, int cols, int ld_out, const double *src, int src_ld, bool transpose,
                         void *hi, void *lo, hipStream_t s) {
    const int rows_pad = (rows + 63) / 64 * 64;
    const size_t tot = transpose ? (size_t)cols * ld_out : (size_t)rows_pad * ld_out;
    const int grid = (int)std::min<size_t>((tot + 255) / 256, 65536);
    if (prec == 1)
        hipLaunchKernelGGL(k_split<1>, dim3(grid), dim3(256), 0, s, rows, cols, ld_out, rows_pad, src, src_ld, transpose,
                           (__bf16 *)hi, (__bf16 *)lo, 1.0);
    else
        hipLaunchKernelGGL(k_split<2>, dim3(grid), dim3(256), 0, s, rows, cols, ld_out, rows_pad, src, src_ld, transpose,
                           (_Float16 *)hi, (_Float16 *)lo, 1.0);
}

// C[row, col] = sum_k A[row, k] B[col, k] with A, B hi/lo split, k-contiguous,
// both with leading dimension ldk (a multiple of 32, zero padded).
// Block = 4 wavefronts, 64x64 output tile, each wave 32x32 (2x2 MFMA tiles).
template <int PREC>
__device__ __forceinline__ void mma_tile(const typename Half<PREC>::T *Ahi, const typename Half<PREC>::T *Alo,
                                         const typename Half<PREC>::T *Bhi, const typename Half<PREC>::T *Blo, int ldk,
                                         int arow0, int amax, int bcol0, int bmax, int kdim, float4_t acc[2][2]) {
    typedef Half<PREC> H;
    typedef typename H::V V;
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, kg = (lane >> 4) * 8;
    for (int k0 = 0; k0 < kdim; k0 += 32) {
        V ah[2], al[2], bh[2], bl[2];
#pragma unroll
        for (int i = 0; i < 2; i++) {
            int row = arow0 + i * 16 + r;
            if (row < amax) {
                ah[i] = *(const V *)(Ahi + (size_t)row * ldk + k0 + kg);
                al[i] = *(const V *)(Alo + (size_t)row * ldk + k0 + kg);
            } else {
                ah[i] = (V){};
                al[i] = (V){};
            }
            int col = bcol0 + i * 16 + r;
            if (col < bmax) {
                bh[i] = *(const V *)(Bhi + (size_t)col * ldk + k0 + kg);
                bl[i] = *(const V *)(Blo + (size_t)col * ldk + k0 + kg);
            } else {
                bh[i] = (V){};
                bl[i] = (V){};
            }
        }
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) {
                acc[i][j] = H::mfma(al[i], bh[j], acc[i][j]);
                acc[i][j] = H::mfma(ah[i], bl[j], acc[i][j]);
                acc[i][j] = H::mfma(ah[i], bh[j], acc[i][j]);
            }
    }
}

// GEMM1: K = X Xs^T; epilogue w K^zeta -> e_part (FP64), W = w zeta K^(zeta-1)
// scaled by wscale, split into hi/lo (nc x Mpad) for GEMM2.
template <int PREC>
__global__ __launch_bounds__(256) void k_gemm1_split(int nc, int M, int dpad, int Mpad,
                                                      const typename Half<PREC>::T *Xhi, const typename Half<PREC>::T *Xlo,
                                                      const typename Half<PREC>::T *Shi, const typename Half<PREC>::T *Slo,
                                                      const double *wvec, double zeta, int izeta, double wscale,
                                                      typename Half<PREC>::T *Whi, typename Half<PREC>::T *Wlo,
                                                      double *e_part) {
    typedef Half<PREC> H;
    __shared__ double red[2][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, wm = w >> 1, wn = w & 1;
    const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
    float4_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = (float4_t){0.f, 0.f, 0.f, 0.f};
    mma_tile<PREC>(Xhi, Xlo, Shi, Slo, dpad, m0 + wm * 32, nc, n0 + wn * 32, M, dpad, acc);
    double rowsum[2][4];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int v = 0; v < 4; v++) rowsum[i][v] = 0.0;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int col = n0 + wn * 32 + j * 16 + (lane & 15);
            const double wv = col < M ? wvec[col] : 0.0;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + v;
                const double kv = (double)acc[i][j][v];
                double kz1, kz;
                if (izeta > 0) {
                    kz1 = 1.0;
                    for (int t = 1; t < izeta; t++) kz1 *= kv;
                    kz = kz1 * kv;
                } else {
                    kz1 = pow(kv, zeta - 1.0);
                    kz = pow(kv, zeta);
                }
                rowsum[i][v] += wv * kz;
                if (row < nc && col < Mpad) {
                    const double wk = col < M ? wv * zeta * kz1 * wscale : 0.0;
                    typename H::T h = H::from(wk);
                    Whi[(size_t)row * Mpad + col] = h;
                    Wlo[(size_t)row * Mpad + col] = H::from(wk - H::to(h));
                }
            }
        }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            double s = rowsum[i][v];
            for (int off = 8; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
            rowsum[i][v] = s;
        }
    if ((lane & 15) == 0)
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int v = 0; v < 4; v++) red[wn][wm * 32 + i * 16 + 4 * (lane >> 4) + v] = rowsum[i][v];
    __syncthreads();
    if (threadIdx.x < 64) {
        const int row = m0 + threadIdx.x;
        if (row < nc) e_part[(size_t)row * gridDim.y + blockIdx.y] = red[0][threadIdx.x] + red[1][threadIdx.x];
    }
}

// GEMM2: G = W Xs = sum_s W[i,s] XsT[q,s]; output FP64 (nc x d), unscaled.
template <int PREC>
__global__ __launch_bounds__(256) void k_gemm2_split(int nc, int d, int Mpad, const typename Half<PREC>::T *Whi,
                                                      const typename Half<PREC>::T *Wlo,
                                                      const typename Half<PREC>::T *Thi,
                                                      const typename Half<PREC>::T *Tlo, double inv_wscale, double *G) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, wm = w >> 1, wn = w & 1;
    const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
    float4_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = (float4_t){0.f, 0.f, 0.f, 0.f};
    mma_tile<PREC>(Whi, Wlo, Thi, Tlo, Mpad, m0 + wm * 32, nc, n0 + wn * 32, d, Mpad, acc);
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int col = n0 + wn * 32 + j * 16 + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + v;
                if (row < nc && col < d) G[(size_t)row * d + col] = (double)acc[i][j][v] * inv_wscale;
            }
        }
}

// ---------------------------------------------------------------------------
// LDS-tiled split GEMMs for large problems (descriptors beyond the fused
// kernel's range, many centres).
//
// Block tile 128 x 128 of C[row, col] = sum_k A[row, k] B[col, k] (A, B hi/lo,
// k-contiguous, leading dimension ldk, a multiple of 32).  256 threads = 2 x 2
// wavefronts of 64 x 64 (4 x 4 MFMA tiles).  Per k step of 32 the block's
// 128 x 32 slices of A and B (hi and lo) are written to LDS from registers
// fetched during the previous step, so every 16-byte fragment read from LDS
// feeds 12 MFMAs.  The three passes per k step run in the order of mma_tile
// (lo*hi, hi*lo, hi*hi), so an element's FP32 result is bitwise the same as
// from the 64 x 64 kernels: a row's result does not depend on which kernel the
// problem size selects.  Tiles are renumbered so that each XCD (blocks are
// dealt to the 8 XCDs round robin) works on a contiguous range of tiles that
// share their A rows in its L2.
// ---------------------------------------------------------------------------
constexpr int kLT = 128, kLLD = 40;   // tile edge; LDS row stride in elements (16-byte reads conflict-free)

__device__ __forceinline__ void lds_tile_coords(int ntn, int &tm, int &tn) {
    const int b = blockIdx.x, nb = gridDim.x;
    const int tile = (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
    tm = tile / ntn;
    tn = tile - tm * ntn;
}

template <int PREC>
__device__ __forceinline__ void lds_gemm_main(const typename Half<PREC>::T *__restrict__ Ahi,
                                              const typename Half<PREC>::T *__restrict__ Alo,
                                              const typename Half<PREC>::T *__restrict__ Bhi,
                                              const typename Half<PREC>::T *__restrict__ Blo, int ldk, int m0,
                                              int amax, int n0, int bmax, typename Half<PREC>::T *As,
                                              typename Half<PREC>::T *Bs, float4_t acc[4][4]) {
    typedef Half<PREC> H;
    typedef typename H::V V;
    constexpr int TS = kLT * kLLD;   // elements of one staged slice
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, wm = w >> 1, wn = w & 1;
    const int r16 = lane & 15, kg = (lane >> 4) * 8;
    // thread t stages 16-byte vectors t and t + 256 of each 128 x 32 slice (row = idx / 4)
    V pa[2][2], pb[2][2];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int idx = t + 256 * u, row = idx >> 2, c8 = (idx & 3) * 8;
            const int ar = m0 + row, br = n0 + row;
            pa[u][0] = pa[u][1] = pb[u][0] = pb[u][1] = (V){};
            if (ar < amax) {
                pa[u][0] = *(const V *)(Ahi + (size_t)ar * ldk + k0 + c8);
                pa[u][1] = *(const V *)(Alo + (size_t)ar * ldk + k0 + c8);
            }
            if (br < bmax) {
                pb[u][0] = *(const V *)(Bhi + (size_t)br * ldk + k0 + c8);
                pb[u][1] = *(const V *)(Blo + (size_t)br * ldk + k0 + c8);
            }
        }
    };
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = (float4_t){0.f, 0.f, 0.f, 0.f};
    fetch(0);
    for (int k0 = 0; k0 < ldk; k0 += 32) {
        __syncthreads();   // the previous step's LDS reads are done
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int idx = t + 256 * u, o = (idx >> 2) * kLLD + (idx & 3) * 8;
            *(V *)(As + o) = pa[u][0];
            *(V *)(As + TS + o) = pa[u][1];
            *(V *)(Bs + o) = pb[u][0];
            *(V *)(Bs + TS + o) = pb[u][1];
        }
        __syncthreads();
        if (k0 + 32 < ldk) fetch(k0 + 32);
        V ah[4], al[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int o = (wm * 64 + 16 * i + r16) * kLLD + kg;
            ah[i] = *(const V *)(As + o);
            al[i] = *(const V *)(As + TS + o);
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int o = (wn * 64 + 16 * j + r16) * kLLD + kg;
            const V bh = *(const V *)(Bs + o), bl = *(const V *)(Bs + TS + o);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                acc[i][j] = H::mfma(al[i], bh, acc[i][j]);
                acc[i][j] = H::mfma(ah[i], bl, acc[i][j]);
                acc[i][j] = H::mfma(ah[i], bh, acc[i][j]);
            }
        }
    }
}

// GEMM1 on 128 x 128 tiles: K = X Xs^T with the epilogue of k_gemm1_split
// (e_part has one column per tile column, ntn in all).  INTZ: zeta is a small
// integer (powers by multiplication); otherwise pow().
template <int PREC, bool INTZ>
__global__ __launch_bounds__(256) void k_gemm1_split_lds(int nc, int M, int dpad, int Mpad, int ntn,
                                                          const typename Half<PREC>::T *Xhi,
                                                          const typename Half<PREC>::T *Xlo,
                                                          const typename Half<PREC>::T *Shi,
                                                          const typename Half<PREC>::T *Slo, const double *wvec,
                                                          double zeta, int izeta, double wscale,
                                                          typename Half<PREC>::T *Whi, typename Half<PREC>::T *Wlo,
                                                          double *e_part) {
    typedef Half<PREC> H;
    __shared__ __attribute__((aligned(16))) typename H::T As[2 * kLT * kLLD];
    __shared__ __attribute__((aligned(16))) typename H::T Bs[2 * kLT * kLLD];
    __shared__ double red[2][kLT];
    int tm, tn;
    lds_tile_coords(ntn, tm, tn);
    const int m0 = tm * kLT, n0 = tn * kLT;
    float4_t acc[4][4];
    lds_gemm_main<PREC>(Xhi, Xlo, Shi, Slo, dpad, m0, nc, n0, M, As, Bs, acc);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, wm = w >> 1, wn = w & 1;
    double wv[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int col = n0 + wn * 64 + j * 16 + (lane & 15);
        wv[j] = col < M ? wvec[col] : 0.0;
    }
    // one row tile at a time keeps the FP64 row sums of four rows live
#pragma unroll
    for (int i = 0; i < 4; i++) {
        double rowsum[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int col = n0 + wn * 64 + j * 16 + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = m0 + wm * 64 + i * 16 + 4 * (lane >> 4) + v;
                const double kv = (double)acc[i][j][v];
                double kz1, kz;
                if constexpr (INTZ) {
                    kz1 = 1.0;
                    for (int q = 1; q < izeta; q++) kz1 *= kv;
                    kz = kz1 * kv;
                } else {
                    kz1 = pow(kv, zeta - 1.0);
                    kz = pow(kv, zeta);
                }
                rowsum[v] += wv[j] * kz;
                if (row < nc && col < Mpad) {
                    const double wk = col < M ? wv[j] * zeta * kz1 * wscale : 0.0;
                    typename H::T h = H::from(wk);
                    Whi[(size_t)row * Mpad + col] = h;
                    Wlo[(size_t)row * Mpad + col] = H::from(wk - H::to(h));
                }
            }
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            double x = rowsum[v];
            for (int off = 8; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
            if ((lane & 15) == 0) red[wn][wm * 64 + i * 16 + 4 * (lane >> 4) + v] = x;
        }
    }
    __syncthreads();
    if (threadIdx.x < kLT) {
        const int row = m0 + threadIdx.x;
        if (row < nc) e_part[(size_t)row * ntn + tn] = red[0][threadIdx.x] + red[1][threadIdx.x];
    }
}

// GEMM2 on 128 x 128 tiles: G = W Xs (FP64 output, unscaled)
template <int PREC>
__global__ __launch_bounds__(256) void k_gemm2_split_lds(int nc, int d, int Mpad, int ntn,
                                                          const typename Half<PREC>::T *Whi,
                                                          const typename Half<PREC>::T *Wlo,
                                                          const typename Half<PREC>::T *Thi,
                                                          const typename Half<PREC>::T *Tlo, double inv_wscale,
                                                          double *G) {
    typedef Half<PREC> H;
    __shared__ __attribute__((aligned(16))) typename H::T As[2 * kLT * kLLD];
    __shared__ __attribute__((aligned(16))) typename H::T Bs[2 * kLT * kLLD];
    int tm, tn;
    lds_tile_coords(ntn, tm, tn);
    const int m0 = tm * kLT, n0 = tn * kLT;
    float4_t acc[4][4];
    lds_gemm_main<PREC>(Whi, Wlo, Thi, Tlo, Mpad, m0, nc, n0, d, As, Bs, acc);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, wm = w >> 1, wn = w & 1;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int col = n0 + wn * 64 + j * 16 + (lane & 15);
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = m0 + wm * 64 + i * 16 + 4 * (lane >> 4) + v;
                if (row < nc && col < d) G[(size_t)row * d + col] = (double)acc[i][j][v] * inv_wscale;
            }
    }
}

// 128 x 128 tiles when they alone fill the GPU (two per CU); 64 x 64 tiles otherwise
// (lds_mode: -1 this rule, 0 the 64 x 64 tiles, 1 the 128 x 128 tiles)
static bool use_lds_tiles(int rows, int cols, int lds_mode = -1) {
    if (lds_mode >= 0) return lds_mode == 1;
    return (long)cdivm(rows, kLT) * cdivm(cols, kLT) >= 512;
}


// LDS rows of the fused split kernel: no padding; 16-byte vectors XOR-swizzled, vector c
// of row r at c ^ (r & XSM), so that the rows one fragment read touches fall on distinct
// 16-byte bank groups in every group of 8 lanes (a 16-byte read moves 128 bytes per
// cycle); the transposing 8-byte reads of four rows x four column words per lane group
// fall on distinct groups too.
//
// One k step of the fused split kernel's products is its own scheduling region: the
// fragment reads of the next step first (NR of them), then the six MFMAs of this one,
// so that the reads are in flight under the MFMAs.  Left to itself the compiler issues
// each step's reads right before its MFMAs and waits for them there (both waves of a
// SIMD leave each barrier together, so neither covers the other's wait).
// Pinned for the zeta = 4 kernels of descriptors up to 256 components, where it was
// measured: the FP64 epilogue of other zeta needs the registers (with the reads held
// ahead it spills), and at 64 components FP16 would lose a wavefront per SIMD.
#define GP_SCHED_STEP(NR)                                         \
    do {                                                          \
        if constexpr (ZI == 4 && DK == 8) {                       \
            __builtin_amdgcn_sched_group_barrier(0x100, NR, 0);   \
            __builtin_amdgcn_sched_group_barrier(0x008, 6, 0);    \
            __builtin_amdgcn_sched_barrier(0);                    \
        }                                                         \
    } while (0)

template <int DK>
struct FusedLds {
    static constexpr int DP = 32 * DK, XSV = DP / 8, XSM = XSV >= 16 ? 15 : XSV - 1;
    static __device__ __forceinline__ int xs_at(int r, int c) { return r * DP + ((c ^ (r & XSM)) << 3); }
};

// Where this lane's 16-byte vectors of a staged pair of chunks (64 Xs rows) come from:
// byte offsets from the pair's first row, the same for every pair.  Wave instruction u
// fills LDS vectors 64u .. 64u+63 of a matrix; lane l loads the global vector that the
// swizzle puts in slot 64u + l.  Instructions u < XSV/2 fill the first chunk's 32 rows.
// Columns past the padded width ldS need no zeros: they meet X components that are zero
// in registers (they take columns ldS - 8 .. ldS - 1 of the same row).
template <int DK>
struct StageOff {
    typedef FusedLds<DK> L;
    static constexpr int XS_U = L::XSV;              // wave instructions per matrix (64 rows)
    static constexpr int NXS = (XS_U + 7) / 8;       // per wave
    unsigned xs[NXS];
};

template <typename T, int DK>
__device__ __forceinline__ StageOff<DK> stage_offsets(int ldS, int wv) {
    typedef FusedLds<DK> L;
    typedef StageOff<DK> O;
    O so;
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < O::NXS; i++) {
        const int vi = 64 * (8 * i + wv) + lane, r = vi / L::XSV, c = (vi % L::XSV) ^ (r & L::XSM);
        so.xs[i] = (unsigned)((r * ldS + min(8 * c, ldS - 8)) * (int)sizeof(T));
    }
    return so;
}

// A raw buffer resource over a whole matrix (stride 0, 32-bit data format word of the
// gfx9 descriptor); its loads are bounded by nbytes.  The chunk loads go through buffer
// instructions rather than global_load_lds: those are FLAT instructions, which may count
// on lgkmcnt too, and while one is outstanding every wait for an LDS read has to drain
// the counter (lgkmcnt(0)); buffer loads count on vmcnt only, so the fragment reads of
// the next k step stay in flight (lgkmcnt(4)..(7)) under the current step's MFMAs.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t mat_rsrc(const void *p, unsigned nbytes) {
    return __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)nbytes, 0x00020000);
}
// 16 bytes per lane from the buffer at voffset + soffset straight into LDS: lane l's
// bytes land at lds + 16 l (lds uniform over the wavefront)
__device__ __forceinline__ void dma16(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, void *lds) {
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void *)lds, 16, (int)voff,
                                             (int)soff, 0, 0);
}

// Stage nch (1 or 2) chunks from sparse point s0 into one LDS buffer: the Xs rows go from
// global memory straight to LDS (buffer loads with the lds flag, 16 bytes per lane: the
// lane's offset plus a wave-uniform offset).  Their weights w_s pass through a register
// (*pw; 0 past M and past the staged chunks).
template <typename T, int DK>
__device__ __forceinline__ void fused_stage(int s0, int nch, int M, int ldS, __amdgpu_buffer_rsrc_t rsh,
                                            __amdgpu_buffer_rsrc_t rsl, const double *wvec,
                                            const StageOff<DK> &so, int wv, T *__restrict__ xs_h,
                                            T *__restrict__ xs_lo, double *pw) {
    typedef StageOff<DK> O;
    const int t = threadIdx.x;
    if (t < 64) *pw = (t < 32 * nch && s0 + t < M) ? wvec[s0 + t] : 0.0;
    const unsigned ss = (unsigned)(s0 * ldS * (int)sizeof(T));
#pragma unroll
    for (int i = 0; i < O::NXS; i++) {
        const int u = 8 * i + wv;
        if ((O::XS_U % 8 == 0 || u < O::XS_U) && (nch == 2 || u < O::XS_U / 2)) {
            dma16(rsh, so.xs[i], ss, xs_h + 512 * u);
            dma16(rsl, so.xs[i], ss, xs_lo + 512 * u);
        }
    }
}

typedef short short4_t __attribute__((ext_vector_type(4)));

// Four 16-bit elements through the transposing LDS read: within each group of 16 lanes,
// lane r receives element r % 4 of the 8-byte words that lanes r/4, 4 + r/4, 8 + r/4 and
// 12 + r/4 of its group address.  So when lane s of the group addresses row R + s/4 at
// column C + 4 (s % 4) of a row-major matrix, lane r gets rows R .. R+3 of column C + r.
__device__ __forceinline__ short4_t lds_read_tr4(const void *p) {
    return __builtin_amdgcn_ds_read_tr16_b64_v4i16(
        (__attribute__((address_space(3))) short4_t *)((__attribute__((address_space(3))) char *)p));
}

// One chunk of 32 sparse points (rows 32h .. 32h+31 of the current buffer).
template <int PREC, int DK, int ZI>
__device__ __forceinline__ void fused_sub(int h, const typename Half<PREC>::T *__restrict__ xs_h,
                                          const typename Half<PREC>::T *__restrict__ xs_lo,
                                          const double *__restrict__ wcur, const float *__restrict__ wfcur,
                                          double zeta, int izeta, double wscale,
                                          const typename Half<PREC>::V (&xh)[DK],
                                          const typename Half<PREC>::V (&xl)[DK], float4_t (&gacc)[2 * DK],
                                          double &rs) {
    typedef Half<PREC> H;
    typedef typename H::T T;
    typedef typename H::V V;
    typedef FusedLds<DK> L;
    const int lane = threadIdx.x & 63;
    const int r16 = lane & 15, g4 = lane >> 4, kg = g4 * 8, rq = 4 * g4;
    const int r0 = 32 * h;
    // K^T tiles: 32 sparse points (two tiles) x 16 centres; the fragments of k step
    // ks + 1 are read before the MFMAs of step ks.
    // Wave priority is raised over the two MFMA phases and dropped over the epilogue,
    // so that of the two waves of a SIMD one runs its MFMAs while the other runs its
    // epilogue's VALU work.
    if constexpr (ZI == 4 && DK == 8) __builtin_amdgcn_s_setprio(1);
    float4_t kacc[2] = {(float4_t){0.f, 0.f, 0.f, 0.f}, (float4_t){0.f, 0.f, 0.f, 0.f}};
    V ah[2], al[2];
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int o = L::xs_at(r0 + 16 * j + r16, kg >> 3);
        ah[j] = *(const V *)(xs_h + o);
        al[j] = *(const V *)(xs_lo + o);
    }
#pragma unroll
    for (int ks = 0; ks < DK; ks++) {
        V nh[2], nl[2];
        if (ks + 1 < DK) {
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const int o = L::xs_at(r0 + 16 * j + r16, 4 * (ks + 1) + (kg >> 3));
                nh[j] = *(const V *)(xs_h + o);
                nl[j] = *(const V *)(xs_lo + o);
            }
        }
#pragma unroll
        for (int j = 0; j < 2; j++) {
            kacc[j] = H::mfma(ah[j], xl[ks], kacc[j]);
            kacc[j] = H::mfma(al[j], xh[ks], kacc[j]);
            kacc[j] = H::mfma(ah[j], xh[ks], kacc[j]);
        }
        if (ks + 1 < DK) {
#pragma unroll
            for (int j = 0; j < 2; j++) {
                ah[j] = nh[j];
                al[j] = nl[j];
            }
        }
        GP_SCHED_STEP(4);
    }
    // B fragment of column tile q of G += W Xs, from the Xs rows: elements 0-3 are sparse
    // points 4 (lane/16) .. +3 and elements 4-7 points 16 + 4 (lane/16) .. +3 of this chunk
    // (the K^T accumulator order), each four by one transposing read: lane s of a group
    // addresses row (group base) + s/4 at column 16q + 4 (s % 4), i.e. 16-byte vector
    // 2q + (s % 4)/2 of the row, half (s % 4) % 2.
    const int tr_c = (r16 & 3) >> 1, tr_e = (r16 & 1) * 4, tr_r = r0 + 4 * g4 + (r16 >> 2);
    auto frag = [&](const T *base, int q) -> V {
        const short4_t x = lds_read_tr4(base + L::xs_at(tr_r, 2 * q + tr_c) + tr_e);
        const short4_t y = lds_read_tr4(base + L::xs_at(tr_r + 16, 2 * q + tr_c) + tr_e);
        return __builtin_bit_cast(V, __builtin_shufflevector(x, y, 0, 1, 2, 3, 4, 5, 6, 7));
    };
    // the first two column tiles, read before the FP32 epilogue (after the FP64 one,
    // whose powers need the registers)
    V bh[2], bl[2];
    auto read_b0 = [&]() {
#pragma unroll
        for (int i = 0; i < 2; i++) {
            bh[i] = frag(xs_h, i);
            bl[i] = frag(xs_lo, i);
        }
    };
    if constexpr (ZI == 4) read_b0();
    if constexpr (ZI == 4 && DK == 8) __builtin_amdgcn_s_setprio(0);
    // epilogue: element v of tile j is centre r16 at sparse point 16 j + 4 (lane/16) + v of
    // the chunk, and element 4 j + v of this lane's A fragment of W.  The energy sum is
    // FP64; K^zeta and W are formed from the FP32 K in FP32 for the common zeta = 4 (W is
    // then split into 16-bit halves), in FP64 otherwise.
    V wh, wl;
#pragma unroll
    for (int j = 0; j < 2; j++) {
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const double wsv = wcur[r0 + 16 * j + rq + v];
            T hv, lo;
            if constexpr (ZI == 4) {
                const float kf = kacc[j][v];
                const float k2 = kf * kf;
                rs += wsv * (double)(k2 * k2);
                const float wk = wfcur[r0 + 16 * j + rq + v] * (k2 * kf);   // (float)(w_s 4 wscale), staged
                hv = (T)wk;
                lo = (T)(wk - (float)hv);
            } else {
                const double kv = (double)kacc[j][v];
                double kz1, kz;
                if (izeta > 0) {
                    kz1 = 1.0;
                    for (int q = 1; q < izeta; q++) kz1 *= kv;
                    kz = kz1 * kv;
                } else {
                    kz1 = pow(kv, zeta - 1.0);
                    kz = pow(kv, zeta);
                }
                rs += wsv * kz;
                const double wk = wsv * zeta * kz1 * wscale;
                hv = H::from(wk);
                lo = H::from(wk - H::to(hv));
            }
            wh[4 * j + v] = hv;
            wl[4 * j + v] = lo;
        }
    }
    if constexpr (ZI != 4) read_b0();
    if constexpr (ZI == 4 && DK == 8) __builtin_amdgcn_s_setprio(1);
    // column tiles in pairs, the next pair's fragments read before this pair's MFMAs
#pragma unroll
    for (int q0 = 0; q0 < 2 * DK; q0 += 2) {
        V nh[2], nl[2];
        if (q0 + 2 < 2 * DK) {
#pragma unroll
            for (int i = 0; i < 2; i++) {
                nh[i] = frag(xs_h, q0 + 2 + i);
                nl[i] = frag(xs_lo, q0 + 2 + i);
            }
        }
#pragma unroll
        for (int i = 0; i < 2; i++) {
            gacc[q0 + i] = H::mfma(wl, bh[i], gacc[q0 + i]);
            gacc[q0 + i] = H::mfma(wh, bl[i], gacc[q0 + i]);
            gacc[q0 + i] = H::mfma(wh, bh[i], gacc[q0 + i]);
        }
        if (q0 + 2 < 2 * DK) {
#pragma unroll
            for (int i = 0; i < 2; i++) {
                bh[i] = nh[i];
                bl[i] = nl[i];
            }
        }
        GP_SCHED_STEP(8);
    }
    if constexpr (ZI == 4 && DK == 8) __builtin_amdgcn_s_setprio(0);
}

// One pair of chunks on the current buffer (read) while the next pair streams into the
// other buffer (written by fused_stage's loads).  Distinct restrict-qualified buffers
// tell the compiler that the reads here do not wait for those loads.
template <int PREC, int DK, int ZI>
__device__ __forceinline__ void fused_pair(int ncur, int nnext, int s0n, int M, int ldS, const StageOff<DK> &so,
                                           int wv, __amdgpu_buffer_rsrc_t rsh, __amdgpu_buffer_rsrc_t rsl,
                                           const double *wvec, double zeta, int izeta, double wscale,
                                           const typename Half<PREC>::T *__restrict__ xs_h,
                                           const typename Half<PREC>::T *__restrict__ xs_lo,
                                           const double *__restrict__ wcur, const float *__restrict__ wfcur,
                                           typename Half<PREC>::T *__restrict__ nxs_h,
                                           typename Half<PREC>::T *__restrict__ nxs_lo, double *__restrict__ wnext,
                                           float *__restrict__ wfnext, const typename Half<PREC>::V (&xh)[DK],
                                           const typename Half<PREC>::V (&xl)[DK], float4_t (&gacc)[2 * DK],
                                           double &rs) {
    typedef typename Half<PREC>::T T;
    const int t = threadIdx.x;
    double pw = 0.0;
    if (nnext > 0) fused_stage<T, DK>(s0n, nnext, M, ldS, rsh, rsl, wvec, so, wv, nxs_h, nxs_lo, &pw);
    fused_sub<PREC, DK, ZI>(0, xs_h, xs_lo, wcur, wfcur, zeta, izeta, wscale, xh, xl, gacc, rs);
    if (ncur > 1) fused_sub<PREC, DK, ZI>(1, xs_h, xs_lo, wcur, wfcur, zeta, izeta, wscale, xh, xl, gacc, rs);
    // the next buffers were last read before the previous barrier
    if (nnext > 0 && t < 64) {
        wnext[t] = pw;
        wfnext[t] = (float)(pw * 4.0 * wscale);
    }
    __builtin_amdgcn_s_waitcnt(0);   // this wave's loads into the next buffers have landed
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Fused GP kernel for descriptors of up to 32*DK components (split forms).
//
// One block = 8 wavefronts = 128 centres; wave w owns centres 16w..16w+15 and
// keeps their X rows (split hi/lo) in registers.  The block walks the sparse points
// of its part in chunks of 32, staged two at a time: per pair the chunks' Xs rows (k =
// descriptor index) are staged in LDS.  Each wave computes the transposed tile K^T =
// Xs X^T of a chunk (32 sparse points x 16 centres; three MFMA passes, FP32
// accumulation): in the accumulator layout lane l holds centre l%16 at sparse points
// 16 j + 4 (l/16) + v (tile j, element v), which is one 8-element A fragment of the
// next product when its k runs over the sparse points in the order
// [0-3, 16-19 | 4-7, 20-23 | 8-11, 24-27 | 12-15, 28-31].  So the epilogue (energy
// w K^zeta summed in FP64, W = w zeta K^(zeta-1) split into hi/lo) leaves W in
// registers ready for G += W Xs (16 x 32DK, in registers): W never leaves the lane.
// That product's B fragments (eight sparse points of one descriptor column, in the
// same order) come from the same Xs rows through transposing LDS reads, so no
// transposed copy of Xs is staged.  The next pair streams from global memory into the
// other of two LDS buffers while the current one is used, one barrier per pair.
// With several parts (gridDim.y) the sparse points are divided between blocks and the
// FP32 partial G tiles are summed in a fixed order afterwards.
//
// Stream-K schedule (sk_units > 0): the work is sk_units = (row blocks) x (chunks)
// units in row-block-major order, and block b of the grid (one or a few per CU) takes
// the contiguous units [b U / B, (b+1) U / B): every block gets the same number of
// chunks to within one, so the last round of blocks is never a partial one.  A block's
// range covers whole row blocks in its middle and possibly part of one at either end;
// a whole row block writes G and its energies directly, a part writes FP32 partials
// to one of the block's two slots (its first row block slot 0, its last slot 1), which
// k_gp_sk_finish sums in block order.
// ---------------------------------------------------------------------------
template <int PREC, int DK, int ZI>
__global__ __launch_bounds__(512) void k_gp_fused_split(int nc, int M, int d, int Mpad, int ldS, int chunks_per_part,
                                                        long sk_units, const double *__restrict__ X,
                                                        const typename Half<PREC>::T *__restrict__ Shi,
                                                        const typename Half<PREC>::T *__restrict__ Slo,
                                                        const double *__restrict__ wvec, double zeta, int izeta,
                                                        double wscale, double inv_wscale, double *__restrict__ G,
                                                        float *__restrict__ Gpart, double *__restrict__ e_part,
                                                        double *__restrict__ e_slot) {
    typedef Half<PREC> H;
    typedef typename H::T T;
    typedef typename H::V V;
    constexpr int DP = 32 * DK;     // descriptor components handled (zero padded)
    __shared__ __attribute__((aligned(16))) T xs0[2][64 * DP], xs1[2][64 * DP];   // [hi, lo] of two buffers
    __shared__ double wv0[64], wv1[64];   // the pair's w_s (0 past M)
    __shared__ float wf0[64], wf1[64];    // and (float)(w_s 4 wscale), the zeta = 4 epilogue's factor
    const int t = threadIdx.x, lane = t & 63, wv = __builtin_amdgcn_readfirstlane(t >> 6);
    const int r16 = lane & 15, g4 = lane >> 4, kg = g4 * 8, rq = 4 * g4;
    const int part = blockIdx.y, nparts = gridDim.y;
    const int C = Mpad / 32;
    const bool sk = sk_units > 0;
    const StageOff<DK> so = stage_offsets<T, DK>(ldS, wv);
    const unsigned nbytes = (unsigned)((size_t)Mpad * ldS * sizeof(T));   // within the copy's allocation
    const __amdgpu_buffer_rsrc_t rsh = mat_rsrc(Shi, nbytes), rsl = mat_rsrc(Slo, nbytes);
    long u, u_end;   // this block's units (row block rb, chunk c) = rb C + c
    if (sk) {
        u = (long)blockIdx.x * sk_units / gridDim.x;
        u_end = (long)(blockIdx.x + 1) * sk_units / gridDim.x;
    } else {
        u = (long)blockIdx.x * C + min(C, part * chunks_per_part);
        u_end = (long)blockIdx.x * C + min(C, (part + 1) * chunks_per_part);
    }
    int seg = 0;
    do {
    const int rb = sk ? (int)(u / C) : (int)blockIdx.x;
    const int c_beg = (int)(u - (long)rb * C), c_end = (int)min((long)C, u_end - (long)rb * C);
    const int row0 = rb * 128 + wv * 16;

    V xh[DK], xl[DK];   // B operand of K^T: X[row0 + r16][32 ks + kg .. + 7]
    {
        const int row = row0 + r16;
#pragma unroll
        for (int ks = 0; ks < DK; ks++) {
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int k = 32 * ks + kg + e;
                const double x = (row < nc && k < d) ? X[(size_t)row * d + k] : 0.0;
                const T h = H::from(x);
                xh[ks][e] = h;
                xl[ks][e] = H::from(x - H::to(h));
            }
        }
    }
    float4_t gacc[2 * DK];
#pragma unroll
    for (int q = 0; q < 2 * DK; q++) gacc[q] = (float4_t){0.f, 0.f, 0.f, 0.f};
    double rs = 0.0;   // energy of centre row0 + r16 over this lane's sparse points

    if (c_beg < c_end) {
        double pw = 0.0;
        fused_stage<T, DK>(32 * c_beg, min(2, c_end - c_beg), M, ldS, rsh, rsl, wvec, so, wv, xs0[0], xs0[1], &pw);
        if (t < 64) {
            wv0[t] = pw;
            wf0[t] = (float)(pw * 4.0 * wscale);
        }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    // pairs of chunks: [ch, ch + 2) on buffer 0, then [ch + 2, ch + 4) on buffer 1
    for (int ch = c_beg; ch < c_end; ch += 4) {
        const int n0 = min(2, c_end - ch), n1 = max(0, min(2, c_end - ch - 2)), n2 = max(0, min(2, c_end - ch - 4));
        fused_pair<PREC, DK, ZI>(n0, n1, 32 * (ch + 2), M, ldS, so, wv, rsh, rsl, wvec, zeta, izeta, wscale, xs0[0],
                                 xs0[1], wv0, wf0, xs1[0], xs1[1], wv1, wf1, xh, xl, gacc, rs);
        if (n1 == 0) break;
        fused_pair<PREC, DK, ZI>(n1, n2, 32 * (ch + 4), M, ldS, so, wv, rsh, rsl, wvec, zeta, izeta, wscale, xs1[0],
                                 xs1[1], wv1, wf1, xs0[0], xs0[1], wv0, wf0, xh, xl, gacc, rs);
    }
    // energy partials of this part: the four lane groups hold different sparse points
    rs += __shfl_xor(rs, 16, 64);
    rs += __shfl_xor(rs, 32, 64);
    // outputs: G and the energies directly (one part, or a whole row block of the
    // stream-K schedule), else the partials of this part or of this block's slot
    const bool whole = sk ? (c_beg == 0 && c_end == C) : nparts == 1;
    const size_t slot = sk ? (size_t)blockIdx.x * 2 + (seg ? 1 : 0) : 0;
    if (g4 == 0 && row0 + r16 < nc) {
        if (!sk) e_part[(size_t)(row0 + r16) * nparts + part] = rs;
        else if (whole) e_part[row0 + r16] = rs;
        else e_slot[slot * 128 + wv * 16 + r16] = rs;
    }
#pragma unroll
    for (int q = 0; q < 2 * DK; q++) {
        const int col = 16 * q + r16;
        if (col >= d) continue;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int row = row0 + rq + v;
            if (row >= nc) continue;
            if (whole) G[(size_t)row * d + col] = (double)gacc[q][v] * inv_wscale;
            else if (sk) Gpart[(slot * 128 + wv * 16 + rq + v) * d + col] = gacc[q][v];
            else Gpart[((size_t)part * nc + row) * d + col] = gacc[q][v];
        }
    }
    u = (long)rb * C + c_end;
    seg++;
    } while (sk && u < u_end);
}

// The block of the stream-K grid (B blocks over U units) whose range holds unit x:
// the largest b with floor(b U / B) <= x.
__device__ __forceinline__ int sk_block_of(long x, int B, long U) { return (int)(((x + 1) * B - 1) / U); }

// Stream-K fix-up.  Blocks [0, gblocks): G of the row blocks that several blocks of the
// grid share, each such row block once (by the grid boundary k that falls first inside
// it), kSkQ blocks per boundary streaming a 1/kSkQ share of the row block's elements each
// (16: batch512 28.9 -> 26.6 us against 4; the sums are the same).
// A shared row block's slots are contiguous 128 x d FP32 tiles and its G rows are
// contiguous, so elements go by flat index, four at a time.  Its first block b0 reached
// it after an earlier row block (slot 1) unless its range starts at the row block
// (slot 0); every later block starts in it (slot 0).  The slots are summed in block
// order.  Blocks from gblocks on: the energy of every centre (whole row blocks from
// e_part, shared ones from the slots), written back to e_part[i] and added to its site.
constexpr int kSkQ = 16;
__global__ __launch_bounds__(256) void k_gp_sk_finish(int nc, int d, int C, int B, long U, int gblocks,
                                                      const float *__restrict__ Gslot,
                                                      const double *__restrict__ e_slot, double inv_wscale,
                                                      double *__restrict__ G, double *__restrict__ e_part,
                                                      const int *__restrict__ centres, double *__restrict__ esite) {
    if ((int)blockIdx.x < gblocks) {
        const int k = blockIdx.x / kSkQ + 1, q = blockIdx.x % kSkQ;   // boundary between blocks k - 1 and k
        const long x = (long)k * U / B;
        if (x % C == 0) return;
        const int rb = (int)(x / C);
        const int b0 = sk_block_of((long)rb * C, B, U), b1 = sk_block_of((long)rb * C + C - 1, B, U);
        if (b0 != k - 1) return;   // an earlier boundary of this row block does it
        const int b0_slot = 2 * b0 + ((long)b0 * U / B < (long)rb * C ? 1 : 0);
        const size_t n = (size_t)min(128, nc - rb * 128) * d;   // elements of the row block's G rows
        const size_t tile = (size_t)128 * d;                    // floats per slot
        double *g = G + (size_t)rb * tile;
        for (size_t i4 = (size_t)q * 256 + threadIdx.x; i4 < n / 4; i4 += 256 * kSkQ) {
            double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
            for (int b = b0; b <= b1; b++) {
                const size_t sl = b == b0 ? b0_slot : 2 * (size_t)b;
                const float4 v = ((const float4 *)(Gslot + sl * tile))[i4];
                a0 += (double)v.x;
                a1 += (double)v.y;
                a2 += (double)v.z;
                a3 += (double)v.w;
            }
            ((double2 *)g)[2 * i4] = make_double2(a0 * inv_wscale, a1 * inv_wscale);
            ((double2 *)g)[2 * i4 + 1] = make_double2(a2 * inv_wscale, a3 * inv_wscale);
        }
        if (q == 0 && threadIdx.x < n % 4) {   // the last row block's tail
            const size_t e = n / 4 * 4 + threadIdx.x;
            double a = 0.0;
            for (int b = b0; b <= b1; b++) a += (double)Gslot[(b == b0 ? b0_slot : 2 * (size_t)b) * tile + e];
            g[e] = a * inv_wscale;
        }
        return;
    }
    const int i = (blockIdx.x - gblocks) * 256 + threadIdx.x;
    if (i >= nc) return;
    const int rb = i / 128, r = i - rb * 128;
    const int b0 = sk_block_of((long)rb * C, B, U), b1 = sk_block_of((long)rb * C + C - 1, B, U);
    double s = 0.0;
    if (b0 == b1) s = e_part[i];
    else {
        const int b0_slot = 2 * b0 + ((long)b0 * U / B < (long)rb * C ? 1 : 0);
        for (int b = b0; b <= b1; b++) s += e_slot[(size_t)(b == b0 ? b0_slot : 2 * b) * 128 + r];
        e_part[i] = s;
    }
    if (centres) esite[centres[i]] += s;
}

// G = inv_wscale * sum over parts (fixed order) of the FP32 partials
__global__ void k_sum_parts_f32(size_t n, int parts, const float *P, double inv_wscale, double *G) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int p = 0; p < parts; p++) s += (double)P[(size_t)p * n + i];
    G[i] = s * inv_wscale;
}

static int int_zeta_m(double z) {
    double r = floor(z + 0.5);
    return (fabs(z - r) == 0.0 && r >= 1.0 && r <= 16.0) ? (int)r : 0;
}


void launch_gemm1_split(int prec, int nc, int M, int d, int dpad, int lds_mode, const double *X, const void *Xs_hi,
                        const void *Xs_lo, const double *wvec, double zeta, double wscale, float *Wf,
                        double *e_part, int *ntile_out, void *xsplit_ws, hipStream_t s) {
    const int Mpad = (M + 31) / 32 * 32;
    dim3 grid(cdivm(nc, 64), cdivm(M, 64));
    const bool lds = use_lds_tiles(nc, M, lds_mode);
    const int ntn = cdivm(M, kLT), nblk = cdivm(nc, kLT) * ntn;
    *ntile_out = lds ? ntn : (int)grid.y;
    if (nc <= 0) return;
    const size_t xsz = (size_t)((nc + 63) / 64 * 64) * dpad;
    char *xs = (char *)xsplit_ws;
    void *Xhi = xs, *Xlo = xs + xsz * 2;
    launch_split_matrix(prec, nc, d, dpad, X, d, false, Xhi, Xlo, s);
    char *wf = (char *)Wf;
    void *Whi = wf, *Wlo = wf + (size_t)nc * Mpad * 2;
    const int iz = int_zeta_m(zeta);
#define G1L(P, T, IZ)                                                                                            \
    hipLaunchKernelGGL((k_gemm1_split_lds<P, IZ>), dim3(nblk), dim3(256), 0, s, nc, M, dpad, Mpad, ntn,          \
                       (const T *)Xhi, (const T *)Xlo, (const T *)Xs_hi, (const T *)Xs_lo, wvec, zeta, iz, wscale, \
                       (T *)Whi, (T *)Wlo, e_part)
    if (lds && prec == 1) {
        if (iz > 0) G1L(1, __bf16, true); else G1L(1, __bf16, false);
    } else if (lds) {
        if (iz > 0) G1L(2, _Float16, true); else G1L(2, _Float16, false);
    } else if (prec == 1)
        hipLaunchKernelGGL(k_gemm1_split<1>, grid, dim3(256), 0, s, nc, M, dpad, Mpad, (const __bf16 *)Xhi,
                           (const __bf16 *)Xlo, (const __bf16 *)Xs_hi, (const __bf16 *)Xs_lo, wvec, zeta,
                           int_zeta_m(zeta), wscale, (__bf16 *)Whi, (__bf16 *)Wlo, e_part);
    else
        hipLaunchKernelGGL(k_gemm1_split<2>, grid, dim3(256), 0, s, nc, M, dpad, Mpad, (const _Float16 *)Xhi,
                           (const _Float16 *)Xlo, (const _Float16 *)Xs_hi, (const _Float16 *)Xs_lo, wvec, zeta,
                           int_zeta_m(zeta), wscale, (_Float16 *)Whi, (_Float16 *)Wlo, e_part);
#undef G1L
}

// d <= 256, and the sparse copies' byte size within the 31-bit bound of the buffer
// resources the kernel reads them through (Mpad < 4M at d = 256)
bool gp_fused_supported(int dpad, int Mpad) { return dpad <= 256 && (size_t)Mpad * dpad * 2 < (1ull << 31); }

int gp_fused_parts(int nc, int Mpad) {
    const int row_blocks = cdivm(nc, 128), chunks = Mpad / 32;
    int parts = std::min(chunks, std::max(1, cdivm(512, row_blocks)));
    const int per = cdivm(chunks, parts);
    return cdivm(chunks, per);
}

// Blocks of one kernel instance that fit a CU at once, and the CU count, per device
// (the stream-K grid is their product)
template <typename K>
static int sk_grid_blocks(K kernel) {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, 512, 0) != hipSuccess) return 0;
    return std::max(1, cus) * std::max(1, per);
}

static int fused_dk(int dpad) { return dpad <= 64 ? 2 : (dpad <= 128 ? 4 : 8); }

// the stream-K grid of a launch: B blocks over U = row blocks x chunks units (B <= U)
static void sk_plan(int prec, int dpad, int nc, int Mpad, int *B, long *U) {
    static std::atomic<int> cache[64][2][3];   // [device][prec][dk]; engines on several threads share it
    int dev = 0;
    hipGetDevice(&dev);
    const int dk = fused_dk(dpad), di = dk == 2 ? 0 : dk == 4 ? 1 : 2, pi = prec == 1 ? 0 : 1;
    int b = cache[dev & 63][pi][di].load(std::memory_order_relaxed);
    if (b == 0) {
#define SKB(P, D) b = sk_grid_blocks(k_gp_fused_split<P, D, 4>)
        if (prec == 1) { if (dk == 2) SKB(1, 2); else if (dk == 4) SKB(1, 4); else SKB(1, 8); }
        else { if (dk == 2) SKB(2, 2); else if (dk == 4) SKB(2, 4); else SKB(2, 8); }
#undef SKB
        if (b <= 0) b = 256;
        cache[dev & 63][pi][di].store(b, std::memory_order_relaxed);
    }
    *U = (long)cdivm(nc, 128) * (Mpad / 32);
    *B = (int)std::max(1L, std::min((long)b, *U));
}

size_t gp_fused_sk_scratch_bytes(int prec, int d, int dpad, int nc, int Mpad) {
    int B = 0;
    long U = 0;
    sk_plan(prec, dpad, nc, Mpad, &B, &U);
    return (size_t)B * 256 * d * 4 + (size_t)B * 256 * 8;
}

template <int PREC, int DK>
static void gp_fused_go(int nc, int M, int d, int dpad, int Mpad, int parts, long sk_units, int sk_blocks,
                        const double *X, const void *Shi, const void *Slo,
                        const double *wvec, double zeta, double wscale, double *G, float *Gpart, double *e_part,
                        double *e_slot, hipStream_t s) {
    typedef typename Half<PREC>::T T;
    const int per = cdivm(Mpad / 32, parts);
    dim3 grid = sk_units > 0 ? dim3(sk_blocks) : dim3(cdivm(nc, 128), parts);
    const int iz = int_zeta_m(zeta);
    if (iz == 4)
        hipLaunchKernelGGL((k_gp_fused_split<PREC, DK, 4>), grid, dim3(512), 0, s, nc, M, d, Mpad, dpad, per,
                           sk_units, X, (const T *)Shi, (const T *)Slo, wvec, zeta,
                           iz, wscale, 1.0 / wscale, G, Gpart, e_part, e_slot);
    else
        hipLaunchKernelGGL((k_gp_fused_split<PREC, DK, 0>), grid, dim3(512), 0, s, nc, M, d, Mpad, dpad, per,
                           sk_units, X, (const T *)Shi, (const T *)Slo, wvec, zeta,
                           iz, wscale, 1.0 / wscale, G, Gpart, e_part, e_slot);
}

int gp_fused_parts_valid(int parts, int Mpad) {
    const int chunks = Mpad / 32;
    parts = std::max(1, std::min(parts, chunks));
    const int per = cdivm(chunks, parts);
    return cdivm(chunks, per);
}

void launch_gp_fused_split(int prec, int nc, int M, int d, int dpad, int Mpad, int parts_req, const double *X,
                           const void *Xs_hi, const void *Xs_lo, const double *wvec, double zeta, double wscale,
                           double *G, float *Gpart, double *e_part, int *ntile_out, hipStream_t s) {
    const bool sk = parts_req < 0;
    const int parts = sk ? 1 : parts_req > 0 ? gp_fused_parts_valid(parts_req, Mpad) : gp_fused_parts(nc, Mpad);
    *ntile_out = sk ? -1 : parts;
    if (nc <= 0) return;
    int B = 0;
    long U = 0;
    if (sk) sk_plan(prec, dpad, nc, Mpad, &B, &U);
    double *e_slot = sk ? (double *)(Gpart + (size_t)B * 256 * d) : nullptr;
    const int dk = fused_dk(dpad);
#define GO(P, D) gp_fused_go<P, D>(nc, M, d, dpad, Mpad, parts, sk ? U : 0, B, X, Xs_hi, Xs_lo, wvec, zeta, wscale, \
                                   G, Gpart, e_part, e_slot, s)
    if (prec == 1) {
        if (dk == 2) GO(1, 2); else if (dk == 4) GO(1, 4); else GO(1, 8);
    } else {
        if (dk == 2) GO(2, 2); else if (dk == 4) GO(2, 4); else GO(2, 8);
    }
#undef GO
}

void launch_gp_sk_finish(int prec, int nc, int d, int dpad, int Mpad, bool want_g, const float *Gpart, double wscale,
                         double *G, double *e_part, const int *centres, double *esite, hipStream_t s) {
    if (nc <= 0) return;
    int B = 0;
    long U = 0;
    sk_plan(prec, dpad, nc, Mpad, &B, &U);
    const double *e_slot = (const double *)(Gpart + (size_t)B * 256 * d);
    const int gb = want_g ? (B - 1) * kSkQ : 0;
    hipLaunchKernelGGL(k_gp_sk_finish, dim3(gb + cdivm(nc, 256)), dim3(256), 0, s, nc, d, Mpad / 32, B, U, gb, Gpart,
                       e_slot, 1.0 / wscale, G, e_part, centres, esite);
}

void launch_gp_fused_reduce(int nc, int d, int parts, const float *Gpart, double wscale, double *G, hipStream_t s) {
    if (parts <= 1 || nc <= 0) return;
    const size_t n = (size_t)nc * d;
    hipLaunchKernelGGL(k_sum_parts_f32, dim3(cdivm((long)n, 256)), dim3(256), 0, s, n, parts, Gpart, 1.0 / wscale, G);
}

void launch_gemm2_split(int prec, int nc, int M, int d, int dpad, int Mpad, int lds_mode, const float *Wf,
                        const void *XsT_hi, const void *XsT_lo, double wscale, double *G, hipStream_t s) {
    (void)M; (void)dpad;
    dim3 grid(cdivm(nc, 64), cdivm(d, 64));
    if (nc <= 0) return;
    const char *wf = (const char *)Wf;
    const void *Whi = wf, *Wlo = wf + (size_t)nc * Mpad * 2;
    const double inv = 1.0 / wscale;
    if (use_lds_tiles(nc, d, lds_mode)) {
        const int ntn = cdivm(d, kLT), nblk = cdivm(nc, kLT) * ntn;
        if (prec == 1)
            hipLaunchKernelGGL(k_gemm2_split_lds<1>, dim3(nblk), dim3(256), 0, s, nc, d, Mpad, ntn,
                               (const __bf16 *)Whi, (const __bf16 *)Wlo, (const __bf16 *)XsT_hi,
                               (const __bf16 *)XsT_lo, inv, G);
        else
            hipLaunchKernelGGL(k_gemm2_split_lds<2>, dim3(nblk), dim3(256), 0, s, nc, d, Mpad, ntn,
                               (const _Float16 *)Whi, (const _Float16 *)Wlo, (const _Float16 *)XsT_hi,
                               (const _Float16 *)XsT_lo, inv, G);
        return;
    }
    if (prec == 1)
        hipLaunchKernelGGL(k_gemm2_split<1>, grid, dim3(256), 0, s, nc, d, Mpad, (const __bf16 *)Whi,
                           (const __bf16 *)Wlo, (const __bf16 *)XsT_hi, (const __bf16 *)XsT_lo, inv, G);
    else
        hipLaunchKernelGGL(k_gemm2_split<2>, grid, dim3(256), 0, s, nc, d, Mpad, (const _Float16 *)Whi,
                           (const _Float16 *)Wlo, (const _Float16 *)XsT_hi, (const _Float16 *)XsT_lo, inv, G);
}

}  // namespace gapx
