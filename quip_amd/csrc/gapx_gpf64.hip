// Fused FP64 GP kernel: energy and gradient of the dot-product covariance in one pass.
//
// For centre rows x_i (descriptor length d <= 256) and sparse points s (M of them):
//   K_is = x_i . X_s,   e_i = sum_s w_s K_is^zeta,   G_i = sum_s w_s zeta K_is^(zeta-1) X_s
// (gp_predict with the dot_product covariance and its gradient, the contraction of
// IPModel_GAP.f95:464-470's gradient terms).  The FP64 two-GEMM form writes W =
// w zeta K^(zeta-1) (nc x M) to HBM and reads it back; here W stays in registers.
//
// One block = 4 wavefronts = 64 centres, one wavefront per SIMD (the register budget of
// a wave is the whole SIMD's).  Wave w owns centres 16w..16w+15 and keeps their X rows in
// registers in the B-operand layout of v_mfma_f64_16x16x4f64.  The block walks its part
// of the sparse points in chunks of 32; a chunk of the padded sparse matrix (32 x DP
// FP64) is staged in LDS, the next one fetched into registers while the current one is
// used (two LDS buffers, one barrier per chunk).  Per chunk each wave computes
// K^T = Xs_chunk X^T (two 16 x 16 tiles): in the accumulator layout lane l, element v
// holds K[centre l%16][sparse point 16j + l/16 + 4v], which is exactly the A-operand
// layout of the second product's k-step (j, v) (k = sparse point).  So the epilogue
// (FP64 pow, energy sum) turns the accumulators into W fragments in place and
// G += W Xs_chunk follows on the same LDS chunk, without any data exchange.
//
// Operand layouts of v_mfma_f64_16x16x4f64 (wave64): A lane l = A[l%16][k = l/16],
// B lane l = B[k = l/16][l%16], C lane l element v = C[l/16 + 4v][l%16].
//
// With several parts (gridDim.y) the sparse points are divided between blocks; the
// partial G rows and the energy partials (e_part columns) are summed in a fixed order
// afterwards (k_gp_finish).
#include "gapx_kernels.h"

#include <algorithm>

namespace gapx {

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef double double2_t __attribute__((ext_vector_type(2)));

static inline int cdivf(long a, long b) { return (int)((a + b - 1) / b); }

constexpr int F64_CH = 32;   // sparse points per chunk

// DK: 32-component groups of the padded descriptor (DP = 32 DK, d <= DP).  ZI: zeta as
// a compile-time integer (powers by multiplication), 0 for the run-time value.
template <int DK, int ZI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_gp_fused_f64(
    int nc, int M, int d, int nchunks, int chunks_per_part, const double *__restrict__ X,
    const double *__restrict__ XsP, const double *__restrict__ wvec, double zeta, int izeta,
    double *__restrict__ G, double *__restrict__ Gpart, double *__restrict__ e_part) {
    constexpr int DP = 32 * DK;       // padded descriptor length: k of K, columns of G
    constexpr int KS = DP / 4;        // k-steps of K^T
    constexpr int NQ = DP / 16;       // 16-column tiles of G
    constexpr int LD = DP + 1;        // LDS row pitch (odd: the 16 rows an A-operand read touches
                                      // fall on distinct bank pairs)
    constexpr int NV = F64_CH * DP / 2 / 256;   // 16-byte vectors of a chunk per thread
    __shared__ double xs[2][F64_CH * LD];
    __shared__ double xw[2][F64_CH];   // the chunk's w_s (0 past M)
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int c16 = lane & 15, g4 = lane >> 4;
    const int row0 = blockIdx.x * 64 + wv * 16;
    const int part = blockIdx.y, nparts = gridDim.y;
    const int c_beg = part * chunks_per_part, c_end = min(nchunks, c_beg + chunks_per_part);

    double xf[KS];   // B operand of K^T: X[row0 + c16][4s + g4]
    {
        const int row = row0 + c16;
#pragma unroll
        for (int s = 0; s < KS; s++) {
            const int k = 4 * s + g4;
            xf[s] = (row < nc && k < d) ? X[(size_t)row * d + k] : 0.0;
        }
    }
    double4_t gacc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) gacc[q] = (double4_t){0.0, 0.0, 0.0, 0.0};
    double esum = 0.0;

    double2_t pf[NV];
    double pw = 0.0;
    auto fetch = [&](int ch) {   // chunk ch of the padded sparse matrix (rows contiguous, DP apart)
        const double2_t *src = (const double2_t *)(XsP + (size_t)ch * F64_CH * DP);
#pragma unroll
        for (int u = 0; u < NV; u++) pf[u] = src[t + 256 * u];
        const int sp = F64_CH * ch + t;
        if (t < F64_CH) pw = sp < M ? wvec[sp] : 0.0;
    };
    auto put = [&](int b) {
#pragma unroll
        for (int u = 0; u < NV; u++) {
            const int idx = t + 256 * u;
            const int r = idx / (DP / 2), c = (idx % (DP / 2)) * 2;
            xs[b][r * LD + c] = pf[u][0];
            xs[b][r * LD + c + 1] = pf[u][1];
        }
        if (t < F64_CH) xw[b][t] = pw;
    };
    if (c_beg < c_end) {
        fetch(c_beg);
        put(0);
    }
    __syncthreads();
    for (int ch = c_beg; ch < c_end; ch++) {
        const int b = (ch - c_beg) & 1;
        if (ch + 1 < c_end) fetch(ch + 1);
        const double *xb = xs[b];
        double4_t ka[2] = {(double4_t){0.0, 0.0, 0.0, 0.0}, (double4_t){0.0, 0.0, 0.0, 0.0}};
#pragma unroll
        for (int s = 0; s < KS; s++)
#pragma unroll
            for (int j = 0; j < 2; j++)
                ka[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(xb[(16 * j + c16) * LD + 4 * s + g4], xf[s], ka[j], 0,
                                                             0, 0);
        // epilogue: element v of tile j is K[centre c16][sparse point 32 ch + 16 j + g4 + 4 v]
        double wf[2][4];
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const double ws = xw[b][16 * j + g4 + 4 * v];
                const double kv = ka[j][v];
                double kz1, kz;
                if constexpr (ZI > 0) {
                    kz1 = 1.0;
#pragma unroll
                    for (int q = 1; q < ZI; q++) kz1 *= kv;
                    kz = kz1 * kv;
                } else if (izeta > 0) {
                    kz1 = 1.0;
                    for (int q = 1; q < izeta; q++) kz1 *= kv;
                    kz = kz1 * kv;
                } else {
                    kz1 = pow(kv, zeta - 1.0);
                    kz = pow(kv, zeta);
                }
                esum += ws * kz;
                wf[j][v] = ws * zeta * kz1;
            }
        // G += W Xs_chunk: k-step (j, v) covers sparse points 16 j + 4 v + 0..3, lane group g4 one of them
#pragma unroll
        for (int q = 0; q < NQ; q++)
#pragma unroll
            for (int j = 0; j < 2; j++)
#pragma unroll
                for (int v = 0; v < 4; v++)
                    gacc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(wf[j][v], xb[(16 * j + 4 * v + g4) * LD + 16 * q + c16],
                                                                   gacc[q], 0, 0, 0);
        if (ch + 1 < c_end) put(b ^ 1);   // the other buffer was last read before the previous barrier
        __syncthreads();
    }
    // energy of centre c16 over this part: the four lane groups hold different sparse points
    esum += __shfl_xor(esum, 16, 64);
    esum += __shfl_xor(esum, 32, 64);
    if (g4 == 0 && row0 + c16 < nc) e_part[(size_t)(row0 + c16) * nparts + part] = esum;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int col = 16 * q + c16;
        if (col >= d) continue;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int row = row0 + g4 + 4 * v;
            if (row >= nc) continue;
            if (nparts == 1) G[(size_t)row * d + col] = gacc[q][v];
            else Gpart[((size_t)part * nc + row) * d + col] = gacc[q][v];
        }
    }
}

bool gp_f64_fused_supported(int d) { return d >= 1 && d <= 256; }

int gp_f64_parts_valid(int parts, int M) {
    const int chunks = cdivf(M, F64_CH);
    parts = std::max(1, std::min(parts, chunks));
    const int per = cdivf(chunks, parts);
    return cdivf(chunks, per);
}

// default: enough blocks for one per CU
int gp_f64_parts_default(int nc, int M) {
    const int rb = cdivf(nc, 64);
    int parts = 1;
    while (rb * parts < 256 && parts < 8) parts *= 2;
    return gp_f64_parts_valid(parts, M);
}

static int int_zeta_f(double z) {
    double r = floor(z + 0.5);
    return (fabs(z - r) == 0.0 && r >= 1.0 && r <= 16.0) ? (int)r : 0;
}

void launch_gp_fused_f64(int nc, int M, int d, int parts_req, const double *X, const double *XsP,
                         const double *wvec, double zeta, double *G, double *Gparts, double *e_part, int *ntile_out,
                         hipStream_t s) {
    const int chunks = cdivf(M, F64_CH);
    const int parts = gp_f64_parts_valid(parts_req, M);
    const int per = cdivf(chunks, parts);
    *ntile_out = parts;
    if (nc <= 0) return;
    const int DK = (d + 31) / 32;
    dim3 grid(cdivf(nc, 64), parts);
    const int iz = int_zeta_f(zeta);
#define GPF(K)                                                                                                  \
    case K:                                                                                                     \
        if (iz == 4)                                                                                            \
            hipLaunchKernelGGL((k_gp_fused_f64<K, 4>), grid, dim3(256), 0, s, nc, M, d, chunks, per, X, XsP, wvec, \
                               zeta, iz, G, Gparts, e_part);                                                    \
        else                                                                                                    \
            hipLaunchKernelGGL((k_gp_fused_f64<K, 0>), grid, dim3(256), 0, s, nc, M, d, chunks, per, X, XsP, wvec, \
                               zeta, iz, G, Gparts, e_part);                                                    \
        break;
    switch (DK) {
        GPF(1) GPF(2) GPF(3) GPF(4) GPF(5) GPF(6) GPF(7) GPF(8)
        default: return;
    }
#undef GPF
}

}  // namespace gapx
