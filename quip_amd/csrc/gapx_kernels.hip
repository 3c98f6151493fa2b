// HIP kernels for the SOAP-GAP energy/force/virial path on gfx950 (MI355X).
//
//  neighbour list   cell binning + per-cell lists + per-atom pair search that
//                   reproduces QUIP's calc_connect tables bit for bit
//                   (src/libAtoms/Connection.f95:1035-1310, 448-634, 1519-1621)
//  soap_fwd         one wavefront per centre atom: radial basis and real
//                   spherical harmonics of a neighbour block staged in LDS,
//                   density coefficients c_(s,b,lm), power spectrum, norm
//  soap_bwd         adjoint pass: dE/dp -> dE/dc in LDS, then one lane per
//                   neighbour re-evaluates the basis and contracts its gradient
//                   (no grad_data is ever materialised)
//  gemm1 / gemm2    FP64 MFMA (v_mfma_f64_16x16x4f64) GEMMs for the dot-product
//                   GP kernel K = X Xs^T with fused (w K^zeta, w zeta K^(zeta-1))
//                   epilogue, and G = W Xs
//  d2b              distance_2b ard_se terms per pair
//  gather           deterministic per-atom segmented reductions for forces,
//                   local virials and local energies (no floating atomics)
#include "gapx_kernels.h"

#include <algorithm>

namespace gapx {

#define GAPX_PI 3.14159265358979323846264338327950288

__constant__ double c_ynorm[(kMaxL + 1) * (kMaxL + 2) / 2];   // N_lm (x sqrt 2 for m>0)
__constant__ double c_dfact[kMaxL + 2];                       // (2m-1)!!

// constant tables are uploaded once per device (each device has its own copy)
static bool g_const_ready[64] = {};
static hipError_t ensure_constants() {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    if (g_const_ready[dev]) return hipSuccess;
    double yn[(kMaxL + 1) * (kMaxL + 2) / 2], df[kMaxL + 2];
    for (int l = 0; l <= kMaxL; l++)
        for (int m = 0; m <= l; m++) {
            double r = 1.0;   // (l-m)!/(l+m)!
            for (int k = l - m + 1; k <= l + m; k++) r /= (double)k;
            double v = sqrt((2.0 * l + 1.0) / (4.0 * GAPX_PI) * r);
            if (m > 0) v *= sqrt(2.0);
            yn[l * (l + 1) / 2 + m] = v;
        }
    df[0] = 1.0;
    for (int m = 1; m < kMaxL + 2; m++) df[m] = df[m - 1] * (2.0 * m - 1.0);
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_ynorm), yn, sizeof(yn))) != hipSuccess) return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_dfact), df, sizeof(df))) != hipSuccess) return e;
    g_const_ready[dev] = true;
    return hipSuccess;
}

hipError_t soap_constants_init() {
    hipError_t e = ensure_constants();
    return e != hipSuccess ? e : soap_fast_constants_init();
}

static inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

// ----------------------------------------------------------------------------
// small fills (graph-safe replacements for memset)
// ----------------------------------------------------------------------------
__global__ void k_fill_i32(int *p, int n, int v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_f64(double *p, size_t n, double v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
void launch_fill_i32(int *p, int n, int v, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_fill_i32, dim3(cdiv(n, 256)), dim3(256), 0, s, p, n, v);
}
void launch_fill_f64(double *p, size_t n, double v, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_fill_f64, dim3(cdiv((long)n, 256)), dim3(256), 0, s, p, n, v);
}

// Zero up to kMultiFill arrays in one launch (blockIdx.y picks the array).
struct ZeroI32Set { int *p[kMultiFill]; int n[kMultiFill]; };
struct ZeroF64Set { double *p[kMultiFill]; size_t n[kMultiFill]; };
__global__ void k_zero_i32_multi(ZeroI32Set z) {
    const int a = blockIdx.y;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < z.n[a]; i += gridDim.x * blockDim.x) z.p[a][i] = 0;
}
__global__ void k_zero_f64_multi(ZeroF64Set z) {
    const int a = blockIdx.y;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < z.n[a]; i += (size_t)gridDim.x * blockDim.x)
        z.p[a][i] = 0.0;
}
void launch_zero_i32_multi(int k, int *const *p, const int *n, hipStream_t s) {
    ZeroI32Set z = {};
    int mx = 0;
    for (int a = 0; a < k && a < kMultiFill; a++) { z.p[a] = p[a]; z.n[a] = p[a] ? n[a] : 0; mx = std::max(mx, z.n[a]); }
    if (mx > 0)
        hipLaunchKernelGGL(k_zero_i32_multi, dim3(std::min(cdiv(mx, 256), 4096), std::min(k, kMultiFill)), dim3(256), 0,
                           s, z);
}
void launch_zero_f64_multi(int k, double *const *p, const size_t *n, hipStream_t s) {
    ZeroF64Set z = {};
    size_t mx = 0;
    for (int a = 0; a < k && a < kMultiFill; a++) { z.p[a] = p[a]; z.n[a] = p[a] ? n[a] : 0; mx = std::max(mx, z.n[a]); }
    if (mx > 0)
        hipLaunchKernelGGL(k_zero_f64_multi, dim3((int)std::min<size_t>((mx + 255) / 256, 8192), std::min(k, kMultiFill)),
                           dim3(256), 0, s, z);
}


// ----------------------------------------------------------------------------
// exclusive prefix sum of n ints into out[0..n] (out[n] = total), three passes:
// per-block sums, a single-block scan of those, then per-block local scans.
// ----------------------------------------------------------------------------
constexpr int SCAN_BS = 256, SCAN_ITEMS = 8, SCAN_TILE = SCAN_BS * SCAN_ITEMS;
__device__ __forceinline__ int block_excl_scan(int v, int *sh, int *total) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_BS; off <<= 1) {
        int a = (t >= off) ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += a;
        __syncthreads();
    }
    int incl = sh[t];
    *total = sh[SCAN_BS - 1];
    __syncthreads();
    return incl - v;
}
// Exclusive scans of one or two int arrays (blockIdx.y or blockIdx.x selects the array);
// out[n] gets each array's total, mx[b] (nullable) its largest item and *tot (nullable)
// the total of array 0.  Large arrays: per-tile sums and maxima, a one-block scan of
// the tile sums, then the tiles.  Small arrays: one block per array.
struct ScanSet { const int *in[2]; int *out[2]; int *part[2]; };   // part: tile sums, then tile maxima
__global__ void k_scan_partials(ScanSet ss, int n, int nb) {
    __shared__ int sh[SCAN_BS];
    const int b = blockIdx.y;
    const int *in = ss.in[b];
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int s = 0, m = 0;
    for (int k = 0; k < SCAN_ITEMS; k++) {
        const int v = (base + k < n) ? in[base + k] : 0;
        s += v;
        m = max(m, v);
    }
    int tot;
    block_excl_scan(s, sh, &tot);
    for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        int bm = sh[0];
        for (int w = 1; w < SCAN_BS / 64; w++) bm = max(bm, sh[w]);
        ss.part[b][blockIdx.x] = tot;
        ss.part[b][nb + blockIdx.x] = bm;
    }
}
__global__ void k_scan_top(ScanSet ss, int n, int nb, int *mx, int *tot_out) {
    __shared__ int sh[SCAN_BS];
    const int b = blockIdx.x;
    int *partials = ss.part[b];
    int carry = 0, m = 0;
    for (int c0 = 0; c0 < nb; c0 += SCAN_BS) {
        int i = c0 + threadIdx.x;
        int v = i < nb ? partials[i] : 0;
        if (i < nb) m = max(m, partials[nb + i]);
        int tot;
        int ex = block_excl_scan(v, sh, &tot);
        if (i < nb) partials[i] = ex + carry;
        carry += tot;
    }
    for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < SCAN_BS / 64; w++) m = max(m, sh[w]);
        ss.out[b][n] = carry;
        if (mx) mx[b] = m;
        if (tot_out && b == 0) *tot_out = carry;
    }
}
__global__ void k_scan_apply(ScanSet ss, int n) {
    __shared__ int sh[SCAN_BS];
    const int b = blockIdx.y;
    const int *in = ss.in[b];
    int *out = ss.out[b];
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS], s = 0;
    for (int k = 0; k < SCAN_ITEMS; k++) { v[k] = (base + k < n) ? in[base + k] : 0; s += v[k]; }
    int tot;
    int ex = block_excl_scan(s, sh, &tot) + ss.part[b][blockIdx.x];
    for (int k = 0; k < SCAN_ITEMS; k++) {
        if (base + k < n) out[base + k] = ex;
        ex += v[k];
    }
}
// One block of 1024 threads per array for n <= kScanOne.  The array passes through LDS
// in chunks of kScanChunk items: loaded and stored coalesced, scanned as 16 contiguous
// items per thread (item i sits at i + i/16, so a wave's 16-item runs fall on distinct
// banks), with a wave scan and a scan of the 16 wave totals joining the runs and a
// carry joining the chunks.
constexpr int kScanOne = 1 << 16;
constexpr int kScanChunk = 16 * 1024;
__device__ __forceinline__ void scan_one_array(const ScanSet &ss, int b, int n, int *mx, int *tot_out, int *buf,
                                               int *wsum, int *wmax) {
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int *in = ss.in[b];
    int *out = ss.out[b];
    int carry = 0, big = 0;
    for (int c0 = 0; c0 < n; c0 += kScanChunk) {
        const int cn = min(kScanChunk, n - c0);
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const int i = t + 1024 * k;
            buf[i + (i >> 4)] = i < cn ? in[c0 + i] : 0;
        }
        __syncthreads();
        int v[16], s = 0, m = 0;
#pragma unroll
        for (int k = 0; k < 16; k++) {
            v[k] = buf[17 * t + k];
            s += v[k];
            m = max(m, v[k]);
        }
        int x = s;   // inclusive wave scan of the run sums
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off, 64));
        if (lane == 63) wsum[wv] = x;
        if (lane == 0) wmax[wv] = m;
        __syncthreads();
        int before = 0, total = 0;
        for (int q = 0; q < 16; q++) {
            before += q < wv ? wsum[q] : 0;
            total += wsum[q];
            big = max(big, wmax[q]);
        }
        int ex = carry + before + x - s;
#pragma unroll
        for (int k = 0; k < 16; k++) {
            buf[17 * t + k] = ex;
            ex += v[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const int i = t + 1024 * k;
            if (i < cn) out[c0 + i] = buf[i + (i >> 4)];
        }
        carry += total;
        __syncthreads();   // buf and the wave totals are rewritten by the next chunk
    }
    if (t == 0) {
        out[n] = carry;
        if (mx) mx[b] = big;
        if (tot_out && b == 0) *tot_out = carry;
    }
}
__global__ __launch_bounds__(1024) void k_scan_one(ScanSet ss, int n, int *mx, int *tot_out) {
    __shared__ int buf[kScanChunk + kScanChunk / 16];
    __shared__ int wsum[16], wmax[16];
    scan_one_array(ss, blockIdx.x, n, mx, tot_out, buf, wsum, wmax);
}
// Both neighbour counts scanned by one block, then the overflow guard of a capacity-sized
// build (k_nl_guard's rule) on the offsets just written: one launch instead of two.
__global__ __launch_bounds__(1024) void k_scan2_guard(ScanSet ss, int n, int *flags, int cap) {
    __shared__ int buf[kScanChunk + kScanChunk / 16];
    __shared__ int wsum[16], wmax[16];
    scan_one_array(ss, 0, n, flags + 1, flags + 3, buf, wsum, wmax);
    __syncthreads();
    scan_one_array(ss, 1, n, flags + 1, nullptr, buf, wsum, wmax);
    __threadfence_block();
    __syncthreads();
    const bool over = flags[3] > cap;
    if (!(over || flags[0] != 0)) return;
    for (int i = threadIdx.x; i <= n; i += blockDim.x) {
        ss.out[0][i] = 0;
        ss.out[1][i] = 0;
    }
    if (threadIdx.x == 0 && over) atomicOr(flags, 2);
}
// partials: 2 * scan_partials_size(n) ints per array (tile sums, then tile maxima)
int scan_partials_size(int n) { return 2 * cdiv(n > 0 ? n : 1, SCAN_TILE); }
static void scan_go(const ScanSet &ss, int narr, int n, int *mx, int *tot, hipStream_t s) {
    if (n <= kScanOne) {
        hipLaunchKernelGGL(k_scan_one, dim3(narr), dim3(1024), 0, s, ss, n, mx, tot);
        return;
    }
    const int nb = cdiv(n, SCAN_TILE);
    hipLaunchKernelGGL(k_scan_partials, dim3(nb, narr), dim3(SCAN_BS), 0, s, ss, n, nb);
    hipLaunchKernelGGL(k_scan_top, dim3(narr), dim3(SCAN_BS), 0, s, ss, n, nb, mx, tot);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb, narr), dim3(SCAN_BS), 0, s, ss, n);
}
void launch_scan_i32(const int *in, int n, int *out, int *partials, hipStream_t s) {
    const ScanSet ss = {{in, in}, {out, out}, {partials, partials}};
    scan_go(ss, 1, n, nullptr, nullptr, s);
}
// Both neighbour counts scanned, their largest values written to mx[0], mx[1] and the
// total of in1 to *tot1.
void launch_scan2_max(const int *in1, const int *in2, int n, int *out1, int *out2, int *mx, int *tot1, int *part1,
                      int *part2, hipStream_t s) {
    const ScanSet ss = {{in1, in2}, {out1, out2}, {part1, part2}};
    scan_go(ss, 2, n, mx, tot1, s);
}
// The same for n <= 8192: the two arrays scanned at once, threads 0-511 the first and
// 512-1023 the second, 16 contiguous items per thread through LDS.
constexpr int kScan2Small = 8192;
__global__ __launch_bounds__(1024) void k_scan2_guard_small(ScanSet ss, int n, int *flags, int cap) {
    __shared__ int buf[2][kScan2Small + kScan2Small / 16];
    __shared__ int wsum[2][8], wmax[2][8];
    const int t = threadIdx.x, b = t >> 9, tl = t & 511, lane = t & 63, wv = tl >> 6;
    const int *in = ss.in[b];
    int *out = ss.out[b];
    int *bb = buf[b];
#pragma unroll
    for (int k = 0; k < 16; k++) {
        const int i = tl + 512 * k;
        bb[i + (i >> 4)] = i < n ? in[i] : 0;
    }
    __syncthreads();
    int v[16], sum = 0, m = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        v[k] = bb[17 * tl + k];
        sum += v[k];
        m = max(m, v[k]);
    }
    int x = sum;
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off, 64));
    if (lane == 63) wsum[b][wv] = x;
    if (lane == 0) wmax[b][wv] = m;
    __syncthreads();
    int before = 0, total = 0, big = 0;
    for (int q = 0; q < 8; q++) {
        before += q < wv ? wsum[b][q] : 0;
        total += wsum[b][q];
        big = max(big, wmax[b][q]);
    }
    int ex = before + x - sum;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        bb[17 * tl + k] = ex;
        ex += v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; k++) {
        const int i = tl + 512 * k;
        if (i < n) out[i] = bb[i + (i >> 4)];
    }
    if (tl == 0) {
        out[n] = total;
        flags[1 + b] = big;
        if (b == 0) flags[3] = total;
    }
    __threadfence_block();
    __syncthreads();
    const bool over = flags[3] > cap;
    if (!(over || flags[0] != 0)) return;
    for (int i = t; i <= n; i += blockDim.x) {
        ss.out[0][i] = 0;
        ss.out[1][i] = 0;
    }
    if (t == 0 && over) atomicOr(flags, 2);
}
bool scan2_guard_fits(int n) { return n <= kScanOne; }
void launch_scan2_guard(const int *in1, const int *in2, int n, int *out1, int *out2, int *flags, int cap,
                        hipStream_t s) {
    const ScanSet ss = {{in1, in2}, {out1, out2}, {nullptr, nullptr}};
    if (n <= kScan2Small) hipLaunchKernelGGL(k_scan2_guard_small, dim3(1), dim3(1024), 0, s, ss, n, flags, cap);
    else hipLaunchKernelGGL(k_scan2_guard, dim3(1), dim3(1024), 0, s, ss, n, flags, cap);
}

// ----------------------------------------------------------------------------
// neighbour list
// All floating point below follows the reference's operation order with no
// contraction, so the in/out decisions and stored distances are bit-exact.
// ----------------------------------------------------------------------------
__device__ __forceinline__ void mult3_exact(const double *g, double p0, double p1, double p2, double *out) {
#pragma clang fp contract(off)
    // matrix_product_vect (linearalgebra.f95:671-686): out(i) = ((0 + g(i,1)p1) + g(i,2)p2) + g(i,3)p3
    for (int i = 0; i < 3; i++) {
        double t = 0.0;
        t = t + g[i] * p0;
        t = t + g[3 + i] * p1;
        t = t + g[6 + i] * p2;
        out[i] = t;
    }
}

// Cell of atom n and its map shift (Connection.f95 cell_of_pos with the lattice shift
// that folds the position into the cell): writes map_shift[3n..3n+2], returns the cell.
__device__ __forceinline__ int nl_bin_atom(int n, const double *__restrict__ pos, const NLGeom *__restrict__ geo,
                                           const int *__restrict__ sid, const int *__restrict__ cell_base,
                                           int *map_shift) {
#pragma clang fp contract(off)
    const int st = sid[n];
    const NLGeom &g = geo[st];
    double lp[3];
    mult3_exact(g.inv_eff, pos[3 * n], pos[3 * n + 1], pos[3 * n + 2], lp);
    int ci[3];
    for (int a = 0; a < 3; a++) {
        double t = lp[a] - g.offset[a];
        int ms = g.pbc[a] ? -(int)floor(t + 0.5) : 0;
        map_shift[3 * n + a] = ms;
        t = t + (double)ms;
        int c = (int)floor((double)g.cells[a] * (t + 0.5)) + 1;
        c = c < 1 ? 1 : (c > g.cells[a] ? g.cells[a] : c);
        ci[a] = c - 1;
    }
    return cell_base[st] + (ci[2] * g.cells[1] + ci[1]) * g.cells[0] + ci[0];
}
__global__ void k_nl_bin(int N, const double *__restrict__ pos, const NLGeom *__restrict__ geo,
                         const int *__restrict__ sid, const int *__restrict__ cell_base, int *cell_of,
                         int *map_shift, int *cell_cnt) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int cell = nl_bin_atom(n, pos, geo, sid, cell_base, map_shift);
    cell_of[n] = cell;
    atomicAdd(&cell_cnt[cell], 1);
}
void launch_nl_bin(int N, const double *pos, const NLGeom *geo, const int *sid, const int *cell_base, int *cell_of,
                   int *map_shift, int *cell_cnt, hipStream_t s) {
    if (N > 0)
        hipLaunchKernelGGL(k_nl_bin, dim3(cdiv(N, 256)), dim3(256), 0, s, N, pos, geo, sid, cell_base, cell_of,
                           map_shift, cell_cnt);
}
__global__ void k_nl_cell_fill(int N, const int *cell_of, const int *start, int *fillc, int *atoms) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    int c = cell_of[n];
    int slot = atomicAdd(&fillc[c], 1);
    atoms[start[c] + slot] = n;
}
void launch_nl_cell_fill(int N, const int *cell_of, const int *start, int *fillc, int *atoms, hipStream_t s) {
    if (N > 0) hipLaunchKernelGGL(k_nl_cell_fill, dim3(cdiv(N, 256)), dim3(256), 0, s, N, cell_of, start, fillc, atoms);
}
// Each cell's list in descending atom index: the order of QUIP's head-inserted
// linked lists (Connection.f95:1580-1584).  Atom n's slot is the number of atoms
// of its cell with a larger index.
__device__ __forceinline__ void put_cell_atom(int slot, int n, const double *pos, const int *map_shift,
                                              CellAtom *cpos, CellShift *cms) {
    cpos[slot] = CellAtom{{pos[3 * n], pos[3 * n + 1], pos[3 * n + 2], 0.0}};
    cms[slot] = CellShift{{map_shift[3 * n], map_shift[3 * n + 1], map_shift[3 * n + 2], n}};
}
__global__ void k_nl_cell_rank(int N, const int *cell_of, const int *start, const int *cnt, const int *unsorted,
                               int *atoms, const double *pos, const int *map_shift, CellAtom *cpos, CellShift *cms) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int c = cell_of[n], st = start[c], m = cnt[c];
    int rank = 0;
    for (int q = 0; q < m; q++) rank += unsorted[st + q] > n ? 1 : 0;
    atoms[st + rank] = n;
    put_cell_atom(st + rank, n, pos, map_shift, cpos, cms);
}
void launch_nl_cell_rank(int N, const int *cell_of, const int *start, const int *cnt, const int *unsorted,
                         int *atoms, const double *pos, const int *map_shift, CellAtom *cpos, CellShift *cms,
                         hipStream_t s) {
    if (N > 0)
        hipLaunchKernelGGL(k_nl_cell_rank, dim3(cdiv(N, 256)), dim3(256), 0, s, N, cell_of, start, cnt, unsorted, atoms,
                           pos, map_shift, cpos, cms);
}

// The front of the neighbour list for small problems in one block of 1024 threads (each
// phase a separate launch otherwise): zero the neighbour2 counters and the flags, bin
// the atoms (cell counts in LDS), scan the counts into cell_start, fill the cells in
// arrival order and rank each cell's atoms by descending index, with the cells and
// their unsorted lists in LDS.  The same tables as the separate kernels: the order
// within a cell is fixed by the ranking, not by the atomics.
constexpr int kFrontThreads = 1024;
__global__ __launch_bounds__(kFrontThreads) void k_nl_front_small(int N, int NC, const double *__restrict__ pos,
                                                                  const NLGeom *__restrict__ geo,
                                                                  const int *__restrict__ sid,
                                                                  const int *__restrict__ cell_base, int *cell_of,
                                                                  int *map_shift, int *cell_cnt, int *cell_start,
                                                                  int *cell_atoms, CellAtom *cpos, CellShift *cms,
                                                                  int *n2_cnt, int *n2_fill, int *flags) {
    __shared__ int scnt[kNlFrontCells], sfill[kNlFrontCells];
    __shared__ int scell[kNlFrontAtoms], suns[kNlFrontAtoms];
    __shared__ int wsum[kFrontThreads / 64];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    for (int c = t; c < NC; c += kFrontThreads) scnt[c] = 0;
    for (int i = t; i < N; i += kFrontThreads) { n2_cnt[i] = 0; n2_fill[i] = 0; }
    if (t < 4) flags[t] = 0;
    __syncthreads();
    for (int n0 = t; n0 < N; n0 += 4 * kFrontThreads) {   // four atoms' loads in flight per thread
        int cl[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int n = n0 + u * kFrontThreads;
            cl[u] = n < N ? nl_bin_atom(n, pos, geo, sid, cell_base, map_shift) : -1;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int n = n0 + u * kFrontThreads;
            if (n < N) {
                scell[n] = cl[u];
                cell_of[n] = cl[u];
                atomicAdd(&scnt[cl[u]], 1);
            }
        }
    }
    __syncthreads();
    // exclusive scan of the NC counts: thread t takes a contiguous run of per items
    const int per = (NC + kFrontThreads - 1) / kFrontThreads;
    const int b0 = t * per, b1 = min(NC, b0 + per);
    int run = 0;
    for (int c = b0; c < b1; c++) run += scnt[c];
    int x = run;
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    int before = 0, total = 0;
    for (int q = 0; q < kFrontThreads / 64; q++) {
        before += q < wv ? wsum[q] : 0;
        total += wsum[q];
    }
    int ex = before + x - run;
    for (int c = b0; c < b1; c++) {
        const int k = scnt[c];
        cell_cnt[c] = k;
        cell_start[c] = ex;
        sfill[c] = ex;   // the fill cursor starts at the cell's first slot
        ex += k;
    }
    if (t == 0) cell_start[NC] = total;
    __syncthreads();
    for (int n = t; n < N; n += kFrontThreads) suns[atomicAdd(&sfill[scell[n]], 1)] = n;
    __syncthreads();
    for (int n = t; n < N; n += kFrontThreads) {
        const int c = scell[n], m = scnt[c], st = sfill[c] - m;   // the cursor ended at the cell's end
        int rank = 0;
        for (int q = 0; q < m; q++) rank += suns[st + q] > n ? 1 : 0;
        cell_atoms[st + rank] = n;
        if (cpos) put_cell_atom(st + rank, n, pos, map_shift, cpos, cms);
    }
}
void launch_nl_front_small(int N, int NC, const double *pos, const NLGeom *geo, const int *sid,
                           const int *cell_base, int *cell_of, int *map_shift, int *cell_cnt, int *cell_start,
                           int *cell_atoms, CellAtom *cpos, CellShift *cms, int *n2_cnt, int *n2_fill, int *flags,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_nl_front_small, dim3(1), dim3(kFrontThreads), 0, s, N, NC, pos, geo, sid, cell_base, cell_of,
                       map_shift, cell_cnt, cell_start, cell_atoms, cpos, cms, n2_cnt, n2_fill, flags);
}

// Wave-parallel form of the same search: a group of lanes per atom1, lanes over the
// neighbour-cell offsets in the reference's loop order (k2 slowest, i2 fastest).  Each
// lane scans one cell list; an exclusive scan of the lane counts places every lane's
// pairs, so neighbour1 keeps the append order.
// Exclusive scan over the W lanes of each W-lane group of the wavefront (W = 64 or 32);
// *total is the group's sum.
template <int W>
__device__ __forceinline__ int group_excl_scan_i(int v, int *total) {
    const int lane = threadIdx.x & (W - 1);
    int x = v;
    for (int off = 1; off < W; off <<= 1) {
        int y = __shfl_up(x, off, W);
        if (lane >= off) x += y;
    }
    *total = __shfl(x, W - 1, W);
    return x - v;
}

// LPA lanes per atom1 (64: one atom per wavefront; 32: two, each half on its own).  Candidates are flattened: the offset cells are
// taken in the reference's loop order (k2 slowest, i2 fastest), and in each
// cell's descending list the atoms with index >= atom1 form a prefix (found by
// binary search).  An exclusive wave scan of the prefix lengths numbers the
// candidates in append order; each lane then tests one candidate, and a ballot
// gives every accepted pair its slot in neighbour1(atom1).
// MODE 0 counts (neighbour1 per atom1, neighbour2 per atom2); MODE 1 fills the
// exact CSR tables; MODE 2 fills per-atom slots of cap1 / cap2 entries
// (neighbour1 of atom a at a*cap1, neighbour2 of atom a at a*cap2), counts both,
// and raises *ovf when a count exceeds its capacity.
// Latency-bound (most of its time waits on the candidates' global loads): the occupancy
// hint gives five wavefronts per SIMD (96 VGPRs, no spills) instead of four.  With two
// atoms per wavefront (large problems) the wavefronts that wait in turn are halved; the
// halves' loops run their own trip counts (each lane reads only its half's ballot bits
// and shuffles stay within the half), and a half past N loads through a valid index and
// runs no iteration.
template <int MODE, bool CC, int LPA>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_nl_pairs_w(int N, const double *__restrict__ pos,
                                                    const int *__restrict__ map_shift, const int *__restrict__ cell_of,
                                                    const int *__restrict__ cell_start, const int *__restrict__ cell_cnt,
                                                    const int *__restrict__ cell_atoms,
                                                    const CellAtom *__restrict__ cpos,
                                                    const CellShift *__restrict__ cms,
                                                    const NLGeom *__restrict__ geo, const int *__restrict__ sid,
                                                    const int *__restrict__ cell_base, int *n1_cnt,
                                                    const int *n1_off, int *n1_j, int *n1_shift, double *n1_r,
                                                    int *n2_cnt, const int *n2_off, int *n2_fill, int *n2_i,
                                                    int *n2_k, int cap1, int cap2, int *ovf) {
#pragma clang fp contract(off)
    constexpr bool FILL = MODE == 1;
    constexpr int APW = 64 / LPA;   // atoms per wavefront
    __shared__ int s_pref[4 * APW][LPA + 1];
    __shared__ int s_start[4 * APW][LPA];
    __shared__ int s_sh[4 * APW][LPA];   // i4, j4, k4 packed (each in [-128, 127]) and the same-cell flag
    const int lane = threadIdx.x & (LPA - 1), wv = threadIdx.x >> 6, row = (threadIdx.x >> 6) * APW + ((threadIdx.x & 63) / LPA);
    const int atom1_r = blockIdx.x * 4 * APW + row;
    if (blockIdx.x * 4 * APW + wv * APW >= N) return;   // the wavefront's first atom: whole wavefront leaves together
    const bool live = atom1_r < N;
    const int atom1 = live ? atom1_r : N - 1;   // a half past N reads a valid atom and runs no iteration
    const int gshift = LPA * ((threadIdx.x & 63) / LPA);
    const unsigned long long gmask = LPA == 64 ? ~0ull : ((1ull << LPA) - 1ull);
    const int st = sid[atom1];
    const NLGeom &g = geo[st];
    const int cb = cell_base[st];
    const int lin = cell_of[atom1] - cb;   // cell index within this structure
    const int cc[3] = {lin % g.cells[0] + 1, (lin / g.cells[0]) % g.cells[1] + 1, lin / (g.cells[0] * g.cells[1]) + 1};
    int lo[3], span[3];
    for (int a = 0; a < 3; a++) {
        int hi;
        if (g.pbc[a]) { lo[a] = -g.img[a]; hi = g.img[a]; }
        else {
            lo[a] = (g.img[a] < cc[a]) ? -g.img[a] : -cc[a] + 1;
            hi = (g.img[a] < g.cells[a] - cc[a]) ? g.img[a] : g.cells[a] - cc[a];
        }
        span[a] = hi - lo[a] + 1;
    }
    const int noff = live ? span[0] * span[1] * span[2] : 0;
    const double p1x = pos[3 * atom1], p1y = pos[3 * atom1 + 1], p1z = pos[3 * atom1 + 2];
    const int ms1x = map_shift[3 * atom1], ms1y = map_shift[3 * atom1 + 1], ms1z = map_shift[3 * atom1 + 2];
    const double rc = g.cutoff;
    const unsigned long long below = (1ull << lane) - 1ull;
    const int base1 = MODE == 1 ? n1_off[atom1] : (MODE == 2 ? atom1 * cap1 : 0);
    int base = base1;
    int total_all = 0;
    for (int o0 = 0; o0 < noff; o0 += LPA) {
        // this lane's offset cell: list start, shift, and the length of its atom >= atom1 prefix
        const int o = o0 + lane;
        int cnt = 0, start = 0, packed = 0;
        if (o < noff) {
            const int i2 = lo[0] + o % span[0], j2 = lo[1] + (o / span[0]) % span[1], k2 = lo[2] + o / (span[0] * span[1]);
            int k3 = 1, j3 = 1, i3 = 1;
            if (g.cells[2] > 1) k3 = (cc[2] + k2 - 1 + g.cells[2]) % g.cells[2] + 1;
            if (g.cells[1] > 1) j3 = (cc[1] + j2 - 1 + g.cells[1]) % g.cells[1] + 1;
            if (g.cells[0] > 1) i3 = (cc[0] + i2 - 1 + g.cells[0]) % g.cells[0] + 1;
            const int k4 = (cc[2] + k2 - k3) / g.cells[2];
            const int j4 = (cc[1] + j2 - j3) / g.cells[1];
            const int i4 = (cc[0] + i2 - i3) / g.cells[0];
            if (k3 >= 1 && k3 <= g.cells[2] && j3 >= 1 && j3 <= g.cells[1] && i3 >= 1 && i3 <= g.cells[0]) {
                const int lin2 = ((k3 - 1) * g.cells[1] + (j3 - 1)) * g.cells[0] + (i3 - 1);
                const int same = (lin2 == lin) ? 1 : 0;
                start = cell_start[lin2 + cb];
                int lo_q = 0, hi_q = cell_cnt[lin2 + cb];
                while (lo_q < hi_q) {   // first position holding an atom < atom1
                    const int mid = (lo_q + hi_q) >> 1;
                    if (cell_atoms[start + mid] >= atom1) lo_q = mid + 1;
                    else hi_q = mid;
                }
                cnt = lo_q;
                packed = ((i4 & 255) | ((j4 & 255) << 8) | ((k4 & 255) << 16)) | (same << 24);
            }
        }
        int tot;
        const int pre = group_excl_scan_i<LPA>(cnt, &tot);
        s_pref[row][lane] = pre;
        s_start[row][lane] = start;
        s_sh[row][lane] = packed;
        if (lane == 0) s_pref[row][LPA] = tot;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int c0 = 0; c0 < tot; c0 += LPA) {
            const int c = c0 + lane;
            bool acc = false;
            int atom2 = 0, sh0 = 0, sh1 = 0, sh2 = 0;
            double d = 0.0;
            if (c < tot) {
                // offset cell of candidate c: the last lane whose prefix start is <= c
                int lo_o = 0, hi_o = LPA - 1;
                while (lo_o < hi_o) {   // prefix starts are non-decreasing over the group's lanes
                    const int mid = (lo_o + hi_o + 1) >> 1;
                    if (s_pref[row][mid] <= c) lo_o = mid;
                    else hi_o = mid - 1;
                }
                const int q = c - s_pref[row][lo_o];
                const int slot = s_start[row][lo_o] + q;
                CellShift cs2;
                CellAtom ca2;
                if constexpr (CC) {   // the cell-ordered copies: no dependent load
                    cs2 = cms[slot];
                    ca2 = cpos[slot];
                } else {              // through the cell's atom list
                    const int a2 = cell_atoms[slot];
                    cs2 = CellShift{{map_shift[3 * a2], map_shift[3 * a2 + 1], map_shift[3 * a2 + 2], a2}};
                    ca2 = CellAtom{{pos[3 * a2], pos[3 * a2 + 1], pos[3 * a2 + 2], 0.0}};
                }
                atom2 = cs2.v[3];
                const int pk = s_sh[row][lo_o];
                const int i4 = (int)(signed char)(pk & 255), j4 = (int)(signed char)((pk >> 8) & 255),
                          k4 = (int)(signed char)((pk >> 16) & 255);
                const bool same_cell = (pk >> 24) & 1;
                if (!(atom2 == atom1 && i4 == 0 && j4 == 0 && k4 == 0 && same_cell)) {
                    sh0 = i4 - ms1x + cs2.v[0];
                    sh1 = j4 - ms1y + cs2.v[1];
                    sh2 = k4 - ms1z + cs2.v[2];
                    double d0 = ca2.p[0] - p1x, d1 = ca2.p[1] - p1y, d2 = ca2.p[2] - p1z;
                    d0 = d0 + g.lat[0] * (double)sh0; d1 = d1 + g.lat[1] * (double)sh0; d2 = d2 + g.lat[2] * (double)sh0;
                    d0 = d0 + g.lat[3] * (double)sh1; d1 = d1 + g.lat[4] * (double)sh1; d2 = d2 + g.lat[5] * (double)sh1;
                    d0 = d0 + g.lat[6] * (double)sh2; d1 = d1 + g.lat[7] * (double)sh2; d2 = d2 + g.lat[8] * (double)sh2;
                    if (!(d0 > rc || d1 > rc || d2 > rc)) {
                        double dsq = d0 * d0;
                        dsq = dsq + d1 * d1;
                        dsq = dsq + d2 * d2;
                        d = sqrt(dsq);
                        acc = d < rc;
                    }
                }
            }
            const unsigned long long mask = (__ballot(acc) >> gshift) & gmask;   // this group's lanes
            if (FILL && acc) {
                const int e = base + __popcll(mask & below);
                n1_j[e] = atom2;
                n1_shift[3 * e] = sh0; n1_shift[3 * e + 1] = sh1; n1_shift[3 * e + 2] = sh2;
                n1_r[e] = d;
                if (atom2 != atom1) {      // mirror entry in neighbour2(atom2), ordered later
                    const int slot = atomicAdd(&n2_fill[atom2], 1);
                    n2_i[n2_off[atom2] + slot] = atom1;
                    n2_k[n2_off[atom2] + slot] = e - n1_off[atom1];
                }
            } else if (MODE == 2 && acc) {
                const int e = base + __popcll(mask & below), k1 = e - base1;
                if (k1 < cap1) {
                    n1_j[e] = atom2;
                    n1_shift[3 * e] = sh0; n1_shift[3 * e + 1] = sh1; n1_shift[3 * e + 2] = sh2;
                    n1_r[e] = d;
                }
                if (atom2 != atom1) {
                    const int slot = atomicAdd(&n2_cnt[atom2], 1);
                    if (slot < cap2 && k1 < cap1) {
                        n2_i[atom2 * cap2 + slot] = atom1;
                        n2_k[atom2 * cap2 + slot] = k1;
                    } else {
                        *ovf = 1;
                    }
                }
            } else if (MODE == 0 && acc && atom2 != atom1) {
                atomicAdd(&n2_cnt[atom2], 1);
            }
            base += __popcll(mask);
            total_all += __popcll(mask);
        }
        __builtin_amdgcn_wave_barrier();   // this group's LDS rows are read before the next group writes them
    }
    if (MODE != 1 && lane == 0 && live) n1_cnt[atom1] = total_all;
    if (MODE == 2 && lane == 0 && live && total_all > cap1) *ovf = 1;
}

constexpr int kNlTwoPerWave = 16384, kNlFourPerWave = 65536;
void launch_nl_pairs(int mode, int N, const double *pos, const int *map_shift, const int *cell_of,
                     const int *cell_start, const int *cell_cnt, const int *cell_atoms, const CellAtom *cpos,
                     const CellShift *cms, const NLGeom *geo, const int *sid, const int *cell_base, int *n1_cnt,
                     const int *n1_off, int *n1_j, int *n1_shift, double *n1_r, int *n2_cnt, const int *n2_off,
                     int *n2_fill, int *n2_i, int *n2_k, int cap1, int cap2, int *ovf, hipStream_t s) {
    if (N <= 0) return;
    // two atoms per wavefront once the one-atom grid would run more than twice the
    // wavefronts the GPU holds at five per SIMD (about 10K atoms on MI355X), four from
    // 64K atoms (batch512: 239 us with one, 178 with two, 170 with four)
    const int lpa = N >= kNlFourPerWave ? 16 : N >= kNlTwoPerWave ? 32 : 64;
#define NLP(M, CC, W)                                                                                           \
    hipLaunchKernelGGL((k_nl_pairs_w<M, CC, W>), dim3(cdiv(N, 4 * (64 / W))), dim3(256), 0, s, N, pos, map_shift, \
                       cell_of, cell_start, cell_cnt, cell_atoms, cpos, cms, geo, sid, cell_base, n1_cnt, n1_off, \
                       n1_j, n1_shift, n1_r, n2_cnt, n2_off, n2_fill, n2_i, n2_k, cap1, cap2, ovf)
#define NLPW(M, CC) \
    if (lpa == 16) NLP(M, CC, 16); else if (lpa == 32) NLP(M, CC, 32); else NLP(M, CC, 64)
    if (cpos) {
        if (mode == 1) NLPW(1, true); else if (mode == 2) NLPW(2, true); else NLPW(0, true);
    } else {
        if (mode == 1) NLPW(1, false); else if (mode == 2) NLPW(2, false); else NLPW(0, false);
    }
#undef NLPW
#undef NLP
}

// Per-atom slots of the one-pass search -> the CSR tables.
__global__ __launch_bounds__(256) void k_nl_compact(int N, int cap1, int cap2, const int *n1_off, const int *n2_off,
                                                    const int *t1_j, const int *t1_shift, const double *t1_r,
                                                    const int *t2_i, const int *t2_k, int *n1_j, int *n1_shift,
                                                    double *n1_r, int *n2_i, int *n2_k) {
    // 16 lanes per atom (a covalent solid has about 15 entries in each table)
    const int lane = threadIdx.x & 15;
    const int a = blockIdx.x * 16 + (threadIdx.x >> 4);
    if (a >= N) return;
    const int b1 = n1_off[a], c1 = n1_off[a + 1] - b1;
    for (int q = lane; q < c1; q += 16) {
        const size_t src = (size_t)a * cap1 + q;
        n1_j[b1 + q] = t1_j[src];
        n1_shift[3 * (b1 + q)] = t1_shift[3 * src];
        n1_shift[3 * (b1 + q) + 1] = t1_shift[3 * src + 1];
        n1_shift[3 * (b1 + q) + 2] = t1_shift[3 * src + 2];
        n1_r[b1 + q] = t1_r[src];
    }
    const int b2 = n2_off[a], c2 = n2_off[a + 1] - b2;
    for (int q = lane; q < c2; q += 16) {
        const size_t src = (size_t)a * cap2 + q;
        n2_i[b2 + q] = t2_i[src];
        n2_k[b2 + q] = t2_k[src];
    }
}
void launch_nl_compact(int N, int cap1, int cap2, const int *n1_off, const int *n2_off, const int *t1_j,
                       const int *t1_shift, const double *t1_r, const int *t2_i, const int *t2_k, int *n1_j,
                       int *n1_shift, double *n1_r, int *n2_i, int *n2_k, hipStream_t s) {
    if (N > 0)
        hipLaunchKernelGGL(k_nl_compact, dim3(cdiv(N, 16)), dim3(256), 0, s, N, cap1, cap2, n1_off, n2_off, t1_j,
                           t1_shift, t1_r, t2_i, t2_k, n1_j, n1_shift, n1_r, n2_i, n2_k);
}

// Largest neighbour1 and neighbour2 counts (atomic max of block maxima into mx[0], mx[1]).

// neighbour2(j) in the reference's append order: the global loop visits atom1
// by (cell linear index, position in its descending cell list), then by its
// neighbour1 entry index.  A group of G lanes per atom (32, or 16 for large problems,
// where halving the wavefronts pays) ranks its entries by that key (all keys distinct;
// up to 32 entries are staged in LDS first) and writes them
// in order; it also records, for every neighbour1 entry, the slot of its mirror
// in neighbour2.
// SLOT: the unsorted entries are read from the per-atom slots (atom j's at j * cap2)
// of a one-pass search, with no compacted copy.
template <bool SLOT, int G>
__global__ __launch_bounds__(256) void k_nl_n2_sort(int N, const int *n2_off, const int *n2_i, const int *n2_k,
                                                    const int *cell_of, const int *n1_off, int *n2_is, int *n2_ks,
                                                    int *n1_rev, int cap2) {
    constexpr int GPB = 256 / G, EPL = 32 / G;   // atoms per block; staged entries per lane
    __shared__ int s_i[GPB][32], s_k[GPB][32], s_c[GPB][32];   // the group's entries when they fit one row
    const int lane = threadIdx.x & (G - 1), wv = threadIdx.x / G;
    const int j = blockIdx.x * GPB + wv;
    if (j >= N) return;   // the whole group leaves together
    const int b = n2_off[j], n = n2_off[j + 1] - b;
    const size_t bs = SLOT ? (size_t)j * cap2 : (size_t)b;   // where the unsorted entries sit
    if (n <= 32) {
        int vi[EPL], vk[EPL], vc[EPL];
#pragma unroll
        for (int u = 0; u < EPL; u++) {
            const int x = lane + u * G;
            vi[u] = 0; vk[u] = 0; vc[u] = 0;
            if (x < n) {
                vi[u] = n2_i[bs + x];
                vk[u] = n2_k[bs + x];
                vc[u] = cell_of[vi[u]];
                s_i[wv][x] = vi[u];
                s_k[wv][x] = vk[u];
                s_c[wv][x] = vc[u];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int u = 0; u < EPL; u++) {
            if (lane + u * G >= n) continue;
            int rank = 0;
            for (int y = 0; y < n; y++) {
                const int yi = s_i[wv][y], yk = s_k[wv][y], yc = s_c[wv][y];
                const bool before = (yc < vc[u]) || (yc == vc[u] && (yi > vi[u] || (yi == vi[u] && yk < vk[u])));
                rank += before ? 1 : 0;
            }
            n2_is[b + rank] = vi[u];
            n2_ks[b + rank] = vk[u];
            n1_rev[n1_off[vi[u]] + vk[u]] = rank;
        }
        return;
    }
    for (int x = lane; x < n; x += G) {
        const int vi = n2_i[bs + x], vk = n2_k[bs + x], vc = cell_of[vi];
        int rank = 0;
        for (int y = 0; y < n; y++) {
            const int yi = n2_i[bs + y], yk = n2_k[bs + y], yc = cell_of[yi];
            const bool before = (yc < vc) || (yc == vc && (yi > vi || (yi == vi && yk < vk)));
            rank += before ? 1 : 0;
        }
        n2_is[b + rank] = vi;
        n2_ks[b + rank] = vk;
        n1_rev[n1_off[vi] + vk] = rank;
    }
}
void launch_nl_n2_sort(int N, const int *n2_off, const int *n2_i, const int *n2_k, const int *cell_of,
                       const int *n1_off, int *n2_is, int *n2_ks, int *n1_rev, int cap2, hipStream_t s) {
    if (N <= 0) return;
    // 16 lanes per atom (two staged entries each) from 64K atoms: half the wavefronts
    const bool g16 = N >= kNlFourPerWave;
#define N2S(SL, GG)                                                                                             \
    hipLaunchKernelGGL((k_nl_n2_sort<SL, GG>), dim3(cdiv(N, 256 / GG)), dim3(256), 0, s, N, n2_off, n2_i, n2_k, \
                       cell_of, n1_off, n2_is, n2_ks, n1_rev, SL ? cap2 : 0)
    if (cap2 > 0) {
        if (g16) N2S(true, 16); else N2S(true, 32);
    } else {
        if (g16) N2S(false, 16); else N2S(false, 32);
    }
#undef N2S
}

// Full per-atom neighbour view in connection_neighbour order (neighbour2 first,
// Connection.f95:2038-2051), with diff = pos_j - pos_i + lattice.shift summed
// m = 1..3 (:2120-2125) and the index of the mirror pair.  A group of 32 lanes
// per atom, lanes over its entries (a covalent solid has about 30).
// SLOT: the neighbour1 data (j, shift, r) are read from the per-atom slots of a one-pass
// search (atom a's entry k at a * cap1 + k) instead of the compacted tables; offsets,
// neighbour2 and the mirror slots n1_rev keep the compacted numbering.
// With pairF non-null the step's accumulators are cleared on the way: pairF (and pairE,
// if given) over every pair written, esite over every atom.
template <bool SLOT>
__global__ __launch_bounds__(256) void k_nl_full(int N, const double *__restrict__ pos,
                                                 const NLGeom *__restrict__ geo, const int *__restrict__ sid,
                                                 const int *n1_off, const int *n1_j, const int *n1_shift,
                                                 const double *n1_r, const int *n2_off, const int *n2_i,
                                                 const int *n2_k, const int *n1_rev, int *nb_off, int *nb_j,
                                                 int *nb_rev, double *nb_r, double *nb_d, int cap1, double *pairF,
                                                 double *pairE, double *esite) {
#pragma clang fp contract(off)
    auto at1 = [&](int a, int k) -> size_t { return SLOT ? (size_t)a * cap1 + k : (size_t)n1_off[a] + k; };
    const int lane = threadIdx.x & 31;
    const int i = blockIdx.x * 8 + (threadIdx.x >> 5);
    if (i > N) return;
    const int base = n1_off[i] + n2_off[i];
    if (lane == 0) nb_off[i] = base;
    if (i == N) return;
    if (pairF && lane == 0) esite[i] = 0.0;
    const NLGeom &g = geo[sid[i]];
    const int n2c = n2_off[i + 1] - n2_off[i];
    const int n1c = n1_off[i + 1] - n1_off[i];
    for (int q = lane; q < n2c + n1c; q += 32) {
        int j, sgn, rev;
        size_t e1;   // the neighbour1 entry's data
        if (q < n2c) {
            j = n2_i[n2_off[i] + q];
            e1 = at1(j, n2_k[n2_off[i] + q]);
            sgn = -1;
            rev = n1_off[j] + n2_off[j] + (n2_off[j + 1] - n2_off[j]) + n2_k[n2_off[i] + q];
        } else {
            e1 = at1(i, q - n2c);
            j = n1_j[e1];
            sgn = 1;
            if (j != i) rev = n1_off[j] + n2_off[j] + n1_rev[n1_off[i] + (q - n2c)];
            else {
                rev = -1;
                for (int k = 0; k < n1c; k++) {
                    const size_t t = at1(i, k);
                    if (n1_j[t] == i && n1_shift[3 * t] == -n1_shift[3 * e1] && n1_shift[3 * t + 1] == -n1_shift[3 * e1 + 1] &&
                        n1_shift[3 * t + 2] == -n1_shift[3 * e1 + 2]) { rev = base + n2c + k; break; }
                }
                if (rev < 0) rev = base + q;   // mirror image missing (cannot happen for an exact cutoff test)
            }
        }
        const int s0 = sgn * n1_shift[3 * e1], s1 = sgn * n1_shift[3 * e1 + 1], s2 = sgn * n1_shift[3 * e1 + 2];
        double d0 = pos[3 * j] - pos[3 * i], d1 = pos[3 * j + 1] - pos[3 * i + 1], d2 = pos[3 * j + 2] - pos[3 * i + 2];
        d0 = d0 + g.lat[0] * (double)s0; d1 = d1 + g.lat[1] * (double)s0; d2 = d2 + g.lat[2] * (double)s0;
        d0 = d0 + g.lat[3] * (double)s1; d1 = d1 + g.lat[4] * (double)s1; d2 = d2 + g.lat[5] * (double)s1;
        d0 = d0 + g.lat[6] * (double)s2; d1 = d1 + g.lat[7] * (double)s2; d2 = d2 + g.lat[8] * (double)s2;
        const int p = base + q;
        nb_j[p] = j;
        nb_rev[p] = rev;
        nb_r[p] = n1_r[e1];
        nb_d[3 * p] = d0; nb_d[3 * p + 1] = d1; nb_d[3 * p + 2] = d2;
        if (pairF) {
            pairF[3 * p] = 0.0; pairF[3 * p + 1] = 0.0; pairF[3 * p + 2] = 0.0;
            if (pairE) pairE[p] = 0.0;
        }
    }
}
void launch_nl_full(int N, const double *pos, const NLGeom *geo, const int *sid, const int *n1_off, const int *n1_j,
                    const int *n1_shift, const double *n1_r, const int *n2_off, const int *n2_i, const int *n2_k,
                    const int *n1_rev, int *nb_off, int *nb_j, int *nb_rev, double *nb_r, double *nb_d, int cap1,
                    hipStream_t s, double *pairF, double *pairE, double *esite) {
    if (cap1 > 0)
        hipLaunchKernelGGL(k_nl_full<true>, dim3(cdiv(N + 1, 8)), dim3(256), 0, s, N, pos, geo, sid, n1_off, n1_j,
                           n1_shift, n1_r, n2_off, n2_i, n2_k, n1_rev, nb_off, nb_j, nb_rev, nb_r, nb_d, cap1, pairF,
                           pairE, esite);
    else
        hipLaunchKernelGGL(k_nl_full<false>, dim3(cdiv(N + 1, 8)), dim3(256), 0, s, N, pos, geo, sid, n1_off, n1_j,
                           n1_shift, n1_r, n2_off, n2_i, n2_k, n1_rev, nb_off, nb_j, nb_rev, nb_r, nb_d, 0, pairF,
                           pairE, esite);
}

// Skin reuse (calc_connect with cutoff_skin, Connection.f95:1085-1127).
// Largest displacement of any atom since the positions `last` of the last build,
// minimum image over periodic axes (fractional displacement wrapped to the
// nearest lattice vector), as the bits of a non-negative double in *out (atomicMax).
__global__ void k_nl_max_disp(int N, const double *__restrict__ pos, const double *__restrict__ last,
                              const NLGeom *__restrict__ geo, const int *__restrict__ sid,
                              unsigned long long *out) {
    __shared__ double red[256];
    double mx = 0.0;
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) {
        const NLGeom &g = geo[sid[n]];
        const double dx = pos[3 * n] - last[3 * n], dy = pos[3 * n + 1] - last[3 * n + 1],
                     dz = pos[3 * n + 2] - last[3 * n + 2];
        // fractional displacement through the inverse of the true lattice (kept in lat_inv order)
        double f[3];
        for (int a = 0; a < 3; a++) f[a] = g.ginv[a] * dx + g.ginv[3 + a] * dy + g.ginv[6 + a] * dz;
        for (int a = 0; a < 3; a++)
            if (g.pbc[a]) f[a] -= rint(f[a]);
        double c[3];
        for (int k = 0; k < 3; k++) c[k] = g.lat[k] * f[0] + g.lat[3 + k] * f[1] + g.lat[6 + k] * f[2];
        mx = fmax(mx, sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]));
    }
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, (unsigned long long)__double_as_longlong(red[0]));
}

// connection_calc_dists (Connection.f95:1321-1445) on the stored neighbour1
// entries: r_ij = norm(pos_j + lattice.shift - pos_i) with j_pos formed as
// pos_j + ((L1 s1 + L2 s2) + L3 s3).  The reference visits every pair from both
// ends and the later visit wins, so for i < j the stored value is computed from
// j's side (shift negated); self images from i's side.  r < 1e-10 sets *overlap.
__global__ void k_nl_dists(int N, const double *__restrict__ pos, const NLGeom *__restrict__ geo,
                           const int *__restrict__ sid, const int *__restrict__ n1_off, const int *__restrict__ n1_j,
                           const int *__restrict__ n1_shift, double *n1_r, int *overlap) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const NLGeom &g = geo[sid[i]];
    for (int e = n1_off[i]; e < n1_off[i + 1]; e++) {
        const int j = n1_j[e];
        int s0 = n1_shift[3 * e], s1 = n1_shift[3 * e + 1], s2 = n1_shift[3 * e + 2];
        int from = i, to = j;          // r = norm(pos(to) + L.s - pos(from))
        if (j != i) { from = j; to = i; s0 = -s0; s1 = -s1; s2 = -s2; }
        double v[3];
        for (int k = 0; k < 3; k++) {
            double t = g.lat[k] * (double)s0;
            t = t + g.lat[3 + k] * (double)s1;
            t = t + g.lat[6 + k] * (double)s2;
            v[k] = (pos[3 * to + k] + t) - pos[3 * from + k];
        }
        double dsq = v[0] * v[0];
        dsq = dsq + v[1] * v[1];
        dsq = dsq + v[2] * v[2];
        const double r = sqrt(dsq);
        if (r < 1e-10) atomicOr(overlap, 1);
        n1_r[e] = r;
    }
}

void launch_nl_skin_check(int N, const double *pos, const double *last, const NLGeom *geo, const int *sid,
                          const int *n1_off, const int *n1_j, const int *n1_shift, double *n1_r,
                          unsigned long long *max_disp, int *overlap, hipStream_t s) {
    if (N <= 0) return;
    hipLaunchKernelGGL(k_nl_max_disp, dim3(std::min(cdiv(N, 256), 1024)), dim3(256), 0, s, N, pos, last, geo, sid,
                       max_disp);
    hipLaunchKernelGGL(k_nl_dists, dim3(cdiv(N, 256)), dim3(256), 0, s, N, pos, geo, sid, n1_off, n1_j, n1_shift,
                       n1_r, overlap);
}

// min / max of fractional coordinates (g .mult. pos) per axis, for calc_lat_eff;
// one block per structure, g = the inverse of that structure's lattice.
__global__ void k_minmax_frac(const int *atom_off, const double *pos, const double *gmat, double *out) {
#pragma clang fp contract(off)
    __shared__ double smin[3][256], smax[3][256];
    const int st = blockIdx.x;
    double g[9];
    for (int k = 0; k < 9; k++) g[k] = gmat[9 * st + k];
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int n = atom_off[st] + threadIdx.x; n < atom_off[st + 1]; n += blockDim.x) {
        double lp[3];
        mult3_exact(g, pos[3 * n], pos[3 * n + 1], pos[3 * n + 2], lp);
        for (int a = 0; a < 3; a++) { mn[a] = fmin(mn[a], lp[a]); mx[a] = fmax(mx[a], lp[a]); }
    }
    for (int a = 0; a < 3; a++) { smin[a][threadIdx.x] = mn[a]; smax[a][threadIdx.x] = mx[a]; }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int a = 0; a < 3; a++) {
                smin[a][threadIdx.x] = fmin(smin[a][threadIdx.x], smin[a][threadIdx.x + s]);
                smax[a][threadIdx.x] = fmax(smax[a][threadIdx.x], smax[a][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int a = 0; a < 3; a++) { out[6 * st + a] = smin[a][0]; out[6 * st + 3 + a] = smax[a][0]; }
}
void launch_minmax_frac(int S, const int *atom_off, const double *pos, const double *g, double *out, hipStream_t s) {
    if (S > 0) hipLaunchKernelGGL(k_minmax_frac, dim3(S), dim3(256), 0, s, atom_off, pos, g, out);
}

// calc_lat_eff (Connection.f95:740-763) and divide_cell (:1674-1691) on the device for
// structures with a non-periodic axis, in the host's operation order (no contraction):
// the fractional extents of this call give lat_eff, its inverse and the offsets in
// geo[st]; the cell counts were fixed on the host from an earlier call's extents, and a
// structure whose counts would differ sets flags[0] |= 4 (the call is then redone).
__device__ __forceinline__ void cross3_d(const double *x, const double *y, double *o) {
#pragma clang fp contract(off)
    o[0] = +(x[1] * y[2] - x[2] * y[1]);
    o[1] = -(x[0] * y[2] - x[2] * y[0]);
    o[2] = +(x[0] * y[1] - x[1] * y[0]);
}
__device__ __forceinline__ double triple3_d(const double *x, const double *y, const double *z) {
#pragma clang fp contract(off)
    return +x[0] * (y[1] * z[2] - y[2] * z[1]) - x[1] * (y[0] * z[2] - y[2] * z[0]) + x[2] * (y[0] * z[1] - y[1] * z[0]);
}
__device__ __forceinline__ double norm3_d(const double *v) {
#pragma clang fp contract(off)
    return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}
__global__ void k_geom_np(int S, const int *atom_off, const double *mm, NLGeom *geo, int *flags) {
#pragma clang fp contract(off)
    const int st = blockIdx.x * blockDim.x + threadIdx.x;
    if (st >= S) return;
    NLGeom &g = geo[st];
    if (g.pbc[0] && g.pbc[1] && g.pbc[2]) return;
    if (atom_off[st + 1] <= atom_off[st]) return;
    double eff[9], off[3];
    for (int k = 0; k < 9; k++) eff[k] = g.lat[k];
    for (int a = 0; a < 3; a++) off[a] = 0.0;
    for (int i = 0; i < 3; i++) {
        if (g.pbc[i]) continue;
        const double mn = mm[6 * st + i], mx = mm[6 * st + 3 + i];
        if (mn <= 0.5 || mx >= 0.5) {
            const double ext = (mx - mn) * 1.1, nv = norm3_d(eff + 3 * i);
            for (int k = 0; k < 3; k++) eff[3 * i + k] = eff[3 * i + k] / nv * ext;
            off[i] = mn - 0.05;
        }
    }
    const double stp = triple3_d(eff, eff + 3, eff + 6);
    double inv[9], c[3];
    for (int r = 0; r < 3; r++) {
        cross3_d(eff + 3 * ((r + 1) % 3), eff + 3 * ((r + 2) % 3), c);
        for (int k = 0; k < 3; k++) inv[3 * k + r] = c[k] / stp;
    }
    const double vol = fabs(stp);
    bool same = true;
    for (int a = 0; a < 3; a++) {
        cross3_d(eff + 3 * ((a + 1) % 3), eff + 3 * ((a + 2) % 3), c);
        const int n = max(1, (int)(vol / (g.cutoff * norm3_d(c))));
        same = same && n == g.cells[a];
    }
    for (int k = 0; k < 9; k++) g.inv_eff[k] = inv[k];
    for (int a = 0; a < 3; a++) g.offset[a] = off[a];
    if (!same) atomicOr(flags, 4);
}
void launch_geom_np(int S, const int *atom_off, const double *mm, NLGeom *geo, int *flags, hipStream_t s) {
    if (S > 0) hipLaunchKernelGGL(k_geom_np, dim3(cdiv(S, 64)), dim3(64), 0, s, S, atom_off, mm, geo, flags);
}

// Overflow guard of a neighbour list built into capacity-sized tables without a host
// read-back: when a per-atom slot overflowed (flags[0] != 0) or the neighbour1 total
// flags[3] exceeds the tables' capacity cap, every offset is set to zero, so the
// kernels after it see empty lists and stay inside their buffers; flags[0] gets bit 2
// for a capacity overflow.  The host redoes such a call.
__global__ void k_nl_guard(int N, int *flags, int cap, int *n1_off, int *n2_off) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int tot = flags[3];
    const bool over = tot > cap;
    if (!(over || flags[0] != 0) || i > N) return;
    n1_off[i] = 0;
    n2_off[i] = 0;
    if (i == 0 && over) atomicOr(flags, 2);
}
void launch_nl_guard(int N, int *flags, int cap, int *n1_off, int *n2_off, hipStream_t s) {
    hipLaunchKernelGGL(k_nl_guard, dim3(cdiv(N + 1, 256)), dim3(256), 0, s, N, flags, cap, n1_off, n2_off);
}

// zero the first mult * count doubles of p, count = *cnt read on the device
__global__ void k_zero_counted(double *p, const int *cnt, int mult) {
    const size_t n = (size_t)(*cnt) * mult;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        p[q] = 0.0;
}
void launch_zero_counted(double *p, const int *cnt, int mult, size_t cap, hipStream_t s) {
    const int blocks = std::max(1, std::min(cdiv((long)cap, 256), 2048));
    hipLaunchKernelGGL(k_zero_counted, dim3(blocks), dim3(256), 0, s, p, cnt, mult);
}

// The per-step clears in one launch: pair forces (3 per pair) and pair energies
// (nullable) of the first *cnt pairs (cnt on the device) or of n_pairs when cnt is
// null, and the N site energies.
__global__ void k_zero_step(double *pairF, double *pairE, double *esite, int N, const int *cnt, int n_pairs) {
    const size_t np = cnt ? (size_t)(*cnt) : (size_t)n_pairs;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (size_t q = i0; q < 3 * np; q += stride) pairF[q] = 0.0;
    if (pairE)
        for (size_t q = i0; q < np; q += stride) pairE[q] = 0.0;
    for (size_t q = i0; q < (size_t)N; q += stride) esite[q] = 0.0;
}
void launch_zero_step(double *pairF, double *pairE, double *esite, int N, const int *cnt, int n_pairs,
                      size_t cap_pairs, hipStream_t s) {
    const size_t mx = std::max<size_t>(3 * cap_pairs, (size_t)N);
    const int blocks = std::max(1, std::min(cdiv((long)mx, 256), 2048));
    hipLaunchKernelGGL(k_zero_step, dim3(blocks), dim3(256), 0, s, pairF, pairE, esite, N, cnt, n_pairs);
}

// ----------------------------------------------------------------------------
// SOAP
// ----------------------------------------------------------------------------
__device__ __forceinline__ double fcut_d(double r, double rc, double w) {
    if (r > rc) return 0.0;
    if (r > rc - w) return 0.5 * (cos(GAPX_PI * (r - rc + w) / w) + 1.0);
    return 1.0;
}
__device__ __forceinline__ double dfcut_d(double r, double rc, double w) {
    if (r > rc) return 0.0;
    if (r > rc - w) return -0.5 * GAPX_PI * sin(GAPX_PI * (r - rc + w) / w) / w;
    return 0.0;
}

__device__ __forceinline__ double wave_sum(double v) {
    // fixed butterfly tree over the 64 lanes of a wavefront (deterministic)
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ int species_of(const SoapDev &sp, int z) {
    for (int s = 0; s < sp.S; s++)
        if (sp.species_Z[s] == z || sp.species_Z[s] == 0) return s;
    return -1;
}

// Radial functions f_l(r; r_a) = exp(-alpha(r^2+r_a^2)) i_l(2 alpha r r_a) by the
// upward recursion i_(l+1) = i_(l-1) - (2l+1)/x i_l started from
// i_(-1) = cosh(x)/x and i_0 = sinh(x)/x.  State per grid point: F_l and F_(l+1).
template <int NMAX>
struct RadialState {
    double F0[NMAX], F1[NMAX], invx[NMAX];
    __device__ __forceinline__ void init(const SoapDev &sp, double r) {
#pragma unroll
        for (int a = 0; a < NMAX; a++) {
            if (a >= sp.n) { F0[a] = F1[a] = invx[a] = 0.0; continue; }
            double ra = sp.rb[a];
            double x = 2.0 * sp.alpha * r * ra;
            if (x == 0.0) {
                F0[a] = exp(-sp.alpha * (ra * ra + r * r));
                F1[a] = 0.0;
                invx[a] = 0.0;
            } else {
                double ep = exp(-sp.alpha * (r + ra) * (r + ra)), em = exp(-sp.alpha * (r - ra) * (r - ra));
                double ix = 1.0 / x;
                double fm = 0.5 * (em + ep) * ix;
                F0[a] = 0.5 * (em - ep) * ix;
                F1[a] = fm - F0[a] * ix;
                invx[a] = ix;
            }
        }
    }
    // advance from (F_l, F_(l+1)) to (F_(l+1), F_(l+2))
    __device__ __forceinline__ void step(int l) {
#pragma unroll
        for (int a = 0; a < NMAX; a++) {
            double n2 = F0[a] - (2.0 * l + 3.0) * invx[a] * F1[a];
            F0[a] = F1[a];
            F1[a] = (invx[a] == 0.0) ? 0.0 : n2;
        }
    }
};

template <int L, int NMAX>
__global__ __launch_bounds__(64) void k_soap_fwd(SoapDev sp, const int *__restrict__ centres, const int *__restrict__ Z,
                                                 const int *__restrict__ nb_off, const int *__restrict__ nb_j,
                                                 const double *__restrict__ nb_r, const double *__restrict__ nb_d,
                                                 double *X, double *norm_out, double *cbuf, int NB, int rows_cap) {
    extern __shared__ double lds[];
    constexpr int LM = (L + 1) * (L + 1);
    const int n = sp.n, KLM = sp.K * LM, RL = n * (L + 1) + LM;
    double *cL = lds;                         // K*LM coefficients
    double *rows = lds + KLM;                 // NB rows of [g (n*(L+1)) | Y (LM)], later the power spectrum
    int *spc = (int *)(rows + rows_cap);
    const int lane = threadIdx.x;
    const int desc = blockIdx.x;
    const int ci = centres[desc];
    for (int e = lane; e < KLM; e += 64) cL[e] = 0.0;
    __syncthreads();
    if (lane == 0) {
        int sc = species_of(sp, Z[ci]);
        if (sc >= 0) cL[(sc * n) * LM] += sp.cw * sp.u00 * c_ynorm[0];
    }
    const int p0 = nb_off[ci], np = nb_off[ci + 1] - p0;
    for (int c0 = 0; c0 < np; c0 += NB) {
        const int nk = min(NB, np - c0);
        __syncthreads();
        if (lane < nk) {
            const int p = p0 + c0 + lane;
            double r = nb_r[p];
            int s = species_of(sp, Z[nb_j[p]]);
            double *row = rows + (size_t)lane * RL;
            if (r > sp.rc || s < 0) {
                s = -1;
            } else {
                double fc = fcut_d(r, sp.rc, sp.ctw);
                RadialState<NMAX> rs;
                rs.init(sp, r);
                for (int l = 0; l <= L; l++) {
                    for (int b = 0; b < n; b++) {
                        double acc = 0.0;
#pragma unroll
                        for (int a = 0; a < NMAX; a++)
                            if (a < n) acc += rs.F0[a] * sp.T[a * n + b];
                        row[b * (L + 1) + l] = fc * acc;
                    }
                    rs.step(l);
                }
                // real spherical harmonics of u = d/r
                double ir = 1.0 / r;
                double ux = nb_d[3 * p] * ir, uy = nb_d[3 * p + 1] * ir, uz = nb_d[3 * p + 2] * ir;
                double Pm1[L + 2], Pm2[L + 2], Pc[L + 2];
                double *Yr = row + n * (L + 1);
#pragma unroll
                for (int l = 0; l <= L; l++) {
#pragma unroll
                    for (int m = 0; m <= l; m++) {
                        if (m == l) Pc[m] = c_dfact[m];
                        else if (m == l - 1) Pc[m] = (2.0 * l - 1.0) * uz * Pm1[m];
                        else Pc[m] = ((2.0 * l - 1.0) * uz * Pm1[m] - (double)(l + m - 1) * Pm2[m]) * (1.0 / (double)(l - m));
                    }
                    double A = 1.0, B = 0.0;
#pragma unroll
                    for (int m = 0; m <= l; m++) {
                        if (m > 0) { double An = A * ux - B * uy; B = A * uy + B * ux; A = An; }
                        double N = c_ynorm[l * (l + 1) / 2 + m] * Pc[m];
                        if (m == 0) Yr[l * l + l] = N;
                        else { Yr[l * l + l + m] = N * A; Yr[l * l + l - m] = N * B; }
                    }
#pragma unroll
                    for (int m = 0; m <= l; m++) { Pm2[m] = Pm1[m]; Pm1[m] = Pc[m]; }
                }
            }
            spc[lane] = s;
        }
        __syncthreads();
        for (int e = lane; e < KLM; e += 64) {
            int s = e / (n * LM), rem = e % (n * LM), b = rem / LM, lm = rem % LM;
            int l = (int)sqrt((double)lm + 0.5);
            if (l * l > lm) l--;
            double acc = 0.0;
            for (int k = 0; k < nk; k++) {
                if (spc[k] != s) continue;
                const double *row = rows + (size_t)k * RL;
                acc += row[b * (L + 1) + l] * row[n * (L + 1) + lm];
            }
            cL[e] += acc;
        }
    }
    __syncthreads();
    // power spectrum -> rows area (reused), then normalise
    double *pb = rows;
    const int d = sp.dim;
    double ss = 0.0;
    for (int q = lane; q < d - 1; q += 64) {
        int code = sp.qmap[q];
        int ia = code >> 16, jb = (code >> 8) & 0xff, l = code & 0xff;
        double v = 0.0;
        for (int m = -l; m <= l; m++) v += cL[ia * LM + l * l + l + m] * cL[jb * LM + l * l + l + m];
        v *= (ia != jb ? 1.4142135623730950488 : 1.0) / sqrt(2.0 * l + 1.0);
        pb[q] = v;
        ss += v * v;
    }
    ss = wave_sum(ss);
    double nrm = sqrt(ss);
    double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
    double *xo = X + (size_t)desc * d;
    for (int q = lane; q < d - 1; q += 64) xo[q] = pb[q] * inv;
    if (lane == 0) { xo[d - 1] = 0.0; norm_out[desc] = nrm; }
    double *co = cbuf + (size_t)desc * KLM;
    for (int e = lane; e < KLM; e += 64) co[e] = cL[e];
}

template <int L, int NMAX>
__global__ __launch_bounds__(64) void k_soap_bwd(SoapDev sp, const int *__restrict__ centres, const int *__restrict__ Z,
                                                 const int *__restrict__ nb_off, const int *__restrict__ nb_j,
                                                 const double *__restrict__ nb_r, const double *__restrict__ nb_d,
                                                 const double *__restrict__ X, const double *__restrict__ G,
                                                 const double *__restrict__ norm, const double *__restrict__ cbuf,
                                                 double *pairF) {
    extern __shared__ double lds[];
    constexpr int LM = (L + 1) * (L + 1);
    const int n = sp.n, K = sp.K, KLM = K * LM, d = sp.dim;
    double *cL = lds;              // K*LM
    double *dcL = lds + KLM;       // K*LM
    double *dpL = lds + 2 * KLM;   // d
    const int lane = threadIdx.x;
    const int desc = blockIdx.x;
    const int ci = centres[desc];
    const double *x = X + (size_t)desc * d, *gg = G + (size_t)desc * d;
    double dot = 0.0;
    for (int q = lane; q < d; q += 64) dot += x[q] * gg[q];
    dot = wave_sum(dot);
    const double nrm = norm[desc];
    const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
    for (int q = lane; q < d; q += 64) dpL[q] = (gg[q] - x[q] * dot) * inv;
    for (int e = lane; e < KLM; e += 64) cL[e] = cbuf[(size_t)desc * KLM + e];
    __syncthreads();
    // dE/dc_(ia,lm) = sum_jb W_l(ia,jb) c_(jb,lm); W_l = dp * (2 on the diagonal, sqrt2 off it) / sqrt(2l+1)
    for (int e = lane; e < KLM; e += 64) {
        int ia = e / LM, lm = e % LM;
        int l = (int)sqrt((double)lm + 0.5);
        if (l * l > lm) l--;
        double acc = 0.0;
        for (int jb = 0; jb < K; jb++) {
            int hi = ia > jb ? ia : jb, lo = ia > jb ? jb : ia;
            int q = (hi * (hi + 1) / 2 + lo) * (L + 1) + l;
            double w = dpL[q] * (ia == jb ? 2.0 : 1.4142135623730950488);
            acc += w * cL[jb * LM + lm];
        }
        dcL[e] = acc / sqrt(2.0 * l + 1.0);
    }
    __syncthreads();
    const int p0 = nb_off[ci], np = nb_off[ci + 1] - p0;
    for (int k = lane; k < np; k += 64) {
        const int p = p0 + k;
        const double r = nb_r[p];
        const int s = species_of(sp, Z[nb_j[p]]);
        if (r > sp.rc || s < 0) continue;
        const double fc = fcut_d(r, sp.rc, sp.ctw), dfc = dfcut_d(r, sp.rc, sp.ctw);
        const double ir = 1.0 / r;
        const double ux = nb_d[3 * p] * ir, uy = nb_d[3 * p + 1] * ir, uz = nb_d[3 * p + 2] * ir;
        const double *dcs = dcL + (size_t)s * n * LM;
        RadialState<NMAX> rs;
        rs.init(sp, r);
        double fx = 0.0, fy = 0.0, fz = 0.0;
        double Pm1[L + 2], Pm2[L + 2], Pc[L + 2];
#pragma unroll
        for (int l = 0; l <= L; l++) {
            double gb[NMAX], dgb[NMAX];
            const double cl = -2.0 * sp.alpha * r + (double)l * ir;
#pragma unroll
            for (int b = 0; b < NMAX; b++) {
                double g0 = 0.0, q1 = 0.0;
                if (b < n) {
#pragma unroll
                    for (int a = 0; a < NMAX; a++)
                        if (a < n) {
                            double t = sp.T[a * n + b];
                            g0 += rs.F0[a] * t;
                            q1 += sp.rb[a] * rs.F1[a] * t;
                        }
                }
                double dg0 = cl * g0 + 2.0 * sp.alpha * q1;
                gb[b] = fc * g0;
                dgb[b] = fc * dg0 + dfc * g0;
            }
            rs.step(l);
#pragma unroll
            for (int m = 0; m <= l; m++) {
                if (m == l) Pc[m] = c_dfact[m];
                else if (m == l - 1) Pc[m] = (2.0 * l - 1.0) * uz * Pm1[m];
                else Pc[m] = ((2.0 * l - 1.0) * uz * Pm1[m] - (double)(l + m - 1) * Pm2[m]) * (1.0 / (double)(l - m));
            }
            Pc[l + 1] = 0.0;
            double A = 1.0, B = 0.0, Am = 0.0, Bm = 0.0;
#pragma unroll
            for (int m = 0; m <= l; m++) {
                if (m > 0) { Am = A; Bm = B; double An = A * ux - B * uy; B = A * uy + B * ux; A = An; }
                const double N = c_ynorm[l * (l + 1) / 2 + m];
                const double P = Pc[m], Pz = Pc[m + 1];
#pragma unroll
                for (int sgn = 0; sgn < 2; sgn++) {
                    if (m == 0 && sgn == 1) continue;
                    double Y, Fx, Fy, Fz;
                    if (m == 0) { Y = N * P; Fx = 0.0; Fy = 0.0; Fz = N * Pz; }
                    else if (sgn == 0) { Y = N * P * A; Fx = N * P * m * Am; Fy = -N * P * m * Bm; Fz = N * Pz * A; }
                    else { Y = N * P * B; Fx = N * P * m * Bm; Fy = N * P * m * Am; Fz = N * Pz * B; }
                    const int lm = l * l + l + (sgn ? -m : m);
                    double s1 = 0.0, s2 = 0.0;
#pragma unroll
                    for (int b = 0; b < NMAX; b++)
                        if (b < n) {
                            double dc = dcs[b * LM + lm];
                            s1 += dc * dgb[b];
                            s2 += dc * gb[b];
                        }
                    // grad of Y(r/|r|) = (grad F - u (u . grad F)) / r
                    const double ud = ux * Fx + uy * Fy + uz * Fz;
                    const double a1 = s1 * Y - s2 * ud * ir, a2 = s2 * ir;
                    fx += a1 * ux + a2 * Fx;
                    fy += a1 * uy + a2 * Fy;
                    fz += a1 * uz + a2 * Fz;
                }
            }
#pragma unroll
            for (int m = 0; m <= l; m++) { Pm2[m] = Pm1[m]; Pm1[m] = Pc[m]; }
        }
        pairF[3 * p] += fx;
        pairF[3 * p + 1] += fy;
        pairF[3 * p + 2] += fz;
    }
}

// instantiation table
#define SOAP_CASES(X) X(2, 8) X(3, 8) X(4, 8) X(5, 8) X(6, 8) X(7, 8) X(8, 8) X(10, 8) X(12, 8) \
                      X(2, 16) X(3, 16) X(4, 16) X(5, 16) X(6, 16) X(7, 16) X(8, 16) X(10, 16) X(12, 16)

int soap_supported(int n, int L) {
    if (n < 1 || n > 16) return 0;
#define CHK(LL, NN) if (L == LL) return 1;
    SOAP_CASES(CHK)
#undef CHK
    return 0;
}

static int pick_nmax(int n) { return n <= 8 ? 8 : 16; }

// generic forward: c' accumulators plus the staged neighbour rows of NB neighbours
static size_t fwd_generic_lds(const SoapDev &sp, int NB) {
    const int LM = (sp.L + 1) * (sp.L + 1);
    const int RL = sp.n * (sp.L + 1) + LM;
    const int KLM = sp.K * LM;
    return (size_t)(KLM + std::max((size_t)NB * RL, (size_t)sp.dim)) * 8 + NB * 4;
}
static int fwd_generic_nb(const SoapDev &sp) {
    int NB = 64;
    while (NB > 8 && fwd_generic_lds(sp, NB) > 60 * 1024) NB /= 2;
    return NB;
}
static size_t bwd_generic_lds(const SoapDev &sp) {
    const int LM = (sp.L + 1) * (sp.L + 1);
    return (size_t)(2 * sp.K * LM + sp.dim) * 8;
}

size_t soap_lds_bytes(const SoapDev &sp) {
    if (soap_fast_supported(sp.n, sp.L)) return soap_fast_lds_bytes(sp);
    return std::max(fwd_generic_lds(sp, fwd_generic_nb(sp)), bwd_generic_lds(sp));
}

void launch_soap_fwd(const SoapDev &sp, int nc, const int *centres, const int *Z, const int *nb_off,
                     const int *nb_j, const double *nb_r, const double *nb_d, double *X, double *norm,
                     double *cbuf, hipStream_t s) {
    if (launch_soap_fwd_fast(sp, nc, centres, Z, nb_off, nb_j, nb_r, nb_d, X, norm, cbuf, s)) return;
    ensure_constants();
    if (nc <= 0) return;
    const int LM = (sp.L + 1) * (sp.L + 1);
    const int RL = sp.n * (sp.L + 1) + LM;
    const int NB = fwd_generic_nb(sp);
    const int rows_cap = (int)std::max((size_t)NB * RL, (size_t)sp.dim);
    size_t sh = fwd_generic_lds(sp, NB);
    const int nm = pick_nmax(sp.n);
#define LAUNCH(LL, NN)                                                                                          \
    if (sp.L == LL && nm == NN) {                                                                               \
        hipLaunchKernelGGL((k_soap_fwd<LL, NN>), dim3(nc), dim3(64), sh, s, sp, centres, Z, nb_off, nb_j, nb_r, \
                           nb_d, X, norm, cbuf, NB, rows_cap);                                                            \
        return;                                                                                                 \
    }
    SOAP_CASES(LAUNCH)
#undef LAUNCH
}

void launch_soap_bwd(const SoapDev &sp, int nc, const int *centres, const int *Z, const int *nb_off,
                     const int *nb_j, const double *nb_r, const double *nb_d, const double *X, const double *G,
                     const double *norm, const double *cbuf, double *pairF, hipStream_t s) {
    if (launch_soap_bwd_fast(sp, nc, centres, Z, nb_off, nb_j, nb_r, nb_d, X, G, norm, cbuf, pairF, s)) return;
    ensure_constants();
    if (nc <= 0) return;
    size_t sh = bwd_generic_lds(sp);
    const int nm = pick_nmax(sp.n);
#define LAUNCH(LL, NN)                                                                                          \
    if (sp.L == LL && nm == NN) {                                                                               \
        hipLaunchKernelGGL((k_soap_bwd<LL, NN>), dim3(nc), dim3(64), sh, s, sp, centres, Z, nb_off, nb_j, nb_r, \
                           nb_d, X, G, norm, cbuf, pairF);                                                      \
        return;                                                                                                 \
    }
    SOAP_CASES(LAUNCH)
#undef LAUNCH
}

// ----------------------------------------------------------------------------
// FP64 MFMA GEMMs.  Block tile 64x64, k-step 16, 4 wavefronts in a 2x2 grid,
// each wavefront owns 32x32 = 2x2 tiles of v_mfma_f64_16x16x4f64.
// Operand layout of the 16x16x4 f64 instruction (wave64):
//   A: lane l holds A[l%16][l/16];  B: lane l holds B[l/16][l%16];
//   C: lane l holds C[(l/16) + 4*v][l%16], v = 0..3 (rows interleaved by 4).
// Tiles are staged k-major in LDS so both fragment reads are unit-stride.
// ----------------------------------------------------------------------------
typedef double double4_t __attribute__((ext_vector_type(4)));

constexpr int GPAD = 1;   // odd: the 16 k-rows one stash write touches fall on distinct bank pairs

// C = A . B (+ fused GP epilogue).  A is row-major (rows x K, k contiguous);
// B is row-major (cols x K) when BT, else (K x cols).  4 wavefronts in a 2x2
// grid, each owning (16 WM) x (16 WN) outputs, block tile (32 WM) x (32 WN).
// The next k tile is fetched into registers while the current one feeds the
// MFMAs from LDS.  blockIdx.z selects a k range (split-K) whose partial result
// goes to C + z * rows * cols.
// The occupancy hint keeps the accumulators in ordinary VGPRs for the whole kernel:
// without it the compiler places them in AGPRs and copies all of them out and back
// around every k tile.
template <bool BT, bool EPI, int WM, int WN, int GBK>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(GBK == 16 ? 3 : 2))) void k_gemm_f64(int Mr, int Nc, int Kd, int kchunk, const double *__restrict__ A,
                                                  const double *__restrict__ B, double *__restrict__ C,
                                                  const double *__restrict__ wvec, double zeta, int izeta,
                                                  double *__restrict__ e_part) {
    constexpr int BM = 32 * WM, BN = 32 * WN;
    constexpr int AL = BM * GBK / 256;   // A elements per thread per k tile
    constexpr int BL = BN * GBK / 256;
    __shared__ double As[GBK][BM + GPAD];
    __shared__ double Bs[GBK][BN + GPAD];
    __shared__ double red[2][BM];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w >> 1, wn = w & 1;
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    const int kbeg = blockIdx.z * kchunk, kend = min(Kd, kbeg + kchunk);
    C += (size_t)blockIdx.z * Mr * Nc;
    double4_t acc[WM][WN];
#pragma unroll
    for (int i = 0; i < WM; i++)
#pragma unroll
        for (int j = 0; j < WN; j++) acc[i][j] = (double4_t){0.0, 0.0, 0.0, 0.0};
    double ra[AL], rb[BL];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int e = 0; e < AL; e++) {      // A: GBK consecutive k of one row per GBK threads
            const int idx = t + 256 * e;
            const int row = idx / GBK, kk = idx % GBK;
            const int gr = m0 + row, gk = k0 + kk;
            ra[e] = (gr < Mr && gk < kend) ? A[(size_t)gr * Kd + gk] : 0.0;
        }
#pragma unroll
        for (int e = 0; e < BL; e++) {
            const int idx = t + 256 * e;
            if (BT) {
                const int col = idx / GBK, kk = idx % GBK;
                const int gc = n0 + col, gk = k0 + kk;
                rb[e] = (gc < Nc && gk < kend) ? B[(size_t)gc * Kd + gk] : 0.0;
            } else {
                const int kk = idx / BN, col = idx % BN;
                const int gc = n0 + col, gk = k0 + kk;
                rb[e] = (gc < Nc && gk < kend) ? B[(size_t)gk * Nc + gc] : 0.0;
            }
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int e = 0; e < AL; e++) {
            const int idx = t + 256 * e;
            As[idx % GBK][idx / GBK] = ra[e];
        }
#pragma unroll
        for (int e = 0; e < BL; e++) {
            const int idx = t + 256 * e;
            if (BT) Bs[idx % GBK][idx / GBK] = rb[e];
            else Bs[idx / BN][idx % BN] = rb[e];
        }
    };
    if (kbeg < kend) fetch(kbeg);
    for (int k0 = kbeg; k0 < kend; k0 += GBK) {
        __syncthreads();
        stash();
        __syncthreads();
        if (k0 + GBK < kend) fetch(k0 + GBK);
#pragma unroll
        for (int kk = 0; kk < GBK; kk += 4) {
            double a[WM], b[WN];
#pragma unroll
            for (int i = 0; i < WM; i++) a[i] = As[kk + (lane >> 4)][wm * 16 * WM + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < WN; j++) b[j] = Bs[kk + (lane >> 4)][wn * 16 * WN + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < WM; i++)
#pragma unroll
                for (int j = 0; j < WN; j++) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
    }
    // epilogue: accumulator lane l, element v holds row (l/16) + 4v, column l%16
    double rowsum[WM][4];
#pragma unroll
    for (int i = 0; i < WM; i++)
#pragma unroll
        for (int v = 0; v < 4; v++) rowsum[i][v] = 0.0;
#pragma unroll
    for (int i = 0; i < WM; i++)
#pragma unroll
        for (int j = 0; j < WN; j++) {
            const int col = n0 + wn * 16 * WN + j * 16 + (lane & 15);
            double wv = 0.0;
            if (EPI) wv = (col < Nc) ? wvec[col] : 0.0;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = m0 + wm * 16 * WM + i * 16 + (lane >> 4) + 4 * v;
                double kv = acc[i][j][v];
                if (EPI) {
                    double kz1, kz;
                    if (izeta > 0) {
                        kz1 = 1.0;
                        for (int q = 1; q < izeta; q++) kz1 *= kv;
                        kz = kz1 * kv;
                    } else {
                        kz1 = pow(kv, zeta - 1.0);
                        kz = pow(kv, zeta);
                    }
                    rowsum[i][v] += wv * kz;
                    kv = wv * zeta * kz1;
                }
                if (row < Mr && col < Nc) C[(size_t)row * Nc + col] = kv;
            }
        }
    if (EPI) {
#pragma unroll
        for (int i = 0; i < WM; i++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                double sm = rowsum[i][v];
                for (int off = 8; off > 0; off >>= 1) sm += __shfl_xor(sm, off, 64);
                rowsum[i][v] = sm;
            }
        if ((lane & 15) == 0)
#pragma unroll
            for (int i = 0; i < WM; i++)
#pragma unroll
                for (int v = 0; v < 4; v++) red[wn][wm * 16 * WM + i * 16 + (lane >> 4) + 4 * v] = rowsum[i][v];
        __syncthreads();
        for (int r = t; r < BM; r += 256) {
            const int row = m0 + r;
            if (row < Mr) e_part[(size_t)row * gridDim.y + blockIdx.y] = red[0][r] + red[1][r];
        }
    }
}

static int int_zeta(double z) {
    double r = floor(z + 0.5);
    return (fabs(z - r) == 0.0 && r >= 1.0 && r <= 16.0) ? (int)r : 0;
}

// Tile shapes the autotuner chooses between: (WM, WN, k tile) per block of 4
// wavefronts, block tile (32 WM) x (32 WN).
static const int kTiles[kGemmTiles][3] = {{4, 2, 16}, {2, 2, 16}, {2, 4, 16}, {4, 2, 32}, {2, 2, 32}, {2, 4, 32}};

int gemm1_ntile(int cfg, int M) { return cdiv(M, 32 * kTiles[cfg][1]); }

template <int WM, int WN, int BK>
static void gemm1_go(int nc, int M, int d, const double *X, const double *Xs, const double *wvec, double zeta,
                     double *W, double *e_part, hipStream_t s) {
    dim3 grid(cdiv(nc, 32 * WM), cdiv(M, 32 * WN), 1);
    hipLaunchKernelGGL((k_gemm_f64<true, true, WM, WN, BK>), grid, dim3(256), 0, s, nc, M, d, d, X, Xs, W, wvec,
                       zeta, int_zeta(zeta), e_part);
}

void launch_gemm1_f64(int cfg, int nc, int M, int d, const double *X, const double *Xs, const double *wvec,
                      double zeta, double *W, double *e_part, int *ntile_out, hipStream_t s) {
    *ntile_out = gemm1_ntile(cfg, M);
    if (nc <= 0) return;
    switch (cfg) {
        case 0: gemm1_go<4, 2, 16>(nc, M, d, X, Xs, wvec, zeta, W, e_part, s); break;
        case 1: gemm1_go<2, 2, 16>(nc, M, d, X, Xs, wvec, zeta, W, e_part, s); break;
        case 2: gemm1_go<2, 4, 16>(nc, M, d, X, Xs, wvec, zeta, W, e_part, s); break;
        case 3: gemm1_go<4, 2, 32>(nc, M, d, X, Xs, wvec, zeta, W, e_part, s); break;
        case 4: gemm1_go<2, 2, 32>(nc, M, d, X, Xs, wvec, zeta, W, e_part, s); break;
        default: gemm1_go<2, 4, 32>(nc, M, d, X, Xs, wvec, zeta, W, e_part, s); break;
    }
}

// split-K partial sums, reduced in a fixed order
__global__ void k_sum_parts(size_t n, int parts, const double *P, double *out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int z = 0; z < parts; z++) s += P[z * n + i];
    out[i] = s;
}

// default split-K factor: enough blocks to fill the chip
int gemm2_parts(int nc, int d) {
    const long tiles = (long)cdiv(nc, 64) * cdiv(d, 64);
    int parts = 1;
    while (tiles * parts < 1024 && parts < 8) parts *= 2;
    return parts;
}

template <int WM, int WN, int BK>
static void gemm2_go(int nc, int M, int d, int parts, const double *W, const double *Xs, double *out,
                     hipStream_t s) {
    const int kchunk = (cdiv(M, parts) + BK - 1) / BK * BK;
    dim3 grid(cdiv(nc, 32 * WM), cdiv(d, 32 * WN), parts);
    hipLaunchKernelGGL((k_gemm_f64<false, false, WM, WN, BK>), grid, dim3(256), 0, s, nc, d, M, kchunk, W, Xs, out,
                       (const double *)nullptr, 0.0, 0, (double *)nullptr);
}

void launch_gemm2_f64(int cfg, int parts, int nc, int M, int d, const double *W, const double *Xs, double *G,
                      double *Gparts, hipStream_t s) {
    if (nc <= 0) return;
    double *out = parts > 1 ? Gparts : G;
    switch (cfg) {
        case 0: gemm2_go<4, 2, 16>(nc, M, d, parts, W, Xs, out, s); break;
        case 1: gemm2_go<2, 2, 16>(nc, M, d, parts, W, Xs, out, s); break;
        case 2: gemm2_go<2, 4, 16>(nc, M, d, parts, W, Xs, out, s); break;
        case 3: gemm2_go<4, 2, 32>(nc, M, d, parts, W, Xs, out, s); break;
        case 4: gemm2_go<2, 2, 32>(nc, M, d, parts, W, Xs, out, s); break;
        default: gemm2_go<2, 4, 32>(nc, M, d, parts, W, Xs, out, s); break;
    }
    if (parts > 1) launch_sum_parts_f64(nc, d, parts, Gparts, G, s);
}

// The end of a fused GP launch in one kernel: G = scale * (sum over parts, fixed order)
// of the partial rows (FP64 or FP32; blocks [0, gblocks)), and the centres' energy rows
// esite[centres[i]] += sum of e_part[i][0..ntile) (the blocks after them).
template <typename T>
__global__ void k_gp_finish(size_t n, int parts, const T *P, double scale, double *G, int gblocks, int nc, int ntile,
                            const int *centres, const double *e_part, double *esite) {
    const int b = blockIdx.x;
    if (b < gblocks) {
        const size_t i = (size_t)b * blockDim.x + threadIdx.x;
        if (i >= n) return;
        double s = 0.0;
        for (int p = 0; p < parts; p++) s += (double)P[(size_t)p * n + i];
        G[i] = s * scale;
        return;
    }
    const int i = (b - gblocks) * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    double s = 0.0;
    for (int t = 0; t < ntile; t++) s += e_part[(size_t)i * ntile + t];
    esite[centres[i]] += s;
}
template <typename T>
static void gp_finish_go(int nc, int d, int parts, const T *Gparts, double scale, double *G, int ntile,
                         const int *centres, const double *e_part, double *esite, hipStream_t s) {
    const size_t n = (size_t)nc * d;
    const int gb = parts > 1 ? cdiv((long)n, 256) : 0, eb = centres ? cdiv(nc, 256) : 0;
    if (nc > 0 && gb + eb > 0)
        hipLaunchKernelGGL(k_gp_finish<T>, dim3(gb + eb), dim3(256), 0, s, n, parts, Gparts, scale, G, gb, nc, ntile,
                           centres, e_part, esite);
}
void launch_gp_finish_f64(int nc, int d, int parts, const double *Gparts, double *G, int ntile, const int *centres,
                          const double *e_part, double *esite, hipStream_t s) {
    gp_finish_go<double>(nc, d, parts, Gparts, 1.0, G, ntile, centres, e_part, esite, s);
}
void launch_gp_finish_f32(int nc, int d, int parts, const float *Gparts, double scale, double *G, int ntile,
                          const int *centres, const double *e_part, double *esite, hipStream_t s) {
    gp_finish_go<float>(nc, d, parts, Gparts, scale, G, ntile, centres, e_part, esite, s);
}

void launch_sum_parts_f64(int nc, int d, int parts, const double *Gparts, double *G, hipStream_t s) {
    const size_t n = (size_t)nc * d;
    if (n > 0) hipLaunchKernelGGL(k_sum_parts, dim3(cdiv((long)n, 256)), dim3(256), 0, s, n, parts, Gparts, G);
}

// e_i = sum over column tiles, fixed order
__global__ void k_energy_rows(int nc, int ntile, const int *centres, const double *e_part, double *esite) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    double s = 0.0;
    for (int t = 0; t < ntile; t++) s += e_part[(size_t)i * ntile + t];
    esite[centres[i]] += s;
}
void launch_energy_rows(int nc, int ntile, const int *centres, const double *e_part, double *esite, hipStream_t s) {
    if (nc > 0) hipLaunchKernelGGL(k_energy_rows, dim3(cdiv(nc, 256)), dim3(256), 0, s, nc, ntile, centres, e_part, esite);
}

// ----------------------------------------------------------------------------
// distance_2b with the ard_se kernel, one thread per atom over its pairs
// ----------------------------------------------------------------------------
__global__ void k_d2b(int N, const int *Z, int Z1, int Z2, double rc, double ctw, int M, const double *xs,
                      const double *wvec, const double *theta, const int *nb_off, const int *nb_j, const double *nb_r,
                      const double *nb_d, double *pairF, double *pairE, const unsigned char *mask,
                      double *e_atom) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (mask && !mask[i]) {
        if (e_atom) e_atom[i] = 0.0;
        return;
    }
    double esum = 0.0;
    const bool zi1 = (Z1 == 0) || (Z[i] == Z1), zi2 = (Z2 == 0) || (Z[i] == Z2);
    const double it2 = 1.0 / (theta[0] * theta[0]);
    for (int p = nb_off[i]; p < nb_off[i + 1]; p++) {
        const double r = nb_r[p];
        if (r > rc) continue;
        const int j = nb_j[p];
        const bool zj1 = (Z1 == 0) || (Z[j] == Z1), zj2 = (Z2 == 0) || (Z[j] == Z2);
        if (!((zi1 && zj2) || (zi2 && zj1))) continue;
        double e = 0.0, de = 0.0;
        for (int s = 0; s < M; s++) {
            double t = r - xs[s];
            double k = wvec[s] * exp(-0.5 * t * t * it2);
            e += k;
            de -= k * t * it2;
        }
        const double fc = fcut_d(r, rc, ctw), dfc = dfcut_d(r, rc, ctw);
        const double dE = de * fc + e * dfc;
        const double ir = 1.0 / r;
        pairE[p] += e * fc;
        esum += e * fc;
        pairF[3 * p] += dE * nb_d[3 * p] * ir;
        pairF[3 * p + 1] += dE * nb_d[3 * p + 1] * ir;
        pairF[3 * p + 2] += dE * nb_d[3 * p + 2] * ir;
    }
    if (e_atom) e_atom[i] = esum;
}
void launch_d2b(int N, const int *Z, int Z1, int Z2, double rc, double ctw, int M, const double *xs,
                const double *wvec, const double *theta, const int *nb_off, const int *nb_j, const double *nb_r,
                const double *nb_d, double *pairF, double *pairE, const unsigned char *mask, double *e_atom,
                hipStream_t s) {
    if (N > 0)
        hipLaunchKernelGGL(k_d2b, dim3(cdiv(N, 128)), dim3(128), 0, s, N, Z, Z1, Z2, rc, ctw, M, xs, wvec, theta,
                           nb_off, nb_j, nb_r, nb_d, pairF, pairE, mask, e_atom);
}

// ----------------------------------------------------------------------------
// fixed-order sum of n doubles by one block (energy per coordinate)
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_sum_fixed(size_t n, const double *v, double *out) {
    __shared__ double red[256];
    double a = 0.0;
    for (size_t q = threadIdx.x; q < n; q += 256) a += v[q];
    red[threadIdx.x] = a;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}
void launch_sum_fixed(size_t n, const double *v, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_sum_fixed, dim3(1), dim3(256), 0, s, n, v, out);
}

// ----------------------------------------------------------------------------
// GAP variance of the dot_product (SOAP) and ard_se (distance_2b) predictions:
//   var(x) = k(x,x) - k_x^T (K_MM + reg I)^-1 k_x,  k_x[s] = delta^2 kernel(x, X_s)
// Cinv = (K_MM + reg I)^-1 is formed on the host when the model is asked for it.
// ----------------------------------------------------------------------------
template <int WM, int WN, int BK>
static void gemm_nt_go(int rows, int cols, int kd, const double *A, const double *B, double *C, hipStream_t s) {
    dim3 grid(cdiv(rows, 32 * WM), cdiv(cols, 32 * WN), 1);
    hipLaunchKernelGGL((k_gemm_f64<true, false, WM, WN, BK>), grid, dim3(256), 0, s, rows, cols, kd, kd, A, B, C,
                       (const double *)nullptr, 0.0, 0, (double *)nullptr);
}
void launch_gemm_nt_f64(int rows, int cols, int kd, const double *A, const double *B, double *C, hipStream_t s) {
    if (rows > 0 && cols > 0) gemm_nt_go<2, 2, 16>(rows, cols, kd, A, B, C, s);
}

__device__ __forceinline__ double powz(double x, double zeta, int iz) {
    if (iz > 0) {
        double r = 1.0;
        for (int q = 0; q < iz; q++) r *= x;
        return r;
    }
    if (iz == 0 && zeta == 0.0) return 1.0;
    return pow(x, zeta);
}

__global__ void k_var_cov(size_t n, const double *K, double delta2, double zeta, int iz, double *kv) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) kv[q] = delta2 * powz(K[q], zeta, iz);
}
void launch_var_cov(int nc, int M, const double *K, double delta2, double zeta, double *kv, hipStream_t s) {
    const size_t n = (size_t)nc * M;
    if (n) hipLaunchKernelGGL(k_var_cov, dim3(cdiv((long)n, 256)), dim3(256), 0, s, n, K, delta2, zeta,
                              int_zeta(zeta), kv);
}

__device__ __forceinline__ double wave_sum_v(double v) {   // fixed butterfly over the 64 lanes
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// one wavefront per centre row
__global__ __launch_bounds__(256) void k_var_rows(int nc, int M, int d, const double *X, const double *K,
                                                  const double *kv, double *V, double delta2, double zeta, int iz,
                                                  int iz1, const int *centres, double *var_inst, double *varsite) {
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= nc) return;
    const double *x = X + (size_t)i * d;
    double xx = 0.0;
    for (int q = lane; q < d; q += 64) xx += x[q] * x[q];
    xx = wave_sum_v(xx);
    const double *Ki = K + (size_t)i * M, *ki = kv + (size_t)i * M;
    double *Vi = V + (size_t)i * M;
    double kck = 0.0;
    const double g = -2.0 * zeta * delta2;
    for (int q = lane; q < M; q += 64) {
        const double v = Vi[q];
        kck += ki[q] * v;
        Vi[q] = g * powz(Ki[q], zeta - 1.0, iz1) * v;
    }
    kck = wave_sum_v(kck);
    if (lane == 0) {
        const double var = delta2 * powz(xx, zeta, iz) - kck;
        var_inst[i] = var;
        varsite[centres[i]] += var;
    }
}
void launch_var_rows(int nc, int M, int d, const double *X, const double *K, const double *kv, double *V,
                     double delta2, double zeta, const int *centres, double *var_inst, double *varsite,
                     hipStream_t s) {
    if (nc > 0)
        hipLaunchKernelGGL(k_var_rows, dim3(cdiv(nc, 4)), dim3(256), 0, s, nc, M, d, X, K, kv, V, delta2, zeta,
                           int_zeta(zeta), int_zeta(zeta - 1.0), centres, var_inst, varsite);
}

__global__ __launch_bounds__(256) void k_var_self_grad(int nc, int d, const double *X, double delta2, double zeta,
                                                       int iz1, double *Gv) {
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= nc) return;
    const double *x = X + (size_t)i * d;
    double xx = 0.0;
    for (int q = lane; q < d; q += 64) xx += x[q] * x[q];
    xx = wave_sum_v(xx);
    const double g = 2.0 * zeta * delta2 * powz(xx, zeta - 1.0, iz1);
    for (int q = lane; q < d; q += 64) Gv[(size_t)i * d + q] += g * x[q];
}
void launch_var_self_grad(int nc, int d, const double *X, double delta2, double zeta, double *Gv, hipStream_t s) {
    if (nc > 0)
        hipLaunchKernelGGL(k_var_self_grad, dim3(cdiv(nc, 4)), dim3(256), 0, s, nc, d, X, delta2, zeta,
                           int_zeta(zeta - 1.0), Gv);
}

// one thread per first atom over its pairs; the covariance vector sits in a per-lane array
__global__ void k_d2b_var(int N, const int *Z, int Z1, int Z2, double rc, double ctw, int M, const double *xs,
                          double delta2, const double *theta, const double *Cinv, const int *nb_off, const int *nb_j,
                          const double *nb_r, const double *nb_d, double *pairFv, double *pairVE, double *var_inst,
                          double *cut2_inst, const unsigned char *mask) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N || (mask && !mask[i])) return;
    const bool zi1 = (Z1 == 0) || (Z[i] == Z1), zi2 = (Z2 == 0) || (Z[i] == Z2);
    const double it2 = 1.0 / (theta[0] * theta[0]);
    double kk[kMaxVarSparse], dk[kMaxVarSparse];
    for (int p = nb_off[i]; p < nb_off[i + 1]; p++) {
        const double r = nb_r[p];
        if (r > rc) continue;
        const int j = nb_j[p];
        const bool zj1 = (Z1 == 0) || (Z[j] == Z1), zj2 = (Z2 == 0) || (Z[j] == Z2);
        if (!((zi1 && zj2) || (zi2 && zj1))) continue;
        for (int a = 0; a < M; a++) {
            const double t = r - xs[a];
            kk[a] = delta2 * exp(-0.5 * t * t * it2);
            dk[a] = -kk[a] * t * it2;
        }
        double var = delta2, dvar = 0.0;
        for (int a = 0; a < M; a++) {
            double y = 0.0;   // (Cinv k)_a
            for (int b = 0; b < M; b++) y += Cinv[(size_t)a * M + b] * kk[b];
            var -= kk[a] * y;
            dvar -= 2.0 * dk[a] * y;
        }
        const double fc = fcut_d(r, rc, ctw), dfc = dfcut_d(r, rc, ctw);
        const double gr = dvar * fc * fc + 2.0 * var * fc * dfc;
        const double ir = 1.0 / r;
        pairVE[p] += var * fc * fc;
        pairFv[3 * p] += gr * nb_d[3 * p] * ir;
        pairFv[3 * p + 1] += gr * nb_d[3 * p + 1] * ir;
        pairFv[3 * p + 2] += gr * nb_d[3 * p + 2] * ir;
        var_inst[p] = var;
        cut2_inst[p] = fc * fc;
    }
}
void launch_d2b_var(int N, const int *Z, int Z1, int Z2, double rc, double ctw, int M, const double *xs,
                    double delta2, const double *theta, const double *Cinv, const int *nb_off, const int *nb_j,
                    const double *nb_r, const double *nb_d, double *pairFv, double *pairVE, double *var_inst,
                    double *cut2_inst, const unsigned char *mask, hipStream_t s) {
    if (N > 0 && M > 0 && M <= kMaxVarSparse)
        hipLaunchKernelGGL(k_d2b_var, dim3(cdiv(N, 64)), dim3(64), 0, s, N, Z, Z1, Z2, rc, ctw, M, xs, delta2,
                           theta, Cinv, nb_off, nb_j, nb_r, nb_d, pairFv, pairVE, var_inst, cut2_inst, mask);
}

__global__ void k_gather_var(int N, const int *nb_off, const int *nb_rev, const double *pairFv,
                             const double *pairVE, const double *varsite, double *gvg, double *lgv) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    double g0 = 0.0, g1 = 0.0, g2 = 0.0, ve = 0.0;
    for (int q = nb_off[j]; q < nb_off[j + 1]; q++) {
        const int r = nb_rev[q];
        g0 += pairFv[3 * q] - pairFv[3 * r];
        g1 += pairFv[3 * q + 1] - pairFv[3 * r + 1];
        g2 += pairFv[3 * q + 2] - pairFv[3 * r + 2];
        if (pairVE) ve += pairVE[q] + pairVE[r];
    }
    if (gvg) {
        gvg[3 * (size_t)j] = -g0;
        gvg[3 * (size_t)j + 1] = -g1;
        gvg[3 * (size_t)j + 2] = -g2;
    }
    if (lgv) lgv[j] = varsite[j] + 0.5 * ve;
}
void launch_gather_var(int N, const int *nb_off, const int *nb_rev, const double *pairFv, const double *pairVE,
                       const double *varsite, double *gvg, double *lgv, hipStream_t s) {
    if (N > 0)
        hipLaunchKernelGGL(k_gather_var, dim3(cdiv(N, 128)), dim3(128), 0, s, N, nb_off, nb_rev, pairFv, pairVE,
                           varsite, gvg, lgv);
}

// ----------------------------------------------------------------------------
// per-atom gather: F_j = sum_q (pairF[q] - pairF[rev q]),
// local_virial_j(a,b) = -sum_q d_(rev q),a pairF_(rev q),b,
// local_e_j = e_site + (pairE[q] + pairE[rev q])/2 + e0(Z_j); block partial sums.
// ----------------------------------------------------------------------------
constexpr int GATHER_BS = 256;   // 4 wavefronts, two atoms each (a half-wave per atom)
__device__ __forceinline__ double half_sum_g(double v) {   // sum over the 32 lanes of a half-wave
    for (int off = 16; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__global__ __launch_bounds__(GATHER_BS) void k_gather(int N, const int *Z, const double *e0, double E_scale,
                                                      const int *nb_off, const int *nb_rev, const double *nb_d,
                                                      const double *pairF, const double *pairE, const double *esite,
                                                      double *f, double *local_e, double *local_virial,
                                                      const unsigned char *mask, double *block_part) {
    __shared__ double sv[8][10];   // the block's atoms' local energy and virial (block partials)
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, hl = lane & 31;
    const int j = (blockIdx.x * 4 + w) * 2 + (lane >> 5);
    const bool live = j < N;
    double v[13];
#pragma unroll
    for (int k = 0; k < 13; k++) v[k] = 0.0;
    // a half-wave over the pairs of atom j; fixed lane assignment and a fixed reduction tree
    const int q0 = live ? nb_off[j] : 0, q1 = live ? nb_off[j + 1] : 0;
    for (int q = q0 + hl; q < q1; q += 32) {
        const int r = nb_rev[q];
        const double gx = pairF[3 * r], gy = pairF[3 * r + 1], gz = pairF[3 * r + 2];
        v[0] += pairF[3 * q] - gx;
        v[1] += pairF[3 * q + 1] - gy;
        v[2] += pairF[3 * q + 2] - gz;
        const double dx = nb_d[3 * r], dy = nb_d[3 * r + 1], dz = nb_d[3 * r + 2];
        v[3] -= dx * gx; v[4] -= dy * gx; v[5] -= dz * gx;
        v[6] -= dx * gy; v[7] -= dy * gy; v[8] -= dz * gy;
        v[9] -= dx * gz; v[10] -= dy * gz; v[11] -= dz * gz;
        if (pairE) v[12] += pairE[q] + pairE[r];   // null without a distance_2b term
    }
#pragma unroll
    for (int k = 0; k < 13; k++) v[k] = half_sum_g(v[k]);
    const int a = (w << 1) | (lane >> 5);   // the atom's place in the block
    if (live) {
        if (hl < 3 && f) f[3 * j + hl] = E_scale * (hl == 0 ? v[0] : (hl == 1 ? v[1] : v[2]));
        if (hl == 0) {
            const double le = E_scale * (esite[j] + 0.5 * v[12] + ((mask && !mask[j]) ? 0.0 : e0[Z[j]]));
            local_e[j] = le;
            if (block_part) sv[a][0] = le;
        }
        if (hl < 9) {
            double x = v[3];
#pragma unroll
            for (int k = 1; k < 9; k++) x = (hl == k) ? v[3 + k] : x;
            local_virial[9 * (size_t)j + hl] = E_scale * x;
            if (block_part) sv[a][1 + hl] = E_scale * x;
        }
    } else if (block_part && hl < 10) {
        sv[a][hl] = 0.0;
    }
    if (!block_part) return;
    __syncthreads();
    if (threadIdx.x < 10) {   // the block's 8 atoms in order
        double x = sv[0][threadIdx.x];
        for (int q = 1; q < 8; q++) x += sv[q][threadIdx.x];
        block_part[10 * (size_t)blockIdx.x + threadIdx.x] = x;
    }
}
void launch_gather(int N, const int *Z, const double *e0, double E_scale, const int *nb_off, const int *nb_rev,
                   const double *nb_d, const double *pairF, const double *pairE, const double *esite, double *f,
                   double *local_e, double *local_virial, const unsigned char *mask, double *block_part,
                   hipStream_t s) {
    if (N > 0)
        hipLaunchKernelGGL(k_gather, dim3(cdiv(N, 8)), dim3(GATHER_BS), 0, s, N, Z, e0, E_scale, nb_off, nb_rev,
                           nb_d, pairF, pairE, esite, f, local_e, local_virial, mask, block_part);
}

// Energy and virial of each structure: E = sum of local energies, V = sum of
// local virials (IPModel_GAP.f95:524-526).  One block of 512 threads per structure;
// thread t sums the structure's atoms t, t+512, ... in that order (four atoms' loads
// issued together), a fixed shuffle tree joins the lanes and threads 0..9 the 8 waves,
// so a structure's totals do not depend on the other structures of a batch.
// With flag_dst (pinned host memory) the first block also copies the four neighbour-list
// build flags there: the step's last kernel hands them to the host without a copy launch.
constexpr int kReduceThreads = 512;
__global__ __launch_bounds__(kReduceThreads) void k_struct_reduce(const int *atom_off, const double *local_e,
                                                                  const double *local_virial, double *e_out,
                                                                  double *virial_out, const int *flag_src,
                                                                  int *flag_dst) {
    __shared__ double sred[10][kReduceThreads / 64];
    const int st = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
    if (flag_dst && st == 0 && t < 4) flag_dst[t] = flag_src[t];
    const int a0 = atom_off[st], a1 = atom_off[st + 1];
    double acc[10];
#pragma unroll
    for (int k = 0; k < 10; k++) acc[k] = 0.0;
    for (int j0 = a0 + t; j0 < a1; j0 += 4 * kReduceThreads) {
        double v[4][10];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int j = j0 + u * kReduceThreads;
            const bool ok = j < a1;
            v[u][0] = ok ? local_e[j] : 0.0;
#pragma unroll
            for (int k = 0; k < 9; k++) v[u][1 + k] = ok ? local_virial[9 * (size_t)j + k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int k = 0; k < 10; k++) acc[k] += v[u][k];
    }
#pragma unroll
    for (int k = 0; k < 10; k++) {
        double x = acc[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if (lane == 0) sred[k][wv] = x;
    }
    __syncthreads();
    if (t < 10) {
        double x = sred[t][0];
        for (int q = 1; q < kReduceThreads / 64; q++) x += sred[t][q];
        if (t == 0) e_out[st] = x;
        else virial_out[9 * (size_t)st + t - 1] = x;
    }
}
// One structure: the totals from the gather's block partials (10 per block, blocks of 8
// atoms in order): thread t sums partials t, t+512, ... in turn, then the fixed tree.
__global__ __launch_bounds__(kReduceThreads) void k_struct_finish(int nb, const double *part, double *e_out,
                                                                  double *virial_out, const int *flag_src,
                                                                  int *flag_dst) {
    __shared__ double sred[10][kReduceThreads / 64];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    if (flag_dst && t < 4) flag_dst[t] = flag_src[t];
    double acc[10];
#pragma unroll
    for (int k = 0; k < 10; k++) acc[k] = 0.0;
    for (int b0 = t; b0 < nb; b0 += 2 * kReduceThreads) {
        double v[2][10];
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int b = b0 + u * kReduceThreads;
#pragma unroll
            for (int k = 0; k < 10; k++) v[u][k] = b < nb ? part[10 * (size_t)b + k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
            for (int k = 0; k < 10; k++) acc[k] += v[u][k];
    }
#pragma unroll
    for (int k = 0; k < 10; k++) {
        double x = acc[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if (lane == 0) sred[k][wv] = x;
    }
    __syncthreads();
    if (t < 10) {
        double x = sred[t][0];
        for (int q = 1; q < kReduceThreads / 64; q++) x += sred[t][q];
        if (t == 0) e_out[0] = x;
        else virial_out[t - 1] = x;
    }
}
void launch_struct_finish(int N, const double *part, double *e_out, double *virial_out, hipStream_t s,
                          const int *flag_src, int *flag_dst) {
    if (N > 0)
        hipLaunchKernelGGL(k_struct_finish, dim3(1), dim3(kReduceThreads), 0, s, cdiv(N, 8), part, e_out, virial_out,
                           flag_src, flag_dst);
}
void launch_struct_reduce(int S, const int *atom_off, const double *local_e, const double *local_virial,
                          double *e_out, double *virial_out, hipStream_t s, const int *flag_src, int *flag_dst) {
    if (S > 0)
        hipLaunchKernelGGL(k_struct_reduce, dim3(S), dim3(kReduceThreads), 0, s, atom_off, local_e, local_virial,
                           e_out, virial_out, flag_src, flag_dst);
}

}  // namespace gapx
