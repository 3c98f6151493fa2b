// SOAP descriptor kernels specialised on (n_max, l_max).
//
// One wavefront per centre atom.  The radial transform T (n x n) and the
// radial grid live in LDS; an empty asm with a memory clobber closes every
// l step so the compiler re-reads T from LDS instead of pinning n^2 doubles in
// VGPRs.  All loops over b, a, l and m are unrolled at compile time, so the
// per-lane recursion state (radial: F_l, F_(l+1), 1/x per grid point;
// Legendre: two previous levels) stays in registers with no scratch.
//
//  forward: per l, lanes = neighbours stage fc F_l(a) and Y_l,m in LDS, lanes
//           = (a, m) accumulate c' in the primitive radial basis; c = T^T c'
//           once per centre; lanes = power-spectrum entries form p, a wave
//           reduction gives |p|.
//  adjoint: dE/dp -> dE/dc per level as a small product staged in LDS ->
//           dE/dc' = T dE/dc, then lanes = neighbours re-evaluate F_l, Y and
//           grad Y and contract them with dE/dc'.
#include "gapx_kernels.h"

#include <algorithm>

namespace gapx {

#define GAPX_PI_S 3.14159265358979323846264338327950288

// N_lm of the real spherical harmonics (x sqrt 2 for m > 0) and (2m-1)!!
static __constant__ double c_ynorm[(kMaxL + 1) * (kMaxL + 2) / 2];
static __constant__ double c_dfact[kMaxL + 2];
static __constant__ double c_rinv[2 * kMaxL + 4];   // 1/k (k = 0 unused)
static __constant__ double c_isq[kMaxL + 1];        // 1/sqrt(2l+1)

static bool g_soap_const_ready[64] = {};   // per device
static hipError_t ensure_soap_constants() {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    if (g_soap_const_ready[dev]) return hipSuccess;
    double yn[(kMaxL + 1) * (kMaxL + 2) / 2], df[kMaxL + 2];
    for (int l = 0; l <= kMaxL; l++)
        for (int m = 0; m <= l; m++) {
            double r = 1.0;   // (l-m)!/(l+m)!
            for (int k = l - m + 1; k <= l + m; k++) r /= (double)k;
            double v = sqrt((2.0 * l + 1.0) / (4.0 * GAPX_PI_S) * r);
            if (m > 0) v *= sqrt(2.0);
            yn[l * (l + 1) / 2 + m] = v;
        }
    df[0] = 1.0;
    for (int m = 1; m < kMaxL + 2; m++) df[m] = df[m - 1] * (2.0 * m - 1.0);
    double ri[2 * kMaxL + 4];
    ri[0] = 0.0;
    for (int k = 1; k < 2 * kMaxL + 4; k++) ri[k] = 1.0 / (double)k;
    double isq[kMaxL + 1];
    for (int l = 0; l <= kMaxL; l++) isq[l] = 1.0 / sqrt(2.0 * l + 1.0);
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_rinv), ri, sizeof(ri))) != hipSuccess) return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_isq), isq, sizeof(isq))) != hipSuccess) return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_ynorm), yn, sizeof(yn))) != hipSuccess) return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_dfact), df, sizeof(df))) != hipSuccess) return e;
    g_soap_const_ready[dev] = true;
    return hipSuccess;
}

hipError_t soap_fast_constants_init() { return ensure_soap_constants(); }

#define RELOAD_BARRIER() __asm__ volatile("" ::: "memory")

// neighbours staged per pass of the forward kernel (one lane each)
constexpr int kSoapChunk = 32;

__device__ __forceinline__ double fc_s(double r, double rc, double w) {
    if (r > rc) return 0.0;
    if (r > rc - w) return 0.5 * (cos(GAPX_PI_S * (r - rc + w) / w) + 1.0);
    return 1.0;
}
// fc and dfc/dr together: one sincos for the transition region
__device__ __forceinline__ void fc_dfc_s(double r, double rc, double w, double &fc, double &dfc) {
    fc = 0.0;
    dfc = 0.0;
    if (r > rc) return;
    if (r > rc - w) {
        double sn, cs;
        sincos(GAPX_PI_S * (r - rc + w) / w, &sn, &cs);
        fc = 0.5 * (cs + 1.0);
        dfc = -0.5 * GAPX_PI_S * sn / w;
        return;
    }
    fc = 1.0;
}
__device__ __forceinline__ double wave_sum_s(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ int species_s(const SoapDev &sp, int z) {
    for (int s = 0; s < sp.S; s++)
        if (sp.species_Z[s] == z || sp.species_Z[s] == 0) return s;
    return -1;
}

// Associated Legendre values P_l^m (m = 0..l) of one neighbour, level by level,
// in two LDS rows: on entry cur holds P_(l-2), prev P_(l-1); on exit cur holds P_l.
__device__ __forceinline__ void legendre_level(int l, double uz, const double *prev, double *cur) {
    for (int m = 0; m <= l; m++) {
        double pc;
        if (m == l) pc = c_dfact[m];
        else if (m == l - 1) pc = (2.0 * l - 1.0) * uz * prev[m];
        else pc = ((2.0 * l - 1.0) * uz * prev[m] - (double)(l + m - 1) * cur[m]) * c_rinv[l - m];
        cur[m] = pc;
    }
}

// i_l recursion state for all grid points: F0 = f_l, F1 = f_(l+1).  The
// recursion factor 1/x_a = 1/(2 alpha r r_a) is ixq / a with ixq = 1/(2 alpha r dr)
// (r_a = a dr), formed from ixq when used (no per-point array: fewer registers).
// rbs: grid r_a = a * dr; gA[a] = exp(-alpha r_a^2); iA[a] = 1/a.
// exp(-alpha (r -+ r_a)^2) = exp(-alpha r^2) gA[a] q^(+-a) with q = exp(2 alpha r dr):
// two exponentials and one division per neighbour.  Outside the range where the
// powers stay finite, every point is direct.
template <int N>
struct Rad {
    double F0[N], F1[N], ixq;
    __device__ __forceinline__ void init(const double *rbs, const double *gA, const double *iA, double alpha,
                                         double r) {
        const double xq = N > 1 ? 2.0 * alpha * r * rbs[1] : 0.0;
        const double rmax = rbs[N - 1];
        ixq = 1.0 / xq;
        if (xq > 0.0 && xq * (N - 1) < 600.0 && alpha * (r * r + rmax * rmax) < 600.0) {
            const double e0 = exp(-alpha * r * r);
            const double q = exp(xq), qi = 1.0 / q;
            double qa = 1.0, qia = 1.0;
            F0[0] = e0;
            F1[0] = 0.0;
#pragma unroll
            for (int a = 1; a < N; a++) {
                qa *= q;
                qia *= qi;
                const double eg = e0 * gA[a];
                const double em = eg * qa, ep = eg * qia;
                const double iv = ixq * iA[a];
                F0[a] = 0.5 * (em - ep) * iv;
                F1[a] = 0.5 * (em + ep) * iv - F0[a] * iv;
            }
            return;
        }
#pragma unroll
        for (int a = 0; a < N; a++) {
            const double ra = rbs[a];
            if (a == 0) {
                F0[a] = exp(-alpha * (ra * ra + r * r));
                F1[a] = 0.0;
            } else {
                const double ep = exp(-alpha * (r + ra) * (r + ra)), em = exp(-alpha * (r - ra) * (r - ra));
                const double iv = ixq * (1.0 / (double)a);
                F0[a] = 0.5 * (em - ep) * iv;
                F1[a] = 0.5 * (em + ep) * iv - F0[a] * iv;
            }
            __builtin_amdgcn_sched_barrier(0);   // one grid point's exponentials at a time
        }
    }
    // Scaled form (the adjoint): F1 holds a f_(l+1)(a), so that the sum over grid points
    // of r_a f_(l+1)(a) v_a is dr times the sum of F1[a] v_a, with no product by r_a.
    // Starting from the state init leaves, step_scaled keeps this form:
    // a f_(l+2) = a f_l - (2l+3) ixq f_(l+1) and f_(l+1) = F1[a] / a.
    __device__ __forceinline__ void scale_f1() {
#pragma unroll
        for (int a = 0; a < N; a++) F1[a] *= (double)a;
    }
    __device__ __forceinline__ void step_scaled(int l) {
        const double c = (2.0 * l + 3.0) * ixq;
        F0[0] = 0.0;   // grid point 0 (r_a = 0): f_(l+1) = 0 for every l
#pragma unroll
        for (int a = 1; a < N; a++) {
            const double t = F1[a] * (1.0 / (double)a);
            F1[a] = (double)a * F0[a] - c * t;
            F0[a] = t;
        }
    }
    __device__ __forceinline__ void step(int l) {
        F0[0] = F1[0];   // grid point 0 (r_a = 0): f_(l+1) = 0 for every l
        F1[0] = 0.0;
#pragma unroll
        for (int a = 1; a < N; a++) {
            const double n2 = F0[a] - (2.0 * l + 3.0) * (ixq * (1.0 / (double)a)) * F1[a];
            F0[a] = F1[a];
            F1[a] = n2;
        }
    }
};

// Forward.  With g_b(r) = fc(r) sum_a F_l(a, r) T[a][b] (F = the i_l recursion
// values of grid point a), c[b][lm] = sum_a T[a][b] c'[a][lm] where
// c'[a][lm] = sum_k fc F_l(a, r_k) Y_lm(k) is accumulated in the primitive
// (Gaussian) radial basis and transformed once per centre.
// Per neighbour chunk: the chunk's neighbours are listed per species (wave
// ballot); per l, lanes = neighbours stage fc F_l(a) and Y_l,m in LDS and
// lanes = (a, m) accumulate c' of each species over that species' list only.
// One level of a chunk: stage fc F_l and Y_l of the neighbours, then accumulate
// c' of level l.  A macro so that both level loops (rolled and unrolled) inline it.
#define SOAP_FWD_LEVEL(l, cur, prv)                                                                                  \
            __syncthreads();   /* the previous level's reads of Rl / Yl are done */                                  \
            if (lane < 32 * CPW) {   /* rows of neighbours outside every list stay zero for the dense loop */        \
_Pragma("unroll")                                                                                                    \
                for (int a = 0; a < N; a++) Rh[a * RT + kl] = s >= 0 ? fc * rs.F0[a] : 0.0;                          \
            }                                                                                                        \
            if (s >= 0) {                                                                                            \
                rs.step(l);                                                                                          \
                legendre_level(l, uz, prv, cur);                                                                     \
                double A = 1.0, B = 0.0;                                                                             \
                double *Yr = Yh + kl;   /* slot m of this neighbour at Yr[m * RT] */                                 \
                for (int m = 0; m <= l; m++) {                                                                       \
                    if (m > 0) {                                                                                     \
                        const double An = A * ux - B * uy;                                                           \
                        B = A * uy + B * ux;                                                                         \
                        A = An;                                                                                      \
                    }                                                                                                \
                    const double Nn = c_ynorm[l * (l + 1) / 2 + m] * cur[m];                                         \
                    if (m == 0) Yr[l * RT] = Nn;                                                                     \
                    else {                                                                                           \
                        Yr[(l + m) * RT] = Nn * A;                                                                   \
                        Yr[(l - m) * RT] = Nn * B;                                                                   \
                    }                                                                                                \
                }                                                                                                    \
            } else if (lane < 32 * CPW) {                                                                            \
                for (int m = 0; m < 2 * l + 1; m++) Yh[m * RT + kl] = 0.0;                                           \
            }                                                                                                        \
            __syncthreads();                                                                                         \
            const int W = 2 * l + 1;                                                                                 \
            for (int it = lane; it < CPW * N * W; it += 64) {                                                        \
                const int hh = CPW == 1 ? 0 : it / (N * W), r_ = it - hh * (N * W);                                 \
                const int a = r_ / W, mi = r_ - a * W;                                                               \
                const double *Rc = Rl + hh * (N * RT), *Yc = Yl + hh * (YW * RT);                                   \
                double *cPc = cP + hh * KLMp;                                                                        \
                if (S == 1) {   /* one species: every row of the chunk in pairs, zero rows included */            \
                    const double2 *R2 = (const double2 *)(Rc + a * RT), *Y2 = (const double2 *)(Yc + mi * RT);      \
                    double acc0 = 0.0, acc1 = 0.0;                                                                   \
_Pragma("unroll 4")                                                                                                  \
                    for (int k2 = 0; k2 < NB / 2; k2++) {                                                            \
                        const double2 rv = R2[k2], yv = Y2[k2];                                                      \
                        acc0 += rv.x * yv.x;                                                                         \
                        acc1 += rv.y * yv.y;                                                                         \
                    }                                                                                                \
                    cPc[a * LM + l * l + mi] += acc0 + acc1;                                                         \
                    continue;                                                                                        \
                }                                                                                                    \
                for (int si = 0; si < S; si++) {                                                                     \
                    const int n = ns[hh * S + si];                                                                   \
                    const int *ki = kidx + (hh * S + si) * NB;                                                       \
                    double acc0 = 0.0, acc1 = 0.0;                                                                   \
                    int q = 0;                                                                                       \
_Pragma("unroll 2")                                                                                                  \
                    for (; q + 1 < n; q += 2) {                                                                      \
                        const int k0 = ki[q], k1 = ki[q + 1];                                                        \
                        acc0 += Rc[a * RT + k0] * Yc[mi * RT + k0];                                                  \
                        acc1 += Rc[a * RT + k1] * Yc[mi * RT + k1];                                                  \
                    }                                                                                                \
                    if (q < n) {                                                                                     \
                        const int k0 = ki[q];                                                                        \
                        acc0 += Rc[a * RT + k0] * Yc[mi * RT + k0];                                                  \
                    }                                                                                                \
                    cPc[(si * N + a) * LM + l * l + mi] += acc0 + acc1;                                              \
                }                                                                                                    \
            }

// PREG: each lane's Legendre levels in registers, with the level loop unrolled (no
// LDS rows, more registers: two wavefronts per SIMD); otherwise two LDS rows per
// lane and a rolled level loop (three wavefronts per SIMD).
// CPW: centres per wavefront.  A centre's chunk of 32 neighbours occupies 32 lanes,
// so with two centres lanes 0-31 and 32-63 stage the neighbours of one centre each
// and every staging instruction serves two centres.
template <int N, int L, bool PREG, int CPW>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(PREG ? 2 : WPE_FWD, 8))) void k_soap_fwd_x(
    SoapDev sp, int nc, const int *__restrict__ centres, const int *__restrict__ Z, const int *__restrict__ nb_off,
    const int *__restrict__ nb_j, const double *__restrict__ nb_r, const double *__restrict__ nb_d, double *X,
    double *norm_out, double *cbuf) {
    extern __shared__ double lds[];
    constexpr int LM = (L + 1) * (L + 1);
    constexpr int YW = 2 * L + 1;
    constexpr int NB = kSoapChunk;
    const int K = sp.K, KLM = K * LM, S = sp.S;
    const int KLMp = (KLM + 1) & ~1;   // per-centre c' rows, 16-byte aligned
    double *Ts = lds;              // N*N
    double *rbs = Ts + N * N;      // N grid points, then exp(-alpha r_a^2) and 1/a (N each)
    double *gA = rbs + N, *iA = gA + N;
    double *cP = rbs + 3 * N;      // CPW x K*LM: c' during accumulation, c after the transform
    // The staged levels are stored one row per radial point a (Rl) and per m slot (Yl),
    // neighbours along the row: the writes of 32 lanes are consecutive, and the dense
    // accumulation reads two neighbours per 16-byte access.  The row stride RT = NB + 2
    // keeps rows 16-byte aligned (N (N + 3) and KLMp are even) and puts successive rows
    // 16 bytes apart in the bank cycle.
    constexpr int RT = NB + 2;     // row stride of Rl and Yl
    constexpr int PR = (2 * (L + 2)) | 1;   // row stride of a lane's two Legendre rows (odd: distinct banks)
    double *Rl = cP + CPW * KLMp;  // CPW x N x RT: fc F_l(a) of neighbour k of centre h at Rl[(h N + a) RT + k]
    double *Yl = Rl + CPW * N * RT;   // CPW x YW x RT: Y_l,m of neighbour k at Yl[(h YW + slot) RT + k]
    double *Pq0 = Yl + CPW * YW * RT; // 32 CPW x PR: the staging lanes' Legendre rows (not PREG)
    int *kidx = (int *)(Pq0 + (PREG ? 0 : 32 * CPW * PR));   // CPW x S x NB: the chunk's neighbours of each species
    int *ns = kidx + CPW * S * NB; // their counts
    const int lane = threadIdx.x;
    const int h = CPW == 2 ? lane >> 5 : 0, kl = lane & 31;   // this lane's centre and neighbour slot
    const int desc = CPW * blockIdx.x + h;
    const bool valid = desc < nc;   // this lane's centre exists (the same for all lanes with CPW 1)
    const int ci = valid ? centres[desc] : 0;
    double *Rh = Rl + h * (N * RT), *Yh = Yl + h * (YW * RT);
    for (int e = lane; e < N * N; e += 64) Ts[e] = sp.T[e];
    if (lane < N) {
        rbs[lane] = sp.rb[lane];
        gA[lane] = exp(-sp.alpha * sp.rb[lane] * sp.rb[lane]);
        iA[lane] = lane > 0 ? 1.0 / (double)lane : 0.0;
    }
    for (int e = lane; e < CPW * KLMp; e += 64) cP[e] = 0.0;
    __syncthreads();
    const int p0 = valid ? nb_off[ci] : 0, np = valid ? nb_off[ci + 1] - p0 : 0;
    int npmax = np;   // chunks of the wave (uniform): those of its larger centre
    if (CPW == 2) npmax = max(np, __shfl_xor(np, 32, 64));
    for (int c0 = 0; c0 < npmax; c0 += NB) {
        const int nk = max(0, min(NB, np - c0));
        __syncthreads();   // the previous chunk's reads of the lists are done
        int s = -1;
        double fc = 0.0, ux = 0.0, uy = 0.0, uz = 0.0;
        Rad<N> rs;
        if (kl < nk && lane < 32 * CPW) {
            const int p = p0 + c0 + kl;
            const double r = nb_r[p];
            s = species_s(sp, Z[nb_j[p]]);
            if (r > sp.rc) s = -1;
            if (s >= 0) {
                fc = fc_s(r, sp.rc, sp.ctw);
                rs.init(rbs, gA, iA, sp.alpha, r);
                const double ir = 1.0 / r;
                ux = nb_d[3 * p] * ir; uy = nb_d[3 * p + 1] * ir; uz = nb_d[3 * p + 2] * ir;
            }
        }
        for (int si = 0; S > 1 && si < S; si++) {   // neighbours beyond the cutoff are in no list
            const bool mine = lane < 32 * CPW && s == si;
            const unsigned long long mk = __ballot(mine);
            const unsigned long long mh = (mk >> (32 * h)) & 0xffffffffull;   // this centre's half
            if (mine) kidx[(h * S + si) * NB + __popcll(mh & ((1ull << kl) - 1ull))] = kl;
            if (kl == 0 && lane < 32 * CPW) ns[h * S + si] = __popcll(mh);
        }
        if constexpr (PREG) {
            double P0[L + 2], P1[L + 2];
#pragma unroll
            for (int l = 0; l <= L; l++) {
                double *cur = (l & 1) ? P1 : P0;
                const double *prv = (l & 1) ? P0 : P1;
                SOAP_FWD_LEVEL(l, cur, prv);
            }
        } else {
            double *Pq = Pq0 + lane * PR;   // this lane's two Legendre rows
#pragma unroll 1
            for (int l = 0; l <= L; l++) {
                double *cur = Pq + (l & 1) * (L + 2);
                const double *prv = Pq + ((l + 1) & 1) * (L + 2);
                SOAP_FWD_LEVEL(l, cur, prv);
            }
        }
    }
    __syncthreads();
    // c = T^T c', one (centre, species, lm) column per lane, in place
    for (int it = lane; it < CPW * S * LM; it += 64) {
        const int hh = CPW == 1 ? 0 : it / (S * LM), r_ = it - hh * (S * LM), si = r_ / LM, lm = r_ - si * LM;
        double *cc = cP + hh * KLMp;
        RELOAD_BARRIER();   // T is re-read from LDS for every column
        double cp[N];
#pragma unroll
        for (int a = 0; a < N; a++) cp[a] = cc[(si * N + a) * LM + lm];
#pragma unroll
        for (int b = 0; b < N; b++) {
            double acc = 0.0;
#pragma unroll
            for (int a = 0; a < N; a++) acc += Ts[a * N + b] * cp[a];
            cc[(si * N + b) * LM + lm] = acc;
            RELOAD_BARRIER();   // one column of T in registers at a time
        }
    }
    __syncthreads();
    if (kl == 0 && valid && lane < 32 * CPW) {   // centre atom term
        const int sc = species_s(sp, Z[ci]);
        if (sc >= 0) cP[h * KLMp + (sc * N) * LM] += sp.cw * sp.u00 * c_ynorm[0];
    }
    __syncthreads();
    // power spectrum, written unnormalised and rescaled in place by the same lanes
    const int d = sp.dim;
    const int d0 = CPW * blockIdx.x;
    double ss[CPW];
#pragma unroll
    for (int hh = 0; hh < CPW; hh++) ss[hh] = 0.0;
    for (int it = lane; it < CPW * (d - 1); it += 64) {
        const int hh = CPW == 1 ? 0 : it / (d - 1), q = it - hh * (d - 1);
        if (d0 + hh >= nc) continue;
        const double *cc = cP + hh * KLMp;
        const int code = sp.qmap[q];
        const int ia = code >> 16, jb = (code >> 8) & 0xff, l = code & 0xff;
        double v = 0.0;
        for (int m = -l; m <= l; m++) v += cc[ia * LM + l * l + l + m] * cc[jb * LM + l * l + l + m];
        v *= (ia != jb ? 1.4142135623730950488 : 1.0) / sqrt(2.0 * l + 1.0);
        X[(size_t)(d0 + hh) * d + q] = v;
#pragma unroll
        for (int g2 = 0; g2 < CPW; g2++)
            if (g2 == hh) ss[g2] += v * v;
    }
    double inv[CPW];
#pragma unroll
    for (int hh = 0; hh < CPW; hh++) {
        const double nrm = sqrt(wave_sum_s(ss[hh]));
        inv[hh] = nrm > 0.0 ? 1.0 / nrm : 0.0;
        if (lane == 0 && d0 + hh < nc) {
            X[(size_t)(d0 + hh) * d + d - 1] = 0.0;
            norm_out[d0 + hh] = nrm;
        }
    }
    for (int it = lane; it < CPW * (d - 1); it += 64) {
        const int hh = CPW == 1 ? 0 : it / (d - 1), q = it - hh * (d - 1);
        if (d0 + hh >= nc) continue;
        double sc = inv[0];
#pragma unroll
        for (int g2 = 1; g2 < CPW; g2++)
            if (g2 == hh) sc = inv[g2];
        X[(size_t)(d0 + hh) * d + q] *= sc;
    }
    for (int hh = 0; hh < CPW; hh++) {
        if (d0 + hh >= nc) break;
        double *co = cbuf + (size_t)(d0 + hh) * KLM;
        for (int e = lane; e < KLM; e += 64) co[e] = cP[hh * KLMp + e];
    }
}

// Adjoint.  dE/dc in the orthonormal basis is transformed once per centre to
// the primitive basis, dc'[a][lm] = sum_b T[a][b] dc[b][lm], so that per
// neighbour and (l, m): sum_b dc g_b = fc S0 and sum_b dc dg_b/dr =
// fc (cl S0 + 2 alpha S1) + dfc S0 with S0 = sum_a F_l(a) dc'[a], S1 =
// sum_a r_a F_(l+1)(a) dc'[a] and cl = -2 alpha r + l / r.
// OCC: the occupancy the register budget is cut for.  The higher one pays off when
// there are many more wavefronts than SIMDs; a small launch runs faster with the
// larger budget.
template <int N, int L, int OCC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(OCC, 8))) void k_soap_bwd_x(
    SoapDev sp, int nc, int cpb, const int *__restrict__ centres, const int *__restrict__ Z,
    const int *__restrict__ nb_off, const int *__restrict__ nb_j, const double *__restrict__ nb_r,
    const double *__restrict__ nb_d, const double *__restrict__ X, const double *__restrict__ G,
    const double *__restrict__ norm, const double *__restrict__ cbuf, double *pairF, int whole) {
    // cpb centres per wavefront: a centre has ~30 neighbours, so one lane per
    // neighbour over 64/cpb lanes keeps the wavefront busy.
    // LDS: T, the grid tables, dE/dc (then dE/dc') of each centre, the centre
    // scalars, and one staging region used first for the per-l blocks of the
    // dE/dc product and then for the lanes' Legendre rows.
    extern __shared__ double lds[];
    constexpr int LM = (L + 1) * (L + 1);
    constexpr int PR = (2 * (L + 2)) | 1;   // odd: a lane's rows on distinct banks
    const int K = sp.K, KLM = K * LM, d = sp.dim, S = sp.S;
    double *Ts = lds;              // N*N
    double *rbs = Ts + N * N;      // N grid points, then exp(-alpha r_a^2) and 1/a (N each)
    double *gA = rbs + N, *iA = gA + N;
    double *dcs_all = rbs + 3 * N; // cpb x K*LM
    double *csc = dcs_all + cpb * KLM;   // cpb x (dot, 1/|p|)
    double *stg = csc + 2 * cpb;   // max(K*K + K*(2L+1), 64*PR), or with whole max(cpb*(L+1)*(K*K+2) | cpb*d, 64*PR)
    const int lane = threadIdx.x;
    const int d0 = cpb * blockIdx.x;
    const int nh = min(cpb, nc - d0);
    for (int e = lane; e < N * N; e += 64) Ts[e] = sp.T[e];
    if (lane < N) {
        rbs[lane] = sp.rb[lane];
        gA[lane] = exp(-sp.alpha * sp.rb[lane] * sp.rb[lane]);
        iA[lane] = lane > 0 ? 1.0 / (double)lane : 0.0;
    }
    for (int h = 0; h < nh; h++) {
        const int desc = d0 + h;
        const double *x = X + (size_t)desc * d, *gg = G + (size_t)desc * d;
        double dot = 0.0;
        for (int q = lane; q < d; q += 64) dot += x[q] * gg[q];
        dot = wave_sum_s(dot);
        if (lane == 0) {
            const double nrm = norm[desc];
            csc[2 * h] = dot;
            csc[2 * h + 1] = nrm > 0.0 ? 1.0 / nrm : 0.0;
        }
        const double nrm = norm[desc];
        const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
        if (whole == 1) {   // dE/dp of every centre staged together; c is read from cbuf
            double *dpL = stg + h * d;
            for (int q = lane; q < d; q += 64) dpL[q] = (gg[q] - x[q] * dot) * inv;
        } else if (whole == 2) {   // the level blocks P_l (below) of every centre, all levels
            // level blocks K*K + 2 doubles apart: rows of different levels read by one group
            // of lanes then fall on different bank groups
            const int KKp = K * K + 2;
            double *Ph = stg + h * (L + 1) * KKp;
            for (int e = lane; e < (L + 1) * K * K; e += 64) {
                const int l = e / (K * K), r = e - l * K * K, ia = r / K, jb = r - ia * K;
                const int hi = ia > jb ? ia : jb, lo = ia > jb ? jb : ia;
                const int q = (hi * (hi + 1) / 2 + lo) * (L + 1) + l;
                Ph[l * KKp + r] = (gg[q] - x[q] * dot) * inv * (ia == jb ? 2.0 : 1.4142135623730950488) * c_isq[l];
            }
        }
    }
    __syncthreads();
    if (whole == 2) {   // dE/dc in one pass over (centre, ia, lm): row ia of P_l times column lm of c
        for (int it = lane; it < nh * KLM; it += 64) {
            const int h = it / KLM, e = it - h * KLM;
            const int ia = e / LM, lm = e - ia * LM;
            const int l = (int)sqrtf((float)lm);   // exact: lm < (l+1)^2 is far from a rounding step
            const double *Pr = stg + h * (L + 1) * (K * K + 2) + l * (K * K + 2) + ia * K;
            const double *cL = cbuf + (size_t)(d0 + h) * KLM + lm;
            double acc = 0.0;
            if (S == 1) {
#pragma unroll
                for (int jb = 0; jb < N; jb++) acc += Pr[jb] * cL[jb * LM];
            } else {
                for (int jb = 0; jb < K; jb++) acc += Pr[jb] * cL[jb * LM];
            }
            dcs_all[h * KLM + e] = acc;
        }
        __syncthreads();
    }
    if (whole == 1) {   // dE/dc in one pass over (centre, ia, lm)
        for (int it = lane; it < nh * KLM; it += 64) {
            const int h = it / KLM, e = it - h * KLM;
            const double *dpL = stg + h * d, *cL = cbuf + (size_t)(d0 + h) * KLM;
            const int ia = e / LM, lm = e - ia * LM;
            int l = (int)sqrt((double)lm + 0.5);
            if (l * l > lm) l--;
            double acc = 0.0;
            for (int jb = 0; jb < K; jb++) {
                const int hi = ia > jb ? ia : jb, lo = ia > jb ? jb : ia;
                const int q = (hi * (hi + 1) / 2 + lo) * (L + 1) + l;
                acc += dpL[q] * (ia == jb ? 2.0 : 1.4142135623730950488) * cL[jb * LM + lm];
            }
            dcs_all[h * KLM + e] = acc / sqrt(2.0 * l + 1.0);
        }
        __syncthreads();
    }
    // otherwise dE/dc per centre and level l: dc_l (K x 2l+1) = P_l (K x K) c_l (K x 2l+1) with
    // P_l[ia][jb] = dE/dp_(ia,jb),l x (2 if ia = jb else sqrt 2) / sqrt(2l+1) and
    // dE/dp = (G - x (x.G)) / |p|; both blocks are staged in LDS one level at a time.
    for (int h = 0; h < (whole ? 0 : nh); h++) {
        const int desc = d0 + h;
        const double *x = X + (size_t)desc * d, *gg = G + (size_t)desc * d;
        const double *cg = cbuf + (size_t)desc * KLM;
        const double dot = csc[2 * h], inv = csc[2 * h + 1];
        double *dcL = dcs_all + h * KLM;
        for (int l = 0; l <= L; l++) {
            const int W = 2 * l + 1;
            const double il = 1.0 / sqrt(2.0 * l + 1.0);
            double *Pl = stg, *Cl = stg + K * K;
            for (int e = lane; e < K * K; e += 64) {
                const int ia = e / K, jb = e - ia * K;
                const int hi = ia > jb ? ia : jb, lo = ia > jb ? jb : ia;
                const int q = (hi * (hi + 1) / 2 + lo) * (L + 1) + l;
                Pl[e] = (gg[q] - x[q] * dot) * inv * (ia == jb ? 2.0 : 1.4142135623730950488) * il;
            }
            for (int e = lane; e < K * W; e += 64) {
                const int jb = e / W, m = e - jb * W;
                Cl[e] = cg[jb * LM + l * l + m];
            }
            __syncthreads();
            for (int e = lane; e < K * W; e += 64) {
                const int ia = e / W, m = e - ia * W;
                double acc = 0.0;
                for (int jb = 0; jb < K; jb++) acc += Pl[ia * K + jb] * Cl[jb * W + m];
                dcL[ia * LM + l * l + m] = acc;
            }
            __syncthreads();
        }
    }
    // dc -> dc' in place, one (centre, species, lm) column per lane
    for (int it = lane; it < nh * S * LM; it += 64) {
        const int h = it / (S * LM), rem = it - h * S * LM, si = rem / LM, lm = rem - si * LM;
        double *col = dcs_all + h * KLM + si * N * LM + lm;
        RELOAD_BARRIER();   // T is re-read from LDS for every column
        double v[N];
#pragma unroll
        for (int b = 0; b < N; b++) v[b] = col[b * LM];
#pragma unroll
        for (int a = 0; a < N; a++) {
            double acc = 0.0;
#pragma unroll
            for (int b = 0; b < N; b++) acc += Ts[a * N + b] * v[b];
            col[a * LM] = acc;
            RELOAD_BARRIER();   // one row of T in registers at a time
        }
    }
    __syncthreads();
    const int lanes_per = 64 / cpb;
    const int h = lane / lanes_per, sub = lane - h * lanes_per;
    if (h >= nh) return;
    const int ci = centres[d0 + h];
    const double *dcP = dcs_all + h * KLM;
    double *Pq = stg + lane * PR;   // this lane's two Legendre rows
    const int p0 = nb_off[ci], np = nb_off[ci + 1] - p0;
    const double al2 = 2.0 * sp.alpha;
    const double dr = rbs[1];   // grid spacing: r_a = a dr
    for (int k = sub; k < np; k += lanes_per) {
        const int p = p0 + k;
        const double r = nb_r[p];
        const int s = species_s(sp, Z[nb_j[p]]);
        if (r > sp.rc || s < 0) continue;
        double fc, dfc;
        fc_dfc_s(r, sp.rc, sp.ctw, fc, dfc);
        const double ir = 1.0 / r;
        const double ux = nb_d[3 * p] * ir, uy = nb_d[3 * p + 1] * ir, uz = nb_d[3 * p + 2] * ir;
        const double *dcs = dcP + s * N * LM;
        Rad<N> rs;
        rs.init(rbs, gA, iA, sp.alpha, r);
        rs.scale_f1();
        // Per (l, m): S0 = sum_a f_l(a) dc'[a], S1 = sum_a r_a f_(l+1)(a) dc'[a] (dr times the
        // scaled sum), s2 = fc S0 and s1 = fc (cl S0 + 2 alpha S1) + dfc S0 with cl = -2 alpha r + l / r;
        // the force on the pair is sum (s1 Y - s2 (u.gY) / r) u + sum s2 gY / r over (l, m), gY the
        // angular gradient of Y.  Only the sums T0 = sum S0 Y, T0L = sum l S0 Y, T1 = sum S1 Y and
        // V = sum S0 gY are kept; a pair of real harmonics (m and -m) is handled together.
        double T0 = 0.0, T0L = 0.0, T1 = 0.0, Vx = 0.0, Vy = 0.0, Vz = 0.0;
#pragma unroll 1
        for (int l = 0; l <= L; l++) {
            double *cur = Pq + (l & 1) * (L + 2);
            legendre_level(l, uz, Pq + ((l + 1) & 1) * (L + 2), cur);
            cur[l + 1] = 0.0;
            double A = 1.0, B = 0.0, Am = 0.0, Bm = 0.0;
            double t0 = 0.0;
            for (int m = 0; m <= l; m++) {
                if (m > 0) {
                    Am = A;
                    Bm = B;
                    const double An = A * ux - B * uy;
                    B = A * uy + B * ux;
                    A = An;
                }
                const double Nn = c_ynorm[l * (l + 1) / 2 + m];
                const double NnP = Nn * cur[m], NnPz = Nn * cur[m + 1];
                const double *dca = dcs + l * l + l + m, *dcb = dcs + l * l + l - m;
                double S0a = 0.0, S1a = 0.0, S0b = 0.0, S1b = 0.0;
#pragma unroll
                for (int a = 0; a < N; a++) {
                    const double x = dca[a * LM];
                    S0a += rs.F0[a] * x;
                    S1a += rs.F1[a] * x;
                }
                if (m > 0) {
#pragma unroll
                    for (int a = 0; a < N; a++) {
                        const double x = dcb[a * LM];
                        S0b += rs.F0[a] * x;
                        S1b += rs.F1[a] * x;
                    }
                }
                // Y = NnP (A, B), gY_z = NnPz (A, B), (gY_x, gY_y) = m NnP (Am, -Bm) and (Bm, Am)
                const double u0 = S0a * A + S0b * B, u1 = S1a * A + S1b * B;
                t0 += NnP * u0;
                T1 += NnP * u1;
                Vz += NnPz * u0;
                if (m > 0) {
                    const double NnPm = NnP * (double)m;
                    Vx += NnPm * (S0a * Am + S0b * Bm);
                    Vy += NnPm * (S0b * Am - S0a * Bm);
                }
            }
            T0 += t0;
            T0L += (double)l * t0;
            rs.step_scaled(l);
        }
        const double sY = (dfc - al2 * r * fc) * T0 + fc * ir * T0L + al2 * fc * dr * T1;
        const double gx = fc * Vx, gy = fc * Vy, gz = fc * Vz;
        const double a1 = sY - (ux * gx + uy * gy + uz * gz) * ir;
        const double fx = a1 * ux + gx * ir, fy = a1 * uy + gy * ir, fz = a1 * uz + gz * ir;
        pairF[3 * p] += fx;
        pairF[3 * p + 1] += fy;
        pairF[3 * p + 2] += fz;
    }
}

#define SOAP_FAST(X) X(4, 4) X(5, 4) X(6, 4) X(8, 4) X(6, 6) X(8, 6) X(10, 6) X(12, 6) X(8, 8) X(10, 8) X(12, 8)

int soap_fast_supported(int n, int L) {
#define CHK(NN, LL) if (n == NN && L == LL) return 1;
    SOAP_FAST(CHK)
#undef CHK
    return 0;
}

// Two centres per wave: single-species models with n <= 8, l <= 6 whose register
// state fits without spilling (n l >= 32: 8/4, 6/6, 8/6).
static constexpr bool fwd_pair_ok(int n, int L) { return n <= 8 && L <= 6 && n * L >= 32; }
// Only the variants the launcher can pick are instantiated.
template <int N, int L>
static void fwd_go(bool preg, int cpw, int nblk, size_t sh, hipStream_t s, const SoapDev &sp, int nc,
                   const int *centres, const int *Z, const int *nb_off, const int *nb_j, const double *nb_r,
                   const double *nb_d, double *X, double *norm, double *cbuf) {
    if constexpr (fwd_pair_ok(N, L)) {
        if (cpw == 2) {
            hipLaunchKernelGGL((k_soap_fwd_x<N, L, true, 2>), dim3(nblk), dim3(64), sh, s, sp, nc, centres, Z,
                               nb_off, nb_j, nb_r, nb_d, X, norm, cbuf);
            return;
        }
    }
    if constexpr (N < 10 && L < 8) {
        if (!preg) {
            hipLaunchKernelGGL((k_soap_fwd_x<N, L, false, 1>), dim3(nblk), dim3(64), sh, s, sp, nc, centres, Z,
                               nb_off, nb_j, nb_r, nb_d, X, norm, cbuf);
            return;
        }
    }
    hipLaunchKernelGGL((k_soap_fwd_x<N, L, true, 1>), dim3(nblk), dim3(64), sh, s, sp, nc, centres, Z, nb_off,
                       nb_j, nb_r, nb_d, X, norm, cbuf);
}

// LDS bytes of the forward block: transform, radial grid, c' accumulators, the staged
// neighbour chunk and (without registers for the Legendre levels) their LDS rows
static size_t fwd_fast_lds(const SoapDev &sp, int cpw, bool preg) {
    const int LM = (sp.L + 1) * (sp.L + 1), KLM = sp.K * LM, NB = kSoapChunk;
    return (size_t)(sp.n * sp.n + 3 * sp.n + cpw * ((KLM + 1) & ~1) + cpw * (NB + 2) * (sp.n + 2 * sp.L + 1) +
                    (preg ? 0 : 32 * cpw * ((2 * (sp.L + 2)) | 1))) * 8 +
           (size_t)cpw * sp.S * (NB + 1) * 4;
}

// LDS bytes of the adjoint block, with its staging choice: whole = 2 stages the level
// blocks P_l of all its centres, 1 dE/dp, 0 one level at a time; cpb centres per wave
static size_t bwd_fast_lds(const SoapDev &sp, int *cpb_out, int *whole_out) {
    const int LM = (sp.L + 1) * (sp.L + 1);
    const int KLM = sp.K * LM, PR = (2 * (sp.L + 2)) | 1;
    const size_t stg0 = std::max((size_t)sp.K * sp.K + (size_t)sp.K * (2 * sp.L + 1), (size_t)64 * PR);
    const size_t pl = (size_t)(sp.L + 1) * (sp.K * sp.K + 2);
    auto stg = [&](int cpb, int whole) {
        return whole == 2 ? std::max(stg0, cpb * pl) : whole == 1 ? std::max(stg0, (size_t)cpb * sp.dim) : stg0;
    };
    auto need = [&](int cpb, int whole) {
        return ((size_t)sp.n * sp.n + 3 * sp.n + (size_t)cpb * (KLM + 2) + stg(cpb, whole)) * 8;
    };
    const int whole = need(2, 2) <= 24 * 1024 ? 2 : need(2, 1) <= 24 * 1024 ? 1 : 0;
    const int cpb = need(2, whole) <= 24 * 1024 ? 2 : 1;   // two centres per wave when LDS stays small
    if (cpb_out) *cpb_out = cpb;
    if (whole_out) *whole_out = whole;
    return need(cpb, whole);
}

size_t soap_fast_lds_bytes(const SoapDev &sp) {
    if (!soap_fast_supported(sp.n, sp.L)) return 0;
    const int cpw = sp.S == 1 && fwd_pair_ok(sp.n, sp.L) ? 2 : 1;
    size_t f = fwd_fast_lds(sp, cpw, true);
    if (cpw == 1 && sp.n < 10 && sp.L < 8) f = std::max(f, fwd_fast_lds(sp, 1, false));
    return std::max(f, bwd_fast_lds(sp, nullptr, nullptr));
}

bool launch_soap_fwd_fast(const SoapDev &sp, int nc, const int *centres, const int *Z, const int *nb_off,
                          const int *nb_j, const double *nb_r, const double *nb_d, double *X, double *norm,
                          double *cbuf, hipStream_t s) {
    if (!soap_fast_supported(sp.n, sp.L)) return false;
    ensure_soap_constants();
    if (nc <= 0) return true;
    const int LM = (sp.L + 1) * (sp.L + 1);
    const int KLM = sp.K * LM;
    // neighbour chunk of 32: covers a centre's neighbours within 5 A in one pass
    // for covalent solids and keeps the block's LDS small
    const int NB = kSoapChunk;
    // Two centres per wave with the Legendre levels in registers for single-species
    // models up to n 8, l 6 (measured on the 8/6 model: batch512 1.22 -> 1.11 ms,
    // si4096 0.049 -> 0.043 ms); one centre per wave otherwise (the two-species 12/8
    // model runs 3.4 ms with one and 5.6 ms with two, its LDS per wave doubling).
    const int cpw = sp.S == 1 && fwd_pair_ok(sp.n, sp.L) ? 2 : 1;
    const bool preg = cpw == 2 || sp.n >= 10 || sp.L >= 8 || nc < 16384;
    const size_t sh = fwd_fast_lds(sp, cpw, preg);
    (void)KLM; (void)NB;
    const int nblk = (nc + cpw - 1) / cpw;
#define LAUNCH(NN, LL)                                                                                          \
    if (sp.n == NN && sp.L == LL) {                                                                             \
        fwd_go<NN, LL>(preg, cpw, nblk, sh, s, sp, nc, centres, Z, nb_off, nb_j, nb_r, nb_d, X, norm, cbuf);     \
        return true;                                                                                            \
    }
    SOAP_FAST(LAUNCH)
#undef LAUNCH
    return false;
}
#undef SOAP_FWD_LEVEL

template <int N, int L>
static void bwd_go(bool occ3, int nblk, size_t sh, hipStream_t s, const SoapDev &sp, int nc, int cpb,
                   const int *centres, const int *Z, const int *nb_off, const int *nb_j, const double *nb_r,
                   const double *nb_d, const double *X, const double *G, const double *norm, const double *cbuf,
                   double *pairF, int whole) {
    if constexpr (N <= 8) {
        if (occ3) {
            hipLaunchKernelGGL((k_soap_bwd_x<N, L, WPE_BWD + 1>), dim3(nblk), dim3(64), sh, s, sp, nc, cpb, centres,
                               Z, nb_off, nb_j, nb_r, nb_d, X, G, norm, cbuf, pairF, whole);
            return;
        }
    }
    hipLaunchKernelGGL((k_soap_bwd_x<N, L, WPE_BWD>), dim3(nblk), dim3(64), sh, s, sp, nc, cpb, centres, Z, nb_off,
                       nb_j, nb_r, nb_d, X, G, norm, cbuf, pairF, whole);
}

bool launch_soap_bwd_fast(const SoapDev &sp, int nc, const int *centres, const int *Z, const int *nb_off,
                          const int *nb_j, const double *nb_r, const double *nb_d, const double *X, const double *G,
                          const double *norm, const double *cbuf, double *pairF, hipStream_t s) {
    if (!soap_fast_supported(sp.n, sp.L)) return false;
    ensure_soap_constants();
    if (nc <= 0) return true;
    // whole: all centres of the block staged at once (small descriptors): 2, the level blocks
    // P_l (K x K each); 1, dE/dp; 0, one level at a time
    int cpb = 1, whole = 0;
    const size_t sh = bwd_fast_lds(sp, &cpb, &whole);
    const int nblk = (nc + cpb - 1) / cpb;
    // three wavefronts per SIMD for n <= 8 (161 VGPRs without spills) once the launch has
    // at least four per SIMD to schedule
    const bool occ3 = sp.n <= 8 && nblk >= 4 * 1024;
#define LAUNCH(NN, LL)                                                                                          \
    if (sp.n == NN && sp.L == LL) {                                                                             \
        bwd_go<NN, LL>(occ3, nblk, sh, s, sp, nc, cpb, centres, Z, nb_off, nb_j, nb_r, nb_d, X, G, norm,   \
                       cbuf, pairF, whole);                                                                     \
        return true;                                                                                            \
    }
    SOAP_FAST(LAUNCH)
#undef LAUNCH
    return false;
}

}  // namespace gapx
