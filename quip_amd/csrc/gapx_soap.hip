// SOAP descriptor kernels specialised on (n_max, l_max).
//
// One wavefront per centre atom.  The radial transform T (n x n) and the
// radial grid live in LDS; an empty asm with a memory clobber closes every
// l step so the compiler re-reads T from LDS instead of pinning n^2 doubles in
// VGPRs.  All loops over b, a, l and m are unrolled at compile time, so the
// per-lane recursion state (radial: F_l, F_(l+1), 1/x per grid point;
// Legendre: two previous levels) stays in registers with no scratch.
//
//  forward: lanes = neighbours of a chunk write [g_(b,l) | Y_lm] rows into LDS,
//           lanes = (species, b, lm) accumulate c, lanes = power-spectrum
//           entries form p, a wave reduction gives |p|.
//  adjoint: dE/dp -> dE/dc (LDS), then lanes = neighbours re-evaluate
//           g, dg/dr, Y and grad Y and contract them with dE/dc.
#include "gapx_kernels.h"

#include <algorithm>

namespace gapx {

#define GAPX_PI_S 3.14159265358979323846264338327950288

// N_lm of the real spherical harmonics (x sqrt 2 for m > 0) and (2m-1)!!
static __constant__ double c_ynorm[(kMaxL + 1) * (kMaxL + 2) / 2];
static __constant__ double c_dfact[kMaxL + 2];

static bool g_soap_const_ready = false;
static void ensure_soap_constants() {
    if (g_soap_const_ready) return;
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
    (void)hipMemcpyToSymbol(HIP_SYMBOL(c_ynorm), yn, sizeof(yn));
    (void)hipMemcpyToSymbol(HIP_SYMBOL(c_dfact), df, sizeof(df));
    g_soap_const_ready = true;
}

#define RELOAD_BARRIER() __asm__ volatile("" ::: "memory")

__device__ __forceinline__ double fc_s(double r, double rc, double w) {
    if (r > rc) return 0.0;
    if (r > rc - w) return 0.5 * (cos(GAPX_PI_S * (r - rc + w) / w) + 1.0);
    return 1.0;
}
__device__ __forceinline__ double dfc_s(double r, double rc, double w) {
    if (r > rc) return 0.0;
    if (r > rc - w) return -0.5 * GAPX_PI_S * sin(GAPX_PI_S * (r - rc + w) / w) / w;
    return 0.0;
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

// i_l recursion state for all grid points: F0 = f_l, F1 = f_(l+1)
template <int N>
struct Rad {
    double F0[N], F1[N], ix[N];
    __device__ __forceinline__ void init(const double *rbs, double alpha, double r) {
#pragma unroll
        for (int a = 0; a < N; a++) {
            const double ra = rbs[a];
            const double x = 2.0 * alpha * r * ra;
            if (x == 0.0) {
                F0[a] = exp(-alpha * (ra * ra + r * r));
                F1[a] = 0.0;
                ix[a] = 0.0;
            } else {
                const double ep = exp(-alpha * (r + ra) * (r + ra)), em = exp(-alpha * (r - ra) * (r - ra));
                const double iv = 1.0 / x;
                F0[a] = 0.5 * (em - ep) * iv;
                F1[a] = 0.5 * (em + ep) * iv - F0[a] * iv;
                ix[a] = iv;
            }
        }
    }
    __device__ __forceinline__ void step(int l) {
#pragma unroll
        for (int a = 0; a < N; a++) {
            const double n2 = F0[a] - (2.0 * l + 3.0) * ix[a] * F1[a];
            F0[a] = F1[a];
            F1[a] = (ix[a] == 0.0) ? 0.0 : n2;
        }
    }
};

template <int N, int L>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE_FWD, 8))) void k_soap_fwd_x(SoapDev sp, const int *__restrict__ centres,
                                                   const int *__restrict__ Z, const int *__restrict__ nb_off,
                                                   const int *__restrict__ nb_j, const double *__restrict__ nb_r,
                                                   const double *__restrict__ nb_d, double *X, double *norm_out,
                                                   double *cbuf, int NB, int rows_cap) {
    extern __shared__ double lds[];
    constexpr int LM = (L + 1) * (L + 1);
    constexpr int GL = N * (L + 1);
    constexpr int RL = GL + LM;
    const int K = sp.K, KLM = K * LM;
    double *Ts = lds;                 // N*N
    double *rbs = Ts + N * N;         // N
    double *cL = rbs + N;             // K*LM
    double *rows = cL + KLM;          // rows_cap
    int *spc = (int *)(rows + rows_cap);
    const int lane = threadIdx.x;
    const int desc = blockIdx.x;
    const int ci = centres[desc];
    for (int e = lane; e < N * N; e += 64) Ts[e] = sp.T[e];
    if (lane < N) rbs[lane] = sp.rb[lane];
    for (int e = lane; e < KLM; e += 64) cL[e] = 0.0;
    __syncthreads();
    if (lane == 0) {
        const int sc = species_s(sp, Z[ci]);
        if (sc >= 0) cL[(sc * N) * LM] += sp.cw * sp.u00 * c_ynorm[0];
    }
    const int p0 = nb_off[ci], np = nb_off[ci + 1] - p0;
    for (int c0 = 0; c0 < np; c0 += NB) {
        const int nk = min(NB, np - c0);
        __syncthreads();
        if (lane < nk) {
            const int p = p0 + c0 + lane;
            const double r = nb_r[p];
            int s = species_s(sp, Z[nb_j[p]]);
            double *row = rows + (size_t)lane * RL;
            if (r > sp.rc) s = -1;
            if (s >= 0) {
                const double fc = fc_s(r, sp.rc, sp.ctw);
                Rad<N> rs;
                rs.init(rbs, sp.alpha, r);
#pragma unroll
                for (int l = 0; l <= L; l++) {
#pragma unroll
                    for (int b = 0; b < N; b++) {
                        double acc = 0.0;
#pragma unroll
                        for (int a = 0; a < N; a++) acc += rs.F0[a] * Ts[a * N + b];
                        row[b * (L + 1) + l] = fc * acc;
                    }
                    rs.step(l);
                    RELOAD_BARRIER();
                }
                const double ir = 1.0 / r;
                const double ux = nb_d[3 * p] * ir, uy = nb_d[3 * p + 1] * ir, uz = nb_d[3 * p + 2] * ir;
                double *Yr = row + GL;
                double Pm1[L + 1], Pm2[L + 1];
#pragma unroll
                for (int l = 0; l <= L; l++) {
                    double Pc[L + 1];
#pragma unroll
                    for (int m = 0; m <= l; m++) {
                        if (m == l) Pc[m] = c_dfact[m];
                        else if (m == l - 1) Pc[m] = (2.0 * l - 1.0) * uz * Pm1[m];
                        else Pc[m] = ((2.0 * l - 1.0) * uz * Pm1[m] - (double)(l + m - 1) * Pm2[m]) * (1.0 / (double)(l - m));
                    }
                    double A = 1.0, B = 0.0;
#pragma unroll
                    for (int m = 0; m <= l; m++) {
                        if (m > 0) {
                            const double An = A * ux - B * uy;
                            B = A * uy + B * ux;
                            A = An;
                        }
                        const double Nn = c_ynorm[l * (l + 1) / 2 + m] * Pc[m];
                        if (m == 0) Yr[l * l + l] = Nn;
                        else {
                            Yr[l * l + l + m] = Nn * A;
                            Yr[l * l + l - m] = Nn * B;
                        }
                    }
#pragma unroll
                    for (int m = 0; m <= l; m++) {
                        Pm2[m] = Pm1[m];
                        Pm1[m] = Pc[m];
                    }
                }
            }
            spc[lane] = s;
        }
        __syncthreads();
        // lanes over lm; each lane accumulates all radial channels of one species at a time.
        // spc[k] is the same for every lane, so the species test is a uniform branch.
        for (int lm = lane; lm < LM; lm += 64) {
            int l = (int)sqrt((double)lm + 0.5);
            if (l * l > lm) l--;
            for (int sp_i = 0; sp_i < K / N; sp_i++) {
                double acc[N];
#pragma unroll
                for (int b = 0; b < N; b++) acc[b] = 0.0;
                bool any = false;
                for (int k = 0; k < nk; k++) {
                    if (spc[k] != sp_i) continue;
                    any = true;
                    const double *row = rows + (size_t)k * RL;
                    const double y = row[GL + lm];
#pragma unroll
                    for (int b = 0; b < N; b++) acc[b] += row[b * (L + 1) + l] * y;
                }
                if (any) {
#pragma unroll
                    for (int b = 0; b < N; b++) cL[(sp_i * N + b) * LM + lm] += acc[b];
                }
            }
        }
    }
    __syncthreads();
    double *pb = rows;
    const int d = sp.dim;
    double ss = 0.0;
    for (int q = lane; q < d - 1; q += 64) {
        const int code = sp.qmap[q];
        const int ia = code >> 16, jb = (code >> 8) & 0xff, l = code & 0xff;
        double v = 0.0;
        for (int m = -l; m <= l; m++) v += cL[ia * LM + l * l + l + m] * cL[jb * LM + l * l + l + m];
        v *= (ia != jb ? 1.4142135623730950488 : 1.0) / sqrt(2.0 * l + 1.0);
        pb[q] = v;
        ss += v * v;
    }
    ss = wave_sum_s(ss);
    const double nrm = sqrt(ss);
    const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
    double *xo = X + (size_t)desc * d;
    for (int q = lane; q < d - 1; q += 64) xo[q] = pb[q] * inv;
    if (lane == 0) {
        xo[d - 1] = 0.0;
        norm_out[desc] = nrm;
    }
    double *co = cbuf + (size_t)desc * KLM;
    for (int e = lane; e < KLM; e += 64) co[e] = cL[e];
}

template <int N, int L>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE_BWD, 8))) void k_soap_bwd_x(
    SoapDev sp, int nc, int cpb, const int *__restrict__ centres, const int *__restrict__ Z, const int *__restrict__ nb_off,
    const int *__restrict__ nb_j, const double *__restrict__ nb_r, const double *__restrict__ nb_d,
    const double *__restrict__ X, const double *__restrict__ G, const double *__restrict__ norm,
    const double *__restrict__ cbuf, double *pairF) {
    // Two centres per wavefront: a centre has ~30 neighbours, so one lane per
    // neighbour over 32 lanes keeps the whole wavefront busy.
    extern __shared__ double lds[];
    constexpr int LM = (L + 1) * (L + 1);
    const int K = sp.K, KLM = K * LM, d = sp.dim;
    const int per = 2 * KLM + d;
    double *Ts = lds;              // N*N
    double *rbs = Ts + N * N;      // N
    double *cen = rbs + N;         // 2 x [cL (K*LM) | dcL (K*LM) | dpL (d)]
    const int lane = threadIdx.x;
    const int d0 = cpb * blockIdx.x;
    const int nh = min(cpb, nc - d0);
    for (int e = lane; e < N * N; e += 64) Ts[e] = sp.T[e];
    if (lane < N) rbs[lane] = sp.rb[lane];
    for (int h = 0; h < nh; h++) {
        const int desc = d0 + h;
        double *cL = cen + h * per, *dpL = cL + 2 * KLM;
        const double *x = X + (size_t)desc * d, *gg = G + (size_t)desc * d;
        double dot = 0.0;
        for (int q = lane; q < d; q += 64) dot += x[q] * gg[q];
        dot = wave_sum_s(dot);
        const double nrm = norm[desc];
        const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
        for (int q = lane; q < d; q += 64) dpL[q] = (gg[q] - x[q] * dot) * inv;
        for (int e = lane; e < KLM; e += 64) cL[e] = cbuf[(size_t)desc * KLM + e];
    }
    __syncthreads();
    for (int it = lane; it < nh * KLM; it += 64) {
        const int h = it / KLM, e = it - h * KLM;
        const double *cL = cen + h * per, *dpL = cL + 2 * KLM;
        double *dcL = cen + h * per + KLM;
        const int ia = e / LM, lm = e - ia * LM;
        int l = (int)sqrt((double)lm + 0.5);
        if (l * l > lm) l--;
        double acc = 0.0;
        for (int jb = 0; jb < K; jb++) {
            const int hi = ia > jb ? ia : jb, lo = ia > jb ? jb : ia;
            const int q = (hi * (hi + 1) / 2 + lo) * (L + 1) + l;
            acc += dpL[q] * (ia == jb ? 2.0 : 1.4142135623730950488) * cL[jb * LM + lm];
        }
        dcL[e] = acc / sqrt(2.0 * l + 1.0);
    }
    __syncthreads();
    const int lanes_per = 64 / cpb;
    const int h = lane / lanes_per, sub = lane - h * lanes_per;
    if (h >= nh) return;
    const int ci = centres[d0 + h];
    const double *dcL = cen + h * per + KLM;
    const int p0 = nb_off[ci], np = nb_off[ci + 1] - p0;
    for (int k = sub; k < np; k += lanes_per) {
        const int p = p0 + k;
        const double r = nb_r[p];
        const int s = species_s(sp, Z[nb_j[p]]);
        if (r > sp.rc || s < 0) continue;
        const double fc = fc_s(r, sp.rc, sp.ctw), dfc = dfc_s(r, sp.rc, sp.ctw);
        const double ir = 1.0 / r;
        const double ux = nb_d[3 * p] * ir, uy = nb_d[3 * p + 1] * ir, uz = nb_d[3 * p + 2] * ir;
        const double *dcs = dcL + s * N * LM;
        Rad<N> rs;
        rs.init(rbs, sp.alpha, r);
        double fx = 0.0, fy = 0.0, fz = 0.0;
        double Pm1[L + 1], Pm2[L + 1];
#pragma unroll
        for (int l = 0; l <= L; l++) {
            double gb[N], dgb[N];
            const double cl = -2.0 * sp.alpha * r + (double)l * ir;
#pragma unroll
            for (int b = 0; b < N; b++) {
                double g0 = 0.0, q1 = 0.0;
#pragma unroll
                for (int a = 0; a < N; a++) {
                    const double t = Ts[a * N + b];
                    g0 += rs.F0[a] * t;
                    q1 += rbs[a] * rs.F1[a] * t;
                }
                gb[b] = fc * g0;
                dgb[b] = fc * (cl * g0 + 2.0 * sp.alpha * q1) + dfc * g0;
            }
            rs.step(l);
            double Pc[L + 2];
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
                if (m > 0) {
                    Am = A;
                    Bm = B;
                    const double An = A * ux - B * uy;
                    B = A * uy + B * ux;
                    A = An;
                }
                const double Nn = c_ynorm[l * (l + 1) / 2 + m];
                const double P = Pc[m], Pz = Pc[m + 1];
#pragma unroll
                for (int sg = 0; sg < 2; sg++) {
                    if (m == 0 && sg == 1) continue;
                    double Y, Fx, Fy, Fz;
                    if (m == 0) { Y = Nn * P; Fx = 0.0; Fy = 0.0; Fz = Nn * Pz; }
                    else if (sg == 0) { Y = Nn * P * A; Fx = Nn * P * m * Am; Fy = -Nn * P * m * Bm; Fz = Nn * Pz * A; }
                    else { Y = Nn * P * B; Fx = Nn * P * m * Bm; Fy = Nn * P * m * Am; Fz = Nn * Pz * B; }
                    const int lm = l * l + l + (sg ? -m : m);
                    double s1 = 0.0, s2 = 0.0;
#pragma unroll
                    for (int b = 0; b < N; b++) {
                        const double dc = dcs[b * LM + lm];
                        s1 += dc * dgb[b];
                        s2 += dc * gb[b];
                    }
                    const double ud = ux * Fx + uy * Fy + uz * Fz;
                    const double a1 = s1 * Y - s2 * ud * ir, a2 = s2 * ir;
                    fx += a1 * ux + a2 * Fx;
                    fy += a1 * uy + a2 * Fy;
                    fz += a1 * uz + a2 * Fz;
                }
            }
#pragma unroll
            for (int m = 0; m <= l; m++) {
                Pm2[m] = Pm1[m];
                Pm1[m] = Pc[m];
            }
            RELOAD_BARRIER();
        }
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

bool launch_soap_fwd_fast(const SoapDev &sp, int nc, const int *centres, const int *Z, const int *nb_off,
                          const int *nb_j, const double *nb_r, const double *nb_d, double *X, double *norm,
                          double *cbuf, hipStream_t s) {
    if (!soap_fast_supported(sp.n, sp.L)) return false;
    ensure_soap_constants();
    if (nc <= 0) return true;
    const int LM = (sp.L + 1) * (sp.L + 1);
    const int RL = sp.n * (sp.L + 1) + LM;
    const int KLM = sp.K * LM;
    const size_t fixed = (size_t)(sp.n * sp.n + sp.n + KLM) * 8;
    auto need = [&](int nb) { return fixed + std::max((size_t)nb * RL, (size_t)sp.dim) * 8 + nb * 4; };
    int NB = 64;
    while (NB > 8 && need(NB) > 40 * 1024) NB /= 2;
    const int rows_cap = (int)std::max((size_t)NB * RL, (size_t)sp.dim);
    const size_t sh = need(NB);
#define LAUNCH(NN, LL)                                                                                          \
    if (sp.n == NN && sp.L == LL) {                                                                             \
        hipLaunchKernelGGL((k_soap_fwd_x<NN, LL>), dim3(nc), dim3(64), sh, s, sp, centres, Z, nb_off, nb_j,     \
                           nb_r, nb_d, X, norm, cbuf, NB, rows_cap);                                            \
        return true;                                                                                            \
    }
    SOAP_FAST(LAUNCH)
#undef LAUNCH
    return false;
}

bool launch_soap_bwd_fast(const SoapDev &sp, int nc, const int *centres, const int *Z, const int *nb_off,
                          const int *nb_j, const double *nb_r, const double *nb_d, const double *X, const double *G,
                          const double *norm, const double *cbuf, double *pairF, hipStream_t s) {
    if (!soap_fast_supported(sp.n, sp.L)) return false;
    ensure_soap_constants();
    if (nc <= 0) return true;
    const int LM = (sp.L + 1) * (sp.L + 1);
    const size_t one = (size_t)(2 * sp.K * LM + sp.dim) * 8;
    const int cpb = (sp.n * sp.n + sp.n) * 8 + 2 * one <= 48 * 1024 ? 2 : 1;   // two centres per wave when LDS allows
    const size_t sh = (size_t)(sp.n * sp.n + sp.n) * 8 + cpb * one;
#define LAUNCH(NN, LL)                                                                                          \
    if (sp.n == NN && sp.L == LL) {                                                                             \
        hipLaunchKernelGGL((k_soap_bwd_x<NN, LL>), dim3((nc + cpb - 1) / cpb), dim3(64), sh, s, sp, nc, cpb,     \
                           centres, Z,                                                                          \
                           nb_off, nb_j, nb_r, nb_d, X, G, norm, cbuf, pairF);                                  \
        return true;                                                                                            \
    }
    SOAP_FAST(LAUNCH)
#undef LAUNCH
    return false;
}

}  // namespace gapx
