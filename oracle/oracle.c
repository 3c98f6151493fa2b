/*
 * oracle.c -- FP64 CPU restatement of QUIP's SOAP-GAP E+F+virial path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg.  The HIP product
 * library does not link or call anything in this directory.
 *
 * What each part restates:
 *   neighbour list   src/libAtoms/Connection.f95:1035-1310 (calc_connect),
 *                    :448-526 (test_form_bond), :585-634 (add_bond),
 *                    :740-763 (calc_lat_eff), :1448-1513 (get_min_max_images),
 *                    :1519-1621 (partition_atoms, cell_of_pos), :1674-1691
 *                    (divide_cell), :2204-2237 (fit_box_in_cell), :2011-2138
 *                    (connection_neighbour), linearalgebra.f95:671-686 (.mult.),
 *                    :6721-6733 (matrix3x3_inverse), :4445-4467 (cross, triple)
 *   cutoff function  linearalgebra.f95:7488-7516 (coordination_function)
 *   Y_lm, grad Y_lm  angular_functions.f95:32-278 (complex, Condon-Shortley)
 *   SOAP             GAP descriptors.f95 soap (module absent from the reference
 *                    tree).  The recipe below was recovered from, and is pinned
 *                    by, tests/Si.soap.cur_points.json (100 vectors, 6e-16) and
 *                    tests/test_descriptor.py:88-190 (C2H gradients, 4e-10):
 *                      alpha = 1/(2 sigma^2)
 *                      grid r_a = a*rb/n_max, a=0..n_max-1,
 *                        rb = cutoff + sigma*sqrt(2*E*ln 10), E = basis_error_exponent
 *                      f_l(r; r_a) = exp(-alpha(r^2+r_a^2)) i_l(2 alpha r r_a)
 *                      C_ab = exp(-alpha (r_a-r_b)^2), S_ab = Gaussian overlap,
 *                      S = L L^T, radial coefficient g_b = sum_a f_a (C^-1 L)_ab
 *                      centre atom: channel (species of centre, b=0), l=m=0,
 *                        value central_weight * L_00 * Y_00
 *                      p_(ia,jb),l = Re sum_m conj(c_ia,lm) c_jb,lm / sqrt(2l+1)
 *                        * sqrt(2) if ia != jb; ia = species*n_max + a; l fastest;
 *                        last element 0; vector normalised to unit length
 *   distance_2b      GAP descriptors.f95 distance_2b (absent); conventions pinned
 *                    by tests/gap_sample.xyz + tests/GAP.xml and
 *                    tests/test_potential_cell.py:40-44
 *   gp_predict       GAP gp_predict.f95 (absent): dot_product and ard_se
 *   assembly         src/Potentials/IPModel_GAP.f95:430-596
 */
#include "oracle.h"
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI 3.14159265358979323846264338327950288

/* ------------------------------------------------------------------ */
/* small growable int/double vectors                                   */
/* ------------------------------------------------------------------ */
typedef struct { int *v; int n, cap; } ivec;
typedef struct { double *v; int n, cap; } dvec;
static void ipush(ivec *a, int x) {
    if (a->n == a->cap) { a->cap = a->cap ? 2 * a->cap : 8; a->v = (int *)realloc(a->v, sizeof(int) * a->cap); }
    a->v[a->n++] = x;
}
static void dpush(dvec *a, double x) {
    if (a->n == a->cap) { a->cap = a->cap ? 2 * a->cap : 8; a->v = (double *)realloc(a->v, sizeof(double) * a->cap); }
    a->v[a->n++] = x;
}

/* ------------------------------------------------------------------ */
/* 3x3 helpers with the reference's operation order                    */
/* ------------------------------------------------------------------ */
static void cross3(const double *x, const double *y, double *out) {   /* linearalgebra.f95:4445-4454 */
    out[0] = +(x[1] * y[2] - x[2] * y[1]);
    out[1] = -(x[0] * y[2] - x[2] * y[0]);
    out[2] = +(x[0] * y[1] - x[1] * y[0]);
}
static double triple3(const double *x, const double *y, const double *z) { /* :4458-4467 */
    return +x[0] * (y[1] * z[2] - y[2] * z[1]) - x[1] * (y[0] * z[2] - y[2] * z[0]) + x[2] * (y[0] * z[1] - y[1] * z[0]);
}
static double norm3(const double *v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

/* matrix3x3_inverse (linearalgebra.f95:6721-6733): g(r,:) = cross(col r+1, col r+2)/stp.
 * m and g are column-major: m[3*col + row]. */
static void inv3(const double *m, double *g) {
    double stp = triple3(m, m + 3, m + 6), c[3];
    for (int r = 0; r < 3; r++) {
        cross3(m + 3 * ((r + 1) % 3), m + 3 * ((r + 2) % 3), c);
        for (int k = 0; k < 3; k++) g[3 * k + r] = c[k] / stp;
    }
}
/* matrix_product_vect (linearalgebra.f95:671-686): prodvect(i) += m(i,j)*v(j), j ascending */
static void mult3(const double *m, const double *v, double *out) {
    for (int i = 0; i < 3; i++) out[i] = 0.0;
    for (int j = 0; j < 3; j++)
        for (int i = 0; i < 3; i++) out[i] = out[i] + m[3 * j + i] * v[j];
}

/* ------------------------------------------------------------------ */
/* neighbour list                                                      */
/* ------------------------------------------------------------------ */
struct orc_nl {
    int N, cells[3];
    int *n1_off, *n1_j, *n1_shift; double *n1_r;
    int *n2_off, *n2_j, *n2_idx;
    int *map_shift;
};

static int imax(int a, int b) { return a > b ? a : b; }

orc_nl *orc_nl_build(int N, const double *pos, const double *lat, const int *pbc, double cutoff, int *err) {
    *err = 0;
    if (cutoff <= 0.0) { *err = 1; return NULL; }   /* Connection.f95:1074-1082 */
    double lat_eff[9], lat_eff_inv[9], lat_offset[3] = {0, 0, 0}, g[9];
    memcpy(lat_eff, lat, sizeof(lat_eff));
    inv3(lat, g);
    /* calc_lat_eff (:740-763) */
    int any_np = !pbc[0] || !pbc[1] || !pbc[2];
    if (any_np && N > 0) {
        for (int i = 0; i < 3; i++) {
            if (pbc[i]) continue;
            double mn = 0, mx = 0, lp[3];
            for (int n = 0; n < N; n++) {
                mult3(g, pos + 3 * n, lp);
                if (n == 0 || lp[i] < mn) mn = lp[i];
                if (n == 0 || lp[i] > mx) mx = lp[i];
            }
            if (mn <= 0.5 || mx >= 0.5) {
                double ext = (mx - mn) * 1.1, nv = norm3(lat_eff + 3 * i);
                for (int k = 0; k < 3; k++) lat_eff[3 * i + k] = lat_eff[3 * i + k] / nv * ext;
                lat_offset[i] = mn - 0.05;
            }
        }
    }
    inv3(lat_eff, lat_eff_inv);
    /* divide_cell (:1674-1691) */
    int cN[3];
    {
        double bc[3], vol = fabs(triple3(lat_eff, lat_eff + 3, lat_eff + 6));
        for (int a = 0; a < 3; a++) {
            cross3(lat_eff + 3 * ((a + 1) % 3), lat_eff + 3 * ((a + 2) % 3), bc);
            cN[a] = imax(1, (int)(vol / (cutoff * norm3(bc))));
        }
    }
    /* fit_box_in_cell (:2204-2237) on the true lattice, then halve (:1145-1151) */
    int img[3];
    {
        double mxs[3] = {0, 0, 0}, lc[3], corner[3];
        for (int k = -1; k <= 1; k += 2)
            for (int j = -1; j <= 1; j += 2)
                for (int i = -1; i <= 1; i += 2) {
                    corner[0] = i * cutoff; corner[1] = j * cutoff; corner[2] = k * cutoff;
                    mult3(g, corner, lc);
                    for (int a = 0; a < 3; a++) if (fabs(lc[a]) > mxs[a]) mxs[a] = fabs(lc[a]);
                }
        for (int a = 0; a < 3; a++) img[a] = imax(1, ((int)ceil(2.0 * mxs[a]) + 1) / 2);
    }
    orc_nl *nl = (orc_nl *)calloc(1, sizeof(orc_nl));
    nl->N = N;
    for (int a = 0; a < 3; a++) nl->cells[a] = cN[a];
    nl->map_shift = (int *)calloc(3 * (size_t)(N > 0 ? N : 1), sizeof(int));
    /* partition_atoms (:1519-1587) + cell_of_pos (:1589-1621) */
    int ncell = cN[0] * cN[1] * cN[2];
    int *head = (int *)malloc(sizeof(int) * ncell), *next = (int *)malloc(sizeof(int) * (N > 0 ? N : 1));
    for (int c = 0; c < ncell; c++) head[c] = -1;
    for (int n = 0; n < N; n++) {
        double lp[3], t[3];
        mult3(lat_eff_inv, pos + 3 * n, lp);
        int ci[3];
        for (int a = 0; a < 3; a++) {
            lp[a] = lp[a] - lat_offset[a];
            int ms = -(int)floor(lp[a] + 0.5);
            if (!pbc[a]) ms = 0;
            nl->map_shift[3 * n + a] = ms;
            t[a] = lp[a] + (double)ms;
            int c = (int)floor((double)cN[a] * (t[a] + 0.5)) + 1;
            if (c < 1) c = 1; else if (c > cN[a]) c = cN[a];
            ci[a] = c;
        }
        int lin = ((ci[2] - 1) * cN[1] + (ci[1] - 1)) * cN[0] + (ci[0] - 1);
        next[n] = head[lin];
        head[lin] = n;
    }
    ivec *n1j = (ivec *)calloc(N > 0 ? N : 1, sizeof(ivec)), *n1s = (ivec *)calloc(N > 0 ? N : 1, sizeof(ivec));
    dvec *n1r = (dvec *)calloc(N > 0 ? N : 1, sizeof(dvec));
    ivec *n2j = (ivec *)calloc(N > 0 ? N : 1, sizeof(ivec)), *n2x = (ivec *)calloc(N > 0 ? N : 1, sizeof(ivec));
    /* main loop (:1194-1290) */
    for (int k = 1; k <= cN[2]; k++)
        for (int j = 1; j <= cN[1]; j++)
            for (int i = 1; i <= cN[0]; i++) {
                int cc[3] = {i, j, k}, lo[3], hi[3];
                for (int a = 0; a < 3; a++) {           /* get_min_max_images (:1448-1513) */
                    if (pbc[a]) { lo[a] = -img[a]; hi[a] = img[a]; }
                    else {
                        lo[a] = (img[a] < cc[a]) ? -img[a] : -cc[a] + 1;
                        hi[a] = (img[a] < cN[a] - cc[a]) ? img[a] : cN[a] - cc[a];
                    }
                }
                int lin1 = ((k - 1) * cN[1] + (j - 1)) * cN[0] + (i - 1);
                for (int atom1 = head[lin1]; atom1 >= 0; atom1 = next[atom1]) {
                    int k3 = 1, k4 = 1, j3 = 1, j4 = 1, i3 = 1, i4 = 1;
                    for (int k2 = lo[2]; k2 <= hi[2]; k2++) {
                        if (cN[2] > 1) k3 = (k + k2 - 1 + cN[2]) % cN[2] + 1;
                        k4 = (k + k2 - k3) / cN[2];
                        for (int j2 = lo[1]; j2 <= hi[1]; j2++) {
                            if (cN[1] > 1) j3 = (j + j2 - 1 + cN[1]) % cN[1] + 1;
                            j4 = (j + j2 - j3) / cN[1];
                            for (int i2 = lo[0]; i2 <= hi[0]; i2++) {
                                if (cN[0] > 1) i3 = (i + i2 - 1 + cN[0]) % cN[0] + 1;
                                i4 = (i + i2 - i3) / cN[0];
                                int lin2 = ((k3 - 1) * cN[1] + (j3 - 1)) * cN[0] + (i3 - 1);
                                for (int atom2 = head[lin2]; atom2 >= 0; atom2 = next[atom2]) {
                                    if (atom1 > atom2) continue;
                                    if (atom1 == atom2 && i4 == 0 && j4 == 0 && k4 == 0 && i == i3 && j == j3 && k == k3) continue;
                                    int sh[3] = {i4 - nl->map_shift[3 * atom1 + 0] + nl->map_shift[3 * atom2 + 0],
                                                 j4 - nl->map_shift[3 * atom1 + 1] + nl->map_shift[3 * atom2 + 1],
                                                 k4 - nl->map_shift[3 * atom1 + 2] + nl->map_shift[3 * atom2 + 2]};
                                    /* test_form_bond (:448-526) */
                                    double dd[3];
                                    for (int q = 0; q < 3; q++) dd[q] = pos[3 * atom2 + q] - pos[3 * atom1 + q];
                                    for (int m = 0; m < 3; m++)
                                        for (int q = 0; q < 3; q++) dd[q] = dd[q] + lat[3 * m + q] * (double)sh[m];
                                    if (dd[0] > cutoff || dd[1] > cutoff || dd[2] > cutoff) continue;
                                    double d = sqrt(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]);
                                    if (!(d < cutoff)) continue;
                                    /* add_bond (:585-634) */
                                    ipush(&n1j[atom1], atom2);
                                    for (int q = 0; q < 3; q++) ipush(&n1s[atom1], sh[q]);
                                    dpush(&n1r[atom1], d);
                                    if (atom1 != atom2) { ipush(&n2j[atom2], atom1); ipush(&n2x[atom2], n1j[atom1].n - 1); }
                                }
                            }
                        }
                    }
                }
            }
    /* compact into CSR */
    nl->n1_off = (int *)malloc(sizeof(int) * (N + 1));
    nl->n2_off = (int *)malloc(sizeof(int) * (N + 1));
    nl->n1_off[0] = nl->n2_off[0] = 0;
    for (int n = 0; n < N; n++) { nl->n1_off[n + 1] = nl->n1_off[n] + n1j[n].n; nl->n2_off[n + 1] = nl->n2_off[n] + n2j[n].n; }
    int t1 = nl->n1_off[N], t2 = nl->n2_off[N];
    nl->n1_j = (int *)malloc(sizeof(int) * (t1 + 1)); nl->n1_shift = (int *)malloc(sizeof(int) * 3 * (t1 + 1));
    nl->n1_r = (double *)malloc(sizeof(double) * (t1 + 1));
    nl->n2_j = (int *)malloc(sizeof(int) * (t2 + 1)); nl->n2_idx = (int *)malloc(sizeof(int) * (t2 + 1));
    for (int n = 0; n < N; n++) {
        int o = nl->n1_off[n];
        for (int q = 0; q < n1j[n].n; q++) {
            nl->n1_j[o + q] = n1j[n].v[q]; nl->n1_r[o + q] = n1r[n].v[q];
            for (int c = 0; c < 3; c++) nl->n1_shift[3 * (o + q) + c] = n1s[n].v[3 * q + c];
        }
        o = nl->n2_off[n];
        for (int q = 0; q < n2j[n].n; q++) { nl->n2_j[o + q] = n2j[n].v[q]; nl->n2_idx[o + q] = n2x[n].v[q]; }
        free(n1j[n].v); free(n1s[n].v); free(n1r[n].v); free(n2j[n].v); free(n2x[n].v);
    }
    free(n1j); free(n1s); free(n1r); free(n2j); free(n2x); free(head); free(next);
    return nl;
}

void orc_nl_counts(const orc_nl *nl, int *n1_total, int *n2_total, int *cells) {
    *n1_total = nl->n1_off[nl->N]; *n2_total = nl->n2_off[nl->N];
    if (cells) for (int a = 0; a < 3; a++) cells[a] = nl->cells[a];
}
void orc_nl_export(const orc_nl *nl, int *n1_off, int *n1_j, int *n1_shift, double *n1_r,
                   int *n2_off, int *n2_j, int *n2_idx, int *map_shift) {
    int N = nl->N, t1 = nl->n1_off[N], t2 = nl->n2_off[N];
    memcpy(n1_off, nl->n1_off, sizeof(int) * (N + 1)); memcpy(n2_off, nl->n2_off, sizeof(int) * (N + 1));
    memcpy(n1_j, nl->n1_j, sizeof(int) * t1); memcpy(n1_shift, nl->n1_shift, sizeof(int) * 3 * t1);
    memcpy(n1_r, nl->n1_r, sizeof(double) * t1);
    memcpy(n2_j, nl->n2_j, sizeof(int) * t2); memcpy(n2_idx, nl->n2_idx, sizeof(int) * t2);
    memcpy(map_shift, nl->map_shift, sizeof(int) * 3 * N);
}
void orc_nl_free(orc_nl *nl) {
    if (!nl) return;
    free(nl->n1_off); free(nl->n1_j); free(nl->n1_shift); free(nl->n1_r);
    free(nl->n2_off); free(nl->n2_j); free(nl->n2_idx); free(nl->map_shift); free(nl);
}

/* connection_neighbour (:2011-2138): the n-th neighbour of atom i, neighbour2
 * entries first.  Returns j; fills shift, distance and diff (i-perspective). */
static int nl_neighbour(const orc_nl *nl, const double *pos, const double *lat, int i, int n,
                        int *shift, double *dist, double *diff) {
    int n2 = nl->n2_off[i + 1] - nl->n2_off[i], j;
    if (n < n2) {
        int e = nl->n2_off[i] + n;
        j = nl->n2_j[e];
        int e1 = nl->n1_off[j] + nl->n2_idx[e];
        *dist = nl->n1_r[e1];
        for (int c = 0; c < 3; c++) shift[c] = -nl->n1_shift[3 * e1 + c];
    } else {
        int e1 = nl->n1_off[i] + (n - n2);
        j = nl->n1_j[e1];
        *dist = nl->n1_r[e1];
        for (int c = 0; c < 3; c++) shift[c] = nl->n1_shift[3 * e1 + c];
    }
    for (int c = 0; c < 3; c++) diff[c] = pos[3 * j + c] - pos[3 * i + c];
    for (int m = 0; m < 3; m++)
        for (int c = 0; c < 3; c++) diff[c] = diff[c] + lat[3 * m + c] * (double)shift[m];
    return j;
}
static int nl_count(const orc_nl *nl, int i) {
    return nl->n1_off[i + 1] - nl->n1_off[i] + nl->n2_off[i + 1] - nl->n2_off[i];
}

/* ------------------------------------------------------------------ */
/* cutoff function (linearalgebra.f95:7488-7516)                        */
/* ------------------------------------------------------------------ */
static double fcut(double r, double rc, double w) {
    if (r > rc) return 0.0;
    if (r > rc - w) return 0.5 * (cos(PI * (r - rc + w) / w) + 1.0);
    return 1.0;
}
static double dfcut(double r, double rc, double w) {
    if (r > rc) return 0.0;
    if (r > rc - w) return -0.5 * PI * sin(PI * (r - rc + w) / w) / w;
    return 0.0;
}

/* ------------------------------------------------------------------ */
/* complex spherical harmonics (angular_functions.f95:32-278)           */
/* Y[l*l + l + m], dY[3*(l*l+l+m) + k]                                   */
/* ------------------------------------------------------------------ */
static double fact(int n) { double f = 1.0; for (int i = 2; i <= n; i++) f *= i; return f; }

static void sph_all(int lmax, const double *x, double complex *Y, double complex *dY) {
    double complex cm = -0.5 * x[0] - 0.5 * x[1] * I, cp = 0.5 * x[0] - 0.5 * x[1] * I;
    double complex cmt[40], cpt[40];
    double x3t[40], fc[40];
    cmt[0] = 1; cpt[0] = 1; x3t[0] = 1;
    for (int p = 1; p <= 2 * lmax + 1; p++) { cmt[p] = cmt[p - 1] * cm; cpt[p] = cpt[p - 1] * cp; x3t[p] = x3t[p - 1] * x[2]; }
    for (int q = 0; q <= 2 * lmax + 1; q++) fc[q] = fact(q);
    double nsq = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
    for (int l = 0; l <= lmax; l++) {
        double rl = pow(nsq, -0.5 * l), nrm = sqrt((2.0 * l + 1.0) / (4.0 * PI));
        for (int m = -l; m <= l; m++) {
            double complex R = 0, g0 = 0, g1 = 0, g2 = 0;
            for (int p = 0; p <= l; p++) {
                int q = p - m, s = l - p - q;
                if (q >= 0 && s >= 0) R += cmt[p] * cpt[q] * x3t[s] / (fc[p] * fc[q] * fc[s]);
                if (dY) {
                    if (p >= 1 && q >= 0 && s >= 0) {
                        double complex tt = cmt[p - 1] * cpt[q] * x3t[s] * 0.5 / (fc[p - 1] * fc[q] * fc[s]);
                        g0 -= tt; g1 -= tt * I;
                    }
                    if (p >= 0 && q >= 1 && s >= 0) {
                        double complex tt = cmt[p] * cpt[q - 1] * x3t[s] * 0.5 / (fc[p] * fc[q - 1] * fc[s]);
                        g0 += tt; g1 -= tt * I;
                    }
                    if (p >= 0 && q >= 0 && s >= 1) g2 += cmt[p] * cpt[q] * x3t[s - 1] / (fc[p] * fc[q] * fc[s - 1]);
                }
            }
            double sc = sqrt(fc[l + m] * fc[l - m]);
            int idx = l * l + l + m;
            Y[idx] = R * sc * nrm * rl;
            if (dY) {
                double gs = sqrt(fc[l + m] * fc[l - m] * (2.0 * l + 1.0) / (4.0 * PI)) * rl;
                dY[3 * idx + 0] = g0 * gs - l * x[0] * Y[idx] / nsq;
                dY[3 * idx + 1] = g1 * gs - l * x[1] * Y[idx] / nsq;
                dY[3 * idx + 2] = g2 * gs - l * x[2] * Y[idx] / nsq;
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* SOAP radial basis                                                   */
/* ------------------------------------------------------------------ */
static void cholesky_lower(int n, const double *S, double *L) {   /* S = L L^T, row-major */
    memset(L, 0, sizeof(double) * n * n);
    for (int j = 0; j < n; j++) {
        double s = S[j * n + j];
        for (int k = 0; k < j; k++) s -= L[j * n + k] * L[j * n + k];
        L[j * n + j] = sqrt(s);
        for (int i = j + 1; i < n; i++) {
            double t = S[i * n + j];
            for (int k = 0; k < j; k++) t -= L[i * n + k] * L[j * n + k];
            L[i * n + j] = t / L[j * n + j];
        }
    }
}

int orc_soap_basis(double cutoff, double sigma, double E, int n, double *rb, double *T, double *u00) {
    if (n < 1 || n > 32 || sigma <= 0) return 1;
    double alpha = 0.5 / (sigma * sigma);
    double cb = cutoff + sigma * sqrt(2.0 * E * log(10.0));
    for (int a = 0; a < n; a++) rb[a] = (double)a * cb / (double)n;
    double *C = (double *)malloc(sizeof(double) * n * n), *S = (double *)malloc(sizeof(double) * n * n);
    double *L = (double *)malloc(sizeof(double) * n * n), *Lc = (double *)malloc(sizeof(double) * n * n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double ri = rb[i], rj = rb[j];
            C[i * n + j] = exp(-alpha * (ri - rj) * (ri - rj));
            S[i * n + j] = exp(-alpha * (ri * ri + rj * rj)) * sqrt(2.0) * pow(alpha, 1.5) * (ri + rj) +
                           alpha * exp(-0.5 * alpha * (ri - rj) * (ri - rj)) * sqrt(PI) *
                               (1.0 + alpha * (ri + rj) * (ri + rj)) * (1.0 + erf(sqrt(alpha / 2.0) * (ri + rj)));
        }
    cholesky_lower(n, S, L);
    cholesky_lower(n, C, Lc);
    /* T = C^-1 L  via C = Lc Lc^T: forward then backward substitution per column */
    for (int b = 0; b < n; b++) {
        double y[64];
        for (int i = 0; i < n; i++) {
            double t = L[i * n + b];
            for (int k = 0; k < i; k++) t -= Lc[i * n + k] * y[k];
            y[i] = t / Lc[i * n + i];
        }
        for (int i = n - 1; i >= 0; i--) {
            double t = y[i];
            for (int k = i + 1; k < n; k++) t -= Lc[k * n + i] * T[k * n + b];
            T[i * n + b] = t / Lc[i * n + i];
        }
    }
    *u00 = L[0];
    free(C); free(S); free(L); free(Lc);
    return 0;
}

/* f_l(r; r_a) for l = 0..lmax+1 and its r-derivative for l = 0..lmax. */
static void radial_fun(double r, const double *rb, int n, int lmax, double alpha, double *f, double *df) {
    double fl[40];
    for (int a = 0; a < n; a++) {
        double x = 2.0 * alpha * r * rb[a];
        if (x == 0.0) {
            fl[0] = exp(-alpha * (rb[a] * rb[a] + r * r));
            for (int l = 1; l <= lmax + 1; l++) fl[l] = 0.0;
        } else {
            double ep = exp(-alpha * (r + rb[a]) * (r + rb[a])), em = exp(-alpha * (r - rb[a]) * (r - rb[a]));
            double fm = 0.5 * (em + ep) / x;
            fl[0] = 0.5 * (em - ep) / x;
            double fprev = fm;
            for (int l = 1; l <= lmax + 1; l++) {
                double fn = fprev - (2.0 * l - 1.0) * fl[l - 1] / x;
                fprev = fl[l - 1];
                fl[l] = fn;
            }
        }
        for (int l = 0; l <= lmax; l++) {
            f[l * n + a] = fl[l];
            if (df) df[l * n + a] = -2.0 * alpha * r * fl[l] + l * fl[l] / r + fl[l + 1] * 2.0 * alpha * rb[a];
        }
    }
}

int orc_soap_dim(const orc_coord *c) {
    int K = c->n_species * c->n_max;
    return K * (K + 1) / 2 * (c->l_max + 1) + 1;
}

static int species_index(const orc_coord *c, int z) {
    for (int s = 0; s < c->n_species; s++) if (c->species_Z[s] == z || c->species_Z[s] == 0) return s;
    return -1;
}
static int is_centre(const orc_coord *c, int z) {
    if (c->n_Z <= 0) return 1;
    for (int s = 0; s < c->n_Z; s++) if (c->Z[s] == z || c->Z[s] == 0) return 1;
    return 0;
}

typedef struct {
    int n, lmax, nsp, K, LM, dim;
    double alpha, u00, rb[32], T[32 * 32];
} soap_setup;

static int soap_prepare(const orc_coord *c, soap_setup *s) {
    s->n = c->n_max; s->lmax = c->l_max; s->nsp = c->n_species; s->K = s->nsp * s->n;
    s->LM = (s->lmax + 1) * (s->lmax + 1); s->dim = orc_soap_dim(c);
    s->alpha = 0.5 / (c->atom_sigma * c->atom_sigma);
    if (c->n_max > 32 || c->l_max > 18) return 1;
    return orc_soap_basis(c->cutoff, c->atom_sigma, c->basis_error_exponent, c->n_max, s->rb, s->T, &s->u00);
}

/* One SOAP instance (centre `ci`): descriptor x (dim) and, if grad != NULL,
 * grad[(e*3+k)*dim + q] for entries e = 0 (centre), 1..nn (neighbours in
 * connection_neighbour order with r <= cutoff).  Returns number of entries. */
static int soap_instance(const orc_coord *c, const soap_setup *s, const orc_nl *nl, const double *pos,
                         const int *Z, const double *lat, int ci, double *x, double *grad, int *ii,
                         double *epos, int max_entries) {
    int n = s->n, lmax = s->lmax, K = s->K, LM = s->LM, dim = s->dim;
    int nnb = nl_count(nl, ci);
    double complex *cc = (double complex *)calloc((size_t)K * LM, sizeof(double complex));
    /* per-neighbour cached data */
    double complex *Yn = (double complex *)malloc(sizeof(double complex) * (size_t)(nnb + 1) * LM);
    double complex *dYn = (double complex *)malloc(sizeof(double complex) * (size_t)(nnb + 1) * LM * 3);
    double *gn = (double *)malloc(sizeof(double) * (size_t)(nnb + 1) * (lmax + 1) * n);
    double *dgn = (double *)malloc(sizeof(double) * (size_t)(nnb + 1) * (lmax + 1) * n);
    double *dvec_ = (double *)malloc(sizeof(double) * 3 * (nnb + 1));
    double *rr = (double *)malloc(sizeof(double) * (nnb + 1));
    int *spc = (int *)malloc(sizeof(int) * (nnb + 1));
    double f[40 * 32], df[40 * 32];
    int ne = 0;
    if (ii) ii[0] = ci;
    if (epos) for (int k = 0; k < 3; k++) epos[k] = pos[3 * ci + k];
    int sc = species_index(c, Z[ci]);
    if (sc >= 0) cc[(size_t)(sc * n + 0) * LM + 0] += c->central_weight * s->u00 * sqrt(1.0 / (4.0 * PI));
    for (int q = 0; q < nnb; q++) {
        int sh[3]; double r, d[3];
        int j = nl_neighbour(nl, pos, lat, ci, q, sh, &r, d);
        if (r > c->cutoff) continue;
        int sj = species_index(c, Z[j]);
        if (sj < 0) continue;
        if (ne + 2 > max_entries) { ne = -1; break; }
        double fc = fcut(r, c->cutoff, c->cutoff_transition_width);
        double dfc = dfcut(r, c->cutoff, c->cutoff_transition_width);
        radial_fun(r, s->rb, n, lmax, s->alpha, f, df);
        double *g = gn + (size_t)ne * (lmax + 1) * n, *dg = dgn + (size_t)ne * (lmax + 1) * n;
        for (int l = 0; l <= lmax; l++)
            for (int b = 0; b < n; b++) {
                double v = 0, dv = 0;
                for (int a = 0; a < n; a++) { v += f[l * n + a] * s->T[a * n + b]; dv += df[l * n + a] * s->T[a * n + b]; }
                g[l * n + b] = fc * v;
                dg[l * n + b] = fc * dv + dfc * v;
            }
        sph_all(lmax, d, Yn + (size_t)ne * LM, grad ? dYn + (size_t)ne * LM * 3 : NULL);
        for (int l = 0; l <= lmax; l++)
            for (int m = -l; m <= l; m++)
                for (int b = 0; b < n; b++)
                    cc[(size_t)(sj * n + b) * LM + l * l + l + m] += g[l * n + b] * Yn[(size_t)ne * LM + l * l + l + m];
        for (int k = 0; k < 3; k++) dvec_[3 * ne + k] = d[k];
        rr[ne] = r; spc[ne] = sj;
        if (ii) ii[ne + 1] = j;
        if (epos) for (int k = 0; k < 3; k++) epos[3 * (ne + 1) + k] = pos[3 * ci + k] + d[k];
        ne++;
    }
    if (ne < 0) { free(cc); free(Yn); free(dYn); free(gn); free(dgn); free(dvec_); free(rr); free(spc); return -1; }
    /* power spectrum */
    int qi = 0;
    for (int ia = 0; ia < K; ia++)
        for (int jb = 0; jb <= ia; jb++)
            for (int l = 0; l <= lmax; l++) {
                double v = 0;
                for (int m = -l; m <= l; m++) v += creal(conj(cc[(size_t)ia * LM + l * l + l + m]) * cc[(size_t)jb * LM + l * l + l + m]);
                v /= sqrt(2.0 * l + 1.0);
                if (ia != jb) v *= sqrt(2.0);
                x[qi++] = v;
            }
    x[qi] = 0.0;
    double nrm = 0;
    for (int q = 0; q < dim; q++) nrm += x[q] * x[q];
    nrm = sqrt(nrm);
    if (nrm > 0) for (int q = 0; q < dim; q++) x[q] /= nrm;
    if (grad) {
        memset(grad, 0, sizeof(double) * (size_t)(ne + 1) * 3 * dim);
        double *dp = (double *)malloc(sizeof(double) * dim);
        for (int e = 0; e < ne; e++) {
            double r = rr[e]; const double *d = dvec_ + 3 * e; int sj = spc[e];
            const double *g = gn + (size_t)e * (lmax + 1) * n, *dg = dgn + (size_t)e * (lmax + 1) * n;
            const double complex *Y = Yn + (size_t)e * LM, *dY = dYn + (size_t)e * LM * 3;
            for (int k = 0; k < 3; k++) {
                /* dc for channel sj only */
                qi = 0;
                for (int ia = 0; ia < K; ia++)
                    for (int jb = 0; jb <= ia; jb++)
                        for (int l = 0; l <= lmax; l++) {
                            double v = 0;
                            for (int m = -l; m <= l; m++) {
                                int lm = l * l + l + m;
                                if (ia / n == sj) {
                                    int a = ia % n;
                                    double complex dca = dg[l * n + a] * Y[lm] * d[k] / r + g[l * n + a] * dY[3 * lm + k];
                                    v += creal(conj(dca) * cc[(size_t)jb * LM + lm]);
                                }
                                if (jb / n == sj) {
                                    int b = jb % n;
                                    double complex dcb = dg[l * n + b] * Y[lm] * d[k] / r + g[l * n + b] * dY[3 * lm + k];
                                    v += creal(conj(cc[(size_t)ia * LM + lm]) * dcb);
                                }
                            }
                            v /= sqrt(2.0 * l + 1.0);
                            if (ia != jb) v *= sqrt(2.0);
                            dp[qi++] = v;
                        }
                dp[qi] = 0.0;
                double xd = 0;
                for (int q = 0; q < dim; q++) xd += x[q] * dp[q];
                double *gr = grad + (size_t)((e + 1) * 3 + k) * dim, *g0 = grad + (size_t)k * dim;
                for (int q = 0; q < dim; q++) {
                    double v = nrm > 0 ? (dp[q] - x[q] * xd) / nrm : 0.0;
                    gr[q] = v;
                    g0[q] -= v;
                }
            }
        }
        free(dp);
    }
    free(cc); free(Yn); free(dYn); free(gn); free(dgn); free(dvec_); free(rr); free(spc);
    return ne + 1;
}

int orc_soap_descriptors(const orc_coord *c, int N, const double *pos, const int *Z, const double *lat,
                         const int *pbc, double nl_cutoff, double *X, int *desc_atom, int *n_desc) {
    soap_setup s;
    if (c->kind != 1 || soap_prepare(c, &s)) return 2;
    int err;
    orc_nl *nl = orc_nl_build(N, pos, lat, pbc, nl_cutoff, &err);
    if (!nl) return 3;
    int nd = 0;
    for (int i = 0; i < N; i++) {
        if (!is_centre(c, Z[i])) continue;
        soap_instance(c, &s, nl, pos, Z, lat, i, X + (size_t)nd * s.dim, NULL, NULL, NULL, 1 << 30);
        desc_atom[nd++] = i;
    }
    *n_desc = nd;
    orc_nl_free(nl);
    return 0;
}

int orc_soap_grad_atom(const orc_coord *c, int N, const double *pos, const int *Z, const double *lat,
                       const int *pbc, double nl_cutoff, int atom, double *x, double *grad, int *ii,
                       int max_entries, int *n_entries) {
    soap_setup s;
    if (c->kind != 1 || soap_prepare(c, &s)) return 2;
    int err;
    orc_nl *nl = orc_nl_build(N, pos, lat, pbc, nl_cutoff, &err);
    if (!nl) return 3;
    int ne = soap_instance(c, &s, nl, pos, Z, lat, atom, x, grad, ii, NULL, max_entries);
    orc_nl_free(nl);
    if (ne < 0) return 4;
    *n_entries = ne;
    return 0;
}

/* ------------------------------------------------------------------ */
/* gp_predict                                                          */
/* ------------------------------------------------------------------ */
/* Covariance of x with sparse point s, without alpha and sparseCutoff:
 * dot_product delta^2 (x.X_s)^zeta, ard_se delta^2 exp(-1/2 sum_d (x_d-X_sd)^2/theta_d^2).
 * dk (nullable) gets dk/dx. */
static double gp_cov(const orc_coord *c, const double *x, int s, double *dk) {
    int d = c->dim;
    const double *xs = c->sparseX + (size_t)s * d;
    double d2 = c->delta * c->delta;
    if (c->covariance == 2) {
        double k = 0;
        for (int q = 0; q < d; q++) k += x[q] * xs[q];
        if (dk) {
            double g = d2 * c->zeta * pow(k, c->zeta - 1.0);
            for (int q = 0; q < d; q++) dk[q] = g * xs[q];
        }
        return d2 * pow(k, c->zeta);
    }
    double a = 0;
    for (int q = 0; q < d; q++) { double t = (x[q] - xs[q]) / c->theta[q]; a += t * t; }
    double k = d2 * exp(-0.5 * a);
    if (dk)
        for (int q = 0; q < d; q++) dk[q] = -k * (x[q] - xs[q]) / (c->theta[q] * c->theta[q]);
    return k;
}

/* e = sum_s alpha_s sparseCutoff_s k_s(x); gradPredict = de/dx (dot_product:
 * k_s = delta^2 (x.X_s)^zeta, gradient alpha delta^2 zeta (x.X_s)^(zeta-1) X_s). */
static double gp_predict(const orc_coord *c, const double *x, double *gradPredict) {
    int d = c->dim;
    double e = 0;
    double *dk = gradPredict ? (double *)malloc(sizeof(double) * d) : NULL;
    if (gradPredict) memset(gradPredict, 0, sizeof(double) * d);
    for (int s = 0; s < c->n_sparse; s++) {
        double w = c->alpha[s] * c->sparse_cutoff[s];
        e += w * gp_cov(c, x, s, dk);
        if (gradPredict)
            for (int q = 0; q < d; q++) gradPredict[q] += w * dk[q];
    }
    free(dk);
    return e;
}

double orc_gp_predict(const orc_coord *c, const double *x, double *gradPredict) {
    return gp_predict(c, x, gradPredict);
}

/* Variance of the sparse GP prediction (projected-process form):
 *   var(x) = k(x,x) - k_x^T (K_MM + reg I)^-1 k_x,   k_x[s] = cov(x, X_s),
 *   K_MM[s][t] = cov(X_s, X_t), k(x,x) = cov(x, x);
 *   d var/dx = dk(x,x)/dx - 2 (dk_x/dx)^T (K_MM + reg I)^-1 k_x.
 * The reference's gpCoordinates_initialise_variance_estimate and the variance
 * branch of gp_predict live in the absent GAP module; this is the textbook
 * estimate that its gap_variance_regularisation key (IPModel_GAP.f95:332)
 * regularises.  The factor L of K_MM + reg I (lower, row-major) is formed once
 * per coordinate. */
typedef struct { int M; double *L; } var_setup;

static int var_prepare(const orc_coord *c, double reg, var_setup *v) {
    int M = c->n_sparse, d = c->dim;
    v->M = M;
    v->L = (double *)calloc((size_t)M * M + 1, sizeof(double));
    double *K = (double *)malloc(sizeof(double) * ((size_t)M * M + 1));
    for (int s = 0; s < M; s++)
        for (int t = 0; t < M; t++) K[(size_t)s * M + t] = gp_cov(c, c->sparseX + (size_t)t * d, s, NULL);
    for (int s = 0; s < M; s++) K[(size_t)s * M + s] += reg;
    for (int j = 0; j < M; j++) {
        double a = K[(size_t)j * M + j];
        for (int k = 0; k < j; k++) a -= v->L[(size_t)j * M + k] * v->L[(size_t)j * M + k];
        if (!(a > 0)) { free(K); return 1; }
        v->L[(size_t)j * M + j] = sqrt(a);
        for (int i = j + 1; i < M; i++) {
            double t = K[(size_t)i * M + j];
            for (int k = 0; k < j; k++) t -= v->L[(size_t)i * M + k] * v->L[(size_t)j * M + k];
            v->L[(size_t)i * M + j] = t / v->L[(size_t)j * M + j];
        }
    }
    free(K);
    return 0;
}

/* var(x); gvar (nullable) gets d var/dx */
static double gp_variance(const orc_coord *c, const var_setup *v, const double *x, double *gvar) {
    int M = v->M, d = c->dim;
    double *k = (double *)malloc(sizeof(double) * (M + 1)), *y = (double *)malloc(sizeof(double) * (M + 1));
    double *dk = (double *)malloc(sizeof(double) * (gvar ? (size_t)M * d : 1));
    for (int s = 0; s < M; s++) k[s] = gp_cov(c, x, s, gvar ? dk + (size_t)s * d : NULL);
    for (int i = 0; i < M; i++) {   /* y = L^-T L^-1 k */
        double t = k[i];
        for (int j = 0; j < i; j++) t -= v->L[(size_t)i * M + j] * y[j];
        y[i] = t / v->L[(size_t)i * M + i];
    }
    for (int i = M - 1; i >= 0; i--) {
        double t = y[i];
        for (int j = i + 1; j < M; j++) t -= v->L[(size_t)j * M + i] * y[j];
        y[i] = t / v->L[(size_t)i * M + i];
    }
    /* k(x,x) and its gradient */
    double kxx;
    if (c->covariance == 2) {
        double xx = 0;
        for (int q = 0; q < d; q++) xx += x[q] * x[q];
        kxx = c->delta * c->delta * pow(xx, c->zeta);
        if (gvar) {
            double g = 2.0 * c->delta * c->delta * c->zeta * pow(xx, c->zeta - 1.0);
            for (int q = 0; q < d; q++) gvar[q] = g * x[q];
        }
    } else {
        kxx = c->delta * c->delta;
        if (gvar) memset(gvar, 0, sizeof(double) * d);
    }
    double var = kxx;
    for (int s = 0; s < M; s++) var -= k[s] * y[s];
    if (gvar)
        for (int s = 0; s < M; s++)
            for (int q = 0; q < d; q++) gvar[q] -= 2.0 * dk[(size_t)s * d + q] * y[s];
    free(k); free(y); free(dk);
    return var;
}

/* ------------------------------------------------------------------ */
/* IPModel_GAP_Calc                                                    */
/* ------------------------------------------------------------------ */
typedef struct { double *f, *vir, *le, *lgv, *gvg; } accum;

/* One descriptor instance (IPModel_GAP.f95:452-498): local energy e_i*cov/size(ci)
 * to each ci; f(:,j) -= gradPredict.grad_data(:,:,n)*cov + e_i*grad_cov(:,n);
 * virial_in(:,:,j) -= (pos_n - pos_first) (x) f_gp.  With a variance (var not NULL):
 * local_gap_variance(ci) += var*cov^2/size(ci) and gap_variance_gradient(:,j) +=
 * gvar.grad_data(:,:,n)*cov^2 + 2 var cov grad_cov(:,n) (:464-470, :484-488). */
static void scatter(int N, accum *A, double e_i, double cov, const double *gcov, int nci, const int *ci,
                    const double *gp, const double *grad, int dim, int ne, const int *ii, const double *epos,
                    const double *var, const double *gvar) {
    double ec = e_i * cov / nci;
    for (int q = 0; q < nci; q++) A->le[ci[q]] += ec;
    if (var)
        for (int q = 0; q < nci; q++) A->lgv[ci[q]] += *var * cov * cov / nci;
    if (!grad) return;
    for (int e = 0; e < ne; e++) {
        double fg[3];
        for (int k = 0; k < 3; k++) {
            double v = 0;
            const double *g = grad + (size_t)(e * 3 + k) * dim;
            for (int q = 0; q < dim; q++) v += gp[q] * g[q];
            fg[k] = v * cov + e_i * (gcov ? gcov[3 * e + k] : 0.0);
        }
        int j = ii[e];
        for (int k = 0; k < 3; k++) A->f[3 * j + k] -= fg[k];
        for (int b = 0; b < 3; b++)
            for (int a = 0; a < 3; a++)
                A->vir[9 * j + 3 * b + a] -= (epos[3 * e + a] - epos[a]) * fg[b];
        if (var && gvar)
            for (int k = 0; k < 3; k++) {
                double v = 0;
                const double *g = grad + (size_t)(e * 3 + k) * dim;
                for (int q = 0; q < dim; q++) v += gvar[q] * g[q];
                A->gvg[3 * j + k] += v * cov * cov + 2.0 * *var * cov * (gcov ? gcov[3 * e + k] : 0.0);
            }
    }
    (void)N;
}

static void record_instance(const orc_calc_opts *o, int coord, double var, double cut2, int a, int b) {
    if (!o || !o->n_inst) return;
    int k = *o->n_inst;
    if (k < o->inst_cap) {
        if (o->inst_var) o->inst_var[k] = var;
        if (o->inst_cut2) o->inst_cut2[k] = cut2;
        if (o->inst_coord) o->inst_coord[k] = coord;
        if (o->inst_atoms) { o->inst_atoms[2 * k] = a; o->inst_atoms[2 * k + 1] = b; }
    }
    *o->n_inst = k + 1;
}

int orc_gap_calc_ex(int n_coord, const orc_coord *coords, const double *e0, double E_scale, int N,
                    const double *pos, const int *Z, const double *lat, const int *pbc, double nl_cutoff,
                    int n_threads, double *e, double *local_e, double *f, double *virial, double *local_virial,
                    const orc_calc_opts *opt) {
    int err;
    orc_nl *nl = orc_nl_build(N, pos, lat, pbc, nl_cutoff, &err);
    if (!nl) return 3;
    const unsigned char *mask = opt ? opt->atom_mask : NULL;
    const int want_var = opt && opt->want_variance;
    const int want_grad = (f || virial || local_virial || (want_var && opt->gap_variance_gradient));
#ifdef _OPENMP
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#else
    n_threads = 1;
#endif
    if (want_var || (opt && opt->n_inst)) n_threads = 1;   /* instances recorded in the reference's order */
    if (opt && opt->n_inst) *opt->n_inst = 0;
    accum *acc = (accum *)calloc(n_threads, sizeof(accum));
    for (int t = 0; t < n_threads; t++) {
        acc[t].f = (double *)calloc(3 * (size_t)N + 1, sizeof(double));
        acc[t].vir = (double *)calloc(9 * (size_t)N + 1, sizeof(double));
        acc[t].le = (double *)calloc((size_t)N + 1, sizeof(double));
        acc[t].lgv = (double *)calloc((size_t)N + 1, sizeof(double));
        acc[t].gvg = (double *)calloc(3 * (size_t)N + 1, sizeof(double));
    }
    int status = 0;
    for (int ic = 0; ic < n_coord; ic++) {
        const orc_coord *c = coords + ic;
        /* only_descriptor selects one coordinate when there are several (IPModel_GAP.f95:398-403) */
        if (opt && opt->only_descriptor > 0 && n_coord > 1 && ic != opt->only_descriptor - 1) continue;
        double epc = 0.0;
        var_setup vs = {0, NULL};
        if (want_var && var_prepare(c, opt->variance_regularisation, &vs)) { status = 6; free(vs.L); break; }
        if (c->kind == 1) {
            soap_setup s;
            if (soap_prepare(c, &s)) { status = 2; free(vs.L); break; }
            int dim = s.dim;
            double *ei_all = (double *)calloc((size_t)N + 1, sizeof(double));
#pragma omp parallel num_threads(n_threads)
            {
                int tid = 0;
#ifdef _OPENMP
                tid = omp_get_thread_num();
#endif
                int maxe = 0;
                double *x = (double *)malloc(sizeof(double) * dim), *gp = (double *)malloc(sizeof(double) * dim);
                double *gv = (double *)malloc(sizeof(double) * dim);
                double *grad = NULL, *epos = NULL; int *ii = NULL;
#pragma omp for schedule(static)
                for (int i = 0; i < N; i++) {
                    if (!is_centre(c, Z[i]) || (mask && !mask[i])) continue;
                    int need = nl_count(nl, i) + 1;
                    if (need > maxe) {
                        maxe = need;
                        free(grad); free(epos); free(ii);
                        grad = want_grad ? (double *)malloc(sizeof(double) * (size_t)maxe * 3 * dim) : NULL;
                        epos = (double *)malloc(sizeof(double) * 3 * maxe); ii = (int *)malloc(sizeof(int) * maxe);
                    }
                    int ne = soap_instance(c, &s, nl, pos, Z, lat, i, x, grad, ii, epos, maxe + 1);
                    double ei = gp_predict(c, x, want_grad ? gp : NULL);
                    ei_all[i] = ei;
                    double var = 0;
                    if (want_var) {
                        var = gp_variance(c, &vs, x, want_grad ? gv : NULL);
                        record_instance(opt, ic, var, 1.0, i, -1);
                    }
                    int ci1[1] = {i};
                    scatter(N, &acc[tid], ei, 1.0, NULL, 1, ci1, gp, grad, dim, ne, ii, epos,
                            want_var ? &var : NULL, want_var ? gv : NULL);
                }
                free(x); free(gp); free(gv); free(grad); free(epos); free(ii);
            }
            for (int i = 0; i < N; i++) epc += ei_all[i];
            free(ei_all);
        } else if (c->kind == 2) {
            double w = c->cutoff_transition_width;
            for (int i = 0; i < N; i++) {
                if (mask && !mask[i]) continue;
                int zi1 = (c->Z1 == 0) || (Z[i] == c->Z1), zi2 = (c->Z2 == 0) || (Z[i] == c->Z2);
                int nnb = nl_count(nl, i);
                for (int q = 0; q < nnb; q++) {
                    int sh[3]; double r, d[3];
                    int j = nl_neighbour(nl, pos, lat, i, q, sh, &r, d);
                    if (r > c->cutoff) continue;
                    int zj1 = (c->Z1 == 0) || (Z[j] == c->Z1), zj2 = (c->Z2 == 0) || (Z[j] == c->Z2);
                    if (!((zi1 && zj2) || (zi2 && zj1))) continue;
                    double u[3] = {d[0] / r, d[1] / r, d[2] / r};
                    double x1 = r, gp1, gv1 = 0;
                    double ei = gp_predict(c, &x1, &gp1);
                    double cov = fcut(r, c->cutoff, w), dcov = dfcut(r, c->cutoff, w);
                    epc += ei * cov;
                    double var = 0;
                    if (want_var) {
                        var = gp_variance(c, &vs, &x1, &gv1);
                        record_instance(opt, ic, var, cov * cov, i, j);
                    }
                    double grad[6] = {-u[0], -u[1], -u[2], u[0], u[1], u[2]};
                    double gcov[6] = {-dcov * u[0], -dcov * u[1], -dcov * u[2], dcov * u[0], dcov * u[1], dcov * u[2]};
                    int ii[2] = {i, j}, ci[2] = {i, j};
                    double epos[6] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], pos[3 * i] + d[0], pos[3 * i + 1] + d[1], pos[3 * i + 2] + d[2]};
                    scatter(N, &acc[0], ei, cov, gcov, 2, ci, &gp1, want_grad ? grad : NULL, 1, 2, ii, epos,
                            want_var ? &var : NULL, want_var ? &gv1 : NULL);
                }
            }
        } else { status = 5; free(vs.L); break; }
        free(vs.L);
        if (opt && opt->energy_per_coordinate) opt->energy_per_coordinate[ic] = epc;
    }
    if (!status) {
        /* fixed-order reduction over thread-private arrays */
        double *F = acc[0].f, *Vv = acc[0].vir, *LE = acc[0].le;
        for (int t = 1; t < n_threads; t++) {
            for (size_t q = 0; q < 3 * (size_t)N; q++) F[q] += acc[t].f[q];
            for (size_t q = 0; q < 9 * (size_t)N; q++) Vv[q] += acc[t].vir[q];
            for (size_t q = 0; q < (size_t)N; q++) LE[q] += acc[t].le[q];
        }
        /* e0 for every atom, or for the masked atoms only (IPModel_GAP.f95:576-590) */
        if (f) for (size_t q = 0; q < 3 * (size_t)N; q++) f[q] = E_scale * F[q];
        if (e) {
            double s = 0;
            for (int i = 0; i < N; i++) s += LE[i];
            for (int i = 0; i < N; i++) if (!mask || mask[i]) s += e0[Z[i]];
            *e = E_scale * s;
        }
        if (local_e) for (int i = 0; i < N; i++) local_e[i] = E_scale * (LE[i] + ((!mask || mask[i]) ? e0[Z[i]] : 0.0));
        if (virial) {
            for (int q = 0; q < 9; q++) {
                double s = 0;
                for (int i = 0; i < N; i++) s += Vv[9 * i + q];
                virial[q] = E_scale * s;
            }
        }
        if (local_virial) for (size_t q = 0; q < 9 * (size_t)N; q++) local_virial[q] = E_scale * Vv[q];
        /* the variance outputs are not scaled by E_scale (:592-596 scales e, f, virials only) */
        if (want_var && opt->local_gap_variance) memcpy(opt->local_gap_variance, acc[0].lgv, sizeof(double) * N);
        if (want_var && opt->gap_variance_gradient) memcpy(opt->gap_variance_gradient, acc[0].gvg, sizeof(double) * 3 * N);
    }
    for (int t = 0; t < n_threads; t++) { free(acc[t].f); free(acc[t].vir); free(acc[t].le); free(acc[t].lgv); free(acc[t].gvg); }
    free(acc);
    orc_nl_free(nl);
    return status;
}

int orc_gap_calc(int n_coord, const orc_coord *coords, const double *e0, double E_scale, int N,
                 const double *pos, const int *Z, const double *lat, const int *pbc, double nl_cutoff,
                 int n_threads, double *e, double *local_e, double *f, double *virial, double *local_virial) {
    return orc_gap_calc_ex(n_coord, coords, e0, E_scale, N, pos, Z, lat, pbc, nl_cutoff, n_threads, e, local_e, f,
                           virial, local_virial, NULL);
}
