// Device-side declarations shared by the kernel and host translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gapx {

constexpr int kMaxSpecies = 16;
constexpr int kMaxL = 12;

// Neighbour-list geometry, prepared on the host from the lattice (3x3 maths only).
struct NLGeom {
    double lat[9];       // lattice(k,m) at lat[3*m+k]: column m is lattice vector m
    double inv_eff[9];   // inverse of the effective lattice, column-major g(i,j) at [3*j+i]
    double ginv[9];      // inverse of the true lattice, same layout
    double offset[3];    // lat_offset of non-periodic axes
    int cells[3];        // cellsNa, cellsNb, cellsNc
    int img[3];          // cell_image_N{a,b,c} after halving
    int pbc[3];
    double cutoff;
};

struct SoapDev {
    double rc, ctw, alpha, cw, u00;
    int n, S, K, LM, dim, L;
    int species_Z[kMaxSpecies];
    const double *rb;    // n grid points
    const double *T;     // n x n transform, T[a*n+b]
    const int *qmap;     // dim-1 entries: (ia<<16)|(jb<<8)|l
};

// launchers (kernels.hip)
int scan_partials_size(int n);   // ints of scan scratch per array
void launch_scan_i32(const int *in, int n, int *out, int *partials, hipStream_t s);
void launch_fill_i32(int *p, int n, int v, hipStream_t s);
void launch_fill_f64(double *p, size_t n, double v, hipStream_t s);
constexpr int kMultiFill = 6;   // arrays one zero-fill launch covers
void launch_zero_i32_multi(int k, int *const *p, const int *n, hipStream_t s);
void launch_zero_f64_multi(int k, double *const *p, const size_t *n, hipStream_t s);
void launch_scan2_max(const int *in1, const int *in2, int n, int *out1, int *out2, int *mx, int *tot1, int *part1,
                      int *part2, hipStream_t s);
// n <= 65536: both scans (maxima to flags[1], flags[2], the in1 total to flags[3]) and the
// overflow guard of launch_nl_guard in one launch
bool scan2_guard_fits(int n);
void launch_scan2_guard(const int *in1, const int *in2, int n, int *out1, int *out2, int *flags, int cap,
                        hipStream_t s);
// bins the atoms and counts each cell's atoms (cell_cnt zeroed by the caller)
void launch_nl_bin(int N, const double *pos, const NLGeom *geo, const int *sid, const int *cell_base, int *cell_of,
                   int *map_shift, int *cell_cnt, hipStream_t s);
void launch_nl_cell_fill(int N, const int *cell_of, const int *start, int *fillc, int *atoms, hipStream_t s);
// The cell-ordered copies the candidate search reads: slot q of the cell lists gets its atom's
// position (cpos, x y z and a pad) and map shift with the atom index (cms: ms0 ms1 ms2 atom).
struct CellAtom { double p[4]; };
struct CellShift { int v[4]; };
void launch_nl_cell_rank(int N, const int *cell_of, const int *start, const int *cnt, const int *unsorted,
                         int *atoms, const double *pos, const int *map_shift, CellAtom *cpos, CellShift *cms,
                         hipStream_t s);
// bin, scan, fill and rank in one block for N <= kNlFrontAtoms atoms in NC <= kNlFrontCells
// cells (also zeroes n2_cnt, n2_fill and the four flags)
constexpr int kNlFrontAtoms = 8192, kNlFrontCells = 8192;
void launch_nl_front_small(int N, int NC, const double *pos, const NLGeom *geo, const int *sid,
                           const int *cell_base, int *cell_of, int *map_shift, int *cell_cnt, int *cell_start,
                           int *cell_atoms, CellAtom *cpos, CellShift *cms, int *n2_cnt, int *n2_fill, int *flags,
                           hipStream_t s);
// mode 0: count; 1: fill the CSR tables; 2: one pass into per-atom slots of cap1 / cap2
// (counts in n1_cnt / n2_cnt, *ovf = 1 when a slot range overflows)
void launch_nl_pairs(int mode, int N, const double *pos, const int *map_shift, const int *cell_of,
                     const int *cell_start, const int *cell_cnt, const int *cell_atoms, const CellAtom *cpos,
                     const CellShift *cms, const NLGeom *geo, const int *sid, const int *cell_base, int *n1_cnt,
                     const int *n1_off, int *n1_j, int *n1_shift, double *n1_r, int *n2_cnt, const int *n2_off,
                     int *n2_fill, int *n2_i, int *n2_k, int cap1, int cap2, int *ovf, hipStream_t s);
void launch_nl_compact(int N, int cap1, int cap2, const int *n1_off, const int *n2_off, const int *t1_j,
                       const int *t1_shift, const double *t1_r, const int *t2_i, const int *t2_k, int *n1_j,
                       int *n1_shift, double *n1_r, int *n2_i, int *n2_k, hipStream_t s);
// cap2 > 0: the unsorted entries are read from the one-pass slots (atom j's at j * cap2)
void launch_nl_n2_sort(int N, const int *n2_off, const int *n2_i, const int *n2_k, const int *cell_of,
                       const int *n1_off, int *n2_is, int *n2_ks, int *n1_rev, int cap2, hipStream_t s);
// cap1 > 0: neighbour1 data from the one-pass slots (atom a's entry k at a * cap1 + k)
void launch_nl_full(int N, const double *pos, const NLGeom *geo, const int *sid, const int *n1_off, const int *n1_j,
                    const int *n1_shift, const double *n1_r, const int *n2_off, const int *n2_i, const int *n2_k,
                    const int *n1_rev, int *nb_off, int *nb_j, int *nb_rev, double *nb_r, double *nb_d, int cap1,
                    hipStream_t s, double *pairF = nullptr,
                    double *pairE = nullptr, double *esite = nullptr);
// skin reuse: max displacement since the last build (bits of a double, atomicMax into a
// zeroed word) and the stored neighbour1 distances recomputed as calc_dists does
void launch_nl_skin_check(int N, const double *pos, const double *last, const NLGeom *geo, const int *sid,
                          const int *n1_off, const int *n1_j, const int *n1_shift, double *n1_r,
                          unsigned long long *max_disp, int *overlap, hipStream_t s);
void launch_minmax_frac(int S, const int *atom_off, const double *pos, const double *g, double *out, hipStream_t s);
// lat_eff of structures with a non-periodic axis from this call's extents mm (6 per
// structure) into geo; flags[0] |= 4 when a structure's cell counts differ from geo's
void launch_geom_np(int S, const int *atom_off, const double *mm, NLGeom *geo, int *flags, hipStream_t s);
// empty every list (offsets zeroed) after a slot or capacity overflow; flags[0] |= 2 for capacity
void launch_nl_guard(int N, int *flags, int cap, int *n1_off, int *n2_off, hipStream_t s);
// p[0 .. mult * (*cnt)) = 0 with *cnt read on the device; cap bounds the grid
void launch_zero_counted(double *p, const int *cnt, int mult, size_t cap, hipStream_t s);
// the per-step clears in one launch: pairF (3 per pair) and pairE (nullable) of *cnt pairs
// (cnt on the device) or n_pairs (cnt null), and esite (N)
void launch_zero_step(double *pairF, double *pairE, double *esite, int N, const int *cnt, int n_pairs,
                      size_t cap_pairs, hipStream_t s);

// upload the constant tables of the SOAP kernels to the current device (once per
// device; done at model load so that no launch sequence contains an upload)
hipError_t soap_constants_init();
hipError_t soap_fast_constants_init();
int soap_supported(int n, int L);
int soap_fast_supported(int n, int L);
// LDS bytes per block the descriptor kernels of this shape use (largest launch variant)
size_t soap_lds_bytes(const SoapDev &sp);
size_t soap_fast_lds_bytes(const SoapDev &sp);
bool launch_soap_fwd_fast(const SoapDev &sp, int nc, const int *centres, const int *Z, const int *nb_off,
                          const int *nb_j, const double *nb_r, const double *nb_d, double *X, double *norm,
                          double *cbuf, hipStream_t s);
bool launch_soap_bwd_fast(const SoapDev &sp, int nc, const int *centres, const int *Z, const int *nb_off,
                          const int *nb_j, const double *nb_r, const double *nb_d, const double *X, const double *G,
                          const double *norm, const double *cbuf, double *pairF, hipStream_t s);
void launch_soap_fwd(const SoapDev &sp, int nc, const int *centres, const int *Z, const int *nb_off,
                     const int *nb_j, const double *nb_r, const double *nb_d, double *X, double *norm,
                     double *cbuf, hipStream_t s);
void launch_soap_bwd(const SoapDev &sp, int nc, const int *centres, const int *Z, const int *nb_off,
                     const int *nb_j, const double *nb_r, const double *nb_d, const double *X, const double *G,
                     const double *norm, const double *cbuf, double *pairF, hipStream_t s);

// GEMMs
// Tile configuration `cfg` in [0, kGemmTiles): (WM, WN, k tile) = (4,2,16), (2,2,16),
// (2,4,16), (4,2,32), (2,2,32), (2,4,32); block tile (32 WM) x (32 WN) of 4 wavefronts.
constexpr int kGemmTiles = 6;
// gemm1: K = X (nc x d) . Xs^T (M x d); W = w_s zeta K^(zeta-1); e_part[i*ntile + t]
int gemm1_ntile(int cfg, int M);
void launch_gemm1_f64(int cfg, int nc, int M, int d, const double *X, const double *Xs, const double *wvec,
                      double zeta, double *W, double *e_part, int *ntile_out, hipStream_t s);
// gemm2: G = W (nc x M) . Xs (M x d); with parts > 1 the k range is split and the
// partials (parts x nc x d in Gparts) are summed in a fixed order
int gemm2_parts(int nc, int d);
void launch_gemm2_f64(int cfg, int parts, int nc, int M, int d, const double *W, const double *Xs, double *G,
                      double *Gparts, hipStream_t s);
// fused FP64 GP kernel for descriptors of up to 256 components: K = X Xs^T, the energy
// partials e_part[i*parts + p] and G = W Xs in one launch (W stays in registers);
// XsP is the sparse matrix zero-padded to ceil(M/32)*32 rows of 32*ceil(d/32) columns.
// With parts > 1 the partial G rows go to Gparts (parts x nc x d); launch_gp_finish_f64
// sums them (with the energy rows).  *ntile_out = parts as the launch rounds them.
bool gp_f64_fused_supported(int d);
int gp_f64_parts_default(int nc, int M);
int gp_f64_parts_valid(int parts, int M);
void launch_gp_fused_f64(int nc, int M, int d, int parts_req, const double *X, const double *XsP,
                         const double *wvec, double zeta, double *G, double *Gparts, double *e_part, int *ntile_out,
                         hipStream_t s);
void launch_sum_parts_f64(int nc, int d, int parts, const double *Gparts, double *G, hipStream_t s);
// G = scale * (fixed-order sum of the parts partial G rows) when parts > 1, and the energy
// rows of launch_energy_rows when centres is not null, in one launch
void launch_gp_finish_f64(int nc, int d, int parts, const double *Gparts, double *G, int ntile, const int *centres,
                          const double *e_part, double *esite, hipStream_t s);
void launch_gp_finish_f32(int nc, int d, int parts, const float *Gparts, double scale, double *G, int ntile,
                          const int *centres, const double *e_part, double *esite, hipStream_t s);
// mixed precision forms: Xs pre-split into hi/lo (bf16 or fp16) copies, row- and column-major
// W is scaled by wscale (a power of two) before its hi/lo split; GEMM2 divides it out.
// lds_mode: -1 the 128 x 128 LDS tiles once they fill the GPU, else 64 x 64; 0 / 1 force
// the 64 x 64 / 128 x 128 tiles (an element's MFMA passes run in the same order in both)
void launch_gemm1_split(int prec, int nc, int M, int d, int dpad, int lds_mode, const double *X, const void *Xs_hi,
                        const void *Xs_lo, const double *wvec, double zeta, double wscale, float *Wf,
                        double *e_part, int *ntile_out, void *xsplit_ws, hipStream_t s);
void launch_gemm2_split(int prec, int nc, int M, int d, int dpad, int Mpad, int lds_mode, const float *Wf,
                        const void *XsT_hi, const void *XsT_lo, double wscale, double *G, hipStream_t s);
// fused K -> epilogue -> G for descriptors of up to 256 components (split forms):
// one launch computes the energy partials e_part[i*parts + p] and, with parts > 1,
// FP32 partial G tiles (parts x nc x d in Gpart) that launch_gp_fused_reduce sums;
// with parts == 1 it writes G directly.  *ntile_out = parts.
bool gp_fused_supported(int dpad, int Mpad);
int gp_fused_parts(int nc, int Mpad);              // the default number of sparse-point parts
int gp_fused_parts_valid(int parts, int Mpad);     // a requested number, as the launch rounds it
// parts_req > 0 asks for that many parts (rounded so that every part has chunks), 0 the
// default, -1 the stream-K schedule (*ntile_out = -1): whole row blocks write G and
// e_part[i] directly, the rest go to slots in Gpart (gp_fused_sk_scratch_bytes) that
// launch_gp_sk_finish sums; it also writes every centre's energy to e_part[i] (one
// column) and adds it to esite[centres[i]] (centres may be null: e_part only)
size_t gp_fused_sk_scratch_bytes(int prec, int d, int dpad, int nc, int Mpad);
void launch_gp_sk_finish(int prec, int nc, int d, int dpad, int Mpad, bool want_g, const float *Gpart, double wscale,
                         double *G, double *e_part, const int *centres, double *esite, hipStream_t s);
void launch_gp_fused_split(int prec, int nc, int M, int d, int dpad, int Mpad, int parts_req, const double *X,
                           const void *Xs_hi, const void *Xs_lo, const double *wvec, double zeta, double wscale,
                           double *G, float *Gpart, double *e_part, int *ntile_out, hipStream_t s);
void launch_gp_fused_reduce(int nc, int d, int parts, const float *Gpart, double wscale, double *G, hipStream_t s);
void launch_split_matrix(int prec, int rows, int cols, int ld_out, const double *src, int src_ld, bool transpose,
                         void *hi, void *lo, hipStream_t s);

void launch_energy_rows(int nc, int ntile, const int *centres, const double *e_part, double *esite, hipStream_t s);
// mask (nullable): only atoms with mask[i] != 0 are the first atom of a pair descriptor.
// e_atom (nullable): e_atom[i] = sum of e * fc over the instances whose first atom is i
// (written for every atom; zero where there are none).
void launch_d2b(int N, const int *Z, int Z1, int Z2, double rc, double ctw, int M, const double *xs,
                const double *wvec, const double *theta, const int *nb_off, const int *nb_j, const double *nb_r,
                const double *nb_d, double *pairF, double *pairE, const unsigned char *mask, double *e_atom,
                hipStream_t s);
// out[0] = v[0] + ... + v[n-1], summed in a fixed order by one block
void launch_sum_fixed(size_t n, const double *v, double *out, hipStream_t s);

// GAP variance (projected-process estimate), FP64.
// C = A (rows x kd) . B^T (cols x kd), no epilogue
void launch_gemm_nt_f64(int rows, int cols, int kd, const double *A, const double *B, double *C, hipStream_t s);
// kv[i][s] = delta2 K[i][s]^zeta
void launch_var_cov(int nc, int M, const double *K, double delta2, double zeta, double *kv, hipStream_t s);
// Per centre row i (V = kv . Cinv): var_i = delta2 (x_i.x_i)^zeta - sum_s kv_is V_is, then
// V_is <- -2 zeta delta2 K_is^(zeta-1) V_is (the variance's weight on X_s);
// var_inst[i] = var_i and varsite[centres[i]] += var_i.
void launch_var_rows(int nc, int M, int d, const double *X, const double *K, const double *kv, double *V,
                     double delta2, double zeta, const int *centres, double *var_inst, double *varsite,
                     hipStream_t s);
// Gv_i += 2 zeta delta2 (x_i.x_i)^(zeta-1) x_i  (the gradient of k(x, x))
void launch_var_self_grad(int nc, int d, const double *X, double delta2, double zeta, double *Gv, hipStream_t s);
// distance_2b variance with the ard_se kernel (M <= kMaxVarSparse): per directed pair p of a
// descriptor instance, var = delta2 - k^T Cinv k; pairVE[p] += var fc^2, pairFv[p] += d(var fc^2)/dr u;
// var_inst[p] = var and cut2_inst[p] = fc^2 (NaN-filled by the caller for non-instances).
constexpr int kMaxVarSparse = 128;
void launch_d2b_var(int N, const int *Z, int Z1, int Z2, double rc, double ctw, int M, const double *xs,
                    double delta2, const double *theta, const double *Cinv, const int *nb_off, const int *nb_j,
                    const double *nb_r, const double *nb_d, double *pairFv, double *pairVE, double *var_inst,
                    double *cut2_inst, const unsigned char *mask, hipStream_t s);
// gap_variance_gradient(:, j) = -sum_q (pairFv[q] - pairFv[rev q]) (the plus-signed gradient);
// local_gap_variance(j) = varsite[j] + (pairVE[q] + pairVE[rev q]) / 2 summed over q
void launch_gather_var(int N, const int *nb_off, const int *nb_rev, const double *pairFv, const double *pairVE,
                       const double *varsite, double *gvg, double *lgv, hipStream_t s);
// mask (nullable): e0 is added to the local energy of masked atoms only
// block_part (nullable): 10 doubles per block of 8 atoms (cdiv(N, 8) blocks), the block's
// local energy and virial sums for launch_struct_finish (the launch covers one structure)
void launch_gather(int N, const int *Z, const double *e0, double E_scale, const int *nb_off, const int *nb_rev,
                   const double *nb_d, const double *pairF, const double *pairE, const double *esite, double *f,
                   double *local_e, double *local_virial, const unsigned char *mask, double *block_part,
                   hipStream_t s);
// the energy and virial of one structure of N atoms from the gather's block partials;
// with flag_dst (pinned host memory, written by the device without a copy launch) the
// four neighbour-list build flags at flag_src are copied there as well
void launch_struct_finish(int N, const double *part, double *e_out, double *virial_out, hipStream_t s,
                          const int *flag_src = nullptr, int *flag_dst = nullptr);
void launch_struct_reduce(int S, const int *atom_off, const double *local_e, const double *local_virial,
                          double *e_out, double *virial_out, hipStream_t s, const int *flag_src = nullptr,
                          int *flag_dst = nullptr);

}  // namespace gapx
