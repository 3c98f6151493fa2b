/*
 * oracle.h -- FP64 CPU restatement of the SOAP-GAP energy/force/virial path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py may load it.  The product library (quip_amd/libgapx.so) never
 * links, loads or calls it.
 *
 * Conventions (same as QUIP's Fortran arrays, see Atoms_types.f95:224-347):
 *   pos[3*i + k]      k-th Cartesian component of atom i        (= pos(k,i))
 *   lat[3*m + k]      k-th component of lattice vector m        (= lattice(k,m))
 *   f[3*i + k]        force component                           (= f(k,i))
 *   virial[3*b + a]   virial(a,b), column-major                 (= virial(a,b))
 *   local_virial[9*i + 3*b + a] = reshape(virial_in(:,:,i),9)   (IPModel_GAP.f95:528-532)
 * All atom indices are 0-based.
 */
#ifndef GAPX_ORACLE_H
#define GAPX_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_SPECIES 16

/* One gpCoordinates block of a GAP model (descriptor + sparse GP). */
typedef struct {
    int kind;                 /* 1 = soap, 2 = distance_2b */
    /* soap */
    double cutoff;
    double cutoff_transition_width;
    double atom_sigma;
    double central_weight;
    double basis_error_exponent;
    int n_max, l_max;
    int n_species;
    int species_Z[ORC_MAX_SPECIES];
    int n_Z;
    int Z[ORC_MAX_SPECIES];
    /* distance_2b */
    int Z1, Z2;
    /* gp */
    int covariance;           /* 1 = ard_se, 2 = dot_product */
    double delta;             /* signal_variance attribute (the kernel uses delta^2) */
    double zeta;
    int n_sparse, dim;
    const double *sparseX;    /* n_sparse x dim, point-major */
    const double *alpha;      /* n_sparse */
    const double *sparse_cutoff; /* n_sparse */
    const double *theta;      /* dim (ard_se) */
} orc_coord;

/* ---- neighbour list: Connection.f95 calc_connect restatement ---- */
typedef struct orc_nl orc_nl;
orc_nl *orc_nl_build(int N, const double *pos, const double *lat, const int *pbc,
                     double cutoff, int *err);
void orc_nl_counts(const orc_nl *nl, int *n1_total, int *n2_total, int *cells);
void orc_nl_export(const orc_nl *nl, int *n1_off, int *n1_j, int *n1_shift, double *n1_r,
                   int *n2_off, int *n2_j, int *n2_idx, int *map_shift);
void orc_nl_free(orc_nl *nl);

/* Radial basis transform of the SOAP descriptor: T (n_max x n_max, row a = grid
 * point, column b = orthonormal channel), grid r_basis(n_max) and U00. */
int orc_soap_basis(double cutoff, double atom_sigma, double basis_error_exponent, int n_max,
                   double *r_basis, double *T, double *u00);

int orc_soap_dim(const orc_coord *c);

/* Descriptor vectors of every instance of coordinate c (no gradients).
 * X: n_desc x dim, desc_atom: centre atom of each instance. */
int orc_soap_descriptors(const orc_coord *c, int N, const double *pos, const int *Z,
                         const double *lat, const int *pbc, double nl_cutoff,
                         double *X, int *desc_atom, int *n_desc);

/* Descriptor of centre atom `atom` plus its gradient block grad[(n*3 + k)*dim + q]
 * = d x_q / d r_k of entry n, and ii[n] (entry 0 = centre). */
int orc_soap_grad_atom(const orc_coord *c, int N, const double *pos, const int *Z,
                       const double *lat, const int *pbc, double nl_cutoff, int atom,
                       double *x, double *grad, int *ii, int max_entries, int *n_entries);

/* gp_predict of one descriptor vector x (dim entries): returns the prediction
 * sum_s alpha_s sparseCutoff_s k(x, X_s), dot_product k = delta^2 (x.X_s)^zeta,
 * ard_se k = delta^2 exp(-1/2 sum_d ((x_d - X_sd)/theta_d)^2); gradPredict
 * (nullable, dim entries) gets its gradient. */
double orc_gp_predict(const orc_coord *c, const double *x, double *gradPredict);

/* Options of IPModel_GAP_Calc's args_str (IPModel_GAP.f95:324-335) and their outputs. */
typedef struct {
    const unsigned char *atom_mask;   /* atom_mask_name: nullable, N bytes; only masked atoms are
                                         descriptor centres / first atoms of pairs and get e0 */
    int only_descriptor;              /* 0: all coordinates; k >= 1: coordinate k alone when n_coord > 1 */
    int want_variance;                /* local_gap_variance / print_gap_variance */
    double variance_regularisation;   /* gap_variance_regularisation */
    double *energy_per_coordinate;    /* nullable, n_coord: sum of e_i * covariance_cutoff */
    double *local_gap_variance;       /* nullable, N */
    double *gap_variance_gradient;    /* nullable, 3N (f(3,N) layout) */
    /* per-instance variances in the reference's order (print_gap_variance), nullable */
    int inst_cap;
    double *inst_var, *inst_cut2;     /* variance, covariance_cutoff^2 */
    int *inst_coord, *inst_atoms;     /* coordinate; first atom, second atom (-1 for soap) */
    int *n_inst;
} orc_calc_opts;

/* Full IPModel_GAP_Calc restatement.  Any output pointer may be NULL. */
int orc_gap_calc(int n_coord, const orc_coord *coords, const double *e0 /* [119] by Z */,
                 double E_scale, int N, const double *pos, const int *Z, const double *lat,
                 const int *pbc, double nl_cutoff, int n_threads,
                 double *e, double *local_e, double *f, double *virial, double *local_virial);

/* As orc_gap_calc with the options above (opt may be NULL). */
int orc_gap_calc_ex(int n_coord, const orc_coord *coords, const double *e0, double E_scale, int N,
                    const double *pos, const int *Z, const double *lat, const int *pbc, double nl_cutoff,
                    int n_threads, double *e, double *local_e, double *f, double *virial, double *local_virial,
                    const orc_calc_opts *opt);

#ifdef __cplusplus
}
#endif
#endif
