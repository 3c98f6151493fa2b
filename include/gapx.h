/*
 * gapx.h -- C ABI of the MI355X SOAP-GAP engine (libgapx.so).
 *
 * Drop-in boundary for QUIP's IP GAP calculator.  Each entry point names the
 * reference interface it stands in for.  Conventions follow QUIP's Fortran
 * arrays so a Fortran host can pass its arrays straight through ISO_C_BINDING:
 *
 *   pos[3*i + k]              pos(k,i)                 (Atoms_types.f95:332)
 *   lattice[3*m + k]          lattice(k,m), column m = lattice vector m
 *   f[3*i + k]                f(k,i)                   (IPModel_GAP.f95:237)
 *   virial[3*b + a]           virial(a,b)
 *   local_virial[9*i + 3*b+a] reshape(virial_in(:,:,i),(/9/)) (IPModel_GAP.f95:528-532)
 *   Atom indices are 0-based; Z are atomic numbers; pbc[3] are 0/1.
 *
 * Every call returns 0 on success and a non-zero GAPX_E_* code otherwise; the
 * library never aborts the host.  gapx_last_error() returns the message of the
 * calling thread's last failure (the role of QUIP's error stack, error.inc:58-71).
 * All outputs are caller-allocated.  A model handle is used by one host thread
 * at a time; it owns one HIP stream and its device workspace.
 */
#ifndef GAPX_H
#define GAPX_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define GAPX_API_VERSION 1

/* status codes */
#define GAPX_OK 0
#define GAPX_E_ARG 1        /* bad argument (sizes, null where required)        */
#define GAPX_E_MODEL 2      /* model XML / sparseX content not usable           */
#define GAPX_E_IO 3         /* file not readable                                */
#define GAPX_E_DEVICE 4     /* HIP runtime error                                */
#define GAPX_E_UNSUPPORTED 5/* descriptor / option outside what is built       */
#define GAPX_E_GEOMETRY 6   /* e.g. atoms on top of each other, cutoff <= 0    */
#define GAPX_E_CAPACITY 7   /* request exceeds the caller-set memory budget    */

/* what to compute (bit mask): the optional outputs of IPModel_GAP_Calc */
#define GAPX_WANT_E 1u
#define GAPX_WANT_F 2u
#define GAPX_WANT_V 4u
#define GAPX_WANT_LOCAL_E 8u
#define GAPX_WANT_LOCAL_V 16u
/* form the GAP variance (gapx_calc_ex / gapx_calc_device_ex; implied by a non-NULL
 * local_gap_variance or gap_variance_gradient) */
#define GAPX_WANT_VARIANCE 32u

/* arithmetic of the descriptor x sparse-point GEMMs */
#define GAPX_PREC_FP64 0    /* plain form: FP64 throughout                          */
#define GAPX_PREC_BF16X3 1  /* BF16 hi+lo split, 3 MFMA passes, FP32 accumulation   */
#define GAPX_PREC_FP16X3 2  /* FP16 hi+lo split, 3 MFMA passes, FP32 accumulation   */

typedef struct gapx_model gapx_model;
typedef struct gapx_nl gapx_nl;
typedef struct gapx_server gapx_server;

/* Library version; a host binding checks it like quip_lammps_wrapper.f95:13,23-28. */
int gapx_api_version(void);

/* Message of the calling thread's last error (NUL-terminated, truncated to len). */
void gapx_last_error(char *buf, int len);

/* Number of visible GPUs. */
int gapx_device_count(int *count);

/* Load a GAP model from the XML text of a potential file.
 * Replaces IPModel_GAP_Initialise_str + IPModel_GAP_read_params_xml + gp_readXML
 * (IPModel_GAP.f95:149-192, 618-944).  `label` selects GAP_params by label
 * (NULL: the first); sparseX files are read relative to `xml_dir`.  Uploads
 * the sparse points and coefficients to device `device`. */
int gapx_model_load(const char *xml_text, int64_t xml_len, const char *label, const char *xml_dir, int device,
                    gapx_model **out);
int gapx_model_free(gapx_model *model);

/* IPModel_GAP%cutoff: the largest descriptor cutoff (IPModel_GAP.f95:182-188, IP.f95:704-705). */
int gapx_model_cutoff(const gapx_model *model, double *cutoff);
/* E_scale (IPModel_GAP.f95:592-596); default 1. */
int gapx_model_set_e_scale(gapx_model *model, double e_scale);
/* GEMM arithmetic, GAPX_PREC_*; default FP64. */
int gapx_model_set_precision(gapx_model *model, int precision);
/* Number of gpCoordinates and the descriptor dimension of one. */
int gapx_model_info(const gapx_model *model, int *n_coord, int coord, int *dim, int *n_sparse);

/* Energy, forces and virials of one structure from host buffers.
 * Replaces IPModel_GAP_Calc(this, at, e, local_e, f, virial, local_virial)
 * (IPModel_GAP.f95:233-605) including the calc_connect it relies on
 * (Potential.f95:844-859 -> Connection.f95:1035-1310).  `nl_cutoff` <= 0 uses the
 * model cutoff (QUIP's at%cutoff after set_cutoff).  Outputs not requested in
 * `want` may be NULL. */
int gapx_calc(gapx_model *model, int n_atoms, const double *pos, const int *Z, const double *lattice,
              const int *pbc, double nl_cutoff, unsigned want, double *e, double *local_e, double *f,
              double *virial, double *local_virial);

/* As gapx_calc with positions and every output in device memory (the inputs
 * are already resident in HBM); Z, lattice and pbc stay host arrays.  Work is
 * queued on the model's stream and the call returns without waiting for it.
 * After the first call on a model the neighbour list is built speculatively, into
 * tables sized from earlier builds, with no host synchronisation; a call whose lists
 * overflow them is detected on the device and redone by gapx_sync.  So the outputs
 * are final only after gapx_sync (or the next calc entry point that cannot run
 * speculatively: options, a skin, the host-buffer entry points), which waits for
 * every queued call of the model and redoes, in order, those that overflowed. */
int gapx_calc_device(gapx_model *model, int n_atoms, const double *d_pos, const int *Z, const double *lattice,
                     const int *pbc, double nl_cutoff, unsigned want, double *d_e, double *d_local_e, double *d_f,
                     double *d_virial, double *d_local_virial);
int gapx_sync(gapx_model *model);

/* gapx_calc_device without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (tight serving loops, code
 * under a capture or harness that forbids hipMalloc).  Every work buffer is carved
 * from one file-scope device array of 160 MiB per GPU, held until the process ends;
 * nothing is allocated or freed, on the device or in pinned host memory (staging is
 * ordinary host memory).  Requests whose buffers do not fit return GAPX_E_CAPACITY
 * (about 4000 atoms with a 2000-point FP64 SOAP model, more with the split forms).
 * All calls on one GPU share the array and calls that overlap on one GPU race: this
 * entry point serialises its own calls within a process and returns only when its
 * work is done.  The neighbour list is rebuilt on every call (no skin reuse), the GEMM
 * tiles are the autotuner's choices where it has made them and the defaults otherwise,
 * and a split precision must be set (gapx_model_set_precision) before the call.
 * Results equal gapx_calc_device's. */
int gapx_calc_device_noalloc(gapx_model *model, int n_atoms, const double *d_pos, const int *Z, const double *lattice,
                             const int *pbc, double nl_cutoff, unsigned want, double *d_e, double *d_local_e,
                             double *d_f, double *d_virial, double *d_local_virial);

/* gapx_calc with the args_str keys of IPModel_GAP_Calc that change what it computes
 * (IPModel_GAP.f95:324-335); each replaces the reference's handling of one key:
 *   atom_mask (nullable, n_atoms bytes)   atom_mask_name (:324-326, :343-347): only atoms
 *       with a non-zero byte are descriptor centres and first atoms of pair descriptors
 *       (the descriptor calc's mask) and get e0 (:576-590); forces still reach every atom;
 *   only_descriptor   (:334, :398-403): 0 evaluates every coordinate; k >= 1 evaluates
 *       coordinate k (1-based) alone when the model has more than one (e0 is still added);
 *   energy_per_coordinate (nullable, one value per coordinate)   (:335, :462, :568): the
 *       sum of e_i * covariance_cutoff per coordinate, before e0 and E_scale;
 *   GAPX_WANT_VARIANCE in want, or a non-NULL local_gap_variance / gap_variance_gradient:
 *       local_gap_variance (:330-332, :464-470, :558-561) and print_gap_variance (:506-516;
 *       the per-instance values from gapx_descriptor_variances), with
 *       variance_regularisation = gap_variance_regularisation (:332, default 0.001);
 *       local_gap_variance[N] gets var * fc^2 / size(ci) per instance and atom of ci;
 *       gap_variance_gradient[3N] (f layout) its gradient, with the reference's plus sign
 *       (:484-488).  The variance is var(x) = k(x,x) - k_x^T (K_MM + reg I)^-1 k_x with
 *       k = delta^2 kernel, formed in FP64 whatever the GEMM precision; models with a
 *       sparseCutoff other than 1 return GAPX_E_UNSUPPORTED for it.  Neither variance
 *       output is scaled by E_scale (:592-596 scales e, f and the virials only). */
int gapx_calc_ex(gapx_model *model, int n_atoms, const double *pos, const int *Z, const double *lattice,
                 const int *pbc, double nl_cutoff, unsigned want, const unsigned char *atom_mask, int only_descriptor,
                 double variance_regularisation, double *e, double *local_e, double *f, double *virial,
                 double *local_virial, double *energy_per_coordinate, double *local_gap_variance,
                 double *gap_variance_gradient);
/* As gapx_calc_ex with positions and every output in device memory (atom_mask stays a
 * host array); queued on the model's stream like gapx_calc_device. */
int gapx_calc_device_ex(gapx_model *model, int n_atoms, const double *d_pos, const int *Z, const double *lattice,
                        const int *pbc, double nl_cutoff, unsigned want, const unsigned char *atom_mask,
                        int only_descriptor, double variance_regularisation, double *d_e, double *d_local_e,
                        double *d_f, double *d_virial, double *d_local_virial, double *d_energy_per_coordinate,
                        double *d_local_gap_variance, double *d_gap_variance_gradient);
/* Per-instance variances of coordinate `coord` from the last calc that formed the
 * variance, in the reference's instance order (the values print_gap_variance prints,
 * IPModel_GAP.f95:506-516): variance, covariance_cutoff^2 and the instance's atoms
 * (soap: the centre and -1; distance_2b: the directed pair i, j).  Up to max_entries are
 * written; n_entries gets the count. */
int gapx_descriptor_variances(gapx_model *model, int coord, int max_entries, double *variance, double *cutoff_sq,
                              int *atoms, int *n_entries);

/* Many structures of uneven size in one pass (struct-of-arrays with offsets):
 * structure s owns atoms [atom_off[s], atom_off[s+1]), atom_off[0] = 0;
 * lattices[9*s..] and pbcs[3*s..] per structure; energies[s], virials[9*s..];
 * forces/local arrays indexed by global atom.  One neighbour-list build, one
 * descriptor launch, one GEMM pair and one gather cover every structure; the
 * physics is IPModel_GAP_Calc per structure (IPModel_GAP.f95:233-605). */
int gapx_calc_batch(gapx_model *model, int n_struct, const int *atom_off, const double *pos, const int *Z,
                    const double *lattices, const int *pbcs, double nl_cutoff, unsigned want, double *energies,
                    double *local_e, double *f, double *virials, double *local_virial);
/* As gapx_calc_batch with positions and every output in device memory; queued on
 * the model's stream like gapx_calc_device, and final after gapx_sync as there. */
int gapx_calc_batch_device(gapx_model *model, int n_struct, const int *atom_off, const double *d_pos, const int *Z,
                           const double *lattices, const int *pbcs, double nl_cutoff, unsigned want, double *d_e,
                           double *d_local_e, double *d_f, double *d_virials, double *d_local_virial);

/* LAMMPS pair style entry.  Replaces quip_lammps_wrapper (quip_lammps_wrapper.f95:30-156)
 * with the model handle in place of the transfer()-encoded Potential
 * (quip_potential, n_quip_potential); the other arguments keep their names,
 * order and meaning.  The n_all = nlocal + nghost atoms are LAMMPS's local atoms
 * followed by its ghosts (periodic images appear as ghosts, so no shifts are
 * used); quip_x is 3 x n_all, atomic_numbers n_all.  ilist (0-based, inum
 * entries) names the local atoms whose full neighbour lists are packed in
 * quip_neigh (1-based atom numbers, quip_num_neigh[ni] of them for ilist[ni],
 * sum_num_neigh in all).  As the reference, the connection keeps a pair once,
 * in neighbour1 of the lower index (:104-144); descriptors are centred on the
 * ilist atoms only (atom_mask_name=local, :147); forces, local energies and
 * local virials are returned for all n_all atoms (ghost entries carry the
 * contributions LAMMPS folds back onto their owners); quip_e and quip_virial
 * are the sums over all of them.  lattice (3 x 3) is accepted for the
 * reference's signature and not used (:111-114).  lmptag may be NULL.
 * Two atoms closer than 1e-10 A are an error, as in the reference. */
int gapx_calc_lammps(gapx_model *model, int nlocal, int nghost, const int *atomic_numbers, const int *lmptag,
                     int inum, int sum_num_neigh, const int *ilist, const int *quip_num_neigh,
                     const int *quip_neigh, const double *lattice, const double *quip_x, double *quip_e,
                     double *quip_local_e, double *quip_virial, double *quip_local_virial, double *quip_force);

/* Serving loop (continuous batching).  No counterpart in the reference, whose
 * IPModel_GAP_Calc serves one structure per call; each request is evaluated
 * with the physics of gapx_calc.
 * gapx_server_create starts n_slots engines (HIP stream + device workspace +
 * host thread each).  A free slot takes queued requests first-in first-out
 * into one batch while their estimated device footprint fits
 * hbm_budget / n_slots and the batch has at most max_batch_atoms atoms, and
 * evaluates the batch in one pass; slots overlap on the device.  The model's
 * precision must not change while a server runs.
 * gapx_server_submit copies the structure and returns a ticket at once
 * (GAPX_E_CAPACITY if it can never fit a slot, or if its footprint exceeds the
 * device memory free now (hipMemGetInfo) plus what a slot's engine already holds).  gapx_server_wait blocks until
 * that request is done, copies the requested outputs (as gapx_calc) and gives
 * its latency from submission to completion; a ticket is waited on once.
 * gapx_server_free finishes the queued requests and stops the threads. */
int gapx_server_create(gapx_model *model, int n_slots, int64_t hbm_budget_bytes, int max_batch_atoms,
                       gapx_server **out);
int gapx_server_submit(gapx_server *server, int n_atoms, const double *pos, const int *Z, const double *lattice,
                       const int *pbc, unsigned want, int64_t *ticket);
int gapx_server_poll(gapx_server *server, int64_t ticket, int *done);
int gapx_server_wait(gapx_server *server, int64_t ticket, double *e, double *local_e, double *f, double *virial,
                     double *local_virial, double *latency_s);
int gapx_server_stats(gapx_server *server, int64_t *n_done, int64_t *n_batches, int64_t *atoms_done,
                      double *busy_s);
/* Sessions: a client that resubmits one structure as it moves (MD) keeps it resident.
 * No counterpart in the reference beyond calc_connect's skin (Connection.f95:1085-1127).
 * gapx_server_session_open gives the session its own engine (stream + workspace) and a
 * skin in Angstrom; gapx_server_session_submit queues a request of the session (waited
 * on with gapx_server_wait like any ticket).  A session's requests are evaluated in
 * order with the neighbour list built at the model cutoff plus the skin and kept while
 * no atom has moved by half the skin (the distances recomputed as calc_dists does);
 * its descriptor and GP buffers stay resident; a resubmission with bitwise identical
 * positions, species, cell and pbc returns the resident results without device work.
 * Sessions run on the slot threads, overlapping on the device with each other and with
 * the batches.  Requests are admitted against the budget's slot share and against
 * hipMemGetInfo (free device memory plus what the session's engine already holds).
 * gapx_server_session_stats counts the requests done, the neighbour lists reused and
 * the results reused; gapx_server_session_close finishes the queued requests and frees
 * the session's device memory. */
int gapx_server_session_open(gapx_server *server, double skin, int64_t *session);
int gapx_server_session_submit(gapx_server *server, int64_t session, int n_atoms, const double *pos, const int *Z,
                               const double *lattice, const int *pbc, unsigned want, int64_t *ticket);
int gapx_server_session_stats(gapx_server *server, int64_t session, int64_t *n_done, int64_t *nl_reused,
                              int64_t *results_reused);
int gapx_server_session_close(gapx_server *server, int64_t session);
int gapx_server_free(gapx_server *server);

/* Skin reuse of the neighbour list (calc_connect with cutoff_skin,
 * Connection.f95:1085-1127; on by default).  When a calc's nl_cutoff exceeds the
 * model cutoff by a skin, the list built by the previous call on the same atoms
 * (same Z, structure layout, lattices, pbc and nl_cutoff) is kept as long as no
 * atom has moved by half the skin since it was built; only its distances are
 * recomputed (connection_calc_dists, Connection.f95:1321-1445).
 * gapx_last_nl_rebuilt tells whether the last calc rebuilt the list. */
int gapx_model_set_nl_reuse(gapx_model *model, int on);
int gapx_last_nl_rebuilt(const gapx_model *model, int *rebuilt);

/* Neighbour tables exactly as QUIP's calc_connect builds them
 * (Connection.f95:1035-1310): neighbour1(i) = (j >= i, shift, r) in append
 * order, neighbour2(i) = (j < i, index into neighbour1(j)).  Built on the GPU. */
int gapx_connect(int device, int n_atoms, const double *pos, const double *lattice, const int *pbc, double cutoff,
                 gapx_nl **out);
int gapx_nl_counts(const gapx_nl *nl, int *n1_total, int *n2_total, int *cells);
int gapx_nl_export(const gapx_nl *nl, int *n1_off, int *n1_j, int *n1_shift, double *n1_r, int *n2_off, int *n2_j,
                   int *n2_idx, int *map_shift);
int gapx_nl_free(gapx_nl *nl);

/* Descriptor vectors of coordinate `coord` (descriptor calc, IPModel_GAP.f95:423-425):
 * X[n_desc x dim] unit-normalised SOAP vectors, desc_atom[n_desc] centre atoms. */
int gapx_descriptors(gapx_model *model, int coord, int n_atoms, const double *pos, const int *Z,
                     const double *lattice, const int *pbc, double nl_cutoff, double *X, int *desc_atom,
                     int *n_desc);

/* FP64 GEMM autotuner (on by default).  No counterpart in the reference.  For
 * each gpCoordinate and centre-count bucket b (2^(b-1) < n_centres <= 2^b) the
 * first evaluation times every GEMM tile shape (tile t: (WM,WN,k tile) =
 * (4,2,16), (2,2,16), (2,4,16), (4,2,32), (2,2,32), (2,4,32); wavefront tiles of
 * 16x16, block 32WM x 32WN) and every GEMM 2 split-K factor in {1,2,4,8} on its
 * own operands, and keeps the fastest.
 * gapx_model_tuning lists the choices (up to max_entries; n_entries gets the
 * count) with the measured milliseconds of one launch. */
int gapx_model_set_autotune(gapx_model *model, int on);

/* Split-form autotuner (on with gapx_model_set_autotune).  No counterpart in the reference.
 * For each gpCoordinate, split precision and centre-count bucket the first evaluation
 * times the fused GP kernel (descriptors up to 256 components) with 1, 2, 4 and 8
 * sparse-point parts and with the stream-K schedule (every block of a CU-filling grid
 * takes an equal run of (centre block, sparse chunk) units), or the split GEMMs (larger
 * descriptors) with 64 x 64 and 128 x 128 LDS tiles, and keeps the fastest.
 * gapx_model_tuning_split lists the choices: fused_parts (-1 the stream-K schedule,
 * 0 where the GEMMs are used), gemm1_lds / gemm2_lds (-1 where the fused
 * kernel is used) and the measured milliseconds of one launch (fused kernel with its
 * partial sums, or GEMM 1; GEMM 2). */
int gapx_model_tuning_split(gapx_model *model, int coord, int max_entries, int *precision, int *bucket,
                            int *fused_parts, int *gemm1_lds, int *gemm2_lds, double *fused_or_gemm1_ms,
                            double *gemm2_ms, int *n_entries);
/* Fixes the schedule of the fused split GP kernel over the autotuner and the default
 * rule: -1 the stream-K schedule, 1..8 that many sparse-point parts (rounded so that
 * every part has chunks), 0 back to the tuner / default rule; other values are
 * GAPX_E_ARG.  No counterpart in the reference (a tuning override, as set_autotune). */
int gapx_model_set_fused_schedule(gapx_model *model, int parts);

/* Replay of the launch sequence after the neighbour count as a HIP graph (off by
 * default).  The sequence is captured on the first call and re-captured when a
 * size, a buffer, an output pointer or E_scale changes; calls with stage timing
 * on run it eagerly.  No counterpart in the reference. */
int gapx_model_set_graphs(gapx_model *model, int on);
int gapx_model_tuning(gapx_model *model, int coord, int max_entries, int *bucket, int *gemm1_tile, int *gemm2_tile,
                      int *gemm2_split, double *gemm1_ms, double *gemm2_ms, int *n_entries);

/* Stage times of the calc entry points (milliseconds, HIP events on the model's
 * stream, read back without extra synchronisation): [0] neighbour list,
 * [1] descriptor forward, [2] GEMM 1 kernel (K with the fused pow/alpha
 * epilogue), [3] per-centre energy rows + GEMM 2 (G), [4] descriptor adjoint,
 * [5] distance_2b + force/virial gather + per-structure sums.
 * gapx_model_set_timing(model, 1) starts recording and zeroes the totals;
 * gapx_last_timings gives the last call, gapx_timing_totals the sums over all
 * calls since timing was switched on and their number.  No counterpart in the
 * reference (QUIP's system_timer is host wall-clock only). */
int gapx_model_set_timing(gapx_model *model, int on);
int gapx_last_timings(gapx_model *model, double *ms6);
int gapx_timing_totals(gapx_model *model, double *ms6, int64_t *n_calls);

#ifdef __cplusplus
}
#endif
#endif
