// Host-side internals of libgapx.so shared by its translation units: error
// plumbing, device / pinned buffers, model and engine state.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gapx.h"
#include "gapx_kernels.h"

namespace gapx {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
// Records the calling thread's error message and returns `code`.
int fail(int code, const char *fmt, ...);
// The calling thread's last error message.
const char *last_error_text();

#define HIPCHK(x)                                                                                  \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) return fail(GAPX_E_DEVICE, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

#define CHK(x)                      \
    do {                            \
        int s_ = (x);               \
        if (s_ != GAPX_OK) return s_; \
    } while (0)

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------
// Incremented whenever a device buffer is (re)allocated: captured launch
// sequences hold buffer addresses and are replayed only while it is unchanged.
inline long long alloc_generation(bool bump = false) {
    static std::atomic<long long> gen{0};
    return bump ? ++gen : gen.load();
}

// A fixed block of device memory that buffers are carved from instead of hipMalloc
// (the allocation-free entry point).  Active for the calling thread between
// ArenaScope's construction and destruction.
struct Arena {
    char *base = nullptr;
    size_t size = 0, used = 0;
};
inline Arena *&active_arena() {
    static thread_local Arena *a = nullptr;
    return a;
}
struct ArenaScope {
    explicit ArenaScope(Arena *a) { active_arena() = a; }
    ~ArenaScope() { active_arena() = nullptr; }
};

struct DBuf {
    void *p = nullptr;
    size_t cap = 0;
    bool owned = true;   // false: a view into an arena, never freed here
    int ensure(size_t bytes) {
        if (bytes <= cap) return GAPX_OK;
        if (Arena *a = active_arena()) {   // carve, 256-byte aligned; nothing is allocated
            const size_t want = (std::max<size_t>(bytes, 256) + 255) & ~(size_t)255;
            if (a->used + want > a->size)
                return fail(GAPX_E_CAPACITY, "the allocation-free entry point's device block (%zu bytes) is too small: "
                            "%zu more bytes needed after %zu", a->size, want, a->used);
            if (p && owned) hipFree(p);
            p = a->base + a->used;
            a->used += want;
            cap = want;
            owned = false;
            return GAPX_OK;
        }
        if (p && owned) hipFree(p);
        p = nullptr;
        cap = 0;
        owned = true;
        size_t want = std::max<size_t>(bytes + bytes / 4, 256);
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) return fail(GAPX_E_DEVICE, "hipMalloc(%zu): %s", want, hipGetErrorString(e));
        cap = want;
        alloc_generation(true);
        return GAPX_OK;
    }
    // forget an arena view (the arena is re-carved by the next call)
    void drop_view() {
        if (!owned) { p = nullptr; cap = 0; owned = true; }
    }
    template <class T> T *as() const { return (T *)p; }
    ~DBuf() { if (p && owned) hipFree(p); }
};

// ---------------------------------------------------------------------------
// model
// ---------------------------------------------------------------------------
struct GemmTune {
    int g1 = 0, g2 = 1, parts = 1;   // GEMM 1 tile, GEMM 2 tile, GEMM 2 split-K factor
    float g1_ms = 0, g2_ms = 0;      // measured times of the chosen configurations
    // split forms: sparse-point parts of the fused kernel (0 = its default rule), and the
    // split GEMMs' tiles (-1 = the default rule, 0 = 64 x 64, 1 = 128 x 128 LDS tiles)
    int fparts = 0, g1lds = -1, g2lds = -1;
    // FP64: sparse-point parts of the fused GP kernel (0 = the two GEMMs) and its time
    int f64parts = 0;
    float f64_ms = 0;
};
constexpr size_t kTuneScratchCap = (size_t)256 << 20;   // largest partial-G buffer a tuning candidate may use
constexpr size_t kNoallocTuneCap = (size_t)32 << 20;    // largest one an allocation-free call sizes for a cached tuning

struct Coord {
    int kind = 0;          // 1 soap, 2 distance_2b
    std::string label, desc;
    // soap
    double rc = 0, ctw = 0.5, sigma = 0.5, cw = 1.0, bee = 10.0;
    int n = 0, L = 0, S = 1, dim = 0;
    std::vector<int> species, centreZ;
    double u00 = 0;
    std::vector<double> rb, T;
    std::vector<int> qmap;
    DBuf d_rb, d_T, d_qmap;
    // d2b
    int Z1 = 0, Z2 = 0;
    // gp
    int cov = 2;
    double delta = 1, zeta = 1, signal_mean = 0;
    int M = 0;
    std::vector<double> Xs, alpha, scut, theta;
    DBuf d_Xs, d_w, d_theta;
    DBuf d_XsP;   // Xs zero-padded to the fused FP64 kernel's chunks (descriptors up to 256 components)
    // mixed-precision copies of Xs (split hi/lo), built on demand
    int split_prec = -1;
    int dpad = 0, Mpad = 0;
    double wscale = 1.0;
    DBuf d_Xs_hi, d_Xs_lo, d_XsT_hi, d_XsT_lo;
    // FP64 GEMM tile choices per centre-count bucket (ceil log2 nc), from the autotuner
    std::mutex tune_mu;
    std::map<int, GemmTune> tune;
    std::map<int, GemmTune> mtune;   // split forms: key (precision << 8) | bucket
    // GAP variance: (K_MM + reg I)^-1 (M x M, FP64) for the regularisation last asked for
    std::mutex var_mu;
    double var_reg = -1.0;
    DBuf d_Cinv;
};

struct Workspace {
    DBuf pos, Z, sid, atom_off, geo, cell_base, ginv, cell_of, map_shift, cell_cnt, cell_start, cell_fill, cell_atoms,
        cell_unsorted, scan_part, scan_part2;
    DBuf cpos, cms;   // cell-ordered positions and map shifts (CellAtom, CellShift) for the candidate search
    DBuf n1_cnt, n1_off, n1_j, n1_shift, n1_r, n1_rev, n2_cnt, n2_off, n2_fill, n2_i, n2_k, n2_is, n2_ks;
    DBuf nb_off, nb_j, nb_rev, nb_r, nb_d, pairF, pairE, esite;
    DBuf X, norm, cbuf, W, Wf, e_part, G, Gparts, split_ws;
    DBuf f, le, lv, e, vir, minmax, spart;
    DBuf last_pos, skin_flags;   // positions of the last build (skin reuse), reuse test results
    DBuf cmask;                  // descriptor-centre mask of a caller-supplied connection
    DBuf t1_j, t1_shift, t1_r, t2_i, t2_k, nl_flags;   // per-atom slots of the one-pass neighbour search
    std::vector<DBuf> centres;   // per coordinate
    DBuf eatom;                  // per-atom distance_2b energies of one coordinate (energy_per_coordinate)
    // GAP variance: K, k = delta^2 K^zeta, V = k Cinv (each nc x M), per-centre G of the
    // variance, pair sums, per-atom variances; per coordinate the instance variances
    DBuf vK, vk, vV, Gv, pairFv, pairVE, varsite;
    std::vector<DBuf> var_inst, cut2_inst;
    DBuf o_epc, o_lgv, o_gvg;    // device outputs of the host-buffer entry points' options
};

// IPModel_GAP_Calc's args_str options that change what is computed (IPModel_GAP.f95:324-335)
struct CalcOpts {
    const unsigned char *mask = nullptr;   // atom_mask_name: host, N bytes
    int only_coord = 0;                    // only_descriptor: 1-based coordinate, 0 = all
    double *d_epc = nullptr;               // energy_per_coordinate: device, one per coordinate
    bool variance = false;                 // local_gap_variance / print_gap_variance
    double reg = 1e-3;                     // gap_variance_regularisation
    double *d_lgv = nullptr, *d_gvg = nullptr;   // device: local_gap_variance[N], gap_variance_gradient[3N]
};

// A connection handed in by the caller instead of calc_connect (a LAMMPS pair
// style): QUIP's neighbour1/neighbour2 tables with zero shifts, built on the
// host, and the atoms that may be descriptor centres.
struct ExtNL {
    std::vector<int> n1_off, n1_j, n1_rev, n2_off, n2_i, n2_k;
    std::vector<double> n1_r;
    std::vector<unsigned char> local;
};

// Pinned host staging (hipHostMalloc), grow-only; `pageable` staging (ordinary host
// memory, for the allocation-free entry point) when set before the first use;
// `coherent` (fine-grained) when kernels write it directly.
struct HBuf {
    void *p = nullptr;
    size_t cap = 0;
    bool pageable = false, coherent = false;
    std::vector<char> host;
    int ensure(size_t bytes) {
        if (bytes <= cap) return GAPX_OK;
        if (pageable) {
            host.resize(std::max<size_t>(bytes + bytes / 4, 4096));
            p = host.data();
            cap = host.size();
            return GAPX_OK;
        }
        if (p) hipHostFree(p);
        p = nullptr;
        cap = 0;
        size_t want = std::max<size_t>(bytes + bytes / 4, 4096);
        hipError_t e = hipHostMalloc(&p, want, coherent ? hipHostMallocCoherent : hipHostMallocDefault);
        if (e != hipSuccess) return fail(GAPX_E_DEVICE, "hipHostMalloc(%zu): %s", want, hipGetErrorString(e));
        cap = want;
        return GAPX_OK;
    }
    template <class T> T *as() const { return (T *)p; }
    ~HBuf() { if (p && !pageable) hipHostFree(p); }
};

// Stage times from HIP events on an engine's stream.  Two event sets alternate
// between calls; a call's set is read after the next call's neighbour-count
// synchronisation (or at query time), so timing adds no synchronisation.
constexpr int kStages = 6;
struct StageTimer {
    bool on = false;
    std::vector<hipEvent_t> pool[2];
    std::vector<int> marks[2];   // triples (stage, event a, event b)
    int used[2] = {0, 0};
    bool pending[2] = {false, false};
    long seq[2] = {0, 0};
    long next_seq = 0;
    int cur = 0;
    double last[kStages] = {0}, total[kStages] = {0};
    long ncalls = 0;
    void begin(hipStream_t) {
        cur ^= 1;
        resolve(cur);   // the call before last: complete by now
        used[cur] = 0;
        marks[cur].clear();
    }
    int mark(hipStream_t s) {
        if (used[cur] == (int)pool[cur].size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return -1;
            pool[cur].push_back(e);
        }
        hipEventRecord(pool[cur][used[cur]], s);
        return used[cur]++;
    }
    void stage(int k, int ea, int eb) {
        if (ea < 0 || eb < 0) return;
        marks[cur].push_back(k); marks[cur].push_back(ea); marks[cur].push_back(eb);
    }
    void end() { pending[cur] = true; seq[cur] = next_seq++; }
    static void add(std::vector<hipEvent_t> &pl, const std::vector<int> &mk, double *t6) {
        for (size_t q = 0; q + 2 < mk.size(); q += 3) {
            float t = 0;
            if (hipEventSynchronize(pl[mk[q + 2]]) != hipSuccess) continue;
            if (hipEventElapsedTime(&t, pl[mk[q + 1]], pl[mk[q + 2]]) != hipSuccess) continue;
            t6[mk[q]] += t;
        }
    }
    void resolve(int set) {
        if (!pending[set]) return;
        pending[set] = false;
        double t6[kStages] = {0};
        add(pool[set], marks[set], t6);
        for (int k = 0; k < kStages; k++) { last[k] = t6[k]; total[k] += t6[k]; }
        ncalls++;
    }
    void resolve_all() {
        int a = seq[0] <= seq[1] ? 0 : 1;
        resolve(a);
        resolve(a ^ 1);
    }
    void reset() {
        resolve_all();
        for (int k = 0; k < kStages; k++) total[k] = 0;
        ncalls = 0;
    }
    ~StageTimer() {
        for (auto &v : pool)
            for (auto e : v) hipEventDestroy(e);
    }
};

// One evaluation engine: a HIP stream, its device workspace, pinned staging and
// the host copies of the last inputs (Z, structure layout, geometry) so that
// unchanged inputs are not uploaded again.
struct Engine {
    hipStream_t stream = nullptr;
    Workspace ws;
    HBuf stage_in;    // Z, structure map, centre lists
    HBuf stage_geo;   // lattice inverses, fractional extents, geometry table
    HBuf stage_io;    // positions in, results out (host-buffer entry points)
    HBuf ctl;         // the neighbour-count read-back
    StageTimer timer;
    std::vector<int> h_Z, h_off, h_sid, h_cell_base;
    std::vector<NLGeom> h_geo;
    std::vector<std::vector<int>> h_centres;   // per coordinate
    bool have_inputs = false, have_geo = false;
    // skin reuse: the last build's cutoff and lattices; valid while the atoms are the same
    bool nl_valid = false, last_rebuilt = true;
    double nl_cut = 0;
    std::vector<double> nl_lats;
    std::vector<int> nl_pbcs;
    hipGraphExec_t gexec = nullptr;   // the captured launch sequence after the neighbour count
    std::vector<long long> gkey;
    int n1_total_last = 0;
    int nl_cap1 = 0, nl_cap2 = 0;   // slot sizes of the one-pass neighbour search (0: count first)
    int ncell_last = 0;
    // the last calc that formed GAP variances: its pair count (instance buffers are per coordinate)
    bool var_valid = false;
    int var_P = 0;
    // Speculative neighbour lists (no read-back inside a call): the capacity of the
    // neighbour1 tables, the inverse lattices and the last read extents of non-periodic
    // axes, and the calls whose overflow flags are not read yet, with what redoes them.
    int nl_capT1 = 0;
    bool force_exact = false;        // redoing calls: count first
    bool staging_dirty = false;      // an upload from pinned staging may still be queued
    std::vector<double> h_mm, h_ginv;
    struct Pending {
        int S = 0;
        std::vector<int> off, Z, pbcs;
        std::vector<double> lats;
        const double *d_pos = nullptr;
        double nl_cutoff = 0;
        unsigned want = 0;
        double *d_e = nullptr, *d_le = nullptr, *d_f = nullptr, *d_v = nullptr, *d_lv = nullptr;
        size_t slot = 0;             // byte offset of its flags (+ extents) in `checks`
        bool extents = false;
    };
    std::vector<Pending> pending;
    HBuf checks;                     // pinned (fine-grained) read-back slots of the pending calls
    size_t checks_used = 0;
    ~Engine() {
        if (gexec) hipGraphExecDestroy(gexec);
        if (stream) {
            hipStreamSynchronize(stream);
            hipStreamDestroy(stream);
        }
    }
};

struct ModelState {
    int device = 0;
    std::string label;
    double e0[119] = {0};
    double E_scale = 1.0;
    double cutoff = 0;
    int precision = GAPX_PREC_FP64;
    std::vector<std::unique_ptr<Coord>> coords;
    DBuf d_e0;
    bool autotune = true;
    int fused_schedule = 0;   // fused split GP kernel: -1 stream-K, 1..8 parts, 0 the tuner / default rule
    bool graphs = false;
    bool nl_reuse = true;   // keep a neighbour list within its skin (calc_connect with cutoff_skin)
    std::unique_ptr<Engine> eng;   // the engine of the plain entry points
    std::unique_ptr<Engine> eng_na;   // the allocation-free entry point's engine (pageable staging)
};

// engine entry points (gapx_api.hip)
int new_engine(gapx_model *m, std::unique_ptr<Engine> &out);
int ensure_split(gapx_model *m, Coord &c, int prec, hipStream_t s);
int check_batch(int S, const int *atom_off, const int *Z, const double *lats, const int *pbcs);
// Device-resident evaluation of S structures on engine eg (IPModel_GAP_Calc per structure).
// noalloc: no autotuning, graphs or timing (nothing that allocates); device buffers come
// from the caller's active arena.
int eval_batch(gapx_model *m, Engine &eg, int S, const int *atom_off, const double *d_pos, const double *h_pos,
               const int *Z, const double *lats, const int *pbcs, double nl_cutoff, unsigned want, double *d_e,
               double *d_le, double *d_f, double *d_v, double *d_lv, const ExtNL *ext = nullptr,
               const CalcOpts *opt = nullptr, bool noalloc = false);
// Forget every arena view of a workspace and the state that refers to its buffers.
void reset_arena_views(Engine &eg);
// Read the overflow flags of the engine's speculative calls (one synchronisation) and
// redo, exactly and in order, every call from the first that overflowed.
int resolve_pending(gapx_model *m, Engine &eg);
// Host-buffer evaluation of S structures on engine eg (pinned staging both ways).
// ext: a caller-supplied connection (one structure), else calc_connect.
// opt (nullable): per-call options; their host outputs h_epc[n_coord], h_lgv[N], h_gvg[3N]
// (nullable) are filled from device scratch.
int calc_host(gapx_model *m, Engine &eg, int S, const int *atom_off, const double *pos, const int *Z,
              const double *lats, const int *pbcs, double nl_cutoff, unsigned want, double *e, double *le, double *f,
              double *v, double *lv, const ExtNL *ext = nullptr, const CalcOpts *opt = nullptr,
              double *h_epc = nullptr, double *h_lgv = nullptr, double *h_gvg = nullptr);

}  // namespace gapx

// the opaque handle of the C ABI
struct gapx_model : gapx::ModelState {};
