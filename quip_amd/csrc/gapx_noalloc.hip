// The allocation-free entry point of libgapx.so: gapx_calc_device with every device
// work buffer carved from one file-scope device array instead of hipMalloc.
//
// Each GPU the library runs on holds its own copy of `scratch` (kScratchBytes) from
// the first launch until the process ends.  Its device address is looked up once per
// device, with that device current, by hipGetSymbolAddress; host code uses only that
// address (the array's name is a host-side stand-in, not a device address).  The
// array keeps the last call's contents: every buffer the kernels need cleared is
// cleared by a fill kernel inside the call, as in the allocating path.  All calls on
// one GPU share the array, so calls that overlap on one GPU race: this entry point
// serialises its own calls within the process (one lock per device) and waits for its
// work before returning.
#include <mutex>

#include "gapx_host.h"

using namespace gapx;

namespace {

constexpr size_t kScratchBytes = (size_t)160 << 20;   // 160 MiB per GPU
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct DeviceScratch {
    std::mutex mu;        // held for a whole call on this device
    void *ptr = nullptr;  // device address of `scratch` on this device
};
DeviceScratch g_scratch[kMaxDevices];
std::mutex g_lookup_mu;

// the device address of `scratch` on `dev` (the current device), looked up once
int scratch_address(int dev, void **out) {
    std::lock_guard<std::mutex> lk(g_lookup_mu);
    if (!g_scratch[dev].ptr) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(GAPX_E_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_scratch[dev].ptr = p;
    }
    *out = g_scratch[dev].ptr;
    return GAPX_OK;
}

}  // namespace

extern "C" int gapx_calc_device_noalloc(gapx_model *m, int N, const double *d_pos, const int *Z, const double *lat,
                                        const int *pbc, double nl_cutoff, unsigned want, double *d_e, double *d_le,
                                        double *d_f, double *d_v, double *d_lv) {
    if (!m) return fail(GAPX_E_ARG, "NULL model");
    if (N < 0) return fail(GAPX_E_ARG, "n_atoms < 0");
    if (m->device < 0 || m->device >= kMaxDevices) return fail(GAPX_E_ARG, "device %d out of range", m->device);
    // a lower bound of the work buffers: per atom the neighbour-list arrays and outputs,
    // per centre the descriptor, its c_nlm and the GP rows; larger requests are refused
    // before anything is queued (the carving refuses what this bound misses)
    double per_atom = 200.0;
    for (auto &cp : m->coords)
        if (cp->kind == 1)
            per_atom += 8.0 * (2.0 * cp->dim + (double)cp->S * cp->n * (cp->L + 1) * (cp->L + 1) +
                               (m->precision == GAPX_PREC_FP64 ? cp->M : 0));
    if (per_atom * N > (double)kScratchBytes)
        return fail(GAPX_E_CAPACITY, "%d atoms need at least %.0f bytes of work space; the allocation-free entry point "
                    "has %zu per device", N, per_atom * N, kScratchBytes);
    HIPCHK(hipSetDevice(m->device));
    DeviceScratch &ds = g_scratch[m->device];
    std::lock_guard<std::mutex> lk(ds.mu);
    void *base = nullptr;
    CHK(scratch_address(m->device, &base));
    Engine &eg = *m->eng_na;
    reset_arena_views(eg);
    Arena arena;
    arena.base = (char *)base;
    arena.size = kScratchBytes;
    int st;
    {
        ArenaScope scope(&arena);
        const int off[2] = {0, N};
        st = eval_batch(m, eg, 1, off, d_pos, nullptr, Z, lat, pbc, nl_cutoff, want, d_e, d_le, d_f, d_v, d_lv,
                        nullptr, nullptr, true);
    }
    // wait for the work in every case: the next call re-carves the same block
    const hipError_t se = hipStreamSynchronize(eg.stream);
    reset_arena_views(eg);
    if (st != GAPX_OK) return st;
    if (se != hipSuccess) return fail(GAPX_E_DEVICE, "hipStreamSynchronize: %s", hipGetErrorString(se));
    return GAPX_OK;
}
