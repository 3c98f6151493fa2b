// Serving loop of libgapx.so: continuous batching of E+F+virial requests.
//
// A server owns n_slots engines (each a HIP stream with its own device
// workspace) and one host thread per engine.  Requests are queued as they are
// submitted.  A free slot takes queued requests first-in first-out into one
// batch while their estimated device footprint fits the slot's share of the
// HBM budget and the batch stays under max_batch_atoms, evaluates the batch in
// one pass (the batched engine), and completes each request.  Requests that
// arrive while other batches are on the GPU are taken by the next free slot,
// so the work of different slots overlaps on the device.  Workspaces grow to
// the largest batch a slot has run and stay resident; after warm-up no request
// allocates device memory.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "gapx_host.h"

using namespace gapx;
using Clock = std::chrono::steady_clock;

namespace {

struct Request {
    int N = 0;
    unsigned want = 0;
    std::vector<double> pos, lv, f, le;
    std::vector<int> Z;
    double lat[9], vir[9], e = 0;
    int pbc[3];
    int status = GAPX_OK;
    std::string err;
    bool done = false;
    Clock::time_point t_submit, t_start, t_done;
    size_t footprint = 0;
};

}  // namespace

struct gapx_server {
    gapx_model *m = nullptr;
    int n_slots = 1;
    size_t slot_budget = 0;
    int max_batch_atoms = 0;
    std::vector<std::unique_ptr<Engine>> engines;
    std::vector<std::thread> threads;
    std::mutex mu;
    std::condition_variable cv_queue, cv_done;
    std::deque<std::shared_ptr<Request>> queue;
    std::unordered_map<int64_t, std::shared_ptr<Request>> tickets;
    int64_t next_ticket = 1;
    bool stop = false;
    int64_t n_done = 0, n_batches = 0, atoms_done = 0;
    double busy_s = 0;
};

// Estimated device bytes of one structure of N atoms inside a batch: per-atom
// neighbour-list, pair, descriptor and GEMM buffers of every coordinate.  Pairs
// per atom are bounded by a density of 0.1 atoms/A^3 inside the cutoff sphere
// (denser than any covalent solid; diamond Si is 0.05, 3C-SiC 0.097).
static size_t footprint_bytes(const gapx_model *m, int N) {
    const double rc = m->cutoff;
    const double nn = 0.1 * 4.18879 * rc * rc * rc + 8;
    double per_atom = 160.0 + 80.0;                   // NL per-atom arrays, outputs
    per_atom += nn * (4 + 4 + 8 + 24 + 24 + 8);        // full pair view + pair forces / energies
    per_atom += 0.5 * nn * (4 + 12 + 8 + 4 + 16);      // neighbour1 / neighbour2 tables
    for (auto &cp : m->coords) {
        const Coord &c = *cp;
        if (c.kind != 1) continue;
        const double KLM = (double)c.S * c.n * (c.L + 1) * (c.L + 1);
        per_atom += 8.0 * (3.0 * c.dim + KLM + c.M + c.M / 32.0 + 2);   // X, G, G partials, c_nlm, W, energy parts
        if (m->precision != GAPX_PREC_FP64) per_atom += 4.0 * c.dim;      // split staging
    }
    return (size_t)(per_atom * N) + (1u << 20);
}

static void worker(gapx_server *srv, int slot) {
    Engine &eg = *srv->engines[slot];
    gapx_model *m = srv->m;
    std::vector<std::shared_ptr<Request>> batch;
    std::vector<int> off, Z, pbcs;
    std::vector<double> pos, lats, E, V, F, LE, LV;
    hipSetDevice(m->device);
    for (;;) {
        batch.clear();
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            srv->cv_queue.wait(lk, [&] { return srv->stop || !srv->queue.empty(); });
            if (srv->queue.empty()) return;   // stop requested and nothing left
            size_t fp = 0;
            long atoms = 0;
            while (!srv->queue.empty()) {
                auto &r = srv->queue.front();
                if (!batch.empty() &&
                    (fp + r->footprint > srv->slot_budget || atoms + r->N > (long)srv->max_batch_atoms))
                    break;
                fp += r->footprint;
                atoms += r->N;
                batch.push_back(r);
                srv->queue.pop_front();
            }
        }
        const auto t0 = Clock::now();
        const int S = (int)batch.size();
        off.assign(S + 1, 0);
        for (int s = 0; s < S; s++) off[s + 1] = off[s] + batch[s]->N;
        const int N = off[S];
        pos.resize((size_t)N * 3); Z.resize(N); lats.resize(9 * (size_t)S); pbcs.resize(3 * (size_t)S);
        E.resize(S); V.resize(9 * (size_t)S); F.resize((size_t)N * 3); LE.resize(N); LV.resize((size_t)N * 9);
        for (int s = 0; s < S; s++) {
            Request &r = *batch[s];
            r.t_start = t0;
            std::copy(r.pos.begin(), r.pos.end(), pos.begin() + 3 * (size_t)off[s]);
            std::copy(r.Z.begin(), r.Z.end(), Z.begin() + off[s]);
            std::copy(r.lat, r.lat + 9, lats.begin() + 9 * s);
            std::copy(r.pbc, r.pbc + 3, pbcs.begin() + 3 * s);
        }
        const unsigned all = GAPX_WANT_E | GAPX_WANT_F | GAPX_WANT_V | GAPX_WANT_LOCAL_E | GAPX_WANT_LOCAL_V;
        std::vector<int> status(S, GAPX_OK);
        std::vector<std::string> msg(S);
        int st = calc_host(m, eg, S, off.data(), pos.data(), Z.data(), lats.data(), pbcs.data(), 0.0, all, E.data(),
                           LE.data(), F.data(), V.data(), LV.data());
        if (st != GAPX_OK && S == 1) {
            status[0] = st;
            msg[0] = last_error_text();
        } else if (st != GAPX_OK) {
            // isolate the failing request(s): every request of the batch alone
            for (int s = 0; s < S; s++) {
                const int o1[2] = {0, batch[s]->N};
                const size_t a = off[s];
                status[s] = calc_host(m, eg, 1, o1, pos.data() + 3 * a, Z.data() + a, lats.data() + 9 * s,
                                      pbcs.data() + 3 * s, 0.0, all, E.data() + s, LE.data() + a, F.data() + 3 * a,
                                      V.data() + 9 * s, LV.data() + 9 * a);
                if (status[s] != GAPX_OK) msg[s] = last_error_text();
            }
        }
        const auto t1 = Clock::now();
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            for (int s = 0; s < S; s++) {
                Request &r = *batch[s];
                r.status = status[s];
                r.err = msg[s];
                if (r.status == GAPX_OK) {
                    const size_t a = off[s], n = r.N;
                    r.e = E[s];
                    std::copy(V.begin() + 9 * s, V.begin() + 9 * s + 9, r.vir);
                    if (r.want & GAPX_WANT_F) r.f.assign(F.begin() + 3 * a, F.begin() + 3 * (a + n));
                    if (r.want & GAPX_WANT_LOCAL_E) r.le.assign(LE.begin() + a, LE.begin() + a + n);
                    if (r.want & GAPX_WANT_LOCAL_V) r.lv.assign(LV.begin() + 9 * a, LV.begin() + 9 * (a + n));
                }
                r.t_done = t1;
                r.done = true;
                srv->atoms_done += r.N;
            }
            srv->n_done += S;
            srv->n_batches += 1;
            srv->busy_s += std::chrono::duration<double>(t1 - t0).count();
        }
        srv->cv_done.notify_all();
    }
}

extern "C" int gapx_server_create(gapx_model *m, int n_slots, int64_t hbm_budget, int max_batch_atoms,
                                  gapx_server **out) {
    if (!m || !out) return fail(GAPX_E_ARG, "NULL argument");
    *out = nullptr;
    if (n_slots < 1 || n_slots > 8) return fail(GAPX_E_ARG, "n_slots must be in [1, 8] (got %d)", n_slots);
    if (hbm_budget <= 0) return fail(GAPX_E_ARG, "hbm_budget must be positive");
    if (max_batch_atoms < 1) return fail(GAPX_E_ARG, "max_batch_atoms must be >= 1");
    HIPCHK(hipSetDevice(m->device));
    // shared read-only model state is completed before any worker starts
    for (auto &cp : m->coords)
        if (cp->kind == 1 && m->precision != GAPX_PREC_FP64) CHK(ensure_split(m, *cp, m->precision, m->eng->stream));
    HIPCHK(hipStreamSynchronize(m->eng->stream));
    std::unique_ptr<gapx_server> srv(new gapx_server());
    srv->m = m;
    srv->n_slots = n_slots;
    srv->slot_budget = (size_t)hbm_budget / n_slots;
    srv->max_batch_atoms = max_batch_atoms;
    for (int k = 0; k < n_slots; k++) {
        std::unique_ptr<Engine> e;
        CHK(new_engine(m, e));
        srv->engines.push_back(std::move(e));
    }
    for (int k = 0; k < n_slots; k++) srv->threads.emplace_back(worker, srv.get(), k);
    *out = srv.release();
    return GAPX_OK;
}

extern "C" int gapx_server_submit(gapx_server *srv, int N, const double *pos, const int *Z, const double *lat,
                                  const int *pbc, unsigned want, int64_t *ticket) {
    if (!srv || !ticket) return fail(GAPX_E_ARG, "NULL argument");
    if (N < 0 || (N > 0 && (!pos || !Z)) || !lat || !pbc) return fail(GAPX_E_ARG, "bad structure arguments");
    const int off[2] = {0, N};
    CHK(check_batch(1, off, Z, lat, pbc));
    const double vol = lat[0] * (lat[4] * lat[8] - lat[5] * lat[7]) - lat[1] * (lat[3] * lat[8] - lat[5] * lat[6]) +
                       lat[2] * (lat[3] * lat[7] - lat[4] * lat[6]);
    if (!(std::fabs(vol) > 0)) return fail(GAPX_E_GEOMETRY, "lattice is singular");
    auto r = std::make_shared<Request>();
    r->N = N;
    r->want = want;
    r->pos.assign(pos, pos + 3 * (size_t)N);
    r->Z.assign(Z, Z + N);
    std::copy(lat, lat + 9, r->lat);
    std::copy(pbc, pbc + 3, r->pbc);
    r->footprint = footprint_bytes(srv->m, N);
    if (r->footprint > srv->slot_budget || N > srv->max_batch_atoms)
        return fail(GAPX_E_CAPACITY, "structure of %d atoms needs ~%zu device bytes; a slot admits %zu bytes and %d atoms",
                    N, r->footprint, srv->slot_budget, srv->max_batch_atoms);
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        r->t_submit = Clock::now();
        *ticket = srv->next_ticket++;
        srv->tickets[*ticket] = r;
        srv->queue.push_back(r);
    }
    srv->cv_queue.notify_one();
    return GAPX_OK;
}

extern "C" int gapx_server_poll(gapx_server *srv, int64_t ticket, int *done) {
    if (!srv || !done) return fail(GAPX_E_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->tickets.find(ticket);
    if (it == srv->tickets.end()) return fail(GAPX_E_ARG, "unknown ticket %lld", (long long)ticket);
    *done = it->second->done ? 1 : 0;
    return GAPX_OK;
}

extern "C" int gapx_server_wait(gapx_server *srv, int64_t ticket, double *e, double *le, double *f, double *v,
                                double *lv, double *latency_s) {
    if (!srv) return fail(GAPX_E_ARG, "NULL server");
    std::shared_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lk(srv->mu);
        auto it = srv->tickets.find(ticket);
        if (it == srv->tickets.end()) return fail(GAPX_E_ARG, "unknown ticket %lld", (long long)ticket);
        r = it->second;
        // the lock is released while waiting: other threads may insert tickets (which
        // can rehash the map) or wait on the same ticket, so it is erased by key
        srv->cv_done.wait(lk, [&] { return r->done; });
        if (srv->tickets.erase(ticket) == 0)
            return fail(GAPX_E_ARG, "ticket %lld was waited on by another thread", (long long)ticket);
    }
    if (latency_s) *latency_s = std::chrono::duration<double>(r->t_done - r->t_submit).count();
    if (r->status != GAPX_OK) return fail(r->status, "%s", r->err.c_str());
    if (e) *e = r->e;
    if (v) std::copy(r->vir, r->vir + 9, v);
    if (f && !r->f.empty()) std::copy(r->f.begin(), r->f.end(), f);
    if (le && !r->le.empty()) std::copy(r->le.begin(), r->le.end(), le);
    if (lv && !r->lv.empty()) std::copy(r->lv.begin(), r->lv.end(), lv);
    return GAPX_OK;
}

extern "C" int gapx_server_stats(gapx_server *srv, int64_t *n_done, int64_t *n_batches, int64_t *atoms_done,
                                 double *busy_s) {
    if (!srv) return fail(GAPX_E_ARG, "NULL server");
    std::lock_guard<std::mutex> lk(srv->mu);
    if (n_done) *n_done = srv->n_done;
    if (n_batches) *n_batches = srv->n_batches;
    if (atoms_done) *atoms_done = srv->atoms_done;
    if (busy_s) *busy_s = srv->busy_s;
    return GAPX_OK;
}

extern "C" int gapx_server_free(gapx_server *srv) {
    if (!srv) return GAPX_OK;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
    }
    srv->cv_queue.notify_all();
    for (auto &t : srv->threads) t.join();   // queued requests are finished first
    delete srv;
    return GAPX_OK;
}
