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
//
// Sessions serve clients that resubmit one structure as it moves (an MD run): a
// session owns an engine whose neighbour list, built with the model cutoff plus the
// session's skin, and whose descriptor buffers stay resident on the device between
// its requests.  A resubmission reuses the list while no atom has moved by half the
// skin (calc_connect with cutoff_skin, Connection.f95:1085-1127; the distances are
// recomputed as calc_dists does), and a resubmission with bitwise identical
// positions, species and cell takes the resident results without device work.  A
// session's requests are evaluated in order by whichever slot thread is free, so
// sessions overlap on the device with each other and with the batches.
//
// Admission: a request is refused (GAPX_E_CAPACITY) when its estimated device
// footprint exceeds the slot's share of the budget, or exceeds the device memory that
// is free now (hipMemGetInfo) plus what the engine that would run it already holds.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
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

namespace {

// A client's resident structure: its engine (stream, workspace, neighbour list kept
// within the skin) and the results of its last evaluated request.
struct Session {
    std::unique_ptr<Engine> eng;
    double skin = 0;
    std::deque<std::shared_ptr<Request>> queue;
    bool busy = false, closing = false;
    bool have_last = false;
    std::vector<double> last_pos, last_lat, last_f, last_le, last_lv;
    std::vector<int> last_Z, last_pbc;
    double last_e = 0, last_vir[9] = {0};
    int64_t n_done = 0, nl_reused = 0, results_reused = 0;
    size_t held = 0;                 // device bytes its engine held after its last request
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
    std::map<int64_t, std::unique_ptr<Session>> sessions;
    std::vector<size_t> slot_held;   // device bytes each slot's engine held after its last batch
    int64_t next_ticket = 1, next_session = 1;
    bool stop = false;
    int64_t n_done = 0, n_batches = 0, atoms_done = 0;
    double busy_s = 0;
};

// device bytes an engine's workspace holds
static size_t engine_bytes(const Engine &eg) {
    const Workspace &w = eg.ws;
    size_t b = 0;
    for (const DBuf *x : {&w.pos, &w.Z, &w.sid, &w.atom_off, &w.geo, &w.cell_base, &w.ginv, &w.cell_of, &w.map_shift,
                          &w.cell_cnt, &w.cell_start, &w.cell_fill, &w.cell_atoms, &w.cell_unsorted, &w.scan_part,
                          &w.scan_part2, &w.n1_cnt, &w.n1_off, &w.n1_j, &w.n1_shift, &w.n1_r, &w.n1_rev, &w.n2_cnt,
                          &w.n2_off, &w.n2_fill, &w.n2_i, &w.n2_k, &w.n2_is, &w.n2_ks, &w.nb_off, &w.nb_j, &w.nb_rev,
                          &w.nb_r, &w.nb_d, &w.pairF, &w.pairE, &w.esite, &w.X, &w.norm, &w.cbuf, &w.W, &w.Wf,
                          &w.e_part, &w.G, &w.Gparts, &w.split_ws, &w.f, &w.le, &w.lv, &w.e, &w.vir, &w.minmax,
                          &w.last_pos, &w.skin_flags, &w.cmask, &w.t1_j, &w.t1_shift, &w.t1_r, &w.t2_i, &w.t2_k,
                          &w.nl_flags, &w.spart, &w.cpos, &w.cms})
        b += x->cap;
    for (const DBuf &x : w.centres) b += x.cap;
    return b;
}

// admission against the device: the footprint must fit the memory free now plus what
// the engine that would run the request already holds (its workspace is reused)
static int admit_device(gapx_server *srv, size_t footprint, size_t held) {
    size_t free_b = 0, total_b = 0;
    HIPCHK(hipSetDevice(srv->m->device));
    HIPCHK(hipMemGetInfo(&free_b, &total_b));
    if (footprint > free_b + held)
        return fail(GAPX_E_CAPACITY, "request needs ~%zu device bytes; %zu are free and the engine holds %zu",
                    footprint, free_b, held);
    return GAPX_OK;
}

// Estimated device bytes of one structure of N atoms inside a batch: per-atom
// neighbour-list, pair, descriptor and GEMM buffers of every coordinate.  Pairs
// per atom are bounded by a density of 0.1 atoms/A^3 inside the cutoff sphere
// (denser than any covalent solid; diamond Si is 0.05, 3C-SiC 0.097).
static size_t footprint_bytes(const gapx_model *m, int N) {
    const double rc = m->cutoff;
    const double nn = 0.1 * 4.18879 * rc * rc * rc + 8;
    double per_atom = 208.0 + 80.0;                   // NL per-atom arrays (cell-ordered copies included), outputs
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

// The session whose oldest queued request is the oldest of all idle sessions (null if none).
static Session *ready_session(gapx_server *srv, Clock::time_point *t) {
    Session *best = nullptr;
    for (auto &kv : srv->sessions) {
        Session &ss = *kv.second;
        if (ss.busy || ss.queue.empty()) continue;
        if (!best || ss.queue.front()->t_submit < best->queue.front()->t_submit) best = &ss;
    }
    if (best && t) *t = best->queue.front()->t_submit;
    return best;
}

// One request of a session on the session's engine: the neighbour list with the
// session's skin is kept while no atom has moved by half of it, and bitwise identical
// inputs take the results of the session's last request.
static void run_session_request(gapx_server *srv, Session &ss, Request &r) {
    gapx_model *m = srv->m;
    const auto t0 = Clock::now();
    r.t_start = t0;
    const size_t n = r.N;
    const bool same = ss.have_last && ss.last_pos == r.pos && ss.last_Z == r.Z &&
                      std::equal(r.lat, r.lat + 9, ss.last_lat.begin()) &&
                      std::equal(r.pbc, r.pbc + 3, ss.last_pbc.begin());
    int st = GAPX_OK;
    std::string msg;
    bool reused_nl = false;
    if (!same) {
        const unsigned all = GAPX_WANT_E | GAPX_WANT_F | GAPX_WANT_V | GAPX_WANT_LOCAL_E | GAPX_WANT_LOCAL_V;
        ss.last_f.resize(3 * n);
        ss.last_le.resize(n);
        ss.last_lv.resize(9 * n);
        const int off[2] = {0, r.N};
        st = calc_host(m, *ss.eng, 1, off, r.pos.data(), r.Z.data(), r.lat, r.pbc, m->cutoff + ss.skin, all,
                       &ss.last_e, ss.last_le.data(), ss.last_f.data(), ss.last_vir, ss.last_lv.data());
        if (st == GAPX_OK) {
            reused_nl = !ss.eng->last_rebuilt;
            ss.last_pos = r.pos;
            ss.last_Z = r.Z;
            ss.last_lat.assign(r.lat, r.lat + 9);
            ss.last_pbc.assign(r.pbc, r.pbc + 3);
            ss.have_last = true;
        } else {
            msg = last_error_text();
            ss.have_last = false;
        }
    }
    const size_t held = engine_bytes(*ss.eng);
    const auto t1 = Clock::now();
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        r.status = st;
        r.err = msg;
        if (st == GAPX_OK) {
            r.e = ss.last_e;
            std::copy(ss.last_vir, ss.last_vir + 9, r.vir);
            if (r.want & GAPX_WANT_F) r.f = ss.last_f;
            if (r.want & GAPX_WANT_LOCAL_E) r.le = ss.last_le;
            if (r.want & GAPX_WANT_LOCAL_V) r.lv = ss.last_lv;
        }
        r.t_done = t1;
        r.done = true;
        ss.busy = false;
        ss.n_done++;
        ss.held = held;
        if (same) ss.results_reused++;
        if (reused_nl) ss.nl_reused++;
        srv->n_done += 1;
        srv->atoms_done += r.N;
        srv->busy_s += std::chrono::duration<double>(t1 - t0).count();
    }
    srv->cv_done.notify_all();
    srv->cv_queue.notify_one();   // the session may have more requests queued
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
        Session *sess = nullptr;
        std::shared_ptr<Request> sreq;
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            srv->cv_queue.wait(lk, [&] { return srv->stop || !srv->queue.empty() || ready_session(srv, nullptr); });
            Clock::time_point ts;
            Session *rs = ready_session(srv, &ts);
            if (srv->queue.empty() && !rs) return;   // stop requested and nothing left
            if (rs && (srv->queue.empty() || ts < srv->queue.front()->t_submit)) {   // oldest first
                sess = rs;
                sreq = rs->queue.front();
                rs->queue.pop_front();
                rs->busy = true;
            }
            size_t fp = 0;
            long atoms = 0;
            while (!sess && !srv->queue.empty()) {
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
        if (sess) {
            run_session_request(srv, *sess, *sreq);
            continue;
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
            srv->slot_held[slot] = engine_bytes(eg);
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
    srv->slot_held.assign(n_slots, 0);
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
    size_t held = 0;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        for (size_t h : srv->slot_held) held = std::max(held, h);
    }
    CHK(admit_device(srv, r->footprint, held));
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

extern "C" int gapx_server_session_open(gapx_server *srv, double skin, int64_t *session) {
    if (!srv || !session) return fail(GAPX_E_ARG, "NULL argument");
    if (!(skin >= 0) || skin > 10.0) return fail(GAPX_E_ARG, "skin must be in [0, 10] A (got %g)", skin);
    HIPCHK(hipSetDevice(srv->m->device));
    auto ss = std::make_unique<Session>();
    CHK(new_engine(srv->m, ss->eng));
    ss->skin = skin;
    std::lock_guard<std::mutex> lk(srv->mu);
    if (srv->sessions.size() >= 256) return fail(GAPX_E_CAPACITY, "at most 256 open sessions");
    *session = srv->next_session++;
    srv->sessions[*session] = std::move(ss);
    return GAPX_OK;
}

extern "C" int gapx_server_session_submit(gapx_server *srv, int64_t session, int N, const double *pos, const int *Z,
                                          const double *lat, const int *pbc, unsigned want, int64_t *ticket) {
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
    if (r->footprint > srv->slot_budget)
        return fail(GAPX_E_CAPACITY, "structure of %d atoms needs ~%zu device bytes; a session admits %zu", N,
                    r->footprint, srv->slot_budget);
    size_t held = 0;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        auto it = srv->sessions.find(session);
        if (it == srv->sessions.end() || it->second->closing)
            return fail(GAPX_E_ARG, "unknown session %lld", (long long)session);
        held = it->second->held;
    }
    CHK(admit_device(srv, r->footprint, held));
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        auto it = srv->sessions.find(session);
        if (it == srv->sessions.end() || it->second->closing)
            return fail(GAPX_E_ARG, "unknown session %lld", (long long)session);
        r->t_submit = Clock::now();
        *ticket = srv->next_ticket++;
        srv->tickets[*ticket] = r;
        it->second->queue.push_back(r);
    }
    srv->cv_queue.notify_one();
    return GAPX_OK;
}

extern "C" int gapx_server_session_stats(gapx_server *srv, int64_t session, int64_t *n_done, int64_t *nl_reused,
                                         int64_t *results_reused) {
    if (!srv) return fail(GAPX_E_ARG, "NULL server");
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->sessions.find(session);
    if (it == srv->sessions.end()) return fail(GAPX_E_ARG, "unknown session %lld", (long long)session);
    if (n_done) *n_done = it->second->n_done;
    if (nl_reused) *nl_reused = it->second->nl_reused;
    if (results_reused) *results_reused = it->second->results_reused;
    return GAPX_OK;
}

extern "C" int gapx_server_session_close(gapx_server *srv, int64_t session) {
    if (!srv) return fail(GAPX_E_ARG, "NULL server");
    std::unique_ptr<Session> gone;
    {
        std::unique_lock<std::mutex> lk(srv->mu);
        auto it = srv->sessions.find(session);
        if (it == srv->sessions.end() || it->second->closing)
            return fail(GAPX_E_ARG, "unknown session %lld", (long long)session);
        Session *ss = it->second.get();
        ss->closing = true;   // no new requests; the queued ones are finished first
        srv->cv_done.wait(lk, [&] { return ss->queue.empty() && !ss->busy; });
        gone = std::move(srv->sessions[session]);
        srv->sessions.erase(session);
    }
    HIPCHK(hipSetDevice(srv->m->device));
    gone.reset();   // the engine: its stream and resident buffers
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
