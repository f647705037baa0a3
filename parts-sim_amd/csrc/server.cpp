// The serving loop of include/psx.h: continuous batching of independent solve requests on one GPU.
//
// Requests queue in arrival order. A pool of max_inflight workers admits the queue head as soon as a
// worker is free and the head's estimated device footprint fits the budget beside what running
// requests (and resident parts) hold; resident parts are evicted, least recently used first, to make
// room. An admitted request builds its hierarchy with the device kernels (or on the host when the
// request asks for it, or takes its resident part) and begins its run. A small part (at most
// pack_max_dof DOFs) then joins the pack: one thread advances every member's run together, each
// step of a batch one shared launch for all members (packed.hpp), members joining and leaving at
// batch boundaries. A larger part solves its load cases on the model's own HIP stream, so its
// kernels overlap with the pack's and other requests'. Either way each request's arithmetic is
// exactly that of a lone solve.
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <list>
#include <map>
#include <mutex>
#include <thread>

#include "abi.hpp"
#include "model.hpp"
#include "packed.hpp"

using namespace psx::abi;
using Clock = std::chrono::steady_clock;

namespace {

double msBetween(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); }

struct Request {
    int64_t ticket = 0;
    // the part, copied
    int nx = 0, ny = 0, nz = 0, maxLevels = 0, onDevice = 0;
    double nu = 0.3;
    int64_t coarsest = 0;
    std::vector<float> density;
    std::vector<uint8_t> bc;
    std::vector<double> diagAdd;
    // the load cases
    int nrhs = 1;
    int64_t nDof = 0;
    std::vector<double> f, x0, u;
    bool x0PerRhs = false;
    int64_t partId = -1;
    psx_solve_opts opts{};
    int64_t estimate = 0;
    // outcome
    bool done = false;
    psx_result res{};
    std::string error;
    Clock::time_point submitted, admitted, built, finished;
};

// device bytes of a part before it is built: masks and coefficients (about 40 B per DOF over the
// hierarchy), per load case the work vectors of every level and the request's own f/u/x0 buffers
// (about 144 B per DOF), and room for own element matrices
int64_t footprint(const Request& r) {
    const int64_t n = r.nDof;
    int64_t b = n * (40 + 144 * int64_t(r.nrhs)) + 64 * (int64_t(1) << 20);
    if (r.onDevice) {
        // device build: the uploaded density and held DOFs, per-level flags, and the coarsest level's
        // two m x m sweep buffers beside its inverse (m <= the coarsest free-DOF target, at most 3000)
        const int64_t m = std::min<int64_t>(std::max<int64_t>(r.coarsest, 1200), 3000);
        b += 3 * n + 3 * 8 * m * m;
    }
    return b;
}

struct Resident {
    std::unique_ptr<psx::Model> model;
    bool busy = false;
    Clock::time_point used;
    int64_t bytes = 0;  // what `held` counts for it
};

// a request between admission and completion: its model and device vectors
struct Job {
    Request* r = nullptr;
    bool reused = false;
    psx::Model* M = nullptr;
    std::unique_ptr<psx::Model> own;
    double* buf[3] = {nullptr, nullptr, nullptr};  // f, u, x0 (nrhs x n each) on M's stream
    std::vector<psx_solve_info> info;
    psx::Model::Run run;
    void freeBuffers() {
        if (!M || !(buf[0] || buf[1] || buf[2])) return;
        for (double*& q : buf)
            if (q) {
                (void)hipFreeAsync(q, M->stream);
                q = nullptr;
            }
        (void)hipStreamSynchronize(M->stream);
    }
    ~Job() { freeBuffers(); }
};

}  // namespace

struct psx_server_s {
    std::mutex mu;
    std::condition_variable wake, finished, packWake;
    std::deque<Request*> queue;
    std::map<int64_t, std::unique_ptr<Request>> requests;
    std::map<int64_t, Resident> parts;
    std::vector<std::thread> workers;
    std::thread packer;
    int64_t budget = 0, held = 0, nextTicket = 1;
    int running = 0, device = 0;
    bool stopping = false, packStop = false;
    // packing: requests of at most packDof DOFs run in the shared-launch pack (up to packMax at once).
    // Measured on the 40-request stream of void-block parts (edges 32-256, 200 requests/s): packing
    // parts up to 100 000 DOFs leaves solves/s and latencies where the lone streams have them, while
    // packing parts up to 3 * 2^20 DOFs kept the solves/s but raised the median latency from about
    // 0.2 s to 0.75 s (the pack's members advance in lockstep, each step as long as its largest
    // member's kernel, beside the big parts' streams); so only tiny, launch-bound parts are packed by
    // default.
    int packMax = 16;
    int64_t packDof = 100000;
    int packed = 0;  // jobs in the pack or its inbox
    std::vector<std::unique_ptr<Job>> inbox;

    // bytes held by resident parts that are not in use (evictable)
    int64_t evictable() const {
        int64_t b = 0;
        for (const auto& [id, p] : parts)
            if (!p.busy) b += p.bytes;
        return b;
    }
    // evict idle resident parts, least recently used first, until need fits (lock held); the evicted
    // models move to `out`, whose owner destroys them after unlocking (a model's teardown waits on its
    // stream, which must not hold up other workers' admissions and completions)
    void makeRoom(int64_t need, std::vector<std::unique_ptr<psx::Model>>& out) {
        while (held + need > budget) {
            auto victim = parts.end();
            for (auto it = parts.begin(); it != parts.end(); ++it)
                if (!it->second.busy && (victim == parts.end() || it->second.used < victim->second.used)) victim = it;
            if (victim == parts.end()) return;
            held -= victim->second.bytes;
            out.push_back(std::move(victim->second.model));
            parts.erase(victim);
        }
    }

    void work() {
        if (hipSetDevice(device) != hipSuccess) return;
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            // admission: the queue head, once its footprint fits beside what is held (a request
            // that could never fit runs alone)
            wake.wait(lk, [&] {
                if (stopping) return true;
                if (queue.empty()) return false;
                const Request* r = queue.front();
                const int64_t need = r->estimate;
                return held + need <= budget || held + need - evictable() <= budget || running == 0;
            });
            if (stopping) return;
            Request* r = queue.front();
            queue.pop_front();
            // a resident part that no other request is using: taken before any eviction, and its
            // model's bytes are already held
            bool reused = false;
            if (r->partId >= 0) {
                auto it = parts.find(r->partId);
                if (it != parts.end() && !it->second.busy) {
                    it->second.busy = true;
                    reused = true;
                    r->estimate = std::max<int64_t>(0, r->estimate - it->second.bytes);
                }
            }
            std::vector<std::unique_ptr<psx::Model>> evicted;
            makeRoom(r->estimate, evicted);
            held += r->estimate;
            running++;
            r->admitted = Clock::now();
            // small parts go to the pack while it has room
            const bool pack = packMax > 1 && r->nDof <= packDof && packed < packMax;
            if (pack) packed++;
            lk.unlock();
            evicted.clear();
            auto job = std::make_unique<Job>();
            job->r = r;
            job->reused = reused;
            bool iterate = false;
            r->res.status = guarded([&] {
                iterate = prepare(*job);
                if (iterate && !pack) {
                    solveLone(*job);
                    iterate = false;
                }
                return PSX_OK;
            });
            if (r->res.status != PSX_OK) {
                r->error = g_error;
                iterate = false;
            }
            if (iterate) {
                // the pack thread runs its batches from here on
                r->res.packed = 1;
                lk.lock();
                inbox.push_back(std::move(job));
                packWake.notify_one();
                continue;
            }
            if (pack) {
                lk.lock();
                packed--;
                lk.unlock();
            }
            complete(std::move(job));
            lk.lock();
        }
    }

    // builds (or takes) the model, uploads the load cases and begins the run; false when the run
    // needs no iterations (its results are then final)
    bool prepare(Job& j) {
        Request& r = *j.r;
        if (j.reused) {
            std::lock_guard<std::mutex> g(mu);
            j.M = parts.at(r.partId).model.get();
        } else {
            psx_model_desc d{};
            d.nx = r.nx;
            d.ny = r.ny;
            d.nz = r.nz;
            d.density = r.density.data();
            d.nu = r.nu;
            d.bc = r.bc.data();
            d.diag_add = r.diagAdd.empty() ? nullptr : r.diagAdd.data();
            d.coarsest_max_dof = r.coarsest;
            d.max_levels = r.maxLevels;
            d.build_on_device = r.onDevice;
            j.own = std::make_unique<psx::Model>(d);
            j.own->estimateSmoothers();
            j.own->sync();
            j.M = j.own.get();
        }
        psx::Model* M = j.M;
        r.built = Clock::now();
        r.res.build_ms = j.reused ? 0 : msBetween(r.admitted, r.built);
        r.res.reused_part = j.reused ? 1 : 0;
        r.res.device_bytes = M->totalBytes();
        const int64_t n = M->nDof();
        if (n != r.nDof) throw std::invalid_argument("part does not match its load vectors");
        // the request's vectors: f, u and x0 (nrhs x n) in HBM for the whole request; all load
        // cases run in one multi-instance solve
        const size_t nk = size_t(n) * size_t(r.nrhs);
        HIPCHK(hipMallocAsync(reinterpret_cast<void**>(&j.buf[0]), nk * 8, M->stream));
        HIPCHK(hipMallocAsync(reinterpret_cast<void**>(&j.buf[1]), nk * 8, M->stream));
        if (!r.x0.empty()) HIPCHK(hipMallocAsync(reinterpret_cast<void**>(&j.buf[2]), nk * 8, M->stream));
        double *df = j.buf[0], *du = j.buf[1], *dx = j.buf[2];
        HIPCHK(hipMemcpyAsync(df, r.f.data(), nk * 8, hipMemcpyHostToDevice, M->stream));
        if (dx) {
            if (r.x0PerRhs) HIPCHK(hipMemcpyAsync(dx, r.x0.data(), nk * 8, hipMemcpyHostToDevice, M->stream));
            else
                for (int k = 0; k < r.nrhs; k++)
                    HIPCHK(hipMemcpyAsync(dx + size_t(k) * n, r.x0.data(), size_t(n) * 8, hipMemcpyHostToDevice, M->stream));
        }
        j.info.assign(size_t(r.nrhs), psx_solve_info{});
        try {
            if (M->begin(j.run, r.nrhs, df, dx, r.opts, du, j.info.data())) return true;
        } catch (...) {
            M->resetRun();
            throw;
        }
        M->resetRun();
        return false;
    }

    // the run's batches on the model's own stream
    void solveLone(Job& j) {
        psx::Model* M = j.M;
        struct Reset {
            psx::Model* m;
            ~Reset() { m->resetRun(); }
        } reset{M};
        for (;;) {
            M->enqueueBatch(j.run);
            M->readState();
            if (M->poll(j.run, false)) break;
        }
        M->end(j.run);
    }

    // results to the host, the model kept resident or freed, the request done
    void complete(std::unique_ptr<Job> j) {
        Request& r = *j->r;
        if (r.res.status == PSX_OK) {
            r.res.status = guarded([&] {
                const size_t nk = size_t(j->M->nDof()) * size_t(r.nrhs);
                r.u.assign(nk, 0.0);
                HIPCHK(hipMemcpyAsync(r.u.data(), j->buf[1], nk * 8, hipMemcpyDeviceToHost, j->M->stream));
                j->M->sync();
                r.res.converged = 1;
                for (const psx_solve_info& i : j->info) {
                    r.res.iterations += i.iterations;
                    r.res.max_iterations = std::max(r.res.max_iterations, i.iterations);
                    r.res.residual = std::max(r.res.residual, i.residual);
                    r.res.converged = r.res.converged && i.converged;
                }
                r.res.solve_ms = msBetween(r.built, Clock::now());
                return PSX_OK;
            });
            if (r.res.status != PSX_OK) r.error = g_error;
        }
        j->freeBuffers();
        std::lock_guard<std::mutex> g(mu);
        j->M = nullptr;  // a resident model may be evicted once the lock is released
        if (j->reused) {
            auto& p = parts.at(r.partId);
            p.busy = false;
            p.used = Clock::now();
            // more load cases than before grow the part's work vectors
            const int64_t now = p.model->totalBytes();
            held += now - p.bytes;
            p.bytes = now;
        } else if (j->own && r.partId >= 0 && r.res.status == PSX_OK && parts.find(r.partId) == parts.end()) {
            // keep the part resident (its bytes count against the budget until evicted)
            const int64_t b = j->own->totalBytes();
            held += b;
            parts[r.partId] = Resident{std::move(j->own), false, Clock::now(), b};
        }
        held -= r.estimate;
        running--;
        r.finished = Clock::now();
        r.res.queue_ms = msBetween(r.submitted, r.admitted);
        r.res.total_ms = msBetween(r.submitted, r.finished);
        r.done = true;
        finished.notify_all();
        wake.notify_all();
        // j, with its model when that is not kept, is destroyed after the lock is released (a model's
        // teardown waits on its stream)
    }

    // the pack: members join from the inbox at batch boundaries and leave when their runs stop
    void packLoop() {
        if (hipSetDevice(device) != hipSuccess) return;
        std::unique_ptr<psx::PackedSolve> pk;
        std::map<psx::Model*, std::unique_ptr<Job>> jobs;
        const bool graphs = !psx::Model::graphsOff();
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            packWake.wait(lk, [&] { return packStop || !inbox.empty() || !jobs.empty(); });
            if (packStop && inbox.empty() && jobs.empty()) return;
            std::vector<std::unique_ptr<Job>> in = std::move(inbox);
            inbox.clear();
            lk.unlock();
            std::vector<std::unique_ptr<Job>> out;
            std::string err;
            const int code = guarded([&] {
                if (!pk) pk = std::make_unique<psx::PackedSolve>();
                for (auto& j : in) {
                    auto mb = std::make_unique<psx::PackedSolve::Member>();
                    mb->m = j->M;
                    mb->run = j->run;
                    psx::Model* key = j->M;
                    jobs[key] = std::move(j);
                    pk->add(std::move(mb));
                }
                in.clear();
                for (auto& mb : pk->step(graphs)) {
                    auto it = jobs.find(mb->m);
                    out.push_back(std::move(it->second));
                    jobs.erase(it);
                }
                return PSX_OK;
            });
            if (code != PSX_OK) {
                // a failed batch fails every member
                err = g_error;
                if (pk) pk->release();
                pk.reset();
                for (auto& j : in) out.push_back(std::move(j));
                for (auto& [m, j] : jobs) out.push_back(std::move(j));
                jobs.clear();
                for (auto& j : out) {
                    j->r->res.status = code;
                    j->r->error = err;
                }
            }
            for (auto& j : out) {
                lk.lock();
                packed--;
                lk.unlock();
                complete(std::move(j));
            }
            lk.lock();
        }
    }
};

extern "C" {

int psx_server_create(const psx_server_opts* opts, psx_server* out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("null output");
        *out = nullptr;
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) throw psx::NoDevice("no usable GPU was found");
        auto s = std::make_unique<psx_server_s>();
        HIPCHK(hipGetDevice(&s->device));
        const int slots = opts && opts->max_inflight > 0 ? opts->max_inflight : 4;
        if (opts && opts->pack_max > 0) s->packMax = std::min(opts->pack_max, psx::PACK_MAX);
        if (opts && opts->pack_max_dof > 0) s->packDof = opts->pack_max_dof;
        if (opts && opts->hbm_budget > 0) s->budget = opts->hbm_budget;
        else {
            size_t freeB = 0, total = 0;
            HIPCHK(hipMemGetInfo(&freeB, &total));
            s->budget = int64_t(0.8 * double(freeB));
        }
        for (int k = 0; k < slots; k++) s->workers.emplace_back([p = s.get()] { p->work(); });
        if (s->packMax > 1) s->packer = std::thread([p = s.get()] { p->packLoop(); });
        *out = s.release();
        return PSX_OK;
    });
}

int psx_server_submit(psx_server s, const psx_request* req, int64_t* ticket) {
    return guarded([&] {
        if (!s || !req || !ticket || !req->f || req->nrhs < 1) throw std::invalid_argument("bad arguments");
        const psx_model_desc& d = req->part;
        if (d.nx <= 0 || d.ny <= 0 || d.nz <= 0 || !d.density || !d.bc) throw std::invalid_argument("Structural grid, density and fixture sizes do not match.");
        auto r = std::make_unique<Request>();
        r->nx = d.nx;
        r->ny = d.ny;
        r->nz = d.nz;
        r->nu = d.nu;
        r->coarsest = d.coarsest_max_dof > 0 ? d.coarsest_max_dof : 1100;
        r->maxLevels = d.max_levels;
        r->onDevice = d.build_on_device;
        const int64_t nVox = int64_t(d.nx) * d.ny * d.nz, nNodes = int64_t(d.nx + 1) * (d.ny + 1) * (d.nz + 1);
        r->nDof = 3 * nNodes;
        r->density.assign(d.density, d.density + nVox);
        r->bc.assign(d.bc, d.bc + r->nDof);
        if (d.diag_add) r->diagAdd.assign(d.diag_add, d.diag_add + r->nDof);
        r->nrhs = req->nrhs;
        r->f.assign(req->f, req->f + size_t(req->nrhs) * r->nDof);
        checkFinite(r->f.data(), int64_t(r->f.size()), "Forces");
        if (req->x0) {
            r->x0PerRhs = req->x0_per_rhs != 0;
            r->x0.assign(req->x0, req->x0 + (r->x0PerRhs ? size_t(req->nrhs) : 1) * r->nDof);
            checkFinite(r->x0.data(), int64_t(r->x0.size()), "Displacements");
        }
        r->partId = req->part_id;
        r->opts = defaults(&req->opts);
        r->opts.on_progress = nullptr;
        r->opts.user = nullptr;
        r->estimate = footprint(*r);
        r->submitted = Clock::now();
        std::lock_guard<std::mutex> g(s->mu);
        if (s->stopping) throw std::runtime_error("server is shutting down");
        r->ticket = s->nextTicket++;
        *ticket = r->ticket;
        s->queue.push_back(r.get());
        s->requests[r->ticket] = std::move(r);
        s->wake.notify_one();
        return PSX_OK;
    });
}

int psx_server_wait(psx_server s, int64_t ticket, double* u, psx_result* out) {
    std::unique_ptr<Request> r;
    const int code = guarded([&] {
        if (!s) throw std::invalid_argument("null server");
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->requests.find(ticket);
        if (it == s->requests.end()) throw std::invalid_argument("unknown ticket");
        Request* p = it->second.get();
        s->finished.wait(lk, [&] { return p->done; });
        r = std::move(it->second);
        s->requests.erase(it);
        return PSX_OK;
    });
    if (code != PSX_OK) return code;
    if (out) *out = r->res;
    if (r->res.status != PSX_OK) return fail(r->res.status, r->error);
    if (u) std::copy(r->u.begin(), r->u.end(), u);
    return PSX_OK;
}

int psx_server_pending(psx_server s) {
    if (!s) return 0;
    std::lock_guard<std::mutex> g(s->mu);
    return int(s->queue.size()) + s->running;
}

void psx_server_destroy(psx_server s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stopping = true;
        for (Request* r : s->queue) {
            r->res.status = PSX_ERR_ARGUMENT;
            r->error = "server destroyed before the request ran";
            r->done = true;
        }
        s->queue.clear();
        s->wake.notify_all();
        s->finished.notify_all();
    }
    for (auto& t : s->workers) t.join();
    // the workers have handed over every request they admitted: the pack finishes them, then stops
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->packStop = true;
        s->packWake.notify_all();
    }
    if (s->packer.joinable()) s->packer.join();
    delete s;
}

}  // extern "C"
