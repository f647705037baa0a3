// Packed solves: the conjugate-gradient runs of several models advance together. One batch of
// iterations of every member is recorded (pack.hpp), and the records are merged greedily into
// shared launches: at each point the kernel at the head of the most members' sequences is launched
// once for all of them (ties go to the longest remaining sequence), so every member's launches stay
// in their own order and the merged sequence is replayed as one HIP graph per batch. Each block of
// a shared launch runs the kernel body with its own member's arguments and block index, and every
// kernel returns at once for an instance that has stopped, so each member's solution is bit for bit
// the one its lone solve (Model::solveMulti) gives. Small parts, whose kernels cannot fill the
// device, then cost one launch per step for the whole pack instead of one each.
#pragma once

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

#include "model.hpp"
#include "pack.hpp"

namespace psx {

class PackedSolve {
public:
    struct Member {
        Model* m = nullptr;
        Model::Run run;
        void* tag = nullptr;  // the caller's
    };

    PackedSolve() { HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking)); }
    ~PackedSolve() {
        dropGraph();
        if (stream) {
            if (table) (void)hipFreeAsync(table, stream);
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
        for (auto& mb : members) mb->m->resetRun();
    }
    PackedSolve(const PackedSolve&) = delete;
    PackedSolve& operator=(const PackedSolve&) = delete;

    size_t size() const { return members.size(); }
    // launches of the last merged batch and the launches its members' batches hold on their own
    size_t mergedLaunches() const { return launches.size(); }
    size_t ownLaunches() const { return recorded; }

    // a member whose run has begun (Model::begin returned true)
    void add(std::unique_ptr<Member> mb) {
        if (members.size() >= size_t(PACK_MAX)) throw std::invalid_argument("a pack holds at most 64 models");
        for (const auto& o : members)
            if (o->m == mb->m) throw std::invalid_argument("a model can be in a pack once");
        members.push_back(std::move(mb));
        dirty = true;
    }

    // one batch for every member; the members whose runs stopped are ended (Model::end) and returned
    std::vector<std::unique_ptr<Member>> step(bool graphs) {
        std::vector<std::unique_ptr<Member>> done;
        if (members.empty()) return done;
        const auto t0 = std::chrono::steady_clock::now();
        const bool rebuilt = dirty;
        if (dirty) rebuild(graphs);
        const auto t1 = std::chrono::steady_clock::now();
        if (graph) HIPCHK(hipGraphLaunch(graph, stream));
        else
            for (const Launch& l : launches) l.fn(l.head, l.args, l.blocks, stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(stream));
        if (logging()) {
            const auto t2 = std::chrono::steady_clock::now();
            std::fprintf(stderr, "pack: %zu members, %zu launches merged into %zu%s, rebuild %.2f ms, batch %.2f ms\n", members.size(), recorded,
                         launches.size(), rebuilt ? " (rebuilt)" : "", std::chrono::duration<double, std::milli>(t1 - t0).count(),
                         std::chrono::duration<double, std::milli>(t2 - t1).count());
        }
        for (size_t i = 0; i < members.size();) {
            Member& mb = *members[i];
            mb.m->readState();
            if (mb.m->poll(mb.run, false)) {
                mb.m->end(mb.run);
                mb.m->resetRun();
                done.push_back(std::move(members[i]));
                members.erase(members.begin() + std::ptrdiff_t(i));
                dirty = true;
            } else {
                i++;
            }
        }
        return done;
    }

    // every member, unfinished (after an error): their models' run settings are reset
    std::vector<std::unique_ptr<Member>> release() {
        dropGraph();
        for (auto& mb : members) mb->m->resetRun();
        std::vector<std::unique_ptr<Member>> out = std::move(members);
        members.clear();
        dirty = true;
        return out;
    }

    // PSX_PACK_LOG=1: one line per batch on stderr (members, launches before and after merging, times)
    static bool logging() {
        static const bool on = [] {
            const char* e = std::getenv("PSX_PACK_LOG");
            return e && e[0] == '1';
        }();
        return on;
    }

private:
    struct Launch {
        PackLaunchFn fn;
        PackHead head;
        const void* args;
        uint32_t blocks;
    };
    hipStream_t stream = nullptr;
    std::vector<std::unique_ptr<Member>> members;
    bool dirty = true;
    std::vector<Launch> launches;
    size_t recorded = 0;
    void* table = nullptr;
    size_t tableBytes = 0;
    hipGraphExec_t graph = nullptr;

    void dropGraph() {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
    }

    void rebuild(bool graphs) {
        dropGraph();
        launches.clear();
        // every member's batch, recorded in member order
        PackRecorder rec;
        const size_t M = members.size();
        std::vector<size_t> head(M), stop(M);
        {
            struct Unset {
                ~Unset() { packRecorder() = nullptr; }
            } unset;
            packRecorder() = &rec;
            for (size_t i = 0; i < M; i++) {
                head[i] = rec.recs.size();
                members[i]->m->enqueueBatch(members[i]->run);
                stop[i] = rec.recs.size();
            }
        }
        recorded = rec.recs.size();
        // greedy merge: the kernel heading the most sequences (the longest remaining one on a tie)
        std::vector<std::vector<size_t>> groups;
        for (;;) {
            const void* best = nullptr;
            size_t bestCount = 0, bestRemain = 0;
            for (size_t i = 0; i < M; i++) {
                if (head[i] == stop[i]) continue;
                const void* key = rec.recs[head[i]].kernel;
                size_t count = 0;
                for (size_t j = 0; j < M; j++) count += head[j] < stop[j] && rec.recs[head[j]].kernel == key ? 1 : 0;
                const size_t remain = stop[i] - head[i];
                if (count > bestCount || (count == bestCount && remain > bestRemain)) {
                    best = key;
                    bestCount = count;
                    bestRemain = remain;
                }
            }
            if (!best) break;
            std::vector<size_t> g;
            for (size_t j = 0; j < M; j++)
                if (head[j] < stop[j] && rec.recs[head[j]].kernel == best) g.push_back(head[j]++);
            groups.push_back(std::move(g));
        }
        // the device table: per shared launch the array of its launches' parameter blocks, 16-byte
        // aligned; the block ranges go into the launch's head (a kernel argument)
        std::vector<unsigned char> host;
        std::vector<size_t> argOff;
        argOff.reserve(groups.size());
        for (const auto& g : groups) {
            Launch l{rec.recs[g[0]].launch, PackHead{}, nullptr, 0};
            l.head.n = int(g.size());
            uint32_t first = 0;
            for (size_t q = 0; q < g.size(); q++) {
                const PackRec& pr = rec.recs[g[q]];
                l.head.first[q] = first;
                l.head.gx[q] = pr.gx;
                const uint64_t nb = uint64_t(pr.gx) * pr.gy;
                if (uint64_t(first) + nb > 0x7fffffffu) throw std::invalid_argument("a packed launch exceeds 2^31 blocks");
                first += uint32_t(nb);
            }
            l.head.first[g.size()] = first;
            l.blocks = first;
            host.resize((host.size() + 15) / 16 * 16);
            argOff.push_back(host.size());
            const uint32_t bytes = rec.recs[g[0]].argBytes;
            host.resize(host.size() + size_t(bytes) * g.size());
            for (size_t q = 0; q < g.size(); q++) {
                const PackRec& pr = rec.recs[g[q]];
                if (pr.argBytes != bytes) throw std::logic_error("packed launch of one kernel with different parameter blocks");
                std::memcpy(host.data() + argOff.back() + q * bytes, rec.pool.data() + pr.argOff, bytes);
            }
            launches.push_back(l);
        }
        if (host.size() > tableBytes) {
            if (table) HIPCHK(hipFreeAsync(table, stream));
            table = nullptr;
            tableBytes = 0;
            HIPCHK(hipMallocAsync(&table, host.size(), stream));
            tableBytes = host.size();
        }
        HIPCHK(hipMemcpyAsync(table, host.data(), host.size(), hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));  // host is a local
        for (size_t gi = 0; gi < launches.size(); gi++) launches[gi].args = static_cast<const unsigned char*>(table) + argOff[gi];
        dirty = false;
        if (!graphs) return;
        HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
        for (const Launch& l : launches) l.fn(l.head, l.args, l.blocks, stream);
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamEndCapture(stream, &g));
        const hipError_t e = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            graph = nullptr;
            throw DeviceError(std::string("hipGraphInstantiate failed: ") + hipGetErrorString(e));
        }
    }
};

}  // namespace psx
