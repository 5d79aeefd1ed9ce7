// serve.hip — the request server declared in include/ursa_hip_serve.h.
//
// Host-side runtime: worker slots (one thread and one HIP stream each), a FIFO
// of waiting queries, same-graph batch formation, admission of each device
// pass against an HBM workspace budget carved from one device allocation, and
// a small pool of hashing threads that digest results off the critical path.
// Device work goes through the same solvers as the plain entry points
// (ppr_solve, k_hop_seq, shortest_paths_seq, triangle_count), so a query's
// result does not depend on the pass it was batched into.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/ursa_hip_serve.h"
#include "ppr.hpp"
#include "topology.hpp"
#include "traverse.hpp"

namespace ursa {

void triangle_prepare(ursa_hip_topo* t);
void triangle_count(ursa_hip_topo* t, uint64_t* out_user);
uint64_t triangle_workspace_bytes(uint64_t n);

namespace {

using Clock = std::chrono::steady_clock;

inline uint64_t sm64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// First-fit allocator of [0, total) in 256-byte units, coalescing on release.
class RangeAlloc {
   public:
    void init(uint64_t total) {
        free_.clear();
        if (total) free_[0] = total;
    }
    static uint64_t round(uint64_t b) { return (b + 255) & ~(uint64_t)255; }
    bool alloc(uint64_t size, uint64_t& off) {
        size = round(size);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < size) continue;
            off = it->first;
            const uint64_t rem = it->second - size;
            free_.erase(it);
            if (rem) free_[off + size] = rem;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t size) {
        size = round(size);
        auto it = free_.emplace(off, size).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

   private:
    std::map<uint64_t, uint64_t> free_;
};

struct Rec {
    ursa_hip_query q{};
    ursa_hip_result r{};
    std::vector<char> data;
    Clock::time_point arrival, start, done;
    bool finished = false;
};

template <class T>
uint64_t bytes_of(const DevBuf<T>& b) {
    return b.bytes();
}

uint64_t topo_device_bytes(const ursa_hip_topo* t) {
    uint64_t b = 0;
    for (const DevCsr* c : {&t->out, &t->in, &t->und}) b += bytes_of(c->off) + bytes_of(c->tgt) + bytes_of(c->eid);
    b += bytes_of(t->out_deg);
    b += bytes_of(t->pr_old_of_new) + bytes_of(t->pr_new_of_old) + bytes_of(t->pr_off) + bytes_of(t->pr_tgt) +
         bytes_of(t->pr_eid) + bytes_of(t->pr_outdeg) + bytes_of(t->pr_head) + bytes_of(t->pr_crow) + bytes_of(t->pr_cflag);
    b += bytes_of(t->tc_off) + bytes_of(t->tc_tgt) + bytes_of(t->tc_ioff) + bytes_of(t->tc_isrc) + bytes_of(t->tc_new) +
         bytes_of(t->tc_iwalk);
    for (const auto& kind : t->tc_items)
        for (const auto& it : kind) b += bytes_of(it.start) + bytes_of(it.end) + bytes_of(it.owner);
    b += bytes_of(t->ppr_hrow) + bytes_of(t->ppr_hsoff) + bytes_of(t->ppr_sbeg) + bytes_of(t->ppr_send);
    for (int d = 0; d < 3; ++d) b += bytes_of(t->ms_pheavy[d]) + bytes_of(t->ms_pcbase[d]);
    return b;
}

}  // namespace
}  // namespace ursa

using namespace ursa;

struct ursa_hip_server {
    ursa_hip_server_config cfg{};
    std::vector<ursa_hip_topo*> graphs;
    uint64_t graph_bytes = 0;
    char* ws_base = nullptr;
    uint64_t ws_bytes = 0;
    RangeAlloc ra;

    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_hash;
    std::vector<Rec> recs;
    std::vector<uint64_t> pending;  // waiting queries, arrival order
    std::vector<uint64_t> hashq;
    uint64_t outstanding = 0;
    uint32_t inflight = 0;
    bool stop = false;
    std::vector<std::thread> workers, hashers;
    std::mutex run_mu;

    bool serial() const { return cfg.serial != 0; }

    // Queries joined to the oldest waiting one (caller holds mu). A PPR query
    // on a small graph takes every waiting PPR query on such graphs (any
    // graph) into one cross-graph pass; otherwise a batch is same-kind,
    // same-graph.
    std::vector<uint64_t> form_batch(bool& cross) {
        cross = false;
        std::vector<uint64_t> b{pending.front()};
        if (serial()) return b;
        const ursa_hip_query& h = recs[pending.front()].q;
        if (h.kind == URSA_HIP_Q_PPR && ppr_multi_candidate(graphs[h.graph])) {
            uint64_t nodes = graphs[h.graph]->n;
            for (size_t i = 1; i < pending.size() && b.size() < cfg.max_batch; ++i) {
                const ursa_hip_query& q = recs[pending[i]].q;
                if (q.kind != URSA_HIP_Q_PPR || !ppr_multi_candidate(graphs[q.graph])) continue;
                const uint64_t n = graphs[q.graph]->n;
                if (nodes + n > PPR_MULTI_MAX_NODES) continue;
                b.push_back(pending[i]);
                nodes += n;
            }
            cross = b.size() > 1;
            if (cross) return b;
        }
        if ((h.kind == URSA_HIP_Q_KHOP || h.kind == URSA_HIP_Q_PATH) && traverse_multi_eligible(graphs[h.graph])) {
            // A workgroup per query: a pass takes many (each runs on its own CU).
            const uint32_t tcap = cfg.max_batch <= 1 ? 1u : std::max(cfg.max_batch, cfg.cross_batch);
            for (size_t i = 1; i < pending.size() && b.size() < tcap; ++i) {
                const ursa_hip_query& q = recs[pending[i]].q;
                if ((q.kind == URSA_HIP_Q_KHOP || q.kind == URSA_HIP_Q_PATH) && traverse_multi_eligible(graphs[q.graph]))
                    b.push_back(pending[i]);
            }
            cross = b.size() > 1;
            if (cross) return b;
            b.resize(1);
        }
        uint32_t cap = cfg.max_batch;
        if (h.kind == URSA_HIP_Q_PPR) cap = std::min<uint32_t>(cap, PPR_MAX_LANES);
        for (size_t i = 1; i < pending.size() && b.size() < cap; ++i) {
            const ursa_hip_query& q = recs[pending[i]].q;
            if (q.kind != h.kind || q.graph != h.graph) continue;
            if ((h.kind == URSA_HIP_Q_KHOP || h.kind == URSA_HIP_Q_PATH) && q.direction != h.direction) continue;
            if (h.kind == URSA_HIP_Q_KHOP && q.k != h.k) continue;
            b.push_back(pending[i]);
        }
        return b;
    }

    // The cross-graph traversal tasks of a batch (results not yet filled).
    std::vector<TravTask> trav_tasks(const std::vector<uint64_t>& b) {
        std::vector<TravTask> ts(b.size());
        for (size_t j = 0; j < b.size(); ++j) {
            const ursa_hip_query& q = recs[b[j]].q;
            ts[j].t = graphs[q.graph];
            ts[j].path = q.kind == URSA_HIP_Q_PATH;
            ts[j].seeds.assign(q.seeds, q.seeds + (ts[j].path ? 2 : q.nseeds));
            ts[j].k = q.k;
            ts[j].dir = q.direction;
        }
        return ts;
    }

    uint64_t workspace_need(const std::vector<uint64_t>& b, bool cross) {
        const ursa_hip_query& h = recs[b[0]].q;
        if (cross && h.kind != URSA_HIP_Q_PPR) {
            const std::vector<TravTask> ts = trav_tasks(b);
            std::vector<const TravTask*> ps;
            for (const TravTask& t : ts) ps.push_back(&t);
            return traverse_multi_workspace_bytes(ps) + 4096;
        }
        if (cross) {
            // The cross-graph pass, or (if a graph stops qualifying) one
            // single-lane solve at a time inside the same reservation.
            std::vector<const ursa_hip_topo*> ts;
            uint64_t single = 0;
            for (uint64_t i : b) {
                ts.push_back(graphs[recs[i].q.graph]);
                single = std::max(single, ppr_workspace_bytes(ts.back(), 1, true));
            }
            return std::max(ppr_multi_workspace_bytes(ts.data(), (int)ts.size()), single) + 4096;
        }
        const ursa_hip_topo* t = graphs[h.graph];
        const uint64_t n = t->n, k = b.size();
        switch (h.kind) {
            case URSA_HIP_Q_PPR:
                return ppr_workspace_bytes(t, (int)k, false) + k * n * 8 + 4096;
            case URSA_HIP_Q_KHOP:
                return traverse_workspace_bytes(n);
            case URSA_HIP_Q_PATH:
                return traverse_workspace_bytes(n) + n * 4 + 4096;
            default:  // the pass's n-sized output, then triangle_count's own temporaries
                return n * 8 + 4096 + triangle_workspace_bytes(n);
        }
    }

    void remove_pending(const std::vector<uint64_t>& b) {
        std::vector<uint64_t> rest;
        rest.reserve(pending.size());
        for (uint64_t i : pending)
            if (std::find(b.begin(), b.end(), i) == b.end()) rest.push_back(i);
        pending.swap(rest);
    }

    void finish_locked(const std::vector<uint64_t>& b) {
        for (uint64_t i : b) hashq.push_back(i);
        cv_hash.notify_all();
    }

    void run_served(const std::vector<uint64_t>& b, bool cross, uint32_t slot, uint64_t off, uint64_t need);
    void run_serial(uint64_t i);
    void worker(uint32_t slot);
    void hasher();
};

void ursa_hip_server::run_served(const std::vector<uint64_t>& b, bool cross, uint32_t slot, uint64_t off,
                                 uint64_t need) {
    const auto t_start = Clock::now();
    for (uint64_t i : b) {
        recs[i].start = t_start;
        recs[i].r.batch = (uint32_t)b.size();
        recs[i].r.slot = slot;
    }
    const ursa_hip_query& h = recs[b[0]].q;
    ursa_hip_topo* t = graphs[h.graph];
    const uint64_t n = t->n;
    int status = URSA_HIP_OK;
    try {
        Arena arena;
        arena.reset(ws_base + off, need);
        ScopedArena scope(&arena);
        hipStream_t s = stream();
        const uint64_t k = b.size();
        if (cross && h.kind != URSA_HIP_Q_PPR) {
            std::vector<TravTask> ts = trav_tasks(b);
            std::vector<TravTask*> ps;
            for (TravTask& t : ts) ps.push_back(&t);
            traverse_multi(ps);
            for (uint64_t j = 0; j < k; ++j) {
                Rec& r = recs[b[j]];
                const TravTask& q = ts[j];
                if (q.path) {
                    r.r.len = q.found.size();
                    r.data.resize(q.found.size() * 4);
                    if (!q.found.empty()) std::memcpy(r.data.data(), q.found.data(), q.found.size() * 4);
                } else {
                    const uint64_t cnt = q.dst.size();
                    r.r.len = cnt;
                    r.data.resize(cnt * 8);
                    uint32_t* src = reinterpret_cast<uint32_t*>(r.data.data());
                    uint64_t at = 0;
                    for (size_t i = 0; i < q.seeds.size(); ++i)
                        for (; at < q.ends[i]; ++at) src[at] = q.seeds[i];
                    if (cnt) std::memcpy(r.data.data() + cnt * 4, q.dst.data(), cnt * 4);
                }
            }
        } else if (cross) {
            std::vector<ursa_hip_topo*> ts(k);
            std::vector<PprQuery> qs(k);
            std::vector<double*> outs(k);
            std::vector<uint32_t> iters(k, 0);
            for (uint64_t j = 0; j < k; ++j) {
                Rec& r = recs[b[j]];
                ts[j] = graphs[r.q.graph];
                qs[j] = PprQuery{r.q.seeds, r.q.nseeds, r.q.damping, r.q.max_iter, r.q.tol};
                r.data.resize(ts[j]->n * 8);
                outs[j] = reinterpret_cast<double*>(r.data.data());
            }
            ppr_solve_many(ts.data(), qs.data(), (int)k, outs.data(), iters.data());
            for (uint64_t j = 0; j < k; ++j) {
                recs[b[j]].r.iters = iters[j];
                recs[b[j]].r.len = ts[j]->n;
            }
        } else if (h.kind == URSA_HIP_Q_PPR) {
            std::vector<PprQuery> qs(k);
            for (uint64_t j = 0; j < k; ++j) {
                const ursa_hip_query& q = recs[b[j]].q;
                qs[j] = PprQuery{q.seeds, q.nseeds, q.damping, q.max_iter, q.tol};
            }
            if (n) {
                DevBuf<double> o(n * k);
                std::vector<double*> outs(k);
                for (uint64_t j = 0; j < k; ++j) outs[j] = o.get() + j * n;
                std::vector<uint32_t> iters(k, 0);
                ppr_solve(t, qs.data(), (int)k, outs.data(), iters.data());
                for (uint64_t j = 0; j < k; ++j) {
                    Rec& r = recs[b[j]];
                    r.data.resize(n * 8);
                    copy_to_host(r.data.data(), outs[j], n * 8, s);
                    r.r.iters = iters[j];
                    r.r.len = n;
                }
            }
        } else if (h.kind == URSA_HIP_Q_KHOP) {
            std::vector<uint32_t> seeds;
            for (uint64_t i : b) seeds.insert(seeds.end(), recs[i].q.seeds, recs[i].q.seeds + recs[i].q.nseeds);
            std::vector<uint32_t> dst;
            std::vector<uint64_t> ends;
            k_hop_seq(t, seeds.data(), seeds.size(), h.k, h.direction, dst, ends);
            uint64_t si = 0, at = 0;
            for (uint64_t i : b) {
                Rec& r = recs[i];
                std::vector<uint32_t> src, dd;
                for (uint32_t j = 0; j < r.q.nseeds; ++j, ++si) {
                    const uint64_t e = ends[si];
                    src.insert(src.end(), e - at, r.q.seeds[j]);
                    dd.insert(dd.end(), dst.begin() + at, dst.begin() + e);
                    at = e;
                }
                r.r.len = src.size();
                r.data.resize(src.size() * 8);
                if (!src.empty()) {
                    std::memcpy(r.data.data(), src.data(), src.size() * 4);
                    std::memcpy(r.data.data() + src.size() * 4, dd.data(), dd.size() * 4);
                }
            }
        } else if (h.kind == URSA_HIP_Q_PATH) {
            std::vector<uint32_t> src(k), tgt(k);
            for (uint64_t j = 0; j < k; ++j) {
                src[j] = recs[b[j]].q.seeds[0];
                tgt[j] = recs[b[j]].q.seeds[1];
            }
            std::vector<std::vector<uint32_t>> paths;
            shortest_paths_seq(t, src.data(), tgt.data(), k, h.direction, paths);
            for (uint64_t j = 0; j < k; ++j) {
                Rec& r = recs[b[j]];
                r.r.len = paths[j].size();
                r.data.resize(paths[j].size() * 4);
                if (!paths[j].empty()) std::memcpy(r.data.data(), paths[j].data(), paths[j].size() * 4);
            }
        } else {
            if (n) {
                DevBuf<uint64_t> o(n);
                triangle_count(t, o.get());
                Rec& r0 = recs[b[0]];
                r0.data.resize(n * 8);
                copy_to_host(r0.data.data(), o.get(), n * 8, s);
                for (uint64_t j = 1; j < k; ++j) recs[b[j]].data = r0.data;
            }
            for (uint64_t i : b) recs[i].r.len = n;
        }
    } catch (const Error& e) {
        status = e.code;
    } catch (const std::bad_alloc&) {
        status = URSA_HIP_ENOMEM;
    } catch (...) {
        status = URSA_HIP_EINTERNAL;
    }
    const auto t_done = Clock::now();
    for (uint64_t i : b) {
        recs[i].done = t_done;
        recs[i].r.status = status;
        if (status != URSA_HIP_OK) recs[i].data.clear();
    }
}

void ursa_hip_server::run_serial(uint64_t i) {
    Rec& r = recs[i];
    r.start = Clock::now();
    r.r.batch = 1;
    r.r.slot = 0;
    const ursa_hip_query& q = r.q;
    ursa_hip_topo* t = graphs[q.graph];
    const uint64_t n = t->n;
    int rc = URSA_HIP_OK;
    if (q.kind == URSA_HIP_Q_PPR) {
        r.data.resize(n * 8);
        uint32_t it = 0;
        rc = ursa_hip_ppr(t, q.seeds, q.nseeds, q.damping, q.max_iter, q.tol, reinterpret_cast<double*>(r.data.data()),
                          &it);
        r.r.iters = it;
        r.r.len = n;
    } else if (q.kind == URSA_HIP_Q_KHOP) {
        ursa_hip_pairs* p = nullptr;
        rc = ursa_hip_k_hop(t, q.seeds, q.nseeds, q.k, q.direction, &p);
        uint64_t cnt = 0;
        if (rc == URSA_HIP_OK) rc = ursa_hip_pairs_size(p, &cnt);
        if (rc == URSA_HIP_OK) {
            r.data.resize(cnt * 8);
            uint32_t* w = reinterpret_cast<uint32_t*>(r.data.data());
            rc = ursa_hip_pairs_export(p, w, w + cnt);
            r.r.len = cnt;
        }
        if (p) ursa_hip_pairs_destroy(p);
    } else if (q.kind == URSA_HIP_Q_PATH) {
        static thread_local std::vector<uint32_t> buf;
        if (buf.size() < n) buf.resize(n);
        uint64_t len = 0;
        rc = ursa_hip_shortest_path(t, q.seeds[0], q.seeds[1], q.direction, buf.data(), buf.size(), &len);
        r.r.len = len;
        r.data.resize(len * 4);
        if (len) std::memcpy(r.data.data(), buf.data(), len * 4);
    } else {
        r.data.resize(n * 8);
        rc = ursa_hip_triangle_count(t, reinterpret_cast<uint64_t*>(r.data.data()));
        r.r.len = n;
    }
    r.done = Clock::now();
    r.r.status = rc;
    if (rc != URSA_HIP_OK) r.data.clear();
}

void ursa_hip_server::worker(uint32_t slot) {
    try {
        ensure_device();
        (void)stream();
    } catch (...) {
    }
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        cv_work.wait(lk, [&] { return stop || !pending.empty(); });
        if (stop) return;
        bool cross = false;
        std::vector<uint64_t> b = form_batch(cross);
        uint64_t need = 0, off = 0;
        if (!serial()) {
            need = workspace_need(b, cross);
            if (!ra.alloc(need, off)) {
                if (inflight == 0) {
                    // Larger than the whole budget: it can never run.
                    remove_pending(b);
                    const auto now = Clock::now();
                    for (uint64_t i : b) {
                        recs[i].start = recs[i].done = now;
                        recs[i].r.status = URSA_HIP_ENOMEM;
                        recs[i].r.batch = (uint32_t)b.size();
                    }
                    finish_locked(b);
                    continue;
                }
                cv_work.wait(lk);  // a running pass will release memory
                continue;
            }
        }
        remove_pending(b);
        ++inflight;
        lk.unlock();
        if (serial())
            run_serial(b[0]);
        else
            run_served(b, cross, slot, off, need);
        lk.lock();
        if (!serial()) ra.release(off, need);
        --inflight;
        finish_locked(b);
        cv_work.notify_all();
    }
}

void ursa_hip_server::hasher() {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        cv_hash.wait(lk, [&] { return stop || !hashq.empty(); });
        if (hashq.empty() && stop) return;
        const uint64_t i = hashq.back();
        hashq.pop_back();
        lk.unlock();
        Rec& r = recs[i];
        r.r.digest = ursa_hip_digest(r.data.data(), r.data.size());
        if (!cfg.keep_results) std::vector<char>().swap(r.data);
        lk.lock();
        r.finished = true;
        --outstanding;
        cv_done.notify_all();
    }
}

extern "C" {

uint64_t ursa_hip_digest(const void* data, uint64_t bytes) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    const uint64_t K = 0xD1B54A32D192ED03ull;
    uint64_t h = sm64(bytes);
    const uint64_t nw = bytes / 8;
    for (uint64_t i = 0; i < nw; ++i) {
        uint64_t w;
        std::memcpy(&w, p + 8 * i, 8);
        h += sm64(w + i * K);
    }
    if (bytes % 8) {
        uint64_t w = 0;
        std::memcpy(&w, p + 8 * nw, bytes % 8);
        h += sm64(w + nw * K);
    }
    return h;
}

int ursa_hip_server_create(const ursa_hip_server_config* cfg, ursa_hip_server** out) {
    return guard([&] {
        URSA_REQUIRE(out != nullptr, URSA_HIP_EINVAL, "out is NULL");
        *out = nullptr;
        ensure_device();
        auto* S = new ursa_hip_server;
        try {
            if (cfg) S->cfg = *cfg;
            if (S->cfg.serial) {
                S->cfg.slots = 1;
                S->cfg.max_batch = 1;
            }
            if (S->cfg.slots == 0) S->cfg.slots = 8;
            if (S->cfg.max_batch == 0) S->cfg.max_batch = 16;
            if (S->cfg.cross_batch == 0) S->cfg.cross_batch = 256;
            URSA_REQUIRE(S->cfg.slots <= 64, URSA_HIP_EINVAL, "at most 64 slots");
            if (!S->serial()) {
                uint64_t ws = S->cfg.workspace_bytes;
                if (ws == 0) {
                    size_t fr = 0, tot = 0;
                    HIP_OK(hipMemGetInfo(&fr, &tot));
                    ws = std::min<uint64_t>(fr / 4, 32ull << 30);
                }
                ws = RangeAlloc::round(ws);
                S->ws_base = static_cast<char*>(dev_alloc(ws));
                S->ws_bytes = ws;
                arena_register(S->ws_base, ws);
                S->ra.init(ws);
            }
            for (uint32_t i = 0; i < S->cfg.slots; ++i) S->workers.emplace_back([S, i] { S->worker(i); });
            for (int i = 0; i < 4; ++i) S->hashers.emplace_back([S] { S->hasher(); });
        } catch (...) {
            ursa_hip_server_destroy(S);
            throw;
        }
        *out = S;
    });
}

int ursa_hip_server_add_graph(ursa_hip_server* S, ursa_hip_topo* t, uint32_t* graph_id) {
    return guard([&] {
        URSA_REQUIRE(S != nullptr && t != nullptr, URSA_HIP_EINVAL, "server/topology is NULL");
        std::lock_guard<std::mutex> rl(S->run_mu);
        {
            NoArena keep_out_of_arena;
            hipStream_t s = stream();
            t->incoming(s);
            t->undirected(s);
            triangle_prepare(t);
            ppr_prepare(t, s);
            HIP_OK(hipStreamSynchronize(s));
        }
        const uint64_t b = topo_device_bytes(t);
        std::lock_guard<std::mutex> lk(S->mu);
        URSA_REQUIRE(S->cfg.graph_bytes == 0 || S->graph_bytes + b <= S->cfg.graph_bytes, URSA_HIP_ENOMEM,
                     "graph does not fit the resident-graph budget");
        S->graph_bytes += b;
        S->graphs.push_back(t);
        if (graph_id) *graph_id = (uint32_t)(S->graphs.size() - 1);
    });
}

int ursa_hip_server_graph_bytes(const ursa_hip_server* S, uint64_t* bytes) {
    return guard([&] {
        URSA_REQUIRE(S != nullptr && bytes != nullptr, URSA_HIP_EINVAL, "NULL argument");
        *bytes = S->graph_bytes;
    });
}

int ursa_hip_server_run(ursa_hip_server* S, const ursa_hip_query* qs, uint64_t nq, int timed_arrivals,
                        ursa_hip_result* res) {
    return guard([&] {
        URSA_REQUIRE(S != nullptr, URSA_HIP_EINVAL, "server is NULL");
        URSA_REQUIRE(nq == 0 || (qs != nullptr && res != nullptr), URSA_HIP_EINVAL, "queries/results is NULL");
        std::lock_guard<std::mutex> rl(S->run_mu);
        std::vector<uint64_t> order(nq);
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->recs.clear();
            S->recs.resize(nq);
            for (uint64_t i = 0; i < nq; ++i) {
                Rec& r = S->recs[i];
                r.q = qs[i];
                const ursa_hip_query& q = r.q;
                bool ok = q.graph < S->graphs.size() && q.kind <= URSA_HIP_Q_TRIANGLES;
                if (ok && (q.kind == URSA_HIP_Q_PPR || q.kind == URSA_HIP_Q_KHOP))
                    ok = q.nseeds >= 1 && q.nseeds <= 16;
                if (ok && q.kind == URSA_HIP_Q_PATH) ok = q.nseeds == 2;
                if (ok && (q.kind == URSA_HIP_Q_KHOP || q.kind == URSA_HIP_Q_PATH))
                    ok = q.direction >= 0 && q.direction <= 2;
                if (ok && q.kind == URSA_HIP_Q_PPR) ok = q.damping >= 0.0 && q.damping <= 1.0;
                r.r.status = ok ? URSA_HIP_OK : URSA_HIP_EINVAL;
                r.finished = !ok;
                order[i] = i;
            }
            S->outstanding = 0;
            for (auto& r : S->recs)
                if (!r.finished) ++S->outstanding;
        }
        if (timed_arrivals)
            std::stable_sort(order.begin(), order.end(),
                             [&](uint64_t a, uint64_t b) { return qs[a].arrival_s < qs[b].arrival_s; });
        const auto t0 = Clock::now();
        for (uint64_t i : order) {
            Rec& r = S->recs[i];
            auto at = t0;
            if (timed_arrivals) {
                at = t0 + std::chrono::duration_cast<Clock::duration>(
                              std::chrono::duration<double>(std::max(0.0, qs[i].arrival_s)));
                std::this_thread::sleep_until(at);
            }
            std::lock_guard<std::mutex> lk(S->mu);
            r.arrival = at;
            if (r.finished) {
                r.start = r.done = at;
                continue;
            }
            S->pending.push_back(i);
            S->cv_work.notify_one();
        }
        std::unique_lock<std::mutex> lk(S->mu);
        S->cv_done.wait(lk, [&] { return S->outstanding == 0; });
        auto sec = [&](Clock::time_point p) { return std::chrono::duration<double>(p - t0).count(); };
        for (uint64_t i = 0; i < nq; ++i) {
            Rec& r = S->recs[i];
            r.r.arrival_s = sec(r.arrival);
            r.r.start_s = sec(r.start);
            r.r.done_s = sec(r.done);
            res[i] = r.r;
        }
    });
}

int ursa_hip_server_result_data(const ursa_hip_server* S, uint64_t index, void* dst, uint64_t cap,
                                uint64_t* bytes) {
    return guard([&] {
        URSA_REQUIRE(S != nullptr, URSA_HIP_EINVAL, "server is NULL");
        URSA_REQUIRE(index < S->recs.size(), URSA_HIP_EINVAL, "query index out of range");
        URSA_REQUIRE(S->cfg.keep_results, URSA_HIP_EINVAL, "results are kept only with keep_results");
        const auto& d = S->recs[index].data;
        if (bytes) *bytes = d.size();
        if (!dst) return;
        URSA_REQUIRE(cap >= d.size(), URSA_HIP_ERANGE, "destination too small");
        if (!d.empty()) std::memcpy(dst, d.data(), d.size());
    });
}

int ursa_hip_server_destroy(ursa_hip_server* S) {
    if (!S) return URSA_HIP_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    S->cv_hash.notify_all();
    for (auto& th : S->workers) th.join();
    for (auto& th : S->hashers) th.join();
    int rc = URSA_HIP_OK;
    if (S->ws_base) {
        rc = guard([&] {
            HIP_OK(hipDeviceSynchronize());
            arena_unregister(S->ws_base);
            dev_free(S->ws_base);
        });
    }
    delete S;
    return rc;
}

}  // extern "C"
