// Continuous-batching serving loop for dam-break requests.
//
// Requests arrive at given wall-clock offsets. A request is admitted (FIFO)
// once it has arrived, a slot is free and its device footprint fits the HBM
// budget next to the requests already running. Device memory is one arena,
// allocated at entry to the size of the budget; each admitted request's
// buffers are ranges of it, returned when it retires (no hipMalloc or
// hipFree while requests run). The footprint is exact: the grid each
// request's upload will build is computed on the host before the clock
// starts.
//
// Mixed-precision requests share grouped step kernels: the requests of one
// dimension, set of pair-term switches and size class form a group on one
// stream, and
// each stage of their step is one launch over all of them (group.hip),
// replayed from a two-step hipGraph that is re-captured when the group's
// membership changes. The scheduler queues a chunk of steps per group; a
// chunk ends at the nearest completion in the group, so a finished request
// leaves at once and the next arrival joins at the next chunk. Other
// requests (FP32, FP64, Shepard filter on) each run on a stream of their
// own, replaying their own step graph. A request's results do not depend on
// what runs beside it: its kernels read and write only its own buffers and
// every kernel is deterministic.
#include <chrono>
#include <cmath>
#include <deque>
#include <memory>
#include <thread>

#include "common.hpp"
#include "group.hpp"
#include "sim.hpp"

namespace {

using clk = std::chrono::steady_clock;
using namespace wcsph;

// Groups by dimension, v_sig rule, beta term and size class. A group steps
// in lockstep at the pace of its largest member, so a small request grouped
// with a large one would take the large one's time per step; size classes
// (particle counts below 150k, below 600k, above) keep members' step costs
// alike, and the groups' streams overlap on the device.
constexpr int kSizeClasses = 3;
constexpr int kGroupKeys = 8 * kSizeClasses;

int size_class(int64_t n) { return n < 150000 ? 0 : (n < 600000 ? 1 : 2); }

struct HostState {
    std::vector<double> pos, vel, dens, mass, h;
    std::vector<int32_t> type;
};

struct Job {
    int idx;
    wcsph_sim* sim;
    int64_t bytes;
    int32_t done_steps;
    int slot;
    int group;  // group key, or -1 for a stream of its own
};

void make_state(const wcsph_request* r, const wcsph_params& p, HostState& st, int64_t& n) {
    int64_t nf = 0, nw = 0;
    if (int rc = wcsph_dam_break_count(&r->geom, &nf, &nw)) throw Error{rc, wcsph_last_error()};
    n = nf + nw;
    const int D = r->geom.dim;
    st.pos.assign((size_t)D * n, 0.0);
    st.vel.assign((size_t)D * n, 0.0);
    st.dens.assign(n, 0.0);
    st.mass.assign(n, 0.0);
    st.h.assign(n, 0.0);
    st.type.assign(n, 0);
    if (int rc = wcsph_dam_break_fill(&r->geom, &p, st.pos.data(), st.vel.data(), st.dens.data(), st.mass.data(),
                                      st.h.data(), st.type.data()))
        throw Error{rc, wcsph_last_error()};
}

int group_key(const wcsph_params& p, int64_t n) {
    return (((p.dim - 2) * 2 + (p.vsig_mode ? 1 : 0)) * 2 + (p.beta != 0.0 ? 1 : 0)) * kSizeClasses + size_class(n);
}

// Device and pinned resources of one wcsph_serve call, released on every exit.
struct Resources {
    char* arena_mem = nullptr;
    char* tables = nullptr;
    volatile uint32_t* flags = nullptr;
    std::vector<hipStream_t> solo, grouped;
    StepGroup* groups[kGroupKeys] = {};
    std::vector<wcsph_sim*> live;

    ~Resources() {
        for (wcsph_sim* s : live) (void)wcsph_sim_destroy(s);
        for (StepGroup* g : groups) group_destroy(g);
        for (hipStream_t s : solo) (void)hipStreamDestroy(s);
        for (hipStream_t s : grouped) (void)hipStreamDestroy(s);
        if (flags) (void)hipHostFree((void*)flags);
        if (tables) (void)hipFree(tables);
        if (arena_mem) (void)hipFree(arena_mem);
    }
};

}  // namespace

extern "C" int wcsph_serve(const wcsph_serve_config* cfg, const wcsph_request* reqs, int32_t nreq,
                           wcsph_request_result* out) {
    return wcsph::guarded([&] {
        WCSPH_REQUIRE(cfg && (nreq == 0 || (reqs && out)), WCSPH_EINVAL, "null argument");
        WCSPH_REQUIRE(cfg->max_slots > 0 && cfg->chunk_steps > 0 && cfg->hbm_budget_bytes > 0, WCSPH_EINVAL,
                      "max_slots, chunk_steps and hbm_budget_bytes must be positive");
        WCSPH_REQUIRE(cfg->max_slots <= kGroupMax, WCSPH_EINVAL, "max_slots above 16");
        // the clients' initial states, the precision each request runs at and
        // its exact device footprint, before the clock starts
        std::vector<HostState> states(nreq);
        std::vector<int64_t> ns(nreq), bytes(nreq);
        std::vector<wcsph_params> params(nreq);
        std::vector<wcsph_sim_config> confs(nreq);
        int device = 0;
        for (int k = 0; k < nreq; ++k) {
            params[k] = reqs[k].params;
            WCSPH_REQUIRE(k == 0 || params[k].device == device, WCSPH_EINVAL, "requests on different devices");
            device = params[k].device;
            make_state(&reqs[k], params[k], states[k], ns[k]);
            wcsph_sim_config& sc = confs[k];
            sc = wcsph_sim_config{};
            sc.precision = reqs[k].precision;
            int32_t nc[3] = {1, 1, 1};
            const int64_t ncell = sim_grid_cells(&params[k], &sc, ns[k], states[k].pos.data(), states[k].h.data(), nc);
            // a mixed request whose grid is beyond the FP16 cell index runs in FP32
            if (sc.precision == WCSPH_PREC_MIXED && nc[0] > 2048) sc.precision = WCSPH_PREC_FP32;
            bytes[k] = sim_footprint(reqs[k].geom.dim, ns[k], sc.precision, 128, ncell);
            WCSPH_REQUIRE(bytes[k] <= cfg->hbm_budget_bytes, WCSPH_ENOMEM,
                          "a request is larger than the whole HBM budget");
            out[k] = wcsph_request_result{};
            out[k].arrival_s = reqs[k].arrival_s;
            out[k].n = ns[k];
            out[k].precision_used = sc.precision;
        }
        WCSPH_HIP(hipSetDevice(device));
        Resources R;
        WCSPH_HIP(hipMalloc((void**)&R.arena_mem, (size_t)cfg->hbm_budget_bytes));
        wcsph_arena arena;
        arena.reset(R.arena_mem, (size_t)cfg->hbm_budget_bytes);
        const size_t tab = group_table_bytes(3);
        WCSPH_HIP(hipMalloc((void**)&R.tables, tab * kGroupKeys));
        WCSPH_HIP(hipHostMalloc((void**)&R.flags, sizeof(uint32_t) * 2 * cfg->max_slots, hipHostMallocDefault));
        R.solo.assign(cfg->max_slots, nullptr);
        for (auto& s : R.solo) WCSPH_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        R.grouped.assign(kGroupKeys, nullptr);
        for (int g = 0; g < kGroupKeys; ++g) {
            WCSPH_HIP(hipStreamCreateWithFlags(&R.grouped[g], hipStreamNonBlocking));
            const int sw = g / kSizeClasses;  // dimension, v_sig rule, beta term
            R.groups[g] = group_create(2 + sw / 4, (sw / 2) % 2, sw % 2, R.grouped[g], R.tables + tab * g);
        }
        std::vector<int> free_slots;
        for (int s = cfg->max_slots - 1; s >= 0; --s) free_slots.push_back(s);

        std::vector<int> order(nreq);
        for (int k = 0; k < nreq; ++k) order[k] = k;
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
        std::deque<int> pending(order.begin(), order.end());
        std::vector<Job> active;
        int64_t used = 0;
        const auto t0 = clk::now();
        auto now = [&] { return std::chrono::duration<double>(clk::now() - t0).count(); };

        auto retire = [&](size_t a, int rc) {
            Job& J = active[a];
            wcsph_request_result& o = out[J.idx];
            const wcsph_request& r = reqs[J.idx];
            if (!rc) {
                wcsph_sim_stats s{};
                rc = wcsph_sim_get_stats(J.sim, &s);
                double d[10] = {0};
                const double band = 2.0 * r.geom.dx;
                if (!rc) rc = wcsph_sim_diagnostics(J.sim, band, band, r.geom.tank_l - band, d);
                o.steps = s.step;
                o.pairs = s.pairs_total;
                o.sim_time = s.time;
                o.front = d[7];
                o.column = d[8];
                o.kinetic = d[4];
            }
            o.status = rc;
            o.end_s = now();
            R.live.erase(std::find(R.live.begin(), R.live.end(), J.sim));
            wcsph_sim_destroy(J.sim);
            used -= J.bytes;
            free_slots.push_back(J.slot);
            active.erase(active.begin() + (std::ptrdiff_t)a);
        };

        while (!pending.empty() || !active.empty()) {
            bool busy = false;
            // admission
            while (!pending.empty() && reqs[pending.front()].arrival_s <= now() && !free_slots.empty() &&
                   used + bytes[pending.front()] <= cfg->hbm_budget_bytes) {
                const int k = pending.front();
                const wcsph_sim_config& sc = confs[k];
                const bool grouped = sc.precision == WCSPH_PREC_MIXED && params[k].shepard_interval == 0;
                const int slot = free_slots.back();
                const int gk = grouped ? group_key(params[k], ns[k]) : -1;
                hipStream_t st = grouped ? R.grouped[gk] : R.solo[slot];
                wcsph_sim* sim = nullptr;
                try {
                    sim = sim_create_in(&params[k], &sc, ns[k], &arena, st, R.flags + 2 * slot);
                } catch (const Error& e) {
                    // the arena is fragmented: wait for a retirement
                    if (e.code == WCSPH_ENOMEM && !active.empty()) break;
                    throw;
                }
                pending.pop_front();
                free_slots.pop_back();
                R.live.push_back(sim);
                used += bytes[k];
                active.push_back(Job{k, sim, bytes[k], 0, slot, gk});
                const HostState& hs = states[k];
                int rc = wcsph_sim_upload(sim, hs.pos.data(), hs.vel.data(), hs.dens.data(), hs.mass.data(),
                                          hs.h.data(), hs.type.data());
                if (!rc) rc = wcsph_sim_init(sim);
                out[k].start_s = now();
                if (!rc && grouped && !group_eligible(sim)) rc = WCSPH_ESTATE;
                if (rc) retire(active.size() - 1, rc);
                busy = true;
            }
            // grouped requests: one chunk per group in flight
            for (int g = 0; g < kGroupKeys; ++g) {
                const hipError_t q = hipStreamQuery(R.grouped[g]);
                if (q == hipErrorNotReady) continue;
                WCSPH_HIP(q);
                std::vector<wcsph_sim*> mem;
                int32_t chunk = cfg->chunk_steps;
                for (size_t a = 0; a < active.size();) {
                    Job& J = active[a];
                    if (J.group != g) {
                        ++a;
                        continue;
                    }
                    int rc = 0;
                    try {
                        sim_check_failed(J.sim);
                    } catch (const Error& e) {
                        rc = e.code;
                    }
                    if (rc || J.done_steps >= reqs[J.idx].steps) {
                        retire(a, rc);
                        busy = true;
                        continue;
                    }
                    mem.push_back(J.sim);
                    chunk = std::min(chunk, reqs[J.idx].steps - J.done_steps);
                    ++a;
                }
                group_set_members(R.groups[g], mem);
                if (mem.empty()) continue;
                if (chunk >= 2) {
                    group_run_pairs(R.groups[g], chunk / 2);
                    for (Job& J : active)
                        if (J.group == g) J.done_steps += chunk / 2 * 2;
                } else {
                    // a chunk of one step (chunk_steps 1, or a member one step
                    // from its end): every member takes one single step
                    for (Job& J : active)
                        if (J.group == g) {
                            group_run_single(R.groups[g], J.sim);
                            J.done_steps += 1;
                        }
                }
                busy = true;
            }
            // requests on streams of their own
            for (size_t a = 0; a < active.size();) {
                Job& J = active[a];
                if (J.group >= 0) {
                    ++a;
                    continue;
                }
                const wcsph_request& r = reqs[J.idx];
                int32_t idle = 0;
                int rc = wcsph_sim_query(J.sim, &idle);
                if (!rc && idle && J.done_steps < r.steps) {
                    const int32_t c = std::min(cfg->chunk_steps, r.steps - J.done_steps);
                    rc = wcsph_sim_step(J.sim, c);
                    J.done_steps += c;
                    busy = true;
                } else if (idle || rc) {
                    retire(a, rc);
                    busy = true;
                    continue;
                }
                ++a;
            }
            if (!busy) std::this_thread::sleep_for(std::chrono::microseconds(50));
        }
    });
}
