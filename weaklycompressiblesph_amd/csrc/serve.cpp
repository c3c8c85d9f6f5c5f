// Continuous-batching serving loop for dam-break requests.
//
// Requests arrive at given wall-clock offsets. A request is admitted (FIFO)
// once it has arrived, a slot is free and its device footprint fits the HBM
// budget next to the requests already running. Each admitted request is a
// device-resident simulation on its own HIP stream whose steps are replayed
// from a captured two-step hipGraph; the scheduler keeps one chunk of steps
// in flight per request, so the requests' kernels overlap on the device. A
// finished request releases its slot and memory to the next arrival at once.
// Requests share no device buffers and every kernel is deterministic, so a
// request's result does not depend on what else runs beside it.
#include <chrono>
#include <cmath>
#include <deque>
#include <thread>

#include "common.hpp"

namespace {

using clk = std::chrono::steady_clock;

struct Active {
    int idx;
    wcsph_sim* sim;
    int64_t bytes;
    int32_t done_steps;
    bool chunk_in_flight;
};

struct HostState {
    std::vector<double> pos, vel, dens, mass, h;
    std::vector<int32_t> type;
};

void make_state(const wcsph_request* r, const wcsph_params& p, HostState& st, int64_t& n) {
    int64_t nf = 0, nw = 0;
    if (int rc = wcsph_dam_break_count(&r->geom, &nf, &nw)) throw wcsph::Error{rc, wcsph_last_error()};
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
        throw wcsph::Error{rc, wcsph_last_error()};
}

}  // namespace

extern "C" int wcsph_serve(const wcsph_serve_config* cfg, const wcsph_request* reqs, int32_t nreq,
                           wcsph_request_result* out) {
    return wcsph::guarded([&] {
        WCSPH_REQUIRE(cfg && (nreq == 0 || (reqs && out)), WCSPH_EINVAL, "null argument");
        WCSPH_REQUIRE(cfg->max_slots > 0 && cfg->chunk_steps > 0 && cfg->hbm_budget_bytes > 0, WCSPH_EINVAL,
                      "max_slots, chunk_steps and hbm_budget_bytes must be positive");
        // the clients' initial states exist before the clock starts
        std::vector<HostState> states(nreq);
        std::vector<int64_t> ns(nreq), bytes(nreq);
        std::vector<wcsph_params> params(nreq);
        for (int k = 0; k < nreq; ++k) {
            params[k] = reqs[k].params;
            make_state(&reqs[k], params[k], states[k], ns[k]);
            wcsph_sim_config sc{};
            sc.precision = reqs[k].precision;
            bytes[k] = wcsph_sim_bytes(reqs[k].geom.dim, ns[k], &sc);
            WCSPH_REQUIRE(bytes[k] <= cfg->hbm_budget_bytes, WCSPH_ENOMEM,
                          "a request is larger than the whole HBM budget");
            out[k] = wcsph_request_result{};
            out[k].arrival_s = reqs[k].arrival_s;
            out[k].n = ns[k];
        }
        std::vector<int> order(nreq);
        for (int k = 0; k < nreq; ++k) order[k] = k;
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
        std::deque<int> pending(order.begin(), order.end());
        std::vector<Active> active;
        int64_t used = 0;
        const auto t0 = clk::now();
        auto now = [&] { return std::chrono::duration<double>(clk::now() - t0).count(); };
        while (!pending.empty() || !active.empty()) {
            bool busy = false;
            // admission
            while (!pending.empty() && reqs[pending.front()].arrival_s <= now() &&
                   (int)active.size() < cfg->max_slots && used + bytes[pending.front()] <= cfg->hbm_budget_bytes) {
                const int k = pending.front();
                pending.pop_front();
                const HostState& st = states[k];
                wcsph_sim* sim = nullptr;
                int rc = 0;
                for (int attempt = 0; attempt < 2; ++attempt) {
                    wcsph_sim_config sc{};
                    // a mixed request whose grid is beyond the FP16 cell index runs in FP32
                    sc.precision = attempt == 0 ? reqs[k].precision : WCSPH_PREC_FP32;
                    out[k].precision_used = sc.precision;
                    rc = wcsph_sim_create(&params[k], &sc, ns[k], &sim);
                    if (!rc)
                        rc = wcsph_sim_upload(sim, st.pos.data(), st.vel.data(), st.dens.data(), st.mass.data(),
                                              st.h.data(), st.type.data());
                    const bool retry = rc == WCSPH_EINVAL && sc.precision == WCSPH_PREC_MIXED &&
                                       std::string(wcsph_last_error()).find("2048 cells") != std::string::npos;
                    if (!retry) break;
                    wcsph_sim_destroy(sim);
                    sim = nullptr;
                }
                if (!rc) rc = wcsph_sim_init(sim);
                out[k].start_s = now();
                if (rc) {
                    out[k].status = rc;
                    out[k].end_s = now();
                    if (sim) wcsph_sim_destroy(sim);
                    continue;
                }
                used += bytes[k];
                active.push_back(Active{k, sim, bytes[k], 0, false});
                busy = true;
            }
            // progress and retirement
            for (size_t a = 0; a < active.size();) {
                Active& A = active[a];
                const wcsph_request& r = reqs[A.idx];
                int32_t idle = 0;
                int rc = wcsph_sim_query(A.sim, &idle);
                if (!rc && idle && A.done_steps < r.steps) {
                    const int32_t c = std::min(cfg->chunk_steps, r.steps - A.done_steps);
                    rc = wcsph_sim_step(A.sim, c);
                    A.done_steps += c;
                    busy = true;
                } else if (!rc && idle) {
                    wcsph_request_result& o = out[A.idx];
                    wcsph_sim_stats s{};
                    rc = wcsph_sim_get_stats(A.sim, &s);
                    double d[10] = {0};
                    const double band = 2.0 * r.geom.dx;
                    if (!rc) rc = wcsph_sim_diagnostics(A.sim, band, band, r.geom.tank_l - band, d);
                    o.end_s = now();
                    o.steps = s.step;
                    o.pairs = s.pairs_total;
                    o.sim_time = s.time;
                    o.front = d[7];
                    o.column = d[8];
                    o.kinetic = d[4];
                    o.status = rc;
                    wcsph_sim_destroy(A.sim);
                    used -= A.bytes;
                    active.erase(active.begin() + a);
                    busy = true;
                    continue;
                }
                if (rc) {
                    out[A.idx].status = rc;
                    out[A.idx].end_s = now();
                    wcsph_sim_destroy(A.sim);
                    used -= A.bytes;
                    active.erase(active.begin() + a);
                    continue;
                }
                ++a;
            }
            if (!busy) std::this_thread::sleep_for(std::chrono::microseconds(50));
        }
    });
}
