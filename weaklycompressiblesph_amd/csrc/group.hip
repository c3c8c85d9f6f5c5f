// Grouped mixed-precision steps for the serving loop.
//
// A group holds up to kGroupMax simulations of one dimension and one set of
// pair-term switches. For each stage of the step (dt commit, kick-drift and
// binning, the two scan passes, scatter, reorder, fused force + final
// stage) one launch covers every member: the grid is the members' grids
// side by side and each block reads its member's argument record from a
// device table (group_kernels.hpp). Two steps are captured into one
// hipGraph per membership: the first step's records take each member from
// its current buffer parity, the second's back, so replays need no new
// arguments. A change of membership rebuilds the tables and re-captures.
#include <cstring>

#include "group.hpp"

namespace wcsph {

namespace {

template <int D>
struct Tables {
    ScalarsArgs sc[kGroupMax];
    BinArgs<D> bin[kGroupMax];
    ScanArgs scan[kGroupMax];
    ScatterArgs scat[kGroupMax];
    ReorderArgs<D, MixedExtra<D>> ro[kGroupMax];
    NFMixedArgs nf[kGroupMax];
};

// device table block: both steps of the group, then the one-member step
template <int D>
constexpr size_t tables_bytes() {
    return 3 * ((sizeof(Tables<D>) + 255) & ~(size_t)255);
}

GroupMap make_map(const std::vector<int32_t>& nblk) {
    GroupMap m{};
    m.n = (int32_t)nblk.size();
    int32_t f = 0;
    for (size_t k = 0; k < nblk.size(); ++k) {
        m.first[k] = f;
        m.nblk[k] = nblk[k];
        f += (nblk[k] + 7) / 8 * 8;  // every member starts on XCD 0 of a cycle of 8
    }
    m.first[nblk.size()] = f;
    return m;
}

struct Maps {
    GroupMap part, reduce, apply;
    int32_t nmem = 0;
};

// A device table address as the kernels' constant-space parameter.
template <class T>
ctab<T> ct(const T* p) {
    return (ctab<T>)(uintptr_t)p;
}

template <int D, int VS, int BT>
void launch_step(hipStream_t st, const Tables<D>* T, const Maps& M) {
    const int n = M.nmem;
    hipLaunchKernelGGL(kg_scalars, dim3(n), dim3(kScalarsThreads), 0, st, ct(T->sc));
    hipLaunchKernelGGL(kg_bin<D>, dim3(M.part.first[n]), dim3(kBlock), 0, st, ct(T->bin), M.part);
    hipLaunchKernelGGL(kg_scan_reduce, dim3(M.reduce.first[n]), dim3(kBlock), 0, st, ct(T->scan), M.reduce);
    hipLaunchKernelGGL(kg_scan_apply, dim3(M.apply.first[n]), dim3(kBlock), 0, st, ct(T->scan), M.apply);
    hipLaunchKernelGGL(kg_scatter, dim3(M.part.first[n]), dim3(kBlock), 0, st, ct(T->scat), M.part);
    hipLaunchKernelGGL((kg_reorder<D, MixedExtra<D>>), dim3(M.part.first[n]), dim3(kBlock), 0, st, ct(T->ro), M.part);
    hipLaunchKernelGGL((kg_nforce_mixed<D, 2, VS, BT>), dim3(M.part.first[n]), dim3(kBlock), 0, st, ct(T->nf), M.part);
}

template <int D>
void launch_step_d(int VS, int BT, hipStream_t st, const Tables<D>* T, const Maps& M) {
    if (VS) BT ? launch_step<D, 1, 1>(st, T, M) : launch_step<D, 1, 0>(st, T, M);
    else BT ? launch_step<D, 0, 1>(st, T, M) : launch_step<D, 0, 0>(st, T, M);
}

}  // namespace

struct StepGroup {
    int D = 3, VS = 0, BT = 0;
    hipStream_t st = nullptr;
    char* d_tab = nullptr;             // device tables (group_table_bytes)
    std::vector<char> h_tab;           // host image of the tables
    std::vector<wcsph_sim*> mem;
    Maps maps;
    hipGraphExec_t graph = nullptr;
    bool dirty = true;

    size_t slot_bytes() const {
        return D == 3 ? ((sizeof(Tables<3>) + 255) & ~(size_t)255) : ((sizeof(Tables<2>) + 255) & ~(size_t)255);
    }
    void drop_graph() {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
    }
};

namespace {

// Fill table slot `slot` (0, 1: the group's two steps; 2: one-member step)
// with the step records of `ms`, each advanced by `phase` steps of parity.
template <int D>
Maps fill_tables(StepGroup* g, int slot, const std::vector<wcsph_sim*>& ms, int phase) {
    Tables<D>* T = reinterpret_cast<Tables<D>*>(g->h_tab.data() + slot * g->slot_bytes());
    std::memset((void*)T, 0, sizeof(Tables<D>));
    std::vector<int32_t> nb, ng, np;
    for (size_t k = 0; k < ms.size(); ++k) {
        wcsph_sim* s = ms[k];
        if (phase) s->cur = 1 - s->cur;
        const MixedStep<D> a = mixed_step_args<D>(s);
        if (phase) s->cur = 1 - s->cur;
        T->sc[k] = a.sc;
        T->bin[k] = a.bin;
        T->scan[k] = a.scan;
        T->scat[k] = a.scat;
        T->ro[k] = a.ro;
        T->nf[k] = a.nf;
        nb.push_back(a.nb);
        ng.push_back(a.ngroups);
        np.push_back(a.npart);
    }
    Maps M;
    M.nmem = (int32_t)ms.size();
    M.part = make_map(nb);
    M.reduce = make_map(ng);
    M.apply = make_map(np);
    return M;
}

void upload_tables(StepGroup* g, int first_slot, int nslots) {
    // the group's stream is idle when tables change (callers synchronise)
    WCSPH_HIP(hipStreamSynchronize(g->st));
    const size_t off = first_slot * g->slot_bytes();
    WCSPH_HIP(hipMemcpy(g->d_tab + off, g->h_tab.data() + off, nslots * g->slot_bytes(), hipMemcpyHostToDevice));
}

template <int D>
const Tables<D>* dev_slot(const StepGroup* g, int slot) {
    return reinterpret_cast<const Tables<D>*>(g->d_tab + slot * g->slot_bytes());
}

void rebuild(StepGroup* g) {
    g->drop_graph();
    if (g->mem.empty()) {
        g->dirty = false;
        return;
    }
    for (wcsph_sim* s : g->mem) WCSPH_REQUIRE(s->st == g->st, WCSPH_EINVAL, "a group member is on another stream");
    if (g->D == 3) {
        g->maps = fill_tables<3>(g, 0, g->mem, 0);
        (void)fill_tables<3>(g, 1, g->mem, 1);
    } else {
        g->maps = fill_tables<2>(g, 0, g->mem, 0);
        (void)fill_tables<2>(g, 1, g->mem, 1);
    }
    upload_tables(g, 0, 2);
    hipGraph_t graph = nullptr;
    WCSPH_HIP(hipStreamBeginCapture(g->st, hipStreamCaptureModeThreadLocal));
    for (int phase = 0; phase < 2; ++phase) {
        if (g->D == 3) launch_step_d<3>(g->VS, g->BT, g->st, dev_slot<3>(g, phase), g->maps);
        else launch_step_d<2>(g->VS, g->BT, g->st, dev_slot<2>(g, phase), g->maps);
    }
    WCSPH_HIP(hipStreamEndCapture(g->st, &graph));
    const hipError_t e = hipGraphInstantiate(&g->graph, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    WCSPH_HIP(e);
    g->dirty = false;
}

}  // namespace

bool group_eligible(const wcsph_sim* s) {
    return s->cfg.precision == WCSPH_PREC_MIXED && s->R == 2 && s->p.shepard_interval == 0 && s->steppable &&
           !s->timing;
}

size_t group_table_bytes(int D) { return D == 3 ? tables_bytes<3>() : tables_bytes<2>(); }

StepGroup* group_create(int D, int VS, int BT, hipStream_t st, void* tables) {
    WCSPH_REQUIRE(D == 2 || D == 3, WCSPH_EINVAL, "dim must be 2 or 3");
    auto g = new StepGroup;
    g->D = D;
    g->VS = VS ? 1 : 0;
    g->BT = BT ? 1 : 0;
    g->st = st;
    g->d_tab = (char*)tables;
    g->h_tab.assign(group_table_bytes(D), 0);
    return g;
}

void group_destroy(StepGroup* g) {
    if (!g) return;
    (void)hipStreamSynchronize(g->st);
    g->drop_graph();
    delete g;
}

void group_set_members(StepGroup* g, const std::vector<wcsph_sim*>& members) {
    WCSPH_REQUIRE((int)members.size() <= kGroupMax, WCSPH_EINVAL, "more than kGroupMax simulations in a group");
    if (members != g->mem) g->dirty = true;
    g->mem = members;
}

void group_run_pairs(StepGroup* g, int pairs) {
    if (g->dirty) rebuild(g);
    if (g->mem.empty() || pairs <= 0) return;
    for (int k = 0; k < pairs; ++k) WCSPH_HIP(hipGraphLaunch(g->graph, g->st));
    for (wcsph_sim* s : g->mem) sim_after_group_steps(s, 2 * (int64_t)pairs);
    WCSPH_HIP(hipGetLastError());
}

void group_run_single(StepGroup* g, wcsph_sim* m) {
    WCSPH_REQUIRE(m->st == g->st, WCSPH_EINVAL, "the simulation is on another stream");
    const std::vector<wcsph_sim*> one{m};
    Maps M;
    if (g->D == 3) M = fill_tables<3>(g, 2, one, 0);
    else M = fill_tables<2>(g, 2, one, 0);
    upload_tables(g, 2, 1);
    if (g->D == 3) launch_step_d<3>(g->VS, g->BT, g->st, dev_slot<3>(g, 2), M);
    else launch_step_d<2>(g->VS, g->BT, g->st, dev_slot<2>(g, 2), M);
    WCSPH_HIP(hipGetLastError());
    m->cur = 1 - m->cur;  // a single step swaps the state buffers
    sim_after_group_steps(m, 1);
    g->dirty = true;  // the member's parity changed: the two-step records are stale
}

}  // namespace wcsph
