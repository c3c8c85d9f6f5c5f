// Grouped mixed-precision steps: several simulations advanced together, one
// launch per stage of the step for all of them (the stage kernels walk a
// table of per-simulation argument records, group_kernels.hpp). Used by the
// serving loop; the request loop being batched is the reference's step,
// src/main.cpp:89-133.
#pragma once
#include <vector>

#include "sim.hpp"

namespace wcsph {

struct StepGroup;

// Whether s can join a group: the mixed path with two cells per support
// length and no Shepard filter (whose schedule is per simulation), after
// wcsph_sim_init.
bool group_eligible(const wcsph_sim* s);

// Device bytes of the argument tables of one group of dimension D.
size_t group_table_bytes(int D);

// A group of dimension D and pair-term switches VS (v_sig rule) and BT
// (beta != 0), working on `st`, with its argument tables at `tables`
// (device memory of group_table_bytes(D), owned by the caller).
StepGroup* group_create(int D, int VS, int BT, hipStream_t st, void* tables);
void group_destroy(StepGroup* g);

// The simulations the next launches advance (at most kGroupMax, all on the
// group's stream). The tables and the two-step graph are rebuilt on the
// next run after a change.
void group_set_members(StepGroup* g, const std::vector<wcsph_sim*>& members);

// Queue `pairs` replays of the two-step graph: every member advances 2 *
// pairs steps.
void group_run_pairs(StepGroup* g, int pairs);

// Queue one step of member m alone (the grouped kernels over a one-member
// table): a step count left odd by the others' chunks.
void group_run_single(StepGroup* g, wcsph_sim* m);

}  // namespace wcsph
