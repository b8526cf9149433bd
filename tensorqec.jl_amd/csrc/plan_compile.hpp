// plan_compile.hpp — host-side compilation of a tqec_plan_desc into the
// per-syndrome step program executed by the kernels.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "plan.hpp"
#include "tqec_hip.h"

namespace tqec {

// Backtrace record of one device step of a max-product plan: the labels of
// its output (memory order), then the contracted and single-side labels in
// the bit order of the step's arg-max record.
struct MapStep {
    int32_t off = 0;      // into CompiledPlan::map_labels
    int32_t nc = 0, nk = 0, nsa = 0, nsb = 0;
    int64_t arg_off = 0;  // arg-max table offset per syndrome (entries)
};

struct CompiledPlan {
    // ---- descriptor copy -------------------------------------------------
    int n_vars = 0, n_leaves = 0, n_steps = 0, n_out = 0, n_syn = 0;
    std::vector<std::vector<int>> raw_labels;   // per leaf, as given
    std::vector<std::vector<double>> raw_data;  // per leaf
    std::vector<int> leaf_ev;                   // one-hot evidence bit or -1
    std::vector<int> lhs, rhs;
    std::vector<int> out_labels;
    std::vector<int> slice_bit_of_label;        // -1 when not sliced

    // ---- derived ---------------------------------------------------------
    std::vector<std::vector<int>> labels;       // effective labels per node
    std::vector<std::vector<int>> bits;         // syndrome bits a node depends on
    std::vector<char> dep;                      // node contracted per syndrome on the device
    std::vector<std::vector<std::vector<double>>> var;  // tabulated nodes: one tensor per bit pattern
    int fold_limit = kMaxEv;                    // tabulate nodes with at most this many bits
    int fold_bits = 0;                          // limit actually used
    std::vector<std::vector<double>> value;     // constant nodes (labels order)
    std::vector<std::vector<int>> mem;          // memory order of operands

    int semiring = TQEC_SEMIRING_SUM;
    std::vector<MapStep> map_steps;             // max-product backtrace, execution order
    std::vector<int32_t> map_labels;
    int64_t arg_entries = 0;                    // arg-max records per syndrome
    std::vector<StepDesc> steps;                // dependent steps, execution order
    std::vector<double> const_pool;             // kConst data (mem order, scaled)
    std::vector<double> sliced_pool;            // kSliced variants
    int64_t ws_entries = 0;                     // per-syndrome workspace (entries)
    int n_slots = 0;                            // scale slots
    int64_t root_off = 0;
    int root_slot = 0;

    std::vector<int> reg_lay;                   // register-resident kernel: two lane bits per step (or empty)
    bool reg_checked = false;                   // reg_lay was decided at plan creation (empty: not offered)
    std::string reg_err;                        // why not, when reg_lay is empty

    // ---- statistics ------------------------------------------------------
    double tc = 0, rw = 0, dep_tc = 0, dep_rw = 0;
    int sc = 0;
};

// Returns TQEC_OK or an error code; `err` receives a message on failure.
int compile_plan(const tqec_plan_desc *d, CompiledPlan &p, std::string &err);

}  // namespace tqec
