// Deck: flattening of an optimised tree into the reference clause list, and
// compilation of that list into the device tape format (tape.hpp).
#pragma once

#include <cstdint>
#include <vector>

#include "tape.hpp"
#include "tree.hpp"

namespace libfive {

// One clause in the reference's layout: opcode and three slot ids
// (libfive/include/libfive/eval/clause.hpp:18-35).
struct RefClause {
    uint32_t op, id, a, b;
};

class Deck {
public:
    // Slot numbering and tape order of libfive/src/eval/deck.cpp:19-94:
    // slot 0 is a dummy, the root is slot 1, ids count down along the walk,
    // constants and free variables own slots but no clause, X/Y/Z own slots
    // (appended if absent), and the clause list is stored root-first.
    explicit Deck(const Tree& root);

    std::vector<RefClause> clauses;                 // root first
    std::vector<std::pair<uint32_t, float>> constants;
    std::vector<uint32_t> vars;                      // VAR_FREE slots
    std::vector<const TreeData*> var_nodes;          // their nodes (tree ids)
    uint32_t X = 0, Y = 0, Z = 0;
    uint32_t root = 1;
    size_t num_clauses = 0;                          // slots, excluding 0
    bool has_oracle = false;

    Tree tree;                                       // the optimised tree
};

// Device tape for a deck.  `choices`, when given, holds one keep decision
// per clause of `in` (tape::KEEP_*); clauses without a choice use
// KEEP_ALWAYS.
struct DeviceTape {
    std::vector<tape::Clause> clauses;   // evaluation order
    uint32_t root_reg = 0;
    uint32_t num_regs = tape::kFirstFreeReg;
    bool terminal = false;               // no MIN/MAX left that could prune
    uint32_t ref_length = 0;             // clause count excluding COPY/LOADI
};

// True when every clause has one of the common opcodes (tape::fast_opcode),
// which the call-free device evaluators handle.
inline bool tape_is_fast(const DeviceTape& t) {
    for (const tape::Clause& c : t.clauses) {
        if (!tape::fast_opcode(tape::op(c))) return false;
    }
    return true;
}

// Compiles a Deck: renames slots to registers (backward pass).
DeviceTape compile_deck(const Deck& deck,
                        const std::vector<float>* var_values = nullptr);

// Backward pruning + register renaming of a device tape, given per-clause
// keep decisions.  This is the host statement of the algorithm the push
// kernel runs per cell; `changed` reports whether any MIN/MAX was pruned and
// `terminal` whether no undecided MIN/MAX remains (libfive/src/eval/
// tape.cpp:23-158).
DeviceTape push_tape(const DeviceTape& in, const std::vector<uint8_t>& keep,
                     bool* changed);

}   // namespace libfive
