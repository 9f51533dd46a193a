// libfive::Evaluator as the meshing entry uses it
// (libfive/include/libfive/eval/evaluator.hpp:20-70): a tree with values
// for its free variables, which Mesh::render(Evaluator*, ...) meshes.
// Point, derivative and interval evaluation go through the C API's array
// entry points (libfive_tree_eval_array and its siblings).
#pragma once

#include <map>

#include "libfive/tree/tree.hpp"

namespace libfive {

class Evaluator {
public:
    explicit Evaluator(const Tree& root) : root(root) {}
    Evaluator(const Tree& root, const std::map<Tree::Id, float>& vars)
        : root(root), vars(vars) {}

    // Updates variable values; true if any value changed.
    bool updateVars(const std::map<Tree::Id, float>& vs) {
        bool changed = false;
        for (const auto& kv : vs) {
            auto it = vars.find(kv.first);
            if (it == vars.end()) {
                changed = true;
                vars.emplace(kv.first, kv.second);
            } else if (it->second != kv.second) {
                changed = true;
                it->second = kv.second;
            }
        }
        return changed;
    }

    const Tree& tree() const { return root; }
    const std::map<Tree::Id, float>& varValues() const { return vars; }

private:
    Tree root;
    std::map<Tree::Id, float> vars;
};

}   // namespace libfive
