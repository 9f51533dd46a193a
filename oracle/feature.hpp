// Oracle internals: FeatureEvaluator restatement (test infrastructure).
#pragma once

#include <array>
#include <memory>
#include <vector>

#include "oracle_impl.hpp"

namespace oracle {

using V3 = std::array<float, 3>;

// A gradient plus the normalised "epsilon" directions that select it
// (libfive/include/libfive/eval/feature.hpp:18-85).
struct Feature {
    V3 deriv = {0, 0, 0};
    std::vector<V3> eps;
    bool check(const V3& e, bool* duplicate) const;
    bool push(const V3& e);
    bool push(const Feature& other);
    bool operator<(const Feature& o) const;
};

struct FeatureEval {
    explicit FeatureEval(Deck& d);
    std::shared_ptr<Tape> value_and_push(const float p[3],
                                         const std::shared_ptr<Tape>& tape,
                                         float* value);
    std::vector<Feature> walk(const Tape& tape);
    bool is_inside(const float p[3], const std::shared_ptr<Tape>& tape);
    std::vector<V3> features(const float p[3], const std::shared_ptr<Tape>& tape);
    Deck& deck;
    ArrayEval arr;
};

int libfive_args(uint32_t op);

}   // namespace oracle
