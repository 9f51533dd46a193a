// Oracle internals (test infrastructure only; see oracle.h).
#pragma once

#include <cstddef>
#include <cstdint>
#include <functional>
#include <memory>
#include <vector>

namespace oracle {

struct Clause {
    uint32_t op, id, a, b;
};

// Tape (libfive/include/libfive/eval/tape.hpp:28-114): root-first clause
// list, root slot, box for INTERVAL tapes, parent link.
struct Tape {
    enum Type { BASE, INTERVAL, SPECIALIZED, FEATURE };
    std::vector<Clause> t;
    uint32_t i = 1;
    Type type = BASE;
    bool terminal = false;
    float box[6] = {0, 0, 0, 0, 0, 0};     // xlo, xhi, ylo, yhi, zlo, zhi
    std::shared_ptr<Tape> parent;
};

struct Deck {
    Deck(const uint32_t* clauses, size_t n, const uint32_t* cids,
         const float* cvals, size_t nc, uint32_t X, uint32_t Y, uint32_t Z,
         uint32_t num_slots);
    uint32_t X, Y, Z;
    uint32_t num_slots;
    std::vector<std::pair<uint32_t, float>> constants;
    std::shared_ptr<Tape> tape;
    std::vector<uint8_t> disabled;
    std::vector<uint32_t> remap;
};

enum Keep { KEEP_BOTH, KEEP_A, KEEP_B, KEEP_ALWAYS };
using KeepFn = std::function<Keep(uint32_t, uint32_t, uint32_t, uint32_t)>;

std::shared_ptr<Tape> push(Deck& deck, const std::shared_ptr<Tape>& tape,
                           const KeepFn& fn, Tape::Type type,
                           const float box[6]);
std::shared_ptr<Tape> get_base(const std::shared_ptr<Tape>& t,
                               const double lo[3], const double hi[3]);

struct ArrayEval {
    static constexpr size_t N = 256;     // LIBFIVE_EVAL_ARRAY_SIZE
    explicit ArrayEval(Deck& d);
    void set(float x, float y, float z, size_t i);
    const float* values(size_t count, const Tape& tape);
    void ambiguous(size_t count, const Tape& tape, uint8_t* out);
    const float* value_row(const Tape& t) const { return &v[size_t(t.i) * N]; }
    float slot(uint32_t s, size_t j) const { return v[size_t(s) * N + j]; }
    Deck& deck;
    std::vector<float> v;                // (num_slots + 1) x N, row-major
};

struct DerivEval : ArrayEval {
    explicit DerivEval(Deck& d);
    void derivs(size_t count, const Tape& tape);
    const float* grad(size_t k, const Tape& tape) const;
    float gslot(uint32_t s, int k, size_t j) const {
        return g[(size_t(s) * 3 + k) * N + j];
    }
    std::vector<float> g;                // (num_slots + 1) x 3 x N
};

struct Iv {
    float lo, hi;
    bool nan;
};
Iv iv_checked(float lo, float hi);
Iv interval_op(uint32_t op, const Iv& a, const Iv& b);

struct IntervalEval {
    explicit IntervalEval(Deck& d);
    Iv eval(const float lo[3], const float hi[3], const Tape& tape);
    std::shared_ptr<Tape> push(const std::shared_ptr<Tape>& tape);
    Deck& deck;
    std::vector<Iv> i;
};

double eval_point_f64(const Deck& deck, const Tape& tape, double x, double y,
                      double z, std::vector<double>& v);
// Value and gradient in double (scratch v is resized as needed).
double eval_deriv_f64(const Deck& deck, const Tape& tape, double x, double y,
                      double z, std::vector<double>& v, double grad[3]);

}   // namespace oracle

struct oracle_deck {
    oracle::Deck deck;
};
