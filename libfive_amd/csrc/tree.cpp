// Expression DAG: construction-time simplification, optimisation, walk and
// printing.  Semantics cite libfive/src/tree/tree.cpp and
// libfive/src/tree/opcode.cpp in the reference.
#include "tree.hpp"

#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstring>
#include <exception>
#include <pthread.h>
#include <iostream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <unordered_set>

#include "flat_map.hpp"
#include "fmath.hpp"

namespace libfive {

////////////////////////////////////////////////////////////////////////////////
// Opcode tables (opcode.cpp:18-273)

int Opcode::args(Opcode op) {
    switch (op) {
        case CONSTANT: case VAR_X: case VAR_Y: case VAR_Z: case VAR_FREE:
        case ORACLE:
            return 0;
        case OP_SQUARE: case OP_SQRT: case OP_NEG: case OP_SIN: case OP_COS:
        case OP_TAN: case OP_ASIN: case OP_ACOS: case OP_ATAN: case OP_EXP:
        case OP_ABS: case OP_LOG: case OP_RECIP: case CONST_VAR:
            return 1;
        case OP_ADD: case OP_MUL: case OP_MIN: case OP_MAX: case OP_SUB:
        case OP_DIV: case OP_ATAN2: case OP_POW: case OP_NTH_ROOT: case OP_MOD:
        case OP_NANFILL: case OP_COMPARE:
            return 2;
        default:
            return -1;
    }
}

static const char* const kOpNames[Opcode::LAST_OP] = {
    "INVALID", "CONSTANT", "VAR_X", "VAR_Y", "VAR_Z", "VAR_FREE", "CONST_VAR",
    "SQUARE", "SQRT", "NEG", "SIN", "COS", "TAN", "ASIN", "ACOS", "ATAN", "EXP",
    "ADD", "MUL", "MIN", "MAX", "SUB", "DIV", "ATAN2", "POW", "NTH_ROOT", "MOD",
    "NANFILL", "ABS", "RECIP", "LOG", "COMPARE", "ORACLE",
};

std::string Opcode::toString(Opcode op) {
    if (op < 0 || op >= LAST_OP) return "";
    return kOpNames[op];
}

std::string Opcode::toScmString(Opcode op) {
    std::string s = toString(op);
    for (auto& c : s) {
        c = (c == '_') ? '-' : char(std::tolower(c));
    }
    return s;
}

Opcode::Opcode Opcode::fromScmString(std::string s) {
    for (auto& c : s) c = char(std::tolower(c));
    for (int i = 0; i < LAST_OP; ++i) {
        if (toScmString(Opcode(i)) == s) return Opcode(i);
    }
    return INVALID;
}

std::string Opcode::toOpString(Opcode op) {
    switch (op) {
        case VAR_X: return "x";
        case VAR_Y: return "y";
        case VAR_Z: return "z";
        case OP_ADD: return "+";
        case OP_MUL: return "*";
        case OP_NEG: case OP_SUB: return "-";
        case OP_RECIP: case OP_DIV: return "/";
        case INVALID: case CONSTANT: case ORACLE: case LAST_OP: return "";
        default: return toScmString(op);
    }
}

bool Opcode::isCommutative(Opcode op) {
    return op == OP_ADD || op == OP_MUL || op == OP_MIN || op == OP_MAX;
}

bool Opcode::isIdempotent(Opcode op) {
    return op == OP_MIN || op == OP_MAX;
}

float fold_constant(Opcode::Opcode op, float a, float b) {
    return fmath::eval_op(int(op), a, b);
}

////////////////////////////////////////////////////////////////////////////////
// Node lifetime

static std::atomic<uint64_t> g_serial{1};

static uint32_t compute_flags(Opcode::Opcode op, const Tree* a, const Tree* b) {
    uint32_t f = 0;
    if (op == Opcode::VAR_X || op == Opcode::VAR_Y || op == Opcode::VAR_Z) {
        f |= TreeData::TREE_FLAG_HAS_XYZ;
    }
    if (a && a->get()) f |= a->get()->flags;
    if (b && b->get()) f |= b->get()->flags;
    return f;
}

TreeData::TreeData(Opcode::Opcode op, float value, const Tree* a,
                   const Tree* b)
    : flags(compute_flags(op, a, b)), op_(op), value_(value),
      lhs_(a ? *a : Tree()), rhs_(b ? *b : Tree()),
      serial_(g_serial.fetch_add(1))
{
}

float TreeData::value() const {
    if (op_ != Opcode::CONSTANT) throw ValueException();
    return value_;
}

const Tree& TreeData::lhs() const {
    if (!lhs_.get()) throw ChildException();
    return lhs_;
}

const Tree& TreeData::rhs() const {
    if (!rhs_.get()) throw ChildException();
    return rhs_;
}

Tree::Tree(const Data* d, bool inc, uint32_t f) : ptr(d), flags(f) {
    if (d && inc) d->refcount++;
}

Tree::~Tree() {
    if (!ptr) return;
    if (--ptr->refcount) return;
    // Iterative teardown so that deep chains do not exhaust the stack.
    std::vector<const Data*> todo{ptr};
    while (!todo.empty()) {
        auto t = const_cast<Data*>(todo.back());
        todo.pop_back();
        for (Tree* k : {&t->lhs_, &t->rhs_}) {
            const Data* c = k->ptr;
            k->ptr = nullptr;
            if (c && !--c->refcount) todo.push_back(c);
        }
        delete t;
    }
}

Tree Tree::make(Opcode::Opcode op, float v, const Tree* a, const Tree* b) {
    return Tree(new Data(op, v, a, b), true, 0);
}

Tree::Tree(double v) : Tree(new Data(Opcode::CONSTANT, float(v), nullptr,
                                     nullptr), true, 0) {}

// X, Y, Z are singletons (tree.cpp:232-244); Y and Z force the earlier axes
// first so that their creation order (and thus their sort order) is fixed.
Tree Tree::X() {
    static Tree x = make(Opcode::VAR_X, 0, nullptr, nullptr);
    return x;
}
Tree Tree::Y() {
    (void)X();
    static Tree y = make(Opcode::VAR_Y, 0, nullptr, nullptr);
    return y;
}
Tree Tree::Z() {
    (void)Y();
    static Tree z = make(Opcode::VAR_Z, 0, nullptr, nullptr);
    return z;
}

Tree Tree::invalid() {
    static Tree i = make(Opcode::INVALID, 0, nullptr, nullptr);
    return i;
}

Tree Tree::one() {
    static Tree o = make(Opcode::CONSTANT, 1.0f, nullptr, nullptr);
    return o;
}

Tree Tree::var() { return nullary(Opcode::VAR_FREE); }

bool Tree::is_valid() const {
    return ptr && ptr->op() != Opcode::INVALID;
}

const Tree::Data* Tree::release() {
    const Data* p = ptr;
    ptr = nullptr;
    return p;
}

Tree Tree::reclaim(const Data* p) { return Tree(p, false, 0); }

////////////////////////////////////////////////////////////////////////////////
// Construction with local simplification (tree.cpp:95-230)

Tree Tree::nullary(Opcode::Opcode op) {
    if (Opcode::args(op) != 0) return invalid();
    switch (op) {
        case Opcode::VAR_X: return X();
        case Opcode::VAR_Y: return Y();
        case Opcode::VAR_Z: return Z();
        default: return make(op, 0, nullptr, nullptr);
    }
}

static bool is_const(const Tree& t) {
    return t.get() && t->op() == Opcode::CONSTANT;
}

Tree Tree::unary(Opcode::Opcode op, const Tree& lhs) {
    if (Opcode::args(op) != 1) return invalid();
    if (is_const(lhs)) {
        // Constant folding: the reference evaluates the one-clause tree at
        // the origin (tree.cpp:101-107).
        return Tree(fold_constant(op, lhs->value(), 0.0f));
    }
    const Opcode::Opcode inner = lhs->op();
    if (op == Opcode::OP_ABS) {
        if (inner == Opcode::OP_ABS || inner == Opcode::OP_SQUARE) return lhs;
    } else if (op == Opcode::OP_NEG) {
        if (inner == Opcode::OP_NEG) return lhs->lhs();
    }
    return make(op, 0, &lhs, nullptr);
}

Tree Tree::binary(Opcode::Opcode op, const Tree& lhs, const Tree& rhs) {
    using namespace Opcode;
    if (::libfive::Opcode::args(op) != 2) return invalid();
    const bool lc = is_const(lhs), rc = is_const(rhs);
    if (lc && rc) {
        return Tree(fold_constant(op, lhs->value(), rhs->value()));
    }
    const bool lneg = args(lhs->op()) == 1 && lhs->op() == OP_NEG;
    const bool rneg = args(rhs->op()) == 1 && rhs->op() == OP_NEG;
    switch (op) {
        case OP_DIV:
            if (rc && rhs->value() == 1.0f) return lhs;
            break;
        case OP_ADD:
            // The reference tests these as an else-if chain on the kind of
            // node (constant lhs, constant rhs, unary lhs, unary rhs).
            if (lc) {
                if (lhs->value() == 0.0f) return rhs;
            } else if (rc) {
                if (rhs->value() == 0.0f) return lhs;
            } else if (args(lhs->op()) == 1) {
                if (lneg) return rhs - lhs->lhs();
            } else if (rneg) {
                return lhs - rhs->lhs();
            }
            break;
        case OP_SUB:
            if (lc) {
                if (lhs->value() == 0.0f) return -rhs;
            } else if (rc) {
                if (rhs->value() == 0.0f) return lhs;
            } else if (rneg) {
                return lhs + rhs->lhs();
            }
            break;
        case OP_MUL:
            if (lc) {
                const float v = lhs->value();
                if (v == 0) return lhs;
                if (v == 1) return rhs;
                if (v == -1) return -rhs;
            } else if (rc) {
                const float v = rhs->value();
                if (v == 0) return rhs;
                if (v == 1) return lhs;
                if (v == -1) return -lhs;
            } else if (lhs.id() == rhs.id()) {
                return square(lhs);
            }
            break;
        case OP_NTH_ROOT:
        case OP_POW:
            if (rc && rhs->value() == 1.0f) return lhs;
            break;
        case OP_MIN:
        case OP_MAX:
            if (lhs.id() == rhs.id()) return lhs;
            break;
        default:
            break;
    }
    return make(op, 0, &lhs, &rhs);
}

////////////////////////////////////////////////////////////////////////////////
// Operators

#define LF_UNARY(NAME, OP) \
    Tree NAME(const Tree& a) { return Tree::unary(Opcode::OP, a); }
#define LF_BINARY(NAME, OP) \
    Tree NAME(const Tree& a, const Tree& b) { \
        return Tree::binary(Opcode::OP, a, b); }
LF_UNARY(square, OP_SQUARE)
LF_UNARY(sqrt, OP_SQRT)
LF_UNARY(abs, OP_ABS)
LF_UNARY(operator-, OP_NEG)
LF_UNARY(sin, OP_SIN)
LF_UNARY(cos, OP_COS)
LF_UNARY(tan, OP_TAN)
LF_UNARY(asin, OP_ASIN)
LF_UNARY(acos, OP_ACOS)
LF_UNARY(atan, OP_ATAN)
LF_UNARY(log, OP_LOG)
LF_UNARY(exp, OP_EXP)
LF_BINARY(operator+, OP_ADD)
LF_BINARY(operator*, OP_MUL)
LF_BINARY(min, OP_MIN)
LF_BINARY(max, OP_MAX)
LF_BINARY(operator-, OP_SUB)
LF_BINARY(operator/, OP_DIV)
LF_BINARY(atan2, OP_ATAN2)
LF_BINARY(mod, OP_MOD)
LF_BINARY(nanfill, OP_NANFILL)
LF_BINARY(compare, OP_COMPARE)
#undef LF_UNARY
#undef LF_BINARY

static void check_integer_constant(const Tree& b) {
    if (!is_const(b) || b->value() != float(int(b->value()))) {
        throw TreeData::PowValueException();
    }
}
Tree pow(const Tree& a, const Tree& b) {
    check_integer_constant(b);
    return Tree::binary(Opcode::OP_POW, a, b);
}
Tree nth_root(const Tree& a, const Tree& b) {
    check_integer_constant(b);
    return Tree::binary(Opcode::OP_NTH_ROOT, a, b);
}

////////////////////////////////////////////////////////////////////////////////
// Printing (tree.cpp:262-331): prefix form, with chains of the same
// commutative operator printed flat.

static void print_node(std::ostream& s, const TreeData* d, Opcode::Opcode parent) {
    const auto op = d->op();
    const int n = Opcode::args(op);
    if (op == Opcode::CONSTANT) {
        s << d->value();
    } else if (n <= 0) {
        s << Opcode::toOpString(op);
    } else if (n == 1) {
        s << "(" << Opcode::toOpString(op) << " ";
        print_node(s, d->lhs().get(), op);
        s << ")";
    } else {
        const bool flat = Opcode::isCommutative(op) && parent == op;
        if (!flat) s << "(" << Opcode::toOpString(op) << " ";
        print_node(s, d->lhs().get(), op);
        s << " ";
        print_node(s, d->rhs().get(), op);
        if (!flat) s << ")";
    }
}

std::ostream& Tree::print_prefix(std::ostream& s) const {
    print_node(s, ptr, Opcode::INVALID);
    return s;
}

std::ostream& operator<<(std::ostream& s, const Tree& t) {
    return t.print_prefix(s);
}

////////////////////////////////////////////////////////////////////////////////
// Size and walk

size_t Tree::size() const {
    std::unordered_set<const Data*> seen;
    std::vector<const Data*> todo{ptr};
    while (!todo.empty()) {
        auto d = todo.back();
        todo.pop_back();
        if (!seen.insert(d).second) continue;
        const int n = Opcode::args(d->op());
        if (n >= 1) todo.push_back(d->lhs().get());
        if (n >= 2) todo.push_back(d->rhs().get());
    }
    return seen.size();
}

// Topological order in which every node appears after all of its children
// (tree.cpp:530-594): count incoming edges, then emit root-first and release
// a child once its last parent has been emitted; the list is reversed.
std::vector<const Tree::Data*> Tree::walk() const {
    FlatMap<const Data*, unsigned> count(1024);
    std::vector<const Data*> todo{ptr};
    while (!todo.empty()) {
        auto d = todo.back();
        todo.pop_back();
        const int n = Opcode::args(d->op());
        if (n >= 1 && count[d->lhs().get()]++ == 0) todo.push_back(d->lhs().get());
        if (n >= 2 && count[d->rhs().get()]++ == 0) todo.push_back(d->rhs().get());
    }
    std::vector<const Data*> flat;
    todo.push_back(ptr);
    while (!todo.empty()) {
        auto d = todo.back();
        todo.pop_back();
        flat.push_back(d);
        const int n = Opcode::args(d->op());
        if (n >= 1 && --*count.find(d->lhs().get()) == 0) todo.push_back(d->lhs().get());
        if (n >= 2 && --*count.find(d->rhs().get()) == 0) todo.push_back(d->rhs().get());
    }
    std::reverse(flat.begin(), flat.end());
    return flat;
}

////////////////////////////////////////////////////////////////////////////////
// Optimiser (tree.cpp:617-961), restated recursively.
//
// Each node is visited with a multiplier ("scale") and the accumulator of
// its parent:
//   - an affine node (NEG, ADD, SUB, MUL or DIV by a constant) merges its
//     scaled children into the parent's affine map if the parent is
//     affine, and otherwise opens a new map that is collapsed to a tree;
//   - a commutative node (MIN, MAX, or MUL of two non-constants) merges its
//     children into the parent's list if the parent is a list of the same
//     operator, and otherwise opens a new list that is sorted and folded;
//   - any other node is rebuilt from its optimised children and hash-consed.
// Children are visited right-hand side first, as in the reference's task
// stack, so that new nodes are created in the same relative order.

struct NodeKey {
    int op;
    uint32_t value_bits;     // constants: value (+0/-0 merged, one NaN)
    const void* a;
    const void* b;
    bool operator==(const NodeKey& o) const {
        return op == o.op && value_bits == o.value_bits && a == o.a && b == o.b;
    }
};
struct NodeKeyHash {
    size_t operator()(const NodeKey& k) const {
        uint64_t h = uint64_t(uint32_t(k.op)) * 0x9e3779b97f4a7c15ull ^ k.value_bits;
        h = h * 0x100000001b3ull ^ uint64_t(uintptr_t(k.a));
        h = h * 0x100000001b3ull ^ uint64_t(uintptr_t(k.b));
        return mix_hash(h);
    }
};

class Optimizer {
public:
    // Key of a node for hash-consing (data.cpp:83-108): constants by value
    // (all NaNs equal, +0 == -0), nullary opcodes by opcode (free variables
    // by identity), operations by opcode and child identities.
    static NodeKey key_of(const TreeData* d) {
        const auto op = d->op();
        if (op == Opcode::CONSTANT) {
            const float v = d->value();
            uint32_t bits = fmath::f2u(v);
            if (std::isnan(v)) bits = 0x7fc00000u;
            else if (v == 0.0f) bits = 0;
            return {op, bits, nullptr, nullptr};
        }
        const int n = Opcode::args(op);
        if (n <= 0) {
            return {op, 0, op == Opcode::VAR_FREE ? d : nullptr, nullptr};
        }
        return {op, 0, d->lhs().get(), n >= 2 ? d->rhs().get() : nullptr};
    }

    Tree uniq(const Tree& t) {
        auto k = key_of(t.get());
        if (const TreeData* const* hit = canonical.find(k)) return Tree(*hit);
        canonical.emplace(k, t.get());
        owned.push_back(t);
        return t;
    }

    // Affine accumulator: term -> multiplier, plus insertion order (the
    // reference iterates an unordered_map, then sorts; the sort makes the
    // iteration order irrelevant except for exact ties, which cannot occur
    // because keys are unique).
    // Most maps hold a handful of terms: they are found by a linear scan,
    // and an index is built only for large maps.
    struct Affine {
        static constexpr size_t kScan = 16;
        std::vector<std::pair<Tree, float>> terms;
        std::unordered_map<const TreeData*, size_t> index;
        void add(const Tree& t, float s) {
            if (terms.size() < kScan) {
                for (auto& p : terms) {
                    if (p.first.get() == t.get()) {
                        p.second += s;
                        return;
                    }
                }
                terms.push_back({t, 0.0f + s});
                return;
            }
            if (index.empty()) {
                for (size_t i = 0; i < terms.size(); ++i) index.emplace(terms[i].first.get(), i);
            }
            auto it = index.find(t.get());
            if (it == index.end()) {
                index.emplace(t.get(), terms.size());
                terms.push_back({t, 0.0f + s});
            } else {
                terms[it->second].second += s;
            }
        }
    };
    struct Comm {
        Opcode::Opcode op;
        std::vector<Tree> list;
    };
    struct Ctx {
        Affine* affine = nullptr;
        Comm* comm = nullptr;
    };

    static bool is_affine(const TreeData* d, float* scale_l, float* scale_r,
                          float s, int* which) {
        // which: bit 0 = visit lhs, bit 1 = visit rhs
        using namespace Opcode;
        switch (d->op()) {
            case OP_NEG: *scale_l = -s; *which = 1; return true;
            case OP_ADD: *scale_l = s; *scale_r = s; *which = 3; return true;
            case OP_SUB: *scale_l = s; *scale_r = -s; *which = 3; return true;
            case OP_MUL:
                if (is_const(d->lhs())) {
                    *scale_r = d->lhs()->value() * s; *which = 2; return true;
                } else if (is_const(d->rhs())) {
                    *scale_l = d->rhs()->value() * s; *which = 1; return true;
                }
                return false;
            case OP_DIV:
                if (is_const(d->rhs())) {
                    *scale_l = s / d->rhs()->value(); *which = 1; return true;
                }
                return false;
            default:
                return false;
        }
    }
    static bool is_comm(const TreeData* d) {
        using namespace Opcode;
        const auto op = d->op();
        return op == OP_MIN || op == OP_MAX ||
               (op == OP_MUL && !is_const(d->lhs()) && !is_const(d->rhs()));
    }

    // Hands a finished subtree to the parent accumulator; returns the tree
    // when there is no accumulator.
    Tree deliver(const Tree& t, float scale, Ctx parent) {
        if (parent.affine) {
            if (is_const(t)) {
                parent.affine->add(Tree::one(), scale * t->value());
            } else {
                parent.affine->add(t, scale);
            }
            return Tree();
        } else if (parent.comm) {
            parent.comm->list.push_back(t);
            return Tree();
        }
        return t;
    }

    // Results that do not depend on the enclosing accumulator: a plain
    // node's rebuilt tree, and the collapsed tree of an affine or
    // commutative group root.  The reference re-traverses shared subtrees
    // every time they are reached; reusing the first result gives the same
    // canonical nodes without the exponential re-traversal of shared DAGs.
    // The tables hold plain node pointers (cheap to rehash); every node in
    // them is canonical and kept alive by `owned`.
    FlatMap<const TreeData*, const TreeData*> memo_plain{1024}, memo_affine{1024},
        memo_comm{256};

    // Recursion guard: when set, visiting below this stack address throws
    // TooDeep (the caller then reruns the pass on a larger stack).
    struct TooDeep {};
    uintptr_t stack_floor = 0;

    Tree visit(const TreeData* d, float scale, Ctx parent) {
        if (stack_floor && uintptr_t(__builtin_frame_address(0)) < stack_floor) throw TooDeep{};
        float sl = 1.0f, sr = 1.0f;
        int which = 0;
        if (is_affine(d, &sl, &sr, scale, &which)) {
            if (parent.affine) {
                // Transparent: accumulate into the enclosing map.
                if (which & 2) visit(d->rhs().get(), sr, parent);
                if (which & 1) visit(d->lhs().get(), sl, parent);
                return Tree();
            }
            if (const TreeData* const* hit = memo_affine.find(d)) {
                if (parent.comm) {
                    parent.comm->list.push_back(Tree(*hit));
                    return Tree();
                }
                return Tree(*hit);
            }
            Affine map;
            Ctx c;
            c.affine = &map;
            if (which & 2) visit(d->rhs().get(), sr, c);
            if (which & 1) visit(d->lhs().get(), sl, c);
            Tree t = collapse_affine(map);
            memo_affine.emplace(d, t.get());
            // An affine root always arrives with scale 1: its parent is
            // either a plain node or a commutative list.
            if (parent.comm) {
                parent.comm->list.push_back(t);
                return Tree();
            }
            return t;
        }
        if (is_comm(d)) {
            if (parent.comm && parent.comm->op == d->op()) {
                visit(d->rhs().get(), 1.0f, parent);
                visit(d->lhs().get(), 1.0f, parent);
                return Tree();
            }
            if (const TreeData* const* hit = memo_comm.find(d)) {
                return deliver(Tree(*hit), scale, parent);
            }
            Comm list{d->op(), {}};
            Ctx c;
            c.comm = &list;
            visit(d->rhs().get(), 1.0f, c);
            visit(d->lhs().get(), 1.0f, c);
            Tree t = fold_comm(list);
            memo_comm.emplace(d, t.get());
            return deliver(t, scale, parent);
        }
        // Plain node.
        if (const TreeData* const* hit = memo_plain.find(d)) {
            return deliver(Tree(*hit), scale, parent);
        }
        const int n = Opcode::args(d->op());
        Tree self;
        if (n == 1) {
            Tree a = visit(d->lhs().get(), 1.0f, Ctx());
            self = (a != d->lhs()) ? uniq(Tree::unary(d->op(), a))
                                   : uniq(Tree(d));
        } else if (n == 2) {
            Tree b = visit(d->rhs().get(), 1.0f, Ctx());
            Tree a = visit(d->lhs().get(), 1.0f, Ctx());
            self = (a != d->lhs() || b != d->rhs())
                ? uniq(Tree::binary(d->op(), a, b))
                : uniq(Tree(d));
        } else {
            self = uniq(Tree(d));
        }
        memo_plain.emplace(d, self.get());
        return deliver(self, scale, parent);
    }

    // UpAffine (tree.cpp:815-913): split into positive and negative terms,
    // sort each by (multiplier, node order), sum runs of equal multipliers,
    // scale each run, and return pos - neg.
    Tree collapse_affine(const Affine& map) {
        std::vector<std::pair<const TreeData*, float>> pos, neg;
        for (auto& p : map.terms) {
            const float m = p.second;
            if (m > 0.0f) pos.push_back({p.first.get(), m});
            else if (m < 0.0f) neg.push_back({p.first.get(), -m});
            else if (std::isnan(m)) pos.push_back({p.first.get(), m});
        }
        auto order = [](const std::pair<const TreeData*, float>& a,
                        const std::pair<const TreeData*, float>& b) {
            if (a.second != b.second) return a.second < b.second;
            return a.first->serial() < b.first->serial();
        };
        std::sort(pos.begin(), pos.end(), order);
        std::sort(neg.begin(), neg.end(), order);
        auto collapse = [&](const std::vector<std::pair<const TreeData*, float>>& v) {
            Tree out;
            size_t i = 0;
            while (i < v.size()) {
                const float m = v[i].second;
                Tree t(v[i].first);
                ++i;
                // NaN multipliers never compare equal, so each forms its
                // own run.
                while (i < v.size() && v[i].second == m) {
                    t = uniq(t + Tree(v[i].first));
                    ++i;
                }
                if (m != 1) {
                    if (t == Tree::one()) {
                        t = uniq(Tree(m));
                    } else {
                        t = uniq(t * uniq(Tree(m)));
                    }
                }
                out = out.get() ? uniq(out + t) : t;
            }
            return out.get() ? out : uniq(Tree(0.0f));
        };
        const Tree p = collapse(pos);
        const Tree n = collapse(neg);
        return uniq(p - n);
    }

    // UpCommutative (tree.cpp:914-958): sort by node order, drop duplicates
    // for min/max, left-fold.
    Tree fold_comm(Comm& c) {
        std::sort(c.list.begin(), c.list.end(),
                  [](const Tree& a, const Tree& b) {
                      return a->serial() < b->serial();
                  });
        auto end = c.list.end();
        if (Opcode::isIdempotent(c.op)) {
            end = std::unique(c.list.begin(), end);
        }
        Tree out = c.list.front();
        for (auto it = c.list.begin() + 1; it != end; ++it) {
            out = uniq(Tree::binary(c.op, out, *it));
        }
        return out;
    }

    FlatMap<NodeKey, const TreeData*, NodeKeyHash> canonical{1024};
    std::vector<Tree> owned;
};

// The optimiser recurses once per tree level; a chain tens of thousands of
// nodes deep would exhaust an ordinary thread's stack, so the pass runs on a
// helper thread with a 512 MiB stack (address space only: pages are touched
// as the recursion reaches them).
namespace {
struct OptimizeJob {
    const TreeData* root;
    Tree out;
    std::exception_ptr error;
};
void* optimize_entry(void* arg) {
    auto job = static_cast<OptimizeJob*>(arg);
    try {
        Optimizer opt;
        job->out = opt.visit(job->root, 1.0f, Optimizer::Ctx());
    } catch (...) {
        job->error = std::current_exception();
    }
    return nullptr;
}
}   // namespace

Tree Tree::optimized() const {
    if (flags & TREE_FLAG_IS_OPTIMIZED) return *this;
    // Most trees are shallow: the pass runs on the calling thread with at
    // most 1 MiB of its stack, and moves to the helper thread only when the
    // recursion goes deeper.
    {
        Optimizer opt;
        opt.stack_floor = uintptr_t(__builtin_frame_address(0)) - (uintptr_t(1) << 20);
        try {
            Tree t = opt.visit(ptr, 1.0f, Optimizer::Ctx());
            return t.with_flags(TREE_FLAG_IS_OPTIMIZED);
        } catch (const Optimizer::TooDeep&) {
        }
    }
    OptimizeJob job{ptr, Tree(), nullptr};
    pthread_attr_t attr;
    pthread_t th;
    bool ran = false;
    if (pthread_attr_init(&attr) == 0) {
        if (pthread_attr_setstacksize(&attr, size_t(512) << 20) == 0 &&
            pthread_create(&th, &attr, optimize_entry, &job) == 0) {
            pthread_join(th, nullptr);
            ran = true;
        }
        pthread_attr_destroy(&attr);
    }
    if (!ran) optimize_entry(&job);   // no thread available: this stack
    if (job.error) std::rethrow_exception(job.error);
    return job.out.with_flags(TREE_FLAG_IS_OPTIMIZED);
}

////////////////////////////////////////////////////////////////////////////////
// Remap (tree.cpp:366-380) with the substitution of Tree::flatten
// (tree.cpp:389-520) done at once.  Nodes without x, y or z are kept; a node
// whose children changed is rebuilt through unary()/binary(), so constant
// folding and the local rewrites apply to the substituted tree exactly as
// they do in flatten.  Shared subtrees are substituted once (memoised); the
// walk is an explicit stack, so trees tens of thousands deep are fine.
Tree Tree::remap(const Tree& rx, const Tree& ry, const Tree& rz) const {
    if (rx == X() && ry == Y() && rz == Z()) return *this;
    if (!(ptr->flags & TreeData::TREE_FLAG_HAS_XYZ)) return *this;
    std::unordered_map<const Data*, Tree> done;
    done.emplace(X().get(), rx);
    done.emplace(Y().get(), ry);
    done.emplace(Z().get(), rz);
    std::vector<std::pair<const Data*, bool>> todo{{ptr, false}};
    while (!todo.empty()) {
        const auto [d, up] = todo.back();
        todo.pop_back();
        if (done.count(d)) continue;
        const int n = Opcode::args(d->op());
        if (n == 0 || !(d->flags & TreeData::TREE_FLAG_HAS_XYZ)) {
            done.emplace(d, Tree(d));
            continue;
        }
        if (!up) {
            // children first; the right operand is substituted first, as in
            // flatten's task stack
            todo.push_back({d, true});
            todo.push_back({d->lhs().get(), false});
            if (n == 2) todo.push_back({d->rhs().get(), false});
            continue;
        }
        const Tree a = done.at(d->lhs().get());
        if (n == 1) {
            done.emplace(d, a.get() == d->lhs().get() ? Tree(d)
                                                      : Tree::unary(d->op(), a));
        } else {
            const Tree b = done.at(d->rhs().get());
            const bool same = a.get() == d->lhs().get() &&
                              b.get() == d->rhs().get();
            done.emplace(d, same ? Tree(d) : Tree::binary(d->op(), a, b));
        }
    }
    return done.at(ptr);
}

}   // namespace libfive
