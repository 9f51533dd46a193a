// Expression DAG for f-rep shapes: the front end that produces tapes.
//
// Source-compatible with the parts of libfive::Tree that the evaluation and
// meshing path uses (libfive/include/libfive/tree/tree.hpp:46-242):
// refcounted nodes handed across the C ABI as raw pointers, local
// simplification on construction (libfive/src/tree/tree.cpp:95-230),
// hash-consed affine/commutative optimisation (tree.cpp:612-961) and the
// leaves-to-root walk that defines tape order (tree.cpp:530-594).
//
// One deliberate difference: where the reference orders terms by heap
// address (tree.cpp:846-854, 926-938), this implementation orders them by a
// per-node creation serial number, so tapes are deterministic run to run.
#pragma once

#include <atomic>
#include <cstdint>
#include <iosfwd>
#include <string>
#include <unordered_map>
#include <vector>

#include "opcode.hpp"

namespace libfive {

struct TreeData;

class Tree {
public:
    using Data = TreeData;
    using Id = const void*;

    Tree() : ptr(nullptr), flags(0) {}
    ~Tree();
    Tree(const Tree& o) : Tree(o.ptr, true, o.flags) {}
    Tree(Tree&& o) noexcept : ptr(o.ptr), flags(o.flags) { o.ptr = nullptr; }
    Tree& operator=(const Tree& o) { return *this = Tree(o.ptr, true, o.flags); }
    Tree& operator=(Tree&& o) noexcept {
        std::swap(ptr, o.ptr);
        flags = o.flags;
        return *this;
    }
    // Borrow a raw pointer from the C API (increments the refcount).
    explicit Tree(const Data* d) : Tree(d, true, 0) {}
    // Constant node (the value is stored as float, tree.cpp:71-75).
    Tree(double v);

    static Tree X();
    static Tree Y();
    static Tree Z();
    static Tree var();
    static Tree invalid();

    static Tree nullary(Opcode::Opcode op);
    static Tree unary(Opcode::Opcode op, const Tree& lhs);
    static Tree binary(Opcode::Opcode op, const Tree& lhs, const Tree& rhs);

    bool is_valid() const;
    Id id() const { return ptr; }
    const Data* get() const { return ptr; }
    const Data* operator->() const { return ptr; }
    bool operator==(const Tree& o) const { return ptr == o.ptr; }
    bool operator!=(const Tree& o) const { return ptr != o.ptr; }

    // Hash-consed, affine-collapsed, commutative-flattened copy.
    Tree optimized() const;
    // q(x, y, z) = this(X(x, y, z), Y(x, y, z), Z(x, y, z)), substituted at
    // once: the result is what the reference's remap node becomes after
    // Tree::flatten (tree.cpp:366-380, 389-520).
    Tree remap(const Tree& X, const Tree& Y, const Tree& Z) const;
    // Number of unique nodes reachable from this tree.
    size_t size() const;
    // Leaves-to-root topological order (tree.cpp:530-594).
    std::vector<const Data*> walk() const;

    // Give up ownership without decrementing (C API handoff) / take it back.
    const Data* release();
    static Tree reclaim(const Data* p);

    std::ostream& print_prefix(std::ostream& s) const;

    bool is_optimized() const { return flags & TREE_FLAG_IS_OPTIMIZED; }

private:
    explicit Tree(const Data* d, bool inc, uint32_t flags);
    static Tree make(Opcode::Opcode op, float v, const Tree* a, const Tree* b);
    static Tree one();
    Tree with_flags(uint32_t f) const { return Tree(ptr, true, flags | f); }

    enum : uint32_t { TREE_FLAG_IS_OPTIMIZED = 1u };
    const Data* ptr;
    uint32_t flags;

    friend struct TreeData;
    friend class Optimizer;
};

// Node storage.  Nodes are immutable after construction apart from the
// reference count.
struct TreeData {
    Opcode::Opcode op() const { return op_; }
    float value() const;               // throws ValueException if not CONSTANT
    const Tree& lhs() const;           // throws ChildException if absent
    const Tree& rhs() const;
    uint64_t serial() const { return serial_; }

    struct ValueException : std::exception {
        const char* what() const noexcept override {
            return "Accessed value of non-constant Tree";
        }
    };
    struct ChildException : std::exception {
        const char* what() const noexcept override {
            return "Accessed missing child";
        }
    };
    struct PowValueException : std::exception {
        const char* what() const noexcept override {
            return "Power and nth-root argument must be a constant integer";
        }
    };

    enum : uint32_t { TREE_FLAG_HAS_XYZ = 1u << 2 };

    mutable std::atomic<uint32_t> refcount{0};
    const uint32_t flags;

    TreeData(Opcode::Opcode op, float value, const Tree* a, const Tree* b);
private:
    friend class Tree;
    Opcode::Opcode op_;
    float value_;
    Tree lhs_;          // owned child references (null Tree if absent)
    Tree rhs_;
    uint64_t serial_;
};


std::ostream& operator<<(std::ostream& s, const Tree& t);

// Operator overloads (libfive/include/libfive/tree/operations.hpp).
Tree square(const Tree& a);
Tree sqrt(const Tree& a);
Tree abs(const Tree& a);
Tree operator-(const Tree& a);
Tree sin(const Tree& a);
Tree cos(const Tree& a);
Tree tan(const Tree& a);
Tree asin(const Tree& a);
Tree acos(const Tree& a);
Tree atan(const Tree& a);
Tree log(const Tree& a);
Tree exp(const Tree& a);
Tree operator+(const Tree& a, const Tree& b);
Tree operator*(const Tree& a, const Tree& b);
Tree min(const Tree& a, const Tree& b);
Tree max(const Tree& a, const Tree& b);
Tree operator-(const Tree& a, const Tree& b);
Tree operator/(const Tree& a, const Tree& b);
Tree atan2(const Tree& a, const Tree& b);
Tree mod(const Tree& a, const Tree& b);
Tree nanfill(const Tree& a, const Tree& b);
Tree compare(const Tree& a, const Tree& b);
Tree pow(const Tree& a, const Tree& b);       // b must be an integer constant
Tree nth_root(const Tree& a, const Tree& b);  // b must be an integer constant

// Scalar semantics of one opcode, used for constant folding; identical to
// what the evaluation kernels compute for the same operands.
float fold_constant(Opcode::Opcode op, float a, float b);

}   // namespace libfive
