// Deck construction (libfive/src/eval/deck.cpp:19-94) and device tape
// compilation.
#include "deck.hpp"

#include <algorithm>
#include <cassert>
#include <stdexcept>

#include "flat_map.hpp"

namespace libfive {

Deck::Deck(const Tree& root_) {
    tree = root_.optimized();
    const auto flat = tree.walk();

    FlatMap<const TreeData*, uint32_t> ids(flat.size() + 4);
    ids[nullptr] = 0;
    uint32_t id = uint32_t(flat.size());
    std::vector<RefClause> rev;
    rev.reserve(flat.size());
    for (const TreeData* m : flat) {
        const auto op = m->op();
        switch (op) {
            case Opcode::CONSTANT:
                constants.push_back({id, m->value()});
                break;
            case Opcode::VAR_FREE:
                vars.push_back(id);
                var_nodes.push_back(m);
                break;
            case Opcode::ORACLE:
                has_oracle = true;
                break;
            case Opcode::VAR_X:
            case Opcode::VAR_Y:
            case Opcode::VAR_Z:
                break;
            default: {
                const int n = Opcode::args(op);
                rev.push_back({uint32_t(op), id,
                               n >= 1 ? *ids.find(m->lhs().get()) : 0u,
                               n >= 2 ? *ids.find(m->rhs().get()) : 0u});
                break;
            }
        }
        ids[m] = id--;
    }
    assert(id == 0);
    clauses.assign(rev.rbegin(), rev.rend());

    const Tree axes[3] = {Tree::X(), Tree::Y(), Tree::Z()};
    for (const auto& a : axes) {
        if (ids.find(a.get()) == nullptr) {
            const uint32_t next = uint32_t(ids.size());
            ids[a.get()] = next;
        }
    }
    num_clauses = ids.size() - 1;
    X = *ids.find(axes[0].get());
    Y = *ids.find(axes[1].get());
    Z = *ids.find(axes[2].get());
    root = *ids.find(tree.get());
    assert(root == 1);
}

////////////////////////////////////////////////////////////////////////////////

static bool is_minmax(uint32_t op) {
    return op == Opcode::OP_MIN || op == Opcode::OP_MAX;
}

DeviceTape push_tape(const DeviceTape& in, const std::vector<uint8_t>& keep,
                     bool* changed_out) {
    using namespace tape;
    DeviceTape out;
    std::vector<int32_t> bind(in.num_regs, -1);
    std::vector<uint32_t> free_regs;
    uint32_t next = kFirstFreeReg;
    auto alloc = [&]() -> uint32_t {
        if (!free_regs.empty()) {
            uint32_t r = free_regs.back();
            free_regs.pop_back();
            return r;
        }
        if (next >= kMaxRegs) {
            throw std::runtime_error("tape needs more than 65535 registers");
        }
        return next++;
    };
    auto release = [&](uint32_t r) { free_regs.push_back(r); };

    if (in.root_reg < kFirstFreeReg) {
        out.root_reg = in.root_reg;
    } else {
        out.root_reg = alloc();
        bind[in.root_reg] = int32_t(out.root_reg);
    }

    bool changed = false, terminal = true;
    std::vector<Clause> rev;
    rev.reserve(in.clauses.size());
    uint32_t ref_len = 0;

    for (size_t i = in.clauses.size(); i-- > 0;) {
        const Clause& c = in.clauses[i];
        const uint32_t d = dst(c);
        const int32_t r = bind[d];
        if (r < 0) continue;                    // not needed by this tape
        bind[d] = -1;
        const uint32_t o = op(c);
        const uint8_t k = keep.empty() ? uint8_t(KEEP_ALWAYS) : keep[i];
        if (is_minmax(o) && k == KEEP_BOTH) terminal = false;

        if (is_minmax(o) && (k == KEEP_A || k == KEEP_B)) {
            changed = true;
            const bool imm_src = (k == KEEP_A) ? (mode(c) == MODE_A_IMM)
                                               : (mode(c) == MODE_B_IMM);
            const uint32_t src = (k == KEEP_A) ? ra(c) : rb(c);
            if (imm_src) {
                rev.push_back(make(OP_LOADI, MODE_A_IMM, uint32_t(r), 0, 0,
                                   imm(c), slot(c)));
                release(uint32_t(r));
            } else if (src < kFirstFreeReg) {
                rev.push_back(make(OP_COPY, MODE_RR, uint32_t(r), src, 0, 0.0f,
                                   slot(c)));
                release(uint32_t(r));
            } else if (bind[src] < 0) {
                bind[src] = r;                  // forward the value, no clause
            } else {
                rev.push_back(make(OP_COPY, MODE_RR, uint32_t(r),
                                   uint32_t(bind[src]), 0, 0.0f, slot(c)));
                release(uint32_t(r));
            }
            continue;
        }

        release(uint32_t(r));
        const int nargs = (o == OP_LOADI) ? 0
                        : (o == OP_COPY) ? 1
                        : Opcode::args(Opcode::Opcode(o));
        uint32_t na = 0, nb = 0;
        auto use = [&](uint32_t reg) -> uint32_t {
            if (reg < kFirstFreeReg) return reg;
            if (bind[reg] < 0) bind[reg] = int32_t(alloc());
            return uint32_t(bind[reg]);
        };
        if (nargs >= 1 && mode(c) != MODE_A_IMM) na = use(ra(c));
        if (nargs >= 2 && mode(c) != MODE_B_IMM) nb = use(rb(c));
        Clause e = c;
        e.w0 = (c.w0 & 0xffffu) | (uint32_t(r) << 16);
        e.w1 = na | (nb << 16);
        rev.push_back(e);
        if (o != OP_LOADI && o != OP_COPY) ref_len++;
    }
    out.clauses.assign(rev.rbegin(), rev.rend());
    out.num_regs = next;
    out.terminal = terminal;
    out.ref_length = ref_len;
    if (changed_out) *changed_out = changed;
    return out;
}

DeviceTape compile_deck(const Deck& deck, const std::vector<float>* var_values) {
    using namespace tape;
    if (deck.has_oracle) {
        throw std::runtime_error("ORACLE clauses cannot be evaluated on the GPU");
    }
    // Pre-registers: x, y, z are 0..2, every other slot s is s + 3.
    std::vector<int8_t> kind(deck.num_clauses + 1, 0);   // 1 = const, 2 = var
    std::vector<float> cval(deck.num_clauses + 1, 0.0f);
    for (auto& c : deck.constants) { kind[c.first] = 1; cval[c.first] = c.second; }
    for (size_t i = 0; i < deck.vars.size(); ++i) {
        kind[deck.vars[i]] = 2;
        cval[deck.vars[i]] =
            (var_values && i < var_values->size()) ? (*var_values)[i] : 0.0f;
    }
    auto prereg = [&](uint32_t s) -> uint32_t {
        if (s == deck.X) return kRegX;
        if (s == deck.Y) return kRegY;
        if (s == deck.Z) return kRegZ;
        return s + kFirstFreeReg;
    };
    uint32_t extra = uint32_t(deck.num_clauses) + 1 + kFirstFreeReg;

    // Evaluation order.  The reference evaluates its clause list backwards
    // (leaves first).  Any topological order gives the same values, so the
    // clauses are re-ordered by Sethi-Ullman labels: the operand that needs
    // more registers is evaluated first, which minimises the register file
    // for tree-shaped expressions.
    std::vector<int32_t> clause_of(deck.num_clauses + 1, -1);
    for (size_t i = 0; i < deck.clauses.size(); ++i) {
        clause_of[deck.clauses[i].id] = int32_t(i);
    }
    std::vector<uint32_t> need(deck.clauses.size(), 0);
    // Labels are computed over the reverse (leaves-first) list, where every
    // operand precedes its users (operands have larger slot ids).
    for (size_t k = deck.clauses.size(); k-- > 0;) {
        const RefClause& rc = deck.clauses[k];
        const int n = Opcode::args(Opcode::Opcode(rc.op));
        auto child_need = [&](uint32_t s) -> uint32_t {
            const int32_t ci = clause_of[s];
            return ci < 0 ? 0u : need[ci];
        };
        if (n == 1) {
            need[k] = std::max<uint32_t>(1, child_need(rc.a));
        } else if (n == 2) {
            const uint32_t na = child_need(rc.a), nb = child_need(rc.b);
            need[k] = (na == nb) ? std::max<uint32_t>(1, na + 1)
                                 : std::max(na, nb);
        } else {
            need[k] = 1;
        }
    }
    std::vector<uint32_t> order;
    order.reserve(deck.clauses.size());
    {
        std::vector<uint8_t> state(deck.clauses.size(), 0);   // 1 open, 2 done
        std::vector<int32_t> stack;
        if (!deck.clauses.empty()) stack.push_back(0);        // the root
        while (!stack.empty()) {
            const int32_t k = stack.back();
            if (state[k] == 2) { stack.pop_back(); continue; }
            const RefClause& rc = deck.clauses[k];
            const int n = Opcode::args(Opcode::Opcode(rc.op));
            int32_t kids[2] = {-1, -1};
            if (n >= 1) kids[0] = clause_of[rc.a];
            if (n >= 2) kids[1] = clause_of[rc.b];
            if (state[k] == 0) {
                state[k] = 1;
                // push so that the larger-need child is evaluated first
                int32_t first = kids[0], second = kids[1];
                if (second >= 0 && (first < 0 || need[second] > need[first])) {
                    std::swap(first, second);
                }
                if (second >= 0 && state[second] == 0) stack.push_back(second);
                if (first >= 0 && state[first] == 0) stack.push_back(first);
                continue;
            }
            state[k] = 2;
            stack.pop_back();
            order.push_back(uint32_t(k));
        }
    }

    DeviceTape pre;
    for (uint32_t v : deck.vars) {
        pre.clauses.push_back(make(OP_LOADI, MODE_A_IMM, prereg(v), 0, 0,
                                   cval[v], v));
    }
    for (uint32_t k : order) {
        const RefClause& rc = deck.clauses[k];
        const int n = Opcode::args(Opcode::Opcode(rc.op));
        uint32_t m = MODE_RR, a = 0, b = 0;
        float im = 0.0f;
        const bool a_const = n >= 1 && kind[rc.a] == 1;
        const bool b_const = n >= 2 && kind[rc.b] == 1;
        if (a_const && b_const) {
            // Both operands constant (not produced by the optimiser, which
            // folds such nodes): materialise a.
            const uint32_t t = extra++;
            pre.clauses.push_back(make(OP_LOADI, MODE_A_IMM, t, 0, 0,
                                       cval[rc.a], rc.a));
            a = t;
            m = MODE_B_IMM;
            im = cval[rc.b];
        } else if (a_const) {
            m = MODE_A_IMM;
            im = cval[rc.a];
            if (n >= 2) b = prereg(rc.b);
        } else if (b_const) {
            m = MODE_B_IMM;
            im = cval[rc.b];
            a = prereg(rc.a);
        } else {
            if (n >= 1) a = prereg(rc.a);
            if (n >= 2) b = prereg(rc.b);
        }
        pre.clauses.push_back(make(rc.op, m, prereg(rc.id), a, b, im, rc.id));
    }
    if (kind[deck.root] == 1) {
        pre.clauses.push_back(make(OP_LOADI, MODE_A_IMM, prereg(deck.root), 0,
                                   0, cval[deck.root], deck.root));
    }
    pre.root_reg = prereg(deck.root);
    pre.num_regs = extra;

    std::vector<uint8_t> keep(pre.clauses.size(), KEEP_ALWAYS);
    for (size_t i = 0; i < pre.clauses.size(); ++i) {
        if (is_minmax(op(pre.clauses[i]))) keep[i] = KEEP_BOTH;
    }
    DeviceTape out = push_tape(pre, keep, nullptr);
    out.terminal = false;    // a base tape is never terminal (tape.hpp:83)
    return out;
}

}   // namespace libfive
