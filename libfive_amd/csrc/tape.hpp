// Device tape format, shared by host code and HIP kernels.
//
// A tape is a straight-line register program in evaluation order
// (leaves first).  Each clause is 16 bytes:
//   w0 = opcode (bits 0-7) | operand mode (bits 8-9) | binary flag (bit 10)
//        | destination register (bits 16-31)
//   w1 = register of operand a (bits 0-15) | register of operand b (16-31)
//   w2 = immediate operand (float bits) when the mode says so
//   w3 = the reference clause id (Deck slot) whose value this clause writes
// Registers 0, 1, 2 always hold x, y, z and are never written by a clause.
//
// The reference keeps one result slot per tree node and evaluates the
// root-first clause list backwards (libfive/src/eval/deck.cpp:19-94,
// libfive/src/eval/eval_array.cpp:107-118).  Here the same clauses, in the
// same order, are renamed onto a small register file by a single backward
// pass (assign a register at a value's last use, release it at its
// definition), and constants become immediates, so that the per-point
// state fits in VGPRs.
#pragma once

#ifndef __HIPCC_RTC__
#include <cstdint>
#endif

#include "fmath.hpp"

namespace libfive {
namespace tape {

struct Clause {
    uint32_t w0, w1, w2, w3;
};
static_assert(sizeof(Clause) == 16, "clause layout");

enum : uint32_t {
    MODE_RR = 0,        // a and b (if any) are registers
    MODE_A_IMM = 1,     // a is the immediate
    MODE_B_IMM = 2,     // b is the immediate
};

// Opcodes that exist only on device tapes.  CONSTANT (1) is reused as
// "load immediate into dst"; INVALID (0) is a no-op used to pad tapes to
// whole fetch batches.
enum : uint32_t {
    OP_NOP = 0,
    OP_LOADI = 1,
    OP_COPY = 40,
};

// Clauses are fetched in batches of this many (one 64-byte scalar load);
// device buffers hold the tape padded with NOPs to a multiple of the batch
// plus one extra batch, so the prefetch of the next batch never reads past
// the allocation.
constexpr uint32_t kFetch = 4;
inline uint32_t padded_length(uint32_t n) {
    return (n + kFetch - 1) / kFetch * kFetch + kFetch;
}

constexpr uint32_t kRegX = 0, kRegY = 1, kRegZ = 2, kFirstFreeReg = 3;

// Value written by the call-free evaluation kernels for a point they could
// not evaluate (a sin / cos argument of 2^17 or above), which a second
// kernel then evaluates with the general rules.  It is a signalling NaN,
// which no arithmetic produces; a point whose true value has these bits
// (an input NaN passed through unchanged) is evaluated twice, with the
// same result.
constexpr uint32_t kRedoMark = 0x7fa5a5a5u;
constexpr uint32_t kMaxRegs = 0xffff;

// Operand count of a device-tape opcode (Opcode::args for tree opcodes,
// libfive/src/tree/opcode.cpp:18-65; 0 for LOADI, 1 for COPY).
LF_HD int nargs(uint32_t o) {
    switch (o) {
        case OP_LOADI: return 0;
        case OP_COPY: return 1;
        case 6: case 7: case 8: case 9: case 10: case 11: case 12: case 13:
        case 14: case 15: case 16: case 28: case 29: case 30:
            return 1;
        case 17: case 18: case 19: case 20: case 21: case 22: case 23: case 24:
        case 25: case 26: case 27: case 31:
            return 2;
        default:
            return 0;
    }
}

LF_HD uint32_t op(const Clause& c) { return c.w0 & 0xffu; }
LF_HD uint32_t mode(const Clause& c) { return (c.w0 >> 8) & 3u; }
LF_HD bool is_binary(const Clause& c) { return (c.w0 >> 10) & 1u; }
LF_HD uint32_t dst(const Clause& c) { return c.w0 >> 16; }
LF_HD uint32_t ra(const Clause& c) { return c.w1 & 0xffffu; }
LF_HD uint32_t rb(const Clause& c) { return c.w1 >> 16; }
LF_HD float imm(const Clause& c) { return fmath::u2f(c.w2); }
LF_HD uint32_t slot(const Clause& c) { return c.w3; }

LF_HD Clause make(uint32_t op_, uint32_t mode_, uint32_t dst_, uint32_t a,
                  uint32_t b, float im, uint32_t slot_) {
    Clause c;
    c.w0 = (op_ & 0xffu) | ((mode_ & 3u) << 8) |
           (nargs(op_) == 2 ? (1u << 10) : 0u) | (dst_ << 16);
    c.w1 = (a & 0xffffu) | (b << 16);
    c.w2 = fmath::f2u(im);
    c.w3 = slot_;
    return c;
}

// The opcodes of the mesher's call-free evaluators (the common ones: IEEE
// arithmetic, min / max, sqrt, sin / cos and the tape's own LOADI / COPY).
LF_HD bool fast_opcode(uint32_t o) {
    switch (o) {
        case 17: case 18: case 19: case 20: case 21: case 9: case 7: case 28:
        case 8: case 10: case 11: case 22: case 6:
        case OP_LOADI: case OP_COPY: case OP_NOP:
            return true;
        default:
            return false;
    }
}

// Keep decision for a MIN/MAX clause (Tape::Keep, tape.hpp:38-41).
enum : uint8_t { KEEP_BOTH = 0, KEEP_A = 1, KEEP_B = 2, KEEP_ALWAYS = 3 };

}   // namespace tape
}   // namespace libfive
