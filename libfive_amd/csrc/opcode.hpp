// Opcode numbering for f-rep expression trees and evaluation tapes.
//
// The numeric values are a stable on-disk / on-tape encoding and match the
// reference's default (unpacked) numbering in
// libfive/include/libfive/tree/opcode.hpp:23-61, including LAST_OP = 33.
// Argument counts follow libfive/src/tree/opcode.cpp:18-65, names and
// commutativity follow libfive/src/tree/opcode.cpp:80-273.
#pragma once
#ifndef __HIPCC_RTC__
#include <cstdint>
#include <string>
#endif

namespace libfive {
namespace Opcode {

enum Opcode : int {
    INVALID = 0,
    CONSTANT = 1,
    VAR_X = 2,
    VAR_Y = 3,
    VAR_Z = 4,
    VAR_FREE = 5,
    CONST_VAR = 6,
    OP_SQUARE = 7,
    OP_SQRT = 8,
    OP_NEG = 9,
    OP_SIN = 10,
    OP_COS = 11,
    OP_TAN = 12,
    OP_ASIN = 13,
    OP_ACOS = 14,
    OP_ATAN = 15,
    OP_EXP = 16,
    OP_ADD = 17,
    OP_MUL = 18,
    OP_MIN = 19,
    OP_MAX = 20,
    OP_SUB = 21,
    OP_DIV = 22,
    OP_ATAN2 = 23,
    OP_POW = 24,
    OP_NTH_ROOT = 25,
    OP_MOD = 26,
    OP_NANFILL = 27,
    OP_ABS = 28,
    OP_RECIP = 29,
    OP_LOG = 30,
    OP_COMPARE = 31,
    ORACLE = 32,
    LAST_OP = 33,
};

#ifndef __HIPCC_RTC__
// Number of tree children for an opcode, or -1 for INVALID / LAST_OP.
int args(Opcode op);

// Bare enum name without the OP_ prefix (e.g. "ATAN2").
std::string toString(Opcode op);

// Lower-case, dash-separated name (e.g. "nth-root"), as used by the
// reference's Scheme bindings and libfive_opcode_enum().
std::string toScmString(Opcode op);
Opcode fromScmString(std::string s);

// Symbol used when printing trees in prefix form ("+", "sqrt", "x", ...).
std::string toOpString(Opcode op);

bool isCommutative(Opcode op);
bool isIdempotent(Opcode op);
#endif

}   // namespace Opcode
}   // namespace libfive
