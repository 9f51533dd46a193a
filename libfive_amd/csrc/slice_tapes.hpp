// Interval-pruned tapes of the slice renderers (the raster and the 2D
// contourer), held in the mesher's device clause pool layout
// (mesh_kernels.hpp): record 0 is the base tape, every other record a tape
// pushed for a box (Tape::push, libfive/src/eval/tape.cpp:23-158).  A host
// copy of the records serves Tape::getBase walks (tape.cpp:168-205).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "deck.hpp"
#include "mesh_kernels.hpp"
#include "runtime.hpp"

namespace libfive {
namespace amd {

class SliceTapes {
public:
    // Uploads the base tape; sx, sy, sz are the deck's X, Y, Z slots.
    SliceTapes(const DeviceTape& base, uint32_t sx, uint32_t sy, uint32_t sz,
               hipStream_t s);

    // IntervalEvaluator::intervalAndPush for each box (xlo, xhi, ylo, yhi,
    // zlo, zhi; 6 floats per box) with the tape of its entry in `tape`.
    // Boxes with equal consecutive tapes are evaluated together.  Per box:
    // bounds (lower, upper), nan (may be NaN) and out (the pushed tape, or
    // the box's own tape where the push would change nothing or is not
    // made).  Pushed when the bounds are neither above nor below zero, and
    // when the result may be NaN only if push_nan.  Synchronises.
    void eval(const std::vector<float>& boxes, const std::vector<uint32_t>& tape,
              bool push_nan, std::vector<float>& bounds, std::vector<uint8_t>& nan,
              std::vector<uint32_t>& out);

    // Tape::getBase: from tape id, the first interval tape whose box holds
    // [lo, hi] (inclusive), else the base tape.  Reads the host copy of the
    // records (refresh_records() after the last eval).
    uint32_t base_for(uint32_t id, const double lo[3], const double hi[3]) const;
    void refresh_records();

    const tape::Clause* pool() const { return pool_.p; }
    const mesh::TapeRec* records() const { return recs_.p; }
    uint32_t root() const { return P_.root_reg; }
    uint32_t nregs() const { return P_.nregs; }
    uint32_t max_len() const { return max_len_; }   // longest tape so far

private:
    hipStream_t s_;
    mesh::Params P_{};
    PoolBuf<tape::Clause> pool_;
    PoolBuf<mesh::TapeRec> recs_;
    PoolBuf<mesh::Counters> counters_;
    mesh::Counters hc_{};
    uint32_t max_len_ = 0;
    std::vector<mesh::TapeRec> host_;
};

}   // namespace amd
}   // namespace libfive
