// Interval-pruned tapes of the slice renderers (see slice_tapes.hpp).
#include "slice_tapes.hpp"

#include <algorithm>

namespace libfive {
namespace amd {

namespace {

// Replaces b by a buffer of at least `need` elements holding its first
// `used` ones.  The stream is idle between evaluations, so the old block
// can go back to the cache at once.
template <typename T>
void grow(PoolBuf<T>& b, size_t used, size_t need, hipStream_t s) {
    if (need <= b.n) return;
    PoolBuf<T> nb(std::max(need, 2 * b.n));
    if (used) {
        hip_check(hipMemcpyAsync(nb.p, b.p, used * sizeof(T), hipMemcpyDeviceToDevice, s),
                  "grow tape storage");
    }
    hip_check(hipStreamSynchronize(s), "grow tape storage");
    b.swap(nb);
}

}   // namespace

SliceTapes::SliceTapes(const DeviceTape& base, uint32_t sx, uint32_t sy, uint32_t sz,
                       hipStream_t s)
    : s_(s) {
    const uint32_t n = uint32_t(base.clauses.size());
    const uint32_t padded = tape::padded_length(n);
    pool_.alloc(std::max<size_t>(size_t(padded) * 16, size_t(1) << 16));
    std::vector<tape::Clause> h(padded, tape::Clause{0, 0, 0, 0});
    std::copy(base.clauses.begin(), base.clauses.end(), h.begin());
    hip_check(hipMemcpyAsync(pool_.p, h.data(), h.size() * sizeof(tape::Clause),
                             hipMemcpyHostToDevice, s), "upload tape");
    recs_.alloc(1024);
    mesh::TapeRec r{};
    r.off = 0;
    r.len = n;
    r.parent = mesh::kNoTape;
    r.flags = 0;
    r.ref_len = base.ref_length;
    hip_check(hipMemcpyAsync(recs_.p, &r, sizeof(r), hipMemcpyHostToDevice, s),
              "upload tape record");
    counters_.alloc(1);
    hc_.pool_top = padded;
    hc_.tape_count = 1;
    hc_.max_len = n;
    hip_check(hipMemcpyAsync(counters_.p, &hc_, sizeof(hc_), hipMemcpyHostToDevice, s),
              "tape counters");
    hip_check(hipStreamSynchronize(s), "upload tape");
    max_len_ = n;
    host_.assign(1, r);

    P_.nregs = std::max<uint32_t>(base.num_regs, tape::kFirstFreeReg);
    P_.root_reg = base.root_reg;
    P_.slot_x = sx;
    P_.slot_y = sy;
    P_.slot_z = sz;
    P_.counters = counters_.p;
}

void SliceTapes::eval(const std::vector<float>& boxes, const std::vector<uint32_t>& tape,
                      bool push_nan, std::vector<float>& bounds, std::vector<uint8_t>& nan,
                      std::vector<uint32_t>& out) {
    const size_t nb = tape.size();
    bounds.resize(2 * nb);
    nan.resize(nb);
    out.resize(nb);
    if (nb == 0) return;
    if (boxes.size() != 6 * nb) throw Error("interval boxes: size mismatch");

    std::vector<mesh::BoxGroup> groups;
    for (size_t k = 0; k < nb;) {
        size_t e = k + 1;
        while (e < nb && e - k < 8 && tape[e] == tape[k]) ++e;
        groups.push_back({tape[k], uint32_t(k), uint32_t(e - k), 0});
        k = e;
    }

    // Every box may push one tape of at most max_len + kPushExtra clauses.
    const uint64_t need_pool = uint64_t(hc_.pool_top) + nb * (uint64_t(max_len_) + mesh::kPushExtra);
    const uint64_t need_recs = uint64_t(hc_.tape_count) + nb;
    if (need_pool > 0xffffffffull || need_recs > 0xffffffffull) {
        throw Error("interval pruning: tape storage exceeds 2^32 clauses");
    }
    grow(pool_, hc_.pool_top, size_t(need_pool), s_);
    grow(recs_, hc_.tape_count, size_t(need_recs), s_);
    P_.pool = pool_.p;
    P_.pool_cap = uint32_t(pool_.n);
    P_.tapes = recs_.p;
    P_.tape_cap = uint32_t(std::min<size_t>(recs_.n, 0xffffffffu));

    PoolBuf<mesh::BoxGroup> dg(groups.size());
    PoolBuf<float> db(6 * nb), dbo(2 * nb);
    PoolBuf<uint8_t> dn(nb), ws(mesh::interval_boxes_workspace_bytes(P_.nregs, max_len_));
    PoolBuf<uint32_t> dt(nb);
    hip_check(hipMemcpyAsync(dg.p, groups.data(), groups.size() * sizeof(mesh::BoxGroup),
                             hipMemcpyHostToDevice, s_), "copy box groups");
    hip_check(hipMemcpyAsync(db.p, boxes.data(), 24 * nb, hipMemcpyHostToDevice, s_),
              "copy boxes");
    hip_check(mesh::launch_interval_boxes(P_, dg.p, uint32_t(groups.size()), db.p, dbo.p, dn.p,
                                          dt.p, push_nan, max_len_, ws.p, ws.n, s_),
              "interval pruning");
    hip_check(hipMemcpyAsync(bounds.data(), dbo.p, 8 * nb, hipMemcpyDeviceToHost, s_),
              "copy bounds");
    hip_check(hipMemcpyAsync(nan.data(), dn.p, nb, hipMemcpyDeviceToHost, s_), "copy nan");
    hip_check(hipMemcpyAsync(out.data(), dt.p, 4 * nb, hipMemcpyDeviceToHost, s_),
              "copy tapes");
    hip_check(hipMemcpyAsync(&hc_, counters_.p, sizeof(hc_), hipMemcpyDeviceToHost, s_),
              "copy tape counters");
    hip_check(hipStreamSynchronize(s_), "interval pruning");
    if (hc_.error & mesh::ERR_POOL) throw Error("interval pruning: tape storage exhausted");
    if (hc_.error & mesh::ERR_TAPES) throw Error("interval pruning: tape records exhausted");
    if (hc_.error) throw Error("interval pruning: device error");
    max_len_ = std::max(max_len_, hc_.max_len);
}

void SliceTapes::refresh_records() {
    host_.resize(hc_.tape_count);
    hip_check(hipMemcpyAsync(host_.data(), recs_.p, host_.size() * sizeof(mesh::TapeRec),
                             hipMemcpyDeviceToHost, s_), "copy tape records");
    hip_check(hipStreamSynchronize(s_), "copy tape records");
}

uint32_t SliceTapes::base_for(uint32_t id, const double lo[3], const double hi[3]) const {
    // A tape's parent always has a smaller id, so the walk ends at the base.
    while (id < host_.size() && host_[id].parent != mesh::kNoTape) {
        const mesh::TapeRec& r = host_[id];
        if ((r.flags & mesh::TAPE_INTERVAL) &&
            lo[0] >= double(r.box[0]) && hi[0] <= double(r.box[1]) &&
            lo[1] >= double(r.box[2]) && hi[1] <= double(r.box[3]) &&
            lo[2] >= double(r.box[4]) && hi[2] <= double(r.box[5])) {
            break;
        }
        id = r.parent;
    }
    if (id >= host_.size()) throw Error("tape walk: unknown tape record");
    return id;
}

}   // namespace amd
}   // namespace libfive
