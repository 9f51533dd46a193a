// Open-addressing hash map for the front end's node tables (node pointer or
// node key -> id / tree).  Keys are never erased; the table doubles at half
// load, so a lookup touches one or two adjacent slots.  Iteration order is
// never used, so results do not depend on the table's layout.
#pragma once

#include <cstddef>
#include <cstdint>
#include <functional>
#include <utility>
#include <vector>

namespace libfive {

// Mixes a pointer or a combined key hash so that aligned addresses spread
// over the low bits the table indexes with.
inline size_t mix_hash(uint64_t h) {
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ull;
    h ^= h >> 33;
    return size_t(h);
}

struct PtrHash {
    size_t operator()(const void* p) const { return mix_hash(uint64_t(uintptr_t(p))); }
};

template <typename K, typename V, typename H = PtrHash>
class FlatMap {
public:
    explicit FlatMap(size_t expected = 16) { rehash(capacity_for(expected)); }

    // The value stored for k, or nullptr.
    V* find(const K& k) {
        size_t i = H()(k) & mask_;
        while (used_[i]) {
            if (slots_[i].first == k) return &slots_[i].second;
            i = (i + 1) & mask_;
        }
        return nullptr;
    }
    const V* find(const K& k) const { return const_cast<FlatMap*>(this)->find(k); }

    // Inserts (k, v) unless k is present; returns the stored value and
    // whether it was inserted.
    std::pair<V*, bool> emplace(const K& k, V v) {
        if (2 * (count_ + 1) > slots_.size()) rehash(2 * slots_.size());
        size_t i = H()(k) & mask_;
        while (used_[i]) {
            if (slots_[i].first == k) return {&slots_[i].second, false};
            i = (i + 1) & mask_;
        }
        used_[i] = 1;
        slots_[i].first = k;
        slots_[i].second = std::move(v);
        ++count_;
        return {&slots_[i].second, true};
    }

    // The value for k, default-constructed on first use.
    V& operator[](const K& k) { return *emplace(k, V()).first; }

    size_t size() const { return count_; }

private:
    static size_t capacity_for(size_t n) {
        size_t c = 16;
        while (c < 2 * n) c *= 2;
        return c;
    }
    void rehash(size_t cap) {
        std::vector<std::pair<K, V>> old;
        std::vector<uint8_t> old_used;
        old.swap(slots_);
        old_used.swap(used_);
        slots_.resize(cap);
        used_.assign(cap, 0);
        mask_ = cap - 1;
        count_ = 0;
        for (size_t j = 0; j < old.size(); ++j) {
            if (!old_used[j]) continue;
            size_t i = H()(old[j].first) & mask_;
            while (used_[i]) i = (i + 1) & mask_;
            used_[i] = 1;
            slots_[i] = std::move(old[j]);
            ++count_;
        }
    }

    std::vector<std::pair<K, V>> slots_;
    std::vector<uint8_t> used_;
    size_t mask_ = 0;
    size_t count_ = 0;
};

}   // namespace libfive
