/* An open-addressing hash map from object pointers to int, for the per-flush passes over tens of
 * thousands of instructions (first-writer flags, unused frees, base ordinals): one flat array,
 * linear probing, no allocation per element. */
#pragma once

#include <cstdint>
#include <utility>
#include <vector>

namespace bohrium {

class FlatPtrMap {
public:
    /* Room for `n` keys without growing. */
    explicit FlatPtrMap(size_t n = 16) { reset(n); }

    void reset(size_t n) {
        size_t cap = 16;
        while (cap < 2 * n) cap <<= 1;
        keys_.assign(cap, nullptr);
        vals_.assign(cap, 0);
        mask_ = cap - 1;
        size_ = 0;
    }

    size_t size() const { return size_; }

    /* The value of `key`, or nullptr when absent. */
    const int *find(const void *key) const {
        for (size_t i = slot(key);; i = (i + 1) & mask_) {
            if (keys_[i] == key) return &vals_[i];
            if (keys_[i] == nullptr) return nullptr;
        }
    }

    /* Inserts `key` with `value` when absent; returns the value stored for `key` and whether it
     * was inserted now. */
    std::pair<int, bool> insert(const void *key, int value) {
        if (2 * (size_ + 1) > keys_.size()) grow();
        for (size_t i = slot(key);; i = (i + 1) & mask_) {
            if (keys_[i] == key) return {vals_[i], false};
            if (keys_[i] == nullptr) {
                keys_[i] = key;
                vals_[i] = value;
                ++size_;
                return {value, true};
            }
        }
    }

private:
    size_t slot(const void *key) const {
        // Fibonacci hashing of the address (allocations are aligned: the low bits carry nothing).
        const uint64_t h = (static_cast<uint64_t>(reinterpret_cast<uintptr_t>(key)) >> 4) * 0x9E3779B97F4A7C15ull;
        return static_cast<size_t>(h >> 20) & mask_;
    }

    void grow() {
        std::vector<const void *> keys;
        std::vector<int> vals;
        keys.swap(keys_);
        vals.swap(vals_);
        reset(keys.size());
        for (size_t i = 0; i < keys.size(); ++i)
            if (keys[i] != nullptr) insert(keys[i], vals[i]);
    }

    std::vector<const void *> keys_;
    std::vector<int> vals_;
    size_t mask_ = 0, size_ = 0;
};

} // namespace bohrium
