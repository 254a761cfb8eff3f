// Operator::Page: a fixed-width tuple buffer with the TupleBuffer contract
// (util/buffer.h:37-337, util/buffer.inl:36-111):
//   allocateTuple() bumps `free` while free + tuplesize <= data + maxsize,
//   getTupleOffset(i) = data + i*tuplesize, or NULL at/after `free`,
//   an overlay page (free == NULL at construction) is full,
//   Iterator::next() walks tuples until NULL.
#pragma once

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <new>

class TupleBuffer {
  public:
    // Owning page of `size` bytes.
    TupleBuffer(uint64_t size, unsigned int tuplesize)
        : data_((char*)std::malloc(size ? size : 1)), maxsize_(size), owner_(true), tuplesize_(tuplesize) {
        if (!data_) throw std::bad_alloc();
        free_ = data_;
    }
    // Overlay of caller memory; free == nullptr means "full".
    TupleBuffer(void* data, uint64_t size, void* free, unsigned int tuplesize)
        : data_((char*)data), maxsize_(size), owner_(false), tuplesize_(tuplesize) {
        free_ = free ? (char*)free : data_ + size;
    }
    ~TupleBuffer() {
        if (owner_) std::free(data_);
    }
    TupleBuffer(const TupleBuffer&) = delete;
    TupleBuffer& operator=(const TupleBuffer&) = delete;

    bool isValidAddress(const void* loc, uint64_t len) const {
        return data_ <= (const char*)loc && (const char*)loc <= data_ + maxsize_ - len;
    }
    bool canStore(uint64_t len) const { return maxsize_ >= len && isValidAddress(free_, len); }
    bool canStoreTuple() const { return canStore(tuplesize_); }
    void* allocate(uint64_t len) {
        if (len == 0) return free_;
        if (!canStore(len)) return nullptr;
        void* r = free_;
        free_ += len;
        return r;
    }
    void* allocateTuple() { return allocate(tuplesize_); }
    void* getTupleOffset(uint64_t pos) const {
        char* r = data_ + pos * tuplesize_;
        return r < free_ ? r : nullptr;
    }
    void clear() { free_ = data_; }
    uint64_t getUsedSpace() const { return (uint64_t)(free_ - data_); }
    uint64_t getNumTuples() const { return getUsedSpace() / tuplesize_; }
    uint64_t capacity() const { return maxsize_; }
    unsigned int getTupleSize() const { return tuplesize_; }
    char* data() const { return data_; }
    // Marks the first `bytes` bytes as used (for bulk fills).
    void setUsed(uint64_t bytes) { free_ = data_ + bytes; }

    class Iterator {
      public:
        Iterator() = default;
        explicit Iterator(const TupleBuffer* p) : page_(p) {}
        void place(const TupleBuffer* p) {
            page_ = p;
            id_ = 0;
        }
        void* next() { return page_ ? page_->getTupleOffset(id_++) : nullptr; }
        void reset() { id_ = 0; }

      private:
        uint64_t id_ = 0;
        const TupleBuffer* page_ = nullptr;
    };
    Iterator createIterator() const { return Iterator(this); }

  private:
    char* data_;
    uint64_t maxsize_;
    bool owner_;
    unsigned int tuplesize_;
    char* free_;
};
