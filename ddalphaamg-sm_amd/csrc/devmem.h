// Owning device and pinned host buffers (allocation in solver.hip).
#pragma once
#include <cstddef>

namespace dd {

// ---- memory ------------------------------------------------------------------------
class DevMem {
   public:
    DevMem() = default;
    explicit DevMem(size_t bytes) { alloc(bytes); }
    ~DevMem() { release(); }
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    DevMem(DevMem&& o) noexcept : p_(o.p_), bytes_(o.bytes_) { o.p_ = nullptr; o.bytes_ = 0; }
    DevMem& operator=(DevMem&& o) noexcept {
        if (this != &o) {
            release();
            p_ = o.p_;
            bytes_ = o.bytes_;
            o.p_ = nullptr;
            o.bytes_ = 0;
        }
        return *this;
    }
    void alloc(size_t bytes);
    void release();
    template <typename T>
    T* as() const { return reinterpret_cast<T*>(p_); }
    size_t bytes() const { return bytes_; }

   private:
    void* p_ = nullptr;
    size_t bytes_ = 0;
};

class PinMem {
   public:
    PinMem() = default;
    explicit PinMem(size_t bytes) { alloc(bytes); }
    ~PinMem() { release(); }
    PinMem(const PinMem&) = delete;
    PinMem& operator=(const PinMem&) = delete;
    void alloc(size_t bytes);
    void release();
    template <typename T>
    T* as() const { return reinterpret_cast<T*>(p_); }

   private:
    void* p_ = nullptr;
};

}  // namespace dd
