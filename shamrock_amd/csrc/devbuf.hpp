// SPDX-License-Identifier: CeCILL-2.1
//
// devbuf.hpp -- grow-only device arrays.  Transient buffers of the step (merged fields, tree,
// neighbour cache, sort ping-pong) are kept across steps and only reallocated when a larger
// size is requested, so a steady-state step performs no hipMalloc.
#pragma once

#include "common.hpp"

#include <stdexcept>
#include <utility>

namespace shamx {

    template<class T>
    class DevArray {
        T *ptr_         = nullptr;
        size_t cap_     = 0;
        size_t size_    = 0;

        public:
        DevArray() = default;
        explicit DevArray(size_t n) { resize(n); }
        DevArray(const DevArray &)            = delete;
        DevArray &operator=(const DevArray &) = delete;
        DevArray(DevArray &&o) noexcept { swap(o); }
        DevArray &operator=(DevArray &&o) noexcept {
            swap(o);
            return *this;
        }
        ~DevArray() { release(); }

        void swap(DevArray &o) noexcept {
            std::swap(ptr_, o.ptr_);
            std::swap(cap_, o.cap_);
            std::swap(size_, o.size_);
        }
        void release() {
            if (ptr_) (void) hipFree(ptr_);
            ptr_ = nullptr;
            cap_ = size_ = 0;
        }
        // contents are not preserved when the capacity grows
        void resize(size_t n) {
            if (n > cap_) {
                release();
                size_t c = n + n / 8 + 64;
                SHAMX_HIP(hipMalloc((void **) &ptr_, c * sizeof(T)));
                cap_ = c;
            }
            size_ = n;
        }
        // grows keeping the first size() elements (device to device copy on the stream)
        void resize_keep(size_t n, hipStream_t s) {
            if (n > cap_) {
                DevArray<T> tmp(n);
                if (size_)
                    SHAMX_HIP(hipMemcpyAsync(tmp.ptr_, ptr_, size_ * sizeof(T),
                                             hipMemcpyDeviceToDevice, s));
                SHAMX_HIP(hipStreamSynchronize(s));
                swap(tmp);
            }
            size_ = n;
        }
        T *data() { return ptr_; }
        const T *data() const { return ptr_; }
        size_t size() const { return size_; }
        size_t capacity() const { return cap_; }
        size_t bytes() const { return size_ * sizeof(T); }
    };

    // N pinned host words for small device-to-host reads (sizes, counts, reductions).  A read
    // into pageable memory is staged through a runtime buffer and blocks the host by itself;
    // a read into pinned memory is queued on the stream like a kernel, so the reads of one
    // decision share a single stream synchronisation.  Allocated on first use.
    template<class T, size_t N>
    class PinnedWords {
        T *ptr_ = nullptr;

        public:
        PinnedWords() = default;
        PinnedWords(const PinnedWords &)            = delete;
        PinnedWords &operator=(const PinnedWords &) = delete;
        PinnedWords(PinnedWords &&o) noexcept : ptr_(std::exchange(o.ptr_, nullptr)) {}
        PinnedWords &operator=(PinnedWords &&o) noexcept {
            std::swap(ptr_, o.ptr_);
            return *this;
        }
        ~PinnedWords() {
            if (ptr_) (void) hipHostFree(ptr_);
        }
        T *data() {
            if (!ptr_) SHAMX_HIP(hipHostMalloc((void **) &ptr_, N * sizeof(T), hipHostMallocDefault));
            return ptr_;
        }
        T &operator[](size_t i) { return data()[i]; }
        static constexpr size_t size() { return N; }
        // queue a read of n <= N words from device memory (synchronise the stream before use)
        void read_async(hipStream_t s, const T *src, size_t n, size_t at = 0) {
            if (at + n > N) throw std::logic_error("PinnedWords: read past the end");
            if (n) SHAMX_HIP(hipMemcpyAsync(data() + at, src, n * sizeof(T), hipMemcpyDeviceToHost, s));
        }
    };

} // namespace shamx
