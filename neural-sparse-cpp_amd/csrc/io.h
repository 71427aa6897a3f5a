// Byte streams for the reference's on-disk layout (io/io.h:18-43,
// io/align.h:43-104): a writer/reader pair that tracks the absolute stream
// position, because the reference pads every array to its element width from
// that position.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace nsparse_amd {

class ByteWriter {
public:
    ByteWriter() = default;
    explicit ByteWriter(std::FILE* f) : file_(f) {}
    void write(const void* p, size_t n) {
        if (n == 0) return;
        if (file_ != nullptr) {
            if (std::fwrite(p, 1, n, file_) != n) throw std::runtime_error("write to index file failed");
        } else {
            const auto* b = static_cast<const uint8_t*>(p);
            buf_.insert(buf_.end(), b, b + n);
        }
        pos_ += n;
    }
    template <class T>
    void scalar(T v) {
        write(&v, sizeof(T));
    }
    // io_align::pad_to: zeros up to the next multiple of `align`.
    void pad(size_t align) {
        if (align == 0) throw std::runtime_error("index file declares a zero element width");
        static const uint8_t zeros[16] = {};
        const size_t n = (align - pos_ % align) % align;
        write(zeros, n);
    }
    // io_align::write_padded: pads unconditionally, even for empty arrays.
    template <class T>
    void padded(const T* p, size_t count, size_t align = alignof(T)) {
        pad(align);
        write(p, count * sizeof(T));
    }
    size_t pos() const { return pos_; }
    std::vector<uint8_t>& buffer() { return buf_; }

private:
    std::FILE* file_ = nullptr;
    std::vector<uint8_t> buf_;
    size_t pos_ = 0;
};

class ByteReader {
public:
    ByteReader(const uint8_t* data, size_t size) : data_(data), size_(size) {}
    void read(void* dst, size_t n) {
        if (n > size_ - pos_) throw std::runtime_error("index file is truncated");
        std::memcpy(dst, data_ + pos_, n);
        pos_ += n;
    }
    template <class T>
    T scalar() {
        T v;
        read(&v, sizeof(T));
        return v;
    }
    void skip_pad(size_t align) {
        if (align == 0) throw std::runtime_error("index file declares a zero element width");
        const size_t n = (align - pos_ % align) % align;
        if (n > size_ - pos_) throw std::runtime_error("index file is truncated");
        pos_ += n;
    }
    template <class T>
    std::vector<T> padded(size_t count, size_t align = alignof(T)) {
        skip_pad(align);
        if (count > (size_ - pos_) / sizeof(T)) throw std::runtime_error("index file is truncated");
        std::vector<T> v(count);
        read(v.data(), count * sizeof(T));
        return v;
    }
    size_t pos() const { return pos_; }
    size_t remaining() const { return size_ - pos_; }

private:
    const uint8_t* data_;
    size_t size_;
    size_t pos_ = 0;
};

}  // namespace nsparse_amd
