// nsparse/io/index_io.h:17-32: write_index / read_index on files, the
// reference's byte layout. An index read back is resident in HBM after its
// first search (or at once, for a SEISMIC index built here).
#ifndef NSPARSE_DROPIN_INDEX_IO_H
#define NSPARSE_DROPIN_INDEX_IO_H

#include "nsparse/index.h"
#include "nsparse/index_factory.h"

namespace nsparse {

enum IndexIoFlag {
    kUseMmap = 0x0001,
};

inline void write_index(Index* index, const char* file_name) {
    if (index == nullptr) throw std::invalid_argument("unexpected nullptr");
    detail::check(nsparse_write_index(index->handle(), file_name));
}
inline void write_index(Index* index, char* file_name) {
    write_index(index, static_cast<const char*>(file_name));
}

inline Index* read_index(const char* file_name, int io_flags = 0) {
    nsparse_index* h = nullptr;
    detail::check(nsparse_read_index(file_name, io_flags, &h));
    return detail::adopt_index(h);
}
inline Index* read_index(char* file_name, int io_flags = 0) {
    return read_index(static_cast<const char*>(file_name), io_flags);
}

}  // namespace nsparse

#endif  // NSPARSE_DROPIN_INDEX_IO_H
