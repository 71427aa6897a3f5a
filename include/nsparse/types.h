// C++ drop-in surface of neural-sparse-cpp over the C ABI (nsparse_c.h): the
// reference's public headers under the same include paths, so code written
// against nsparse::Index compiles unchanged and runs on the device library.
// This header: nsparse/types.h:16-33 (idx_t, term_t, weight_t, Residency).
#ifndef NSPARSE_DROPIN_TYPES_H
#define NSPARSE_DROPIN_TYPES_H

#include <cstdint>
#include <utility>
#include <vector>

namespace nsparse {

using idx_t = int32_t;
using term_t = uint16_t;
using weight_t = float;

// Whether a loaded index owns its buffers or borrows them from a file mapping.
enum class Residency : uint8_t {
    kInMemory,
    kMmap,
};

using pair_of_score_id_vectors_t =
    std::pair<std::vector<std::vector<float>>, std::vector<std::vector<idx_t>>>;

namespace detail {
constexpr idx_t INVALID_IDX = -1;
}

}  // namespace nsparse

#endif  // NSPARSE_DROPIN_TYPES_H
