// nsparse/utils/scalar_quantizer.h:21-24: the quantizer widths (8-bit and
// 16-bit codes).
#ifndef NSPARSE_DROPIN_SCALAR_QUANTIZER_H
#define NSPARSE_DROPIN_SCALAR_QUANTIZER_H

#include <cstdint>

namespace nsparse {

enum class QuantizerType : uint8_t {
    QT_8bit,
    QT_16bit,
};

}  // namespace nsparse

#endif  // NSPARSE_DROPIN_SCALAR_QUANTIZER_H
