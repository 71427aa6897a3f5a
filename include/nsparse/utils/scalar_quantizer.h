// nsparse/utils/scalar_quantizer.h:21-24: the quantizer widths. The device
// path implements QT_8bit; an index asked for QT_16bit is refused with
// std::runtime_error when it is made.
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
