// Explicit instantiation of the hip:: MD driver, the counterpart of
// `template class Driver<md::Phase, cuda::device_tag>;`
// (src/cuda/driver/driver.cu:202-204). Linked into programs that use
// ATV::Driver<md::Phase, hip::device_tag>.
#include "atomiverse/hip/driver.hpp"

namespace ATV {
template class Driver<md::Phase, hip::device_tag>;
} // namespace ATV
