// One-wave-per-world step kernel instantiated for 32 padded dofs (see sapx_wave_impl.h).
#include "sapx_wave_impl.h"

namespace sapx {
hipError_t launch_wave_np32(const StepParams& P, int cb, hipStream_t stream) {
    if (cb == 16) return wv::launch_variant<32, 16>(P, stream);
    if (cb == 32) return wv::launch_variant<32, 32>(P, stream);
    if (cb == 64) return wv::launch_variant<32, 64>(P, stream);
    return hipErrorInvalidValue;
}
}  // namespace sapx
