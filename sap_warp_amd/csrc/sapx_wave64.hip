// One-wave-per-world step kernel instantiated for 64 padded dofs (see sapx_wave_impl.h).
#include "sapx_wave_impl.h"

namespace sapx {
hipError_t launch_wave_np64(const StepParams& P, int cb, hipStream_t stream) {
    if (cb == 16) return wv::launch_variant<64, 16>(P, stream);
    if (cb == 32) return wv::launch_variant<64, 32>(P, stream);
    if (cb == 64) return wv::launch_variant<64, 64>(P, stream);
    return hipErrorInvalidValue;
}
}  // namespace sapx
