// wavefront.hip — wavefront pipeline (placeholder until the queue kernels land).
#include "wavefront.h"

namespace rt {

struct WavefrontState {
    int unused;
};

WavefrontState* wavefront_create() { return new WavefrontState{0}; }
void wavefront_destroy(WavefrontState* s) { delete s; }

int launch_wavefront(WavefrontState*, const FrameArgs&, hipStream_t, bool, bool, std::string* err) {
    *err = "wavefront pipeline not available";
    return kWavefrontUnsupported;
}

}  // namespace rt
