// wavefront.h — the wavefront pipeline (per-bounce ray queues) behind RT_MODE_WAVEFRONT.
#pragma once

#include <string>

#include "render.h"

namespace rt {

struct WavefrontState;

WavefrontState* wavefront_create();
void wavefront_destroy(WavefrontState* s);
// Renders the frame described by `fa` with queue kernels; same image as the
// megakernel.  Returns 0 or a hipError_t code (kWavefrontUnsupported when the
// wavefront path is not available for this frame).
constexpr int kWavefrontUnsupported = -1000;
// `variant` selects the traversal kernels' LDS stack depth and register target
// (0 .. wavefront_num_variants() - 1).
int launch_wavefront(WavefrontState* s, const FrameArgs& fa, hipStream_t st, bool any_hit, bool stats, int variant,
                     StageTimer* tm, std::string* err);
int wavefront_num_variants();
// Replay each chunk of queue iterations as one captured HIP graph.
void wavefront_set_graph(WavefrontState* s, bool on);
// Busy-lane thresholds below which a wave of the extend / shadow kernel takes new
// rays (1 .. 64).
void wavefront_set_refill(WavefrontState* s, uint32_t extend, uint32_t shadow);
// Number of slot sets (1 .. wavefront_max_streams()) traced concurrently on their own
// streams; stage timing uses one.
void wavefront_set_streams(WavefrontState* s, int n);
int wavefront_max_streams();
// Live-path count below which the rest of a slot set's paths are run to completion by
// one launch of megakernel-style threads (0: never).  Same images.
void wavefront_set_finish(WavefrontState* s, uint32_t live_paths);
// Renders `count` frames (host array) in one set of queue kernels: their slots are
// laid end to end and every kernel serves all of them.  Each frame's image is
// the one launch_wavefront gives it alone.  Frames with no samples or bounces
// are skipped (the caller renders them with the megakernel).
int launch_wavefront_batch(WavefrontState* s, const FrameArgs* frames, uint32_t count, hipStream_t st, bool any_hit,
                           bool stats, int variant, StageTimer* tm, std::string* err);

}  // namespace rt
