// Serialises HIP graph captures with the HIP calls that must not overlap a
// capture in another host thread.  HIP refuses a null-stream copy while a
// stream of the process is capturing, and allocations during a capture are
// not stream-ordered, so every capture in the library holds this lock from
// hipStreamBeginCapture to hipGraphInstantiate, and so do the allocations
// and null-stream copies on the solve paths that two host threads may run at
// once (the batch solve and result, the WaveHoltz fixed-point step).
#pragma once

#include <mutex>

namespace wdg {

std::mutex& capture_mutex();

}  // namespace wdg
