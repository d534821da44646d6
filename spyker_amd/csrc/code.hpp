// Rank-order spike times shared by the dense and the sparse coders.
#pragma once

#include "common.hpp"

namespace spk
{
// times[i][k] = the U16 spike time of element k of sample i under rank
// coding (sort = true: by rank, src/spyker/cpu/code.cpp:86-118; false:
// latency, code.cpp:43-55); values >= time mean "no spike".
void rank_times_launch(const float *in, int64_t n, int64_t len, int64_t time, bool sort, uint16_t *times,
                       hipStream_t s);
}  // namespace spk
