// fs_internal.hpp -- host-side declarations shared by the library's C++ translation units.
#pragma once

#include <mutex>
#include <string>

namespace fsim {

// Sets the message fs_last_error() returns on this thread.
void set_last_error(const std::string& message);

// Held around device allocation, freeing, synchronous null-stream copies and fills, and hipGraph
// capture, so that threads driving different grids never overlap one thread's capture with another's
// implicitly synchronising call.
std::mutex& device_setup_mutex();

// Checks a W x H solid mask (nonzero = solid): at least one fluid cell, and the fluid cells one
// 4-connected region.  Returns an empty string, or the reason it is refused.
std::string solid_mask_problem(const unsigned char* mask, int W, int H);

}  // namespace fsim
