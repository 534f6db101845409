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

// Device memory cache.  Buffers that grids release are kept per device and size class and handed out
// again, so creating and retiring instances (the serving loop does both continually) reaches neither
// hipMalloc nor hipFree -- whose implicit device-wide synchronisation would otherwise stall every lane --
// once the cache is warm.  A caller releases memory only when no queued work still uses it (grids drain
// their stream first).  Zero-fills and host uploads of set-up data run on a per-device non-blocking
// stream and are waited for, so they neither wait for nor block the lanes' streams.
void* device_alloc(size_t bytes);
void device_free(void* p, size_t bytes);
void device_zero(void* p, size_t bytes);
void device_upload(void* dst, const void* src, size_t bytes);
// Small pinned host blocks (scalar read-back slots), cached the same way: hipHostFree also synchronises.
void* pinned_alloc(size_t bytes);
void pinned_free(void* p, size_t bytes);

// Checks a W x H solid mask (nonzero = solid): at least one fluid cell, and the fluid cells one
// 4-connected region.  Returns an empty string, or the reason it is refused.
std::string solid_mask_problem(const unsigned char* mask, int W, int H);

}  // namespace fsim
