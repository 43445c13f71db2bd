// common.hpp -- launch configuration, warp/buffer constants and error
// handling of the C++ drop-in layer.  Counterpart of src/common.cuh:8-209.
//
// The drop-in layer is plain C++ (built with g++): it reaches the GPU only
// through the C ABI of libmgv.so (include/mgv.h).
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "../../../../include/mgv.h"

// Lanes per wavefront on CDNA (the reference hard-codes 32, common.cuh:25).
#define WARPSIZE (64)

#ifndef DEFAULT_BUFF_CAP
#define DEFAULT_BUFF_CAP (4 * 1024 * 1024)
#endif

namespace mgv {

struct error : std::runtime_error {
    int code;
    error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// Throws mgv::error for a failed C-ABI call (the reference's gpu() printed
// and called exit(), common.cuh:12-23).
inline void check(int rc, const char *what = "") {
    if (rc != MGV_OK) throw error(rc, std::string(what) + ": " + mgv_last_error_string());
}

}  // namespace mgv

#define gpu(ans) ::mgv::check((ans), #ans)

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned x_ = 1, unsigned y_ = 1, unsigned z_ = 1) : x(x_), y(y_), z(z_) {}
};

// launch_conf{gridDim, blockDim, shared_mem, device} (common.cuh:183-209).
// total_num_of_warps is blocks x warps per block (the reference multiplied
// threads by warps, common.cuh:206-208).
struct launch_conf {
    dim3 gridDim;
    dim3 blockDim;
    int shared_mem;
    int device;

    launch_conf() : shared_mem(0), device(-5) {}
    launch_conf(const dim3 &g, const dim3 &b, int shm, int dev) : gridDim(g), blockDim(b), shared_mem(shm), device(dev) {}

    int get_blocks_per_grid() const { return (int)(gridDim.x * gridDim.y * gridDim.z); }
    int get_threads_per_block() const { return (int)(blockDim.x * blockDim.y * blockDim.z); }
    int get_warps_per_block() const { return (get_threads_per_block() + WARPSIZE - 1) / WARPSIZE; }
    int total_num_of_warps() const { return get_blocks_per_grid() * get_warps_per_block(); }

    // The launch shape handed to libmgv: blocks of 64, 256 or 1024 threads.
    mgv_launch to_launch(int compact = -1) const {
        const int t = get_threads_per_block();
        const uint32_t b = t >= 1024 ? 1024u : (t >= 256 ? 256u : 64u);
        return mgv_launch{(uint32_t)get_blocks_per_grid(), b, compact, 0u};
    }
};
