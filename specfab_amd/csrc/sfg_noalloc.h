// Internal hooks of the no-allocation entry point (sfg_noalloc.hip), defined
// in sfg_api.cpp.
#pragma once
#include <cstddef>

#include "../../include/specfab_gpu.h"

namespace sfg {
// the last-error message of the C-ABI; returns code
int set_error(int code, const char* msg);
int ctx_device(const sfg_ctx* ctx);
// sfg_evolve with the caller's workspace ws (ws_bytes bytes, device memory):
// allocates nothing; fails if the coupling tables of desc->L are not on the
// device yet or the batch needs more workspace than ws_bytes
int evolve_in_workspace(sfg_ctx* ctx, const sfg_evolve_desc* desc, void* stream, void* ws, size_t ws_bytes);
// (sfg_noalloc.hip) the device address of the no-allocation workspace on
// device dev, taken on the first call per device (out may be null)
int noalloc_setup(int dev, double** out);
}  // namespace sfg
