// bcd_ctx.hpp -- the opaque context of the C ABI (one system on one device).
#pragma once

#include <hip/hip_runtime.h>

#include <memory>

#include "bcd_batch.hpp"
#include "bcd_twobody.hpp"

struct bcd_ctx {
  int device = 0;
  bcd::HostSystem sys;
  hipStream_t stream = nullptr;
  bcd::DevArena scratch;
  std::unique_ptr<bcd::T2System> t2;   // device tables of the two-body matrices, uploaded on first use
};
