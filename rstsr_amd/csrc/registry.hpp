// Extent checks of device operands against the allocations this library made.
#pragma once

#include <cstddef>
#include <cstdint>
#include <utility>

#include "layout.hpp"

namespace rstsr {
namespace hip {

void registry_add(void* p, size_t bytes);
bool registry_remove(void* p);
std::pair<uintptr_t, size_t> registry_find(const void* p);
// True when x and y lie in one registered buffer, or when either is not found.
bool same_buffer(const void* x, const void* y);
// Raise unless every element `l` addresses (elements of `elem` bytes from ptr)
// lies inside one buffer returned by rstsr_hip_malloc.
void check_extent(const void* ptr, const Layout& l, size_t elem, const char* name);
void check_extent_bytes(const void* ptr, size_t bytes, const char* name);
// Operand given by an element pointer and per-axis strides (possibly negative):
// every element ptr + sum_d i_d stride_d must lie inside one registered buffer.
void check_extent_rel(const void* ptr, size_t nd, const size_t* shape, const int64_t* stride, size_t elem,
                      const char* name);

}  // namespace hip
}  // namespace rstsr
