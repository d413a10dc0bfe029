// plan_io.hpp -- a plan written to a file and read back (plan reuse across processes).
#pragma once

#include <cstdint>
#include <string>

#include "plan.hpp"

namespace ifgf
{
// what the file records beside the plan: the operator parameters it was made for and the
// launch shape in use when it was written (48 bytes, no implicit padding)
struct PlanFileInfo
{
    int64_t leaf = 0;
    int32_t order = 0, n_elem = 0;
    int32_t shape[6] = {0, 0, 0, 0, 0, 0};
    uint8_t same_points = 0, merged = 0, had_cutoff = 0, host_plan = 0;
    int32_t reserved = 0;
};
static_assert(sizeof(PlanFileInfo) == 48, "file layout");

// "" on success, else the reason; plan_load leaves P untouched unless the whole file checks
std::string plan_save(const Plan& P, const PlanFileInfo& info, const char* path);
std::string plan_load(Plan& P, PlanFileInfo& info, const char* path);
// consistency of a plan's array sizes and indices (a file that passed its checksum but was
// not written by plan_save must not reach the upload): "" or the first problem found
std::string plan_check(const Plan& P);
} // namespace ifgf
