// Entry points the library's own runtime uses between its operators; not part of the public C ABI
// in include/millgpu.h (no reference counterpart).
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "../../include/millgpu.h"

extern "C" {
// The DEM workspace's impact records: population d's last step's records start at d * stride.
const mg_impact_record* mg_dem_records_device(const mg_dem* dem, uint32_t* stride);
// StressAccum::update with drum d's records d_records[d_beg[d] .. d_end[d]) (device arrays).
int mg_stress_update_ranges(mg_stress* stress, const mg_impact_record* d_records, const uint32_t* d_beg,
                            const uint32_t* d_end, const mg_stress_step* steps, char* err, size_t errlen);
// One drum's accumulator values, debounce table and clock as an opaque blob (snapshots).
size_t mg_stress_blob_bytes(const mg_stress* stress);
int mg_stress_save(mg_stress* stress, uint32_t drum, void* out);
int mg_stress_load(mg_stress* stress, uint32_t drum, const void* in);
// StressAccum::default for one drum (the debounce factor is kept).
int mg_stress_reset(mg_stress* stress, uint32_t drum);
}
