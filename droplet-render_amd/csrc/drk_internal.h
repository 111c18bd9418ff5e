// Host-internal declarations shared by the library's translation units.
#pragma once

#include <stdint.h>
#include "droplet_render.h"

namespace drk {

// Records the thread-local error text and returns code.
int set_error(int code, const char *fmt, ...);

int scene_build_dense(const float *fog, const float *sdf, uint32_t nx, uint32_t ny, uint32_t nz,
                      float voxel_size, uint32_t maxd, drk_host_scene **out);
void host_scene_free(drk_host_scene *hs);

// Hosek-Wilkie RGB state (double), see hosek.cpp.
void hosek_state(double turbidity, double albedo, double elevation, double cfg[3][9], double rad[3]);

}  // namespace drk
