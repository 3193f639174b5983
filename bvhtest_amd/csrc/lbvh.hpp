// Device-side parameters of one LBVH build (scene box and Morton mapping).
#pragma once

#include "runtime.hpp"

namespace bvht {

struct BuildParams {
  float sceneMin[3];
  float sceneMax[3];
  float offset[3];  // 0 - sceneMin          (LBVH_CUDA.cu:172)
  float scale[3];   // 1 / (sceneMax + offset) (LBVH_CUDA.cu:178-180)
  uint32_t done;    // finished blocks of the scene-box reduction (staged calls)
  uint32_t pad[3];
};

}  // namespace bvht
