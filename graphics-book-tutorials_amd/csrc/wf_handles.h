// Opaque handle types behind the C ABI (include/wfamd.h) and the shared last-error slot.
#pragma once
#include <cstdint>
#include <string>

#include "../../include/wfamd.h"
#include "wf_scene.h"

namespace wf {
// A number no other scene handle of the process has (scene caches key on it, not on the address,
// which a later scene may reuse).
uint64_t next_scene_uid();
}  // namespace wf

struct wf_scene {
    wf::Scene s;
    const uint64_t uid = wf::next_scene_uid();
    // wf_scene_info of s, computed on first request (the BVH statistics and the scene hash walk
    // every node and primitive); cleared when the scene changes (wf_scene_orbit).
    mutable bool info_valid = false;
    mutable wf_scene_info info{};
};

namespace wf {
// Records the message that wf_last_error() returns on the calling thread; returns `code`.
int set_last_error(int code, const std::string& msg);
}  // namespace wf
