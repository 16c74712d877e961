// Opaque handle types behind the C ABI (include/wfamd.h) and the shared last-error slot.
#pragma once
#include <string>

#include "wf_scene.h"

struct wf_scene {
    wf::Scene s;
};

namespace wf {
// Records the message that wf_last_error() returns on the calling thread; returns `code`.
int set_last_error(int code, const std::string& msg);
}  // namespace wf
