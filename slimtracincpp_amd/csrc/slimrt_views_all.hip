// The views kernel's instantiations for scenes with lights or emissive quads (slimrt_views_trace.h).
#include "slimrt_views_trace.h"

namespace slimrt_views_detail {

hipError_t launch_views_all(const ViewsLaunch &l) { return launch_views<kLightsAll>(l); }

}  // namespace slimrt_views_detail
