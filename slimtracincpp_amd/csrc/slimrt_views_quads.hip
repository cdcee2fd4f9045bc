// The views kernel's instantiations for scenes without point or directional lights (slimrt_views_trace.h).
#include "slimrt_views_trace.h"

namespace slimrt_views_detail {

hipError_t launch_views_quads(const ViewsLaunch &l) { return launch_views<kLightsQuads>(l); }

}  // namespace slimrt_views_detail
