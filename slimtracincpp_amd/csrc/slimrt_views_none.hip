// The views kernel's instantiations for scenes without lights and without emissive quads (slimrt_views_trace.h).
#include "slimrt_views_trace.h"

namespace slimrt_views_detail {

hipError_t launch_views_none(const ViewsLaunch &l) { return launch_views<kLightsNone>(l); }

}  // namespace slimrt_views_detail
