// Mixed-precision form (see mixed.hpp).
#include "mixed.hpp"

namespace nnpx {

void setup_mixed(MixedModel& mm, const HostModel& hm)
{
    (void)hm;
    mm.ready = false;
}

void free_mixed(MixedModel& mm)
{
    if (mm.d_mem) (void)hipFree(mm.d_mem);
    if (mm.d_w) (void)hipFree(mm.d_w);
    if (mm.d_b) (void)hipFree(mm.d_b);
    mm = MixedModel{};
}

void run_mixed(const MixedModel& mm, const ModelT&, const SysView&, const MixedBuffers&, bool, bool,
               hipStream_t, const TimedFn&)
{
    if (!mm.ready) throw Error{7, "mixed-precision form not available in this build"};
}

}  // namespace nnpx
