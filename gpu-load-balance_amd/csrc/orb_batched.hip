// orb_batched.hip -- the batched (device-driven) build.
#include "orb_internal.hpp"

namespace orb {

struct BatchedState {
};

int build_batched(Ctx &c, int32_t levels_mode, orb_build_info *info)
{
    (void)c;
    (void)levels_mode;
    (void)info;
    return ORB_ERR_STATE;
}

void batched_release(Ctx &c)
{
    delete c.batched;
    c.batched = nullptr;
}

}  // namespace orb
