#include "wt_rt.h"

std::shared_mutex& wt_rt_mutex()
{
    static std::shared_mutex m;
    return m;
}
