// nova-mi355x: the ownership registry declared in nova_common.h.
#include <mutex>
#include <unordered_set>

#include "nova_common.h"

namespace nova_amd {

namespace {
std::mutex g_mu;
std::unordered_set<const void *> &owned()
{
    static std::unordered_set<const void *> set;
    return set;
}
} // namespace

void registry_track(const void *p)
{
    if (p == nullptr)
        return;
    std::lock_guard<std::mutex> lk(g_mu);
    owned().insert(p);
}

bool registry_untrack(const void *p)
{
    std::lock_guard<std::mutex> lk(g_mu);
    return owned().erase(p) > 0;
}

} // namespace nova_amd
