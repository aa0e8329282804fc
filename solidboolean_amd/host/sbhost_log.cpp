#include "sbhost_log.h"

#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <stdexcept>
#include <string>

namespace sbhost {

void diag(const char *msg)
{
    static const bool verbose = std::getenv("SBHOST_VERBOSE") != nullptr;
    if (verbose)
        std::fprintf(stderr, "%s\n", msg);
}

void check(int rc, const char *what)
{
    if (rc != SBG_OK)
        throw std::runtime_error(std::string(what) + ": " + sbg_last_error());
}

std::recursive_mutex &gpuMutex()
{
    static std::recursive_mutex m;
    return m;
}

sbg_context gpu()
{
    static std::once_flag once;
    static sbg_context ctx = nullptr;
    static int rc = SBG_OK;
    static std::string err;
    std::call_once(once, [] {
        const char *dev = std::getenv("SBG_DEVICE");
        rc = sbg_context_create(dev ? std::atoi(dev) : 0, &ctx);
        if (rc != SBG_OK)
            err = sbg_last_error();
    });
    if (rc != SBG_OK)
        throw std::runtime_error("solidboolean: no usable GPU context: " + err);
    return ctx;
}

}  // namespace sbhost
