// RAII launch timer used around kernel launches (see prof.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace pycbf {
class ProfScope {
public:
    ProfScope(const char *name, hipStream_t s);
    ~ProfScope();
    ProfScope(const ProfScope &) = delete;
    ProfScope &operator=(const ProfScope &) = delete;

private:
    const char *name_;
    hipStream_t s_;
    hipEvent_t a_{}, b_{};
    bool live_ = false;
};
}  // namespace pycbf
