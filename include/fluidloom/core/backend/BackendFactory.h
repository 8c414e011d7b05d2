// Backend selection (src/core/backend/BackendFactory.cpp:9-59 of the
// reference): FL_BACKEND=HIP|MOCK; AUTO picks HIP when a GPU is visible.
#pragma once
#include <memory>

#include "fluidloom/core/backend/IBackend.h"

namespace fluidloom {

enum class BackendChoice { AUTO, MOCK, OPENCL, HIP };

class BackendFactory {
public:
    static std::unique_ptr<IBackend> createBackend(BackendChoice choice = BackendChoice::AUTO);
    static IBackend& getInstance();
    static void setBackend(std::unique_ptr<IBackend> backend);

private:
    static std::unique_ptr<IBackend>& instance();
};

}  // namespace fluidloom
