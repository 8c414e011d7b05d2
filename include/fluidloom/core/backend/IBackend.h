// Device abstraction of the C++ API: the reference's IBackend
// (include/fluidloom/core/backend/IBackend.h:25-132) with BackendType::HIP
// added.  Signatures, argument meaning and error behaviour are unchanged.
#pragma once
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace fluidloom {

enum class BackendType { MOCK, OPENCL, HIP };

enum class MemoryLocation { HOST, DEVICE };

class DeviceBuffer;
using DeviceBufferPtr = std::unique_ptr<DeviceBuffer>;

class IBackend {
public:
    virtual ~IBackend() = default;

    virtual BackendType getType() const = 0;
    virtual std::string getName() const = 0;
    virtual int getDeviceCount() const = 0;
    virtual void initialize(int device_id = 0) = 0;
    virtual void shutdown() = 0;

    virtual DeviceBufferPtr allocateBuffer(size_t size_in_bytes, const void* initial_data = nullptr) = 0;
    virtual void copyHostToDevice(const void* host_src, DeviceBuffer& device_dst, size_t size) = 0;
    virtual void copyDeviceToHost(const DeviceBuffer& device_src, void* host_dst, size_t size) = 0;
    virtual void copyDeviceToDevice(const DeviceBuffer& src, DeviceBuffer& dst, size_t size) = 0;

    virtual void flush() = 0;
    virtual void finish() = 0;

    struct KernelHandle {
        void* handle;
        explicit KernelHandle(void* h) : handle(h) {}
    };

    virtual KernelHandle compileKernel(const std::string& source_file, const std::string& kernel_name,
                                       const std::string& build_options = "") = 0;

    struct KernelArg {
        std::vector<uint8_t> data;
        template <typename T>
        static KernelArg fromScalar(const T& val) {
            KernelArg a;
            a.data.resize(sizeof(T));
            std::memcpy(a.data.data(), &val, sizeof(T));
            return a;
        }
        static KernelArg fromBuffer(const void* ptr) {
            KernelArg a;
            a.data.resize(sizeof(void*));
            std::memcpy(a.data.data(), &ptr, sizeof(void*));
            return a;
        }
    };

    virtual void launchKernel(const KernelHandle& kernel, size_t global_work_size, size_t local_work_size,
                              const std::vector<KernelArg>& args) = 0;
    virtual void releaseKernel(const KernelHandle& kernel) = 0;

    virtual size_t getMaxAllocationSize() const = 0;
    virtual size_t getTotalMemory() const = 0;
    virtual bool isInitialized() const = 0;

protected:
    IBackend() = default;
};

class DeviceBuffer {
public:
    virtual ~DeviceBuffer() = default;
    virtual void* getDevicePointer() = 0;
    virtual const void* getDevicePointer() const = 0;
    virtual size_t getSize() const = 0;
    virtual MemoryLocation getLocation() const = 0;

protected:
    DeviceBuffer() = default;
};

}  // namespace fluidloom
