// MI355X backend: device memory and the hot-path kernels of
// libfluidloom_hip.so behind the reference's IBackend interface (the role of
// OpenCLBackend, src/core/backend/OpenCLBackend.cpp:38-332).
//
// Kernels are compiled ahead of time for gfx950, so compileKernel() resolves
// the reference's (source file, kernel name) pairs to built-in launchers:
//   kernels/hilbert.cl            test_hilbert_encode, test_hilbert_decode
//   kernels/halo_pack.cl          pack_fields, unpack_fields
//   kernels/hashmap/hash_build.cl hash_build, hash_clear
//   kernels/hashmap/hash_query.cl hash_validate
// launchKernel() reads the reference kernels' argument lists (byte blobs in
// the same order) and runs the HIP kernel on the backend's stream.  Any
// other pair throws BackendError, as a failed clBuildProgram does.
#pragma once
#include <mutex>
#include <string>

#include "fluidloom/core/backend/IBackend.h"

namespace fluidloom {

class HIPBuffer : public DeviceBuffer {
public:
    HIPBuffer(void* p, size_t n) : ptr_(p), size_(n) {}
    ~HIPBuffer() override;
    void* getDevicePointer() override { return ptr_; }
    const void* getDevicePointer() const override { return ptr_; }
    size_t getSize() const override { return size_; }
    MemoryLocation getLocation() const override { return MemoryLocation::DEVICE; }

private:
    void* ptr_;
    size_t size_;
};

class HIPBackend : public IBackend {
public:
    HIPBackend() = default;
    ~HIPBackend() override;

    BackendType getType() const override { return BackendType::HIP; }
    std::string getName() const override;
    int getDeviceCount() const override;
    void initialize(int device_id = 0) override;
    void shutdown() override;

    DeviceBufferPtr allocateBuffer(size_t size_in_bytes, const void* initial_data = nullptr) override;
    void copyHostToDevice(const void* host_src, DeviceBuffer& device_dst, size_t size) override;
    void copyDeviceToHost(const DeviceBuffer& device_src, void* host_dst, size_t size) override;
    void copyDeviceToDevice(const DeviceBuffer& src, DeviceBuffer& dst, size_t size) override;

    void flush() override {}
    void finish() override;

    KernelHandle compileKernel(const std::string& source_file, const std::string& kernel_name,
                               const std::string& build_options = "") override;
    void launchKernel(const KernelHandle& kernel, size_t global_work_size, size_t local_work_size,
                      const std::vector<KernelArg>& args) override;
    void releaseKernel(const KernelHandle& kernel) override {}

    size_t getMaxAllocationSize() const override;
    size_t getTotalMemory() const override;
    bool isInitialized() const override { return initialized_; }

    // HIP stream (hipStream_t) all work of this backend is ordered on.
    void* stream() const { return stream_; }

private:
    bool initialized_ = false;
    int device_ = 0;
    void* stream_ = nullptr;
    std::string name_;
    size_t total_mem_ = 0;
};

// Host-memory backend for CPU-only tests (MockBackend.cpp:56-149 of the
// reference): allocations are host vectors, kernel compile/launch are
// logged no-ops, a 16 GB simulated limit.
class MockBuffer : public DeviceBuffer {
public:
    explicit MockBuffer(size_t n) : data_(n) {}
    void* getDevicePointer() override { return data_.data(); }
    const void* getDevicePointer() const override { return data_.data(); }
    size_t getSize() const override { return data_.size(); }
    MemoryLocation getLocation() const override { return MemoryLocation::HOST; }

private:
    std::vector<uint8_t> data_;
};

class MockBackend : public IBackend {
public:
    static constexpr size_t kMemoryLimit = size_t(16) << 30;
    BackendType getType() const override { return BackendType::MOCK; }
    std::string getName() const override { return "Mock Backend (host memory)"; }
    int getDeviceCount() const override { return 1; }
    void initialize(int device_id = 0) override;
    void shutdown() override;
    DeviceBufferPtr allocateBuffer(size_t size_in_bytes, const void* initial_data = nullptr) override;
    void copyHostToDevice(const void* host_src, DeviceBuffer& device_dst, size_t size) override;
    void copyDeviceToHost(const DeviceBuffer& device_src, void* host_dst, size_t size) override;
    void copyDeviceToDevice(const DeviceBuffer& src, DeviceBuffer& dst, size_t size) override;
    void flush() override {}
    void finish() override {}
    KernelHandle compileKernel(const std::string&, const std::string&, const std::string& = "") override {
        return KernelHandle(reinterpret_cast<void*>(0x1));
    }
    void launchKernel(const KernelHandle&, size_t, size_t, const std::vector<KernelArg>&) override {}
    void releaseKernel(const KernelHandle&) override {}
    size_t getMaxAllocationSize() const override { return kMemoryLimit; }
    size_t getTotalMemory() const override { return kMemoryLimit; }
    bool isInitialized() const override { return initialized_; }
    size_t allocated() const { return allocated_; }

private:
    bool initialized_ = false;
    size_t allocated_ = 0;
};

}  // namespace fluidloom
