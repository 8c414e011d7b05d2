// C++ API: HIPBackend / MockBackend / BackendFactory and the field layer
// (FieldDescriptor, SOAFieldManager, FieldRegistry) over the C ABI.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <unordered_map>

#include "fluidloom/common/FluidLoomError.h"
#include "fluidloom/core/backend/BackendFactory.h"
#include "fluidloom/core/backend/HIPBackend.h"
#include "fluidloom/core/fields/SOAFieldManager.h"
#include "fluidloom/core/registry/FieldRegistry.h"
#include "fluidloom_hip.h"

namespace fluidloom {

// ------------------------------------------------------------------ HIP
HIPBuffer::~HIPBuffer() {
    if (ptr_) (void)fl_free(ptr_);
}

HIPBackend::~HIPBackend() {
    if (initialized_) {
        (void)fl_stream_sync(stream_);
        (void)fl_stream_destroy(stream_);
    }
}

std::string HIPBackend::getName() const { return name_.empty() ? std::string("HIP (gfx950)") : name_; }

int HIPBackend::getDeviceCount() const {
    int n = 0;
    FL_CHECK_STATUS(fl_device_count(&n));
    return n;
}

void HIPBackend::initialize(int device_id) {
    if (initialized_) FL_THROW(BackendError, "HIPBackend already initialized");
    FL_CHECK_STATUS(fl_init(device_id));
    char name[64];
    int cus = 0, arch = 0;
    FL_CHECK_STATUS(fl_device_info(device_id, name, sizeof(name), &total_mem_, &cus, &arch));
    name_ = std::string("HIP ") + name + " (" + std::to_string(cus) + " CUs)";
    FL_CHECK_STATUS(fl_stream_create(&stream_));
    device_ = device_id;
    initialized_ = true;
}

void HIPBackend::shutdown() {
    if (!initialized_) FL_THROW(BackendError, "HIPBackend not initialized");
    FL_CHECK_STATUS(fl_stream_sync(stream_));
    FL_CHECK_STATUS(fl_stream_destroy(stream_));
    stream_ = nullptr;
    initialized_ = false;
}

DeviceBufferPtr HIPBackend::allocateBuffer(size_t size, const void* init) {
    if (!initialized_) FL_THROW(BackendError, "HIPBackend not initialized");
    if (size == 0) FL_THROW(BackendError, "Cannot allocate zero-size buffer");
    void* p = nullptr;
    FL_CHECK_STATUS(fl_malloc(&p, size));
    auto buf = std::make_unique<HIPBuffer>(p, size);
    if (init) FL_CHECK_STATUS(fl_memcpy_h2d(p, init, size, stream_));
    return buf;
}

void HIPBackend::copyHostToDevice(const void* src, DeviceBuffer& dst, size_t size) {
    if (size > dst.getSize()) FL_THROW(BackendError, "copy larger than destination buffer");
    FL_CHECK_STATUS(fl_memcpy_h2d(dst.getDevicePointer(), src, size, stream_));
}

void HIPBackend::copyDeviceToHost(const DeviceBuffer& src, void* dst, size_t size) {
    if (size > src.getSize()) FL_THROW(BackendError, "copy larger than source buffer");
    FL_CHECK_STATUS(fl_memcpy_d2h(dst, src.getDevicePointer(), size, stream_));
}

void HIPBackend::copyDeviceToDevice(const DeviceBuffer& src, DeviceBuffer& dst, size_t size) {
    if (size > src.getSize() || size > dst.getSize()) FL_THROW(BackendError, "copy larger than buffer");
    FL_CHECK_STATUS(fl_memcpy_d2d(dst.getDevicePointer(), src.getDevicePointer(), size, stream_));
    FL_CHECK_STATUS(fl_stream_sync(stream_));
}

void HIPBackend::finish() {
    if (initialized_) FL_CHECK_STATUS(fl_stream_sync(stream_));
}

size_t HIPBackend::getMaxAllocationSize() const { return total_mem_; }
size_t HIPBackend::getTotalMemory() const { return total_mem_; }

namespace {
enum class Builtin : int {
    HilbertEncode = 1,
    HilbertDecode,
    PackFields,
    UnpackFields,
    HashBuild,
    HashClear,
    HashValidate
};

struct KernelEntry {
    const char* file;
    const char* name;
    Builtin id;
};

const KernelEntry kKernels[] = {
    {"hilbert.cl", "test_hilbert_encode", Builtin::HilbertEncode},
    {"hilbert.cl", "test_hilbert_decode", Builtin::HilbertDecode},
    {"halo_pack.cl", "pack_fields", Builtin::PackFields},
    {"halo_pack.cl", "unpack_fields", Builtin::UnpackFields},
    {"hash_build.cl", "hash_build", Builtin::HashBuild},
    {"hash_build.cl", "hash_clear", Builtin::HashClear},
    {"hash_query.cl", "hash_validate", Builtin::HashValidate},
};

bool ends_with(const std::string& s, const std::string& t) {
    return s.size() >= t.size() && s.compare(s.size() - t.size(), t.size(), t) == 0;
}

template <typename T>
T arg(const std::vector<IBackend::KernelArg>& a, size_t i) {
    if (i >= a.size() || a[i].data.size() < sizeof(T))
        FL_THROW(BackendError, "kernel argument " + std::to_string(i) + " missing or too small");
    T v;
    std::memcpy(&v, a[i].data.data(), sizeof(T));
    return v;
}
}  // namespace

IBackend::KernelHandle HIPBackend::compileKernel(const std::string& source_file, const std::string& kernel_name,
                                                 const std::string&) {
    for (const auto& k : kKernels)
        if (kernel_name == k.name && ends_with(source_file, k.file))
            return KernelHandle(reinterpret_cast<void*>(static_cast<intptr_t>(k.id)));
    FL_THROW(BackendError, "kernel " + kernel_name + " from " + source_file +
                               " has no gfx950 implementation in this library");
}

void HIPBackend::launchKernel(const KernelHandle& kernel, size_t global, size_t, const std::vector<KernelArg>& a) {
    if (!initialized_) FL_THROW(BackendError, "HIPBackend not initialized");
    const auto id = static_cast<Builtin>(reinterpret_cast<intptr_t>(kernel.handle));
    fl_stream s = stream_;
    switch (id) {
        case Builtin::HilbertEncode: {  // (int3* coords, ulong* out, uint n, uchar level)
            uint32_t n = arg<uint32_t>(a, 2);
            FL_CHECK_STATUS(fl_hilbert_encode_aos4(arg<const int32_t*>(a, 0), arg<uint8_t>(a, 3),
                                                   arg<uint64_t*>(a, 1), std::min<size_t>(n, global), s));
            break;
        }
        case Builtin::HilbertDecode: {  // (ulong* in, int3* coords, uint n, uchar level)
            uint32_t n = arg<uint32_t>(a, 2);
            FL_CHECK_STATUS(fl_hilbert_decode_aos4(arg<const uint64_t*>(a, 0), arg<uint8_t>(a, 3),
                                                   arg<int32_t*>(a, 1), std::min<size_t>(n, global), s));
            break;
        }
        case Builtin::PackFields:  // (float* field, uchar* packed, ulong n, ulong offset)
            FL_CHECK_STATUS(fl_pack_fields(arg<const float*>(a, 0), nullptr, std::min<size_t>(arg<uint64_t>(a, 2), global),
                                           arg<uint8_t*>(a, 1), arg<uint64_t>(a, 3), s));
            break;
        case Builtin::UnpackFields:  // (uchar* packed, float* field, ulong n, ulong offset)
            FL_CHECK_STATUS(fl_unpack_fields(arg<const uint8_t*>(a, 0), arg<uint64_t>(a, 3), nullptr,
                                             std::min<size_t>(arg<uint64_t>(a, 2), global), arg<float*>(a, 1), s));
            break;
        case Builtin::HashBuild:  // (tkeys, tvals, keys, vals, ulong cap, uint n, uint* probes)
            FL_CHECK_STATUS(fl_hash_build(arg<uint64_t*>(a, 0), arg<uint32_t*>(a, 1), arg<uint64_t>(a, 4),
                                          arg<const uint64_t*>(a, 2), arg<const uint32_t*>(a, 3),
                                          std::min<size_t>(arg<uint32_t>(a, 5), global), arg<uint32_t*>(a, 6), s));
            break;
        case Builtin::HashClear:  // (tkeys, tvals, ulong cap)
            FL_CHECK_STATUS(fl_hash_clear(arg<uint64_t*>(a, 0), arg<uint32_t*>(a, 1), arg<uint64_t>(a, 2), s));
            break;
        case Builtin::HashValidate:  // (tkeys, tvals, ulong cap, queries, results, probes, uint n)
            FL_CHECK_STATUS(fl_hash_query(arg<const uint64_t*>(a, 0), arg<const uint32_t*>(a, 1), arg<uint64_t>(a, 2),
                                          arg<const uint64_t*>(a, 3), arg<uint32_t*>(a, 4),
                                          std::min<size_t>(arg<uint32_t>(a, 6), global), arg<uint32_t*>(a, 5), s));
            break;
        default:
            FL_THROW(BackendError, "unknown kernel handle");
    }
}

// ------------------------------------------------------------------ Mock
void MockBackend::initialize(int) {
    if (initialized_) FL_THROW(BackendError, "MockBackend already initialized");
    initialized_ = true;
}
void MockBackend::shutdown() {
    if (!initialized_) FL_THROW(BackendError, "MockBackend not initialized");
    initialized_ = false;
}
DeviceBufferPtr MockBackend::allocateBuffer(size_t size, const void* init) {
    if (size == 0) FL_THROW(BackendError, "Cannot allocate zero-size buffer");
    if (allocated_ + size > kMemoryLimit) FL_THROW(BackendError, "Mock out of memory");
    auto b = std::make_unique<MockBuffer>(size);
    if (init) std::memcpy(b->getDevicePointer(), init, size);
    allocated_ += size;
    return b;
}
void MockBackend::copyHostToDevice(const void* src, DeviceBuffer& dst, size_t size) {
    if (size > dst.getSize()) FL_THROW(BackendError, "copy larger than destination buffer");
    std::memcpy(dst.getDevicePointer(), src, size);
}
void MockBackend::copyDeviceToHost(const DeviceBuffer& src, void* dst, size_t size) {
    if (size > src.getSize()) FL_THROW(BackendError, "copy larger than source buffer");
    std::memcpy(dst, src.getDevicePointer(), size);
}
void MockBackend::copyDeviceToDevice(const DeviceBuffer& src, DeviceBuffer& dst, size_t size) {
    if (size > src.getSize() || size > dst.getSize()) FL_THROW(BackendError, "copy larger than buffer");
    std::memcpy(dst.getDevicePointer(), src.getDevicePointer(), size);
}

// ------------------------------------------------------------------ factory
std::unique_ptr<IBackend>& BackendFactory::instance() {
    static std::unique_ptr<IBackend> inst;
    return inst;
}

std::unique_ptr<IBackend> BackendFactory::createBackend(BackendChoice choice) {
    if (choice == BackendChoice::AUTO) {
        const char* env = std::getenv("FL_BACKEND");
        std::string e = env ? env : "";
        if (e == "MOCK") choice = BackendChoice::MOCK;
        else if (e == "HIP" || e == "OPENCL") choice = BackendChoice::HIP;
        else {
            int n = 0;
            choice = (fl_device_count(&n) == FL_OK && n > 0) ? BackendChoice::HIP : BackendChoice::MOCK;
        }
    }
    if (choice == BackendChoice::MOCK) return std::make_unique<MockBackend>();
    if (choice == BackendChoice::OPENCL)
        FL_THROW(BackendError, "OpenCL is not part of this build; use BackendChoice::HIP");
    return std::make_unique<HIPBackend>();
}

IBackend& BackendFactory::getInstance() {
    auto& i = instance();
    if (!i) {
        i = createBackend(BackendChoice::AUTO);
        i->initialize(0);
    }
    return *i;
}

void BackendFactory::setBackend(std::unique_ptr<IBackend> backend) { instance() = std::move(backend); }

// ------------------------------------------------------------------ fields
namespace fields {

uint64_t fnv1a_64(const std::string& s) {
    uint64_t h = 0xcbf29ce484222325ull;
    for (unsigned char c : s) {
        h ^= c;
        h *= 0x100000001b3ull;
    }
    return h;
}

size_t fieldTypeSize(FieldType t) {
    switch (t) {
        case FieldType::FLOAT32: return 4;
        case FieldType::FLOAT64: return 8;
        case FieldType::INT32: return 4;
        case FieldType::INT64: return 8;
        case FieldType::UINT8: return 1;
        case FieldType::FLOAT16: return 2;
    }
    return 0;
}

FieldDescriptor::FieldDescriptor(const std::string& field_name, FieldType field_type, uint16_t components,
                                 uint8_t halo, AveragingRule avg, SolidScheme solid)
    : name(field_name), type(field_type), num_components(components), halo_depth(halo), avg_rule(avg),
      solid_scheme(solid) {
    if (field_name.empty() || field_name.size() > 64)
        throw std::invalid_argument("Field name must be 1-64 characters");
    if (components == 0 || components > 32) throw std::invalid_argument("num_components must be in [1,32]");
    if (halo > 2) throw std::invalid_argument("halo_depth must be 0, 1, or 2");
    id = fnv1a_64(field_name);
    default_value.assign(components, 0.0);
}

size_t FieldDescriptor::bytesPerCell() const { return fieldTypeSize(type) * num_components; }

bool FieldDescriptor::isValid() const {
    return id != 0 && !name.empty() && name.size() <= 64 && num_components >= 1 && num_components <= 32 &&
           halo_depth <= 2 && fieldTypeSize(type) > 0;
}

SOAFieldManager::SOAFieldManager(IBackend* backend) : backend_(backend) {
    if (!backend) throw std::invalid_argument("Backend must not be null");
}

size_t SOAFieldManager::pitchFor(const FieldDescriptor& d, size_t n) {
    size_t b = fieldTypeSize(d.type) * (n ? n : 1);
    return (b + 255) & ~size_t(255);
}

SOAFieldManager::FieldState& SOAFieldManager::state(FieldHandle h) {
    auto it = fields_.find(h.id);
    if (it == fields_.end()) throw std::out_of_range("Unknown field handle");
    return it->second;
}
const SOAFieldManager::FieldState& SOAFieldManager::state(FieldHandle h) const {
    auto it = fields_.find(h.id);
    if (it == fields_.end()) throw std::out_of_range("Unknown field handle");
    return it->second;
}

FieldHandle SOAFieldManager::allocate(const FieldDescriptor& desc, size_t num_cells) {
    std::lock_guard<std::mutex> g(mutex_);
    if (!desc.isValid()) throw std::invalid_argument("Invalid field descriptor");
    if (fields_.count(desc.id)) throw std::runtime_error("Field already allocated");
    FieldState st;
    st.descriptor = desc;
    st.num_cells = num_cells;
    st.pitch = pitchFor(desc, num_cells);
    st.device_buffer = backend_->allocateBuffer(st.pitch * desc.num_components);
    char* base = static_cast<char*>(st.device_buffer->getDevicePointer());
    for (uint16_t c = 0; c < desc.num_components; ++c) st.device_ptrs.push_back(base + c * st.pitch);
    st.descriptor.gpu_state.device_ptr = base;
    st.descriptor.gpu_state.bytes_allocated = st.pitch * desc.num_components;
    st.descriptor.gpu_state.version = 1;
    fields_.emplace(desc.id, std::move(st));
    return FieldHandle(desc.id);
}

void SOAFieldManager::resize(FieldHandle handle, size_t new_num_cells) {
    std::lock_guard<std::mutex> g(mutex_);
    FieldState& st = state(handle);
    if (st.num_cells == new_num_cells) return;
    size_t np = pitchFor(st.descriptor, new_num_cells);
    auto nb = backend_->allocateBuffer(np * st.descriptor.num_components);
    size_t keep = fieldTypeSize(st.descriptor.type) * std::min(st.num_cells, new_num_cells);
    char* base = static_cast<char*>(nb->getDevicePointer());
    std::vector<uint8_t> tmp(st.pitch * st.descriptor.num_components);
    backend_->copyDeviceToHost(*st.device_buffer, tmp.data(), tmp.size());
    std::vector<uint8_t> out(np * st.descriptor.num_components, 0);
    for (uint16_t c = 0; c < st.descriptor.num_components; ++c)
        std::memcpy(out.data() + c * np, tmp.data() + c * st.pitch, keep);
    backend_->copyHostToDevice(out.data(), *nb, out.size());
    st.device_ptrs.clear();
    for (uint16_t c = 0; c < st.descriptor.num_components; ++c) st.device_ptrs.push_back(base + c * np);
    st.device_buffer = std::move(nb);
    st.pitch = np;
    st.num_cells = new_num_cells;
    st.version++;
    st.descriptor.gpu_state.is_dirty = true;
    st.descriptor.gpu_state.device_ptr = base;
    st.descriptor.gpu_state.bytes_allocated = np * st.descriptor.num_components;
}

void SOAFieldManager::deallocate(FieldHandle handle) {
    std::lock_guard<std::mutex> g(mutex_);
    fields_.erase(handle.id);
}

void* SOAFieldManager::getDevicePtr(FieldHandle handle, uint16_t component) const {
    std::lock_guard<std::mutex> g(mutex_);
    const FieldState& st = state(handle);
    if (component >= st.device_ptrs.size()) throw std::out_of_range("Component index out of range");
    return st.device_ptrs[component];
}

const FieldDescriptor& SOAFieldManager::getDescriptor(FieldHandle handle) const {
    std::lock_guard<std::mutex> g(mutex_);
    return state(handle).descriptor;
}

size_t SOAFieldManager::getAllocationSize(FieldHandle handle) const {
    std::lock_guard<std::mutex> g(mutex_);
    return state(handle).num_cells;
}

size_t SOAFieldManager::getComponentPitch(FieldHandle handle) const {
    std::lock_guard<std::mutex> g(mutex_);
    return state(handle).pitch;
}

void SOAFieldManager::markDirty(FieldHandle handle) {
    std::lock_guard<std::mutex> g(mutex_);
    FieldState& st = state(handle);
    st.descriptor.gpu_state.is_dirty = true;
    st.version++;
}

void SOAFieldManager::markClean(FieldHandle handle) {
    std::lock_guard<std::mutex> g(mutex_);
    state(handle).descriptor.gpu_state.is_dirty = false;
}

bool SOAFieldManager::isDirty(FieldHandle handle) const {
    std::lock_guard<std::mutex> g(mutex_);
    return state(handle).descriptor.gpu_state.is_dirty;
}

uint64_t SOAFieldManager::getVersion(FieldHandle handle) const {
    std::lock_guard<std::mutex> g(mutex_);
    return state(handle).version;
}

size_t SOAFieldManager::getTotalMemoryUsage() const {
    std::lock_guard<std::mutex> g(mutex_);
    size_t t = 0;
    for (const auto& kv : fields_) t += kv.second.device_buffer->getSize();
    return t;
}

size_t SOAFieldManager::getMemoryUsage(FieldHandle handle) const {
    std::lock_guard<std::mutex> g(mutex_);
    return state(handle).device_buffer->getSize();
}

}  // namespace fields

// ------------------------------------------------------------------ registry
namespace registry {

FieldRegistry& FieldRegistry::instance() {
    static FieldRegistry r;
    return r;
}

bool FieldRegistry::registerField(const fields::FieldDescriptor& d) {
    std::unique_lock<std::shared_mutex> g(mutex_);
    if (by_name_.count(d.name) || by_id_.count(d.id)) return false;
    by_name_.emplace(d.name, d);
    by_id_.emplace(d.id, d);
    return true;
}

std::optional<fields::FieldDescriptor> FieldRegistry::lookupByName(const std::string& name) const {
    std::shared_lock<std::shared_mutex> g(mutex_);
    auto it = by_name_.find(name);
    if (it == by_name_.end()) return std::nullopt;
    return it->second;
}

std::optional<fields::FieldDescriptor> FieldRegistry::lookupById(fields::FieldHandle h) const {
    std::shared_lock<std::shared_mutex> g(mutex_);
    auto it = by_id_.find(h.id);
    if (it == by_id_.end()) return std::nullopt;
    return it->second;
}

bool FieldRegistry::exists(const std::string& name) const {
    std::shared_lock<std::shared_mutex> g(mutex_);
    return by_name_.count(name) != 0;
}

std::vector<std::string> FieldRegistry::getAllNames() const {
    std::shared_lock<std::shared_mutex> g(mutex_);
    std::vector<std::string> v;
    for (const auto& kv : by_name_) v.push_back(kv.first);
    return v;
}

void FieldRegistry::clear() {
    std::unique_lock<std::shared_mutex> g(mutex_);
    by_name_.clear();
    by_id_.clear();
}

}  // namespace registry
}  // namespace fluidloom
