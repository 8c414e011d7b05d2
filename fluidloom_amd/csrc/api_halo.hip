// C++ API: HaloExchanger on one device (pack -> unpack through a device
// pack buffer laid out by PackBufferLayout).
#include <hip/hip_runtime.h>

#include "fluidloom/common/FluidLoomError.h"
#include "fluidloom/core/backend/HIPBackend.h"
#include "fluidloom/halo/HaloExchanger.h"
#include "fluidloom_hip.h"

namespace fluidloom {
namespace halo {

HaloExchanger::HaloExchanger(std::shared_ptr<IBackend> backend, std::shared_ptr<fields::SOAFieldManager> fm)
    : m_backend(std::move(backend)), m_field_manager(std::move(fm)) {
    if (!m_backend || !m_field_manager) throw std::invalid_argument("backend and field manager required");
}

HaloExchanger::~HaloExchanger() = default;

void* HaloExchanger::stream() const {
    if (m_backend->getType() != BackendType::HIP) FL_THROW(BackendError, "HaloExchanger requires the HIP backend");
    return static_cast<HIPBackend*>(m_backend.get())->stream();
}

void HaloExchanger::registerField(fields::FieldHandle h) {
    const auto& d = m_field_manager->getDescriptor(h);
    if (d.type != fields::FieldType::FLOAT32) FL_THROW(BackendError, "halo exchange packs FLOAT32 fields");
    m_fields.push_back(h);
    m_initialized = false;
}

void HaloExchanger::setGhostMap(const std::vector<uint32_t>& send, const std::vector<uint32_t>& recv) {
    if (send.size() != recv.size()) throw std::invalid_argument("send/recv size mismatch");
    m_send_host = send;
    m_recv_host = recv;
    m_initialized = false;
}

void HaloExchanger::initialize() {
    m_layout = PackBufferLayout{};
    for (auto h : m_fields) {
        const auto& d = m_field_manager->getDescriptor(h);
        m_layout.addField(d.name, d.num_components, 4);
    }
    const size_t n = m_send_host.size();
    m_layout.capacity_bytes = std::max<size_t>(1, n) * std::max<size_t>(m_layout.cell_size_bytes, 4);
    m_layout.used_bytes = n * m_layout.cell_size_bytes;
    m_pack = m_backend->allocateBuffer(m_layout.capacity_bytes);
    m_send_idx = n ? m_backend->allocateBuffer(n * 4, m_send_host.data()) : nullptr;
    m_recv_idx = n ? m_backend->allocateBuffer(n * 4, m_recv_host.data()) : nullptr;
    m_initialized = true;
}

void HaloExchanger::startExchange() {
    if (!m_initialized) initialize();
    const size_t n = m_send_host.size();
    if (!n) return;
    fl_stream s = stream();
    for (size_t f = 0; f < m_fields.size(); ++f) {
        const auto& d = m_field_manager->getDescriptor(m_fields[f]);
        for (uint16_t c = 0; c < d.num_components; ++c)
            FL_CHECK_STATUS(fl_pack_fields(static_cast<const float*>(m_field_manager->getDevicePtr(m_fields[f], c)),
                                           static_cast<const uint32_t*>(m_send_idx->getDevicePointer()), n,
                                           static_cast<uint8_t*>(m_pack->getDevicePointer()), m_layout.getOffset(f, c, 0),
                                           s));
    }
}

void HaloExchanger::finishExchange() {
    const size_t n = m_send_host.size();
    if (!n) return;
    fl_stream s = stream();
    for (size_t f = 0; f < m_fields.size(); ++f) {
        const auto& d = m_field_manager->getDescriptor(m_fields[f]);
        for (uint16_t c = 0; c < d.num_components; ++c)
            FL_CHECK_STATUS(fl_unpack_fields(static_cast<const uint8_t*>(m_pack->getDevicePointer()),
                                             m_layout.getOffset(f, c, 0),
                                             static_cast<const uint32_t*>(m_recv_idx->getDevicePointer()), n,
                                             static_cast<float*>(m_field_manager->getDevicePtr(m_fields[f], c)), s));
        m_field_manager->markDirty(m_fields[f]);
    }
    FL_CHECK_STATUS(fl_stream_sync(s));
}

}  // namespace halo
}  // namespace fluidloom
