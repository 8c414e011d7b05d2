// Ghost exchange on one GPU (HaloExchanger, src/halo/HaloExchanger.cpp:27-187
// of the reference).  With a single device the MPI send/receive between
// ranks collapses to an intra-device copy: startExchange() packs the source
// cells of every registered float field into a device pack buffer laid out
// by PackBufferLayout (pack_fields with an index list), finishExchange()
// unpacks it into the ghost cells (unpack_fields).  Coarse/fine ghost fill of
// the lattice populations is done inside the solver (fl_lbm_*).
#pragma once
#include <memory>
#include <vector>

#include "fluidloom/core/backend/IBackend.h"
#include "fluidloom/core/fields/SOAFieldManager.h"
#include "fluidloom/halo/PackBufferLayout.h"

namespace fluidloom {
namespace halo {

class HaloExchanger {
public:
    HaloExchanger(std::shared_ptr<IBackend> backend, std::shared_ptr<fields::SOAFieldManager> field_manager);
    ~HaloExchanger();

    void registerField(fields::FieldHandle handle);
    // ghost cell recv_cells[k] receives the value of send_cells[k]
    void setGhostMap(const std::vector<uint32_t>& send_cells, const std::vector<uint32_t>& recv_cells);

    void initialize();
    void startExchange();
    void finishExchange();

    const PackBufferLayout& layout() const { return m_layout; }
    DeviceBuffer* packBuffer() { return m_pack.get(); }

private:
    std::shared_ptr<IBackend> m_backend;
    std::shared_ptr<fields::SOAFieldManager> m_field_manager;
    std::vector<fields::FieldHandle> m_fields;
    std::vector<uint32_t> m_send_host, m_recv_host;
    DeviceBufferPtr m_send_idx, m_recv_idx, m_pack;
    PackBufferLayout m_layout;
    bool m_initialized = false;
    void* stream() const;
};

}  // namespace halo
}  // namespace fluidloom
