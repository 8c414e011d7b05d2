// C++ API tests mirroring the reference's gtest suites (no gtest offline):
//   --mock : host-only tests (field layer, layouts, MockBackend errors)
//   --hip  : device tests through HIPBackend and the adaptation engines
// Reference tests followed: tests/unit/fields/test_field_manager.cpp,
// tests/unit/halo/test_pack_layout.cpp, tests/unit/test_mock_backend.cpp,
// tests/unit/hilbert/test_hilbert_opencl.cpp, tests/unit/hashmap/test_hash_table.cpp,
// tests/unit/adaptation/{Split,Merge}EngineTest.cpp, BalanceEnforcerTest.cpp,
// tests/integration/adaptation/AdaptationIntegrationTest.cpp,
// tests/unit/halo/test_halo_exchanger.cpp.
#include <cstdio>
#include <cstring>
#include <functional>
#include <random>
#include <string>
#include <vector>

#include "fluidloom/adaptation/Engines.h"
#include "fluidloom/common/FluidLoomError.h"
#include "fluidloom/core/backend/BackendFactory.h"
#include "fluidloom/core/backend/HIPBackend.h"
#include "fluidloom/core/fields/SOAFieldManager.h"
#include "fluidloom/core/hashmap/HashTableManager.h"
#include "fluidloom/core/registry/FieldRegistry.h"
#include "fluidloom/halo/HaloExchanger.h"
#include "fluidloom/halo/PackBufferLayout.h"
#include "fluidloom_hip.h"

extern "C" uint64_t or_hilbert_encode(int32_t x, int32_t y, int32_t z, int level);

using namespace fluidloom;
static int g_fail = 0, g_pass = 0;
#define CHECK(c)                                                              \
    do {                                                                      \
        if (!(c)) {                                                           \
            std::printf("  FAILED %s:%d: %s\n", __FILE__, __LINE__, #c);      \
            ++g_fail;                                                         \
            return;                                                           \
        }                                                                     \
    } while (0)
#define CHECK_THROWS(stmt, E)                                                 \
    do {                                                                      \
        bool t_ = false;                                                      \
        try { stmt; } catch (const E&) { t_ = true; }                         \
        CHECK(t_);                                                            \
    } while (0)

static void run(const char* name, const std::function<void()>& f) {
    int before = g_fail;
    try {
        f();
    } catch (const std::exception& e) {
        std::printf("  FAILED %s: exception %s\n", name, e.what());
        ++g_fail;
    }
    if (g_fail == before) ++g_pass;
    std::printf("%s %s\n", g_fail == before ? "[ OK ]" : "[FAIL]", name);
}

using namespace fluidloom::fields;

static void mock_tests() {
    run("FieldDescriptor.validation", [] {
        FieldDescriptor d("velocity", FieldType::FLOAT32, 3, 1);
        CHECK(d.num_components == 3 && d.halo_depth == 1 && d.isValid());
        CHECK_THROWS(FieldDescriptor("", FieldType::FLOAT32, 1), std::invalid_argument);
        CHECK_THROWS(FieldDescriptor(std::string(65, 'a'), FieldType::FLOAT32, 1), std::invalid_argument);
        CHECK_THROWS(FieldDescriptor("t", FieldType::FLOAT32, 0), std::invalid_argument);
        CHECK_THROWS(FieldDescriptor("t", FieldType::FLOAT32, 33), std::invalid_argument);
        CHECK_THROWS(FieldDescriptor("t", FieldType::FLOAT32, 1, 3), std::invalid_argument);
    });
    run("FieldDescriptor.bytesPerCellAndIds", [] {
        CHECK(FieldDescriptor("a", FieldType::FLOAT32, 1).bytesPerCell() == 4);
        CHECK(FieldDescriptor("b", FieldType::FLOAT64, 1).bytesPerCell() == 8);
        CHECK(FieldDescriptor("c", FieldType::FLOAT32, 3).bytesPerCell() == 12);
        CHECK(FieldDescriptor("d", FieldType::UINT8, 1).bytesPerCell() == 1);
        CHECK(FieldDescriptor("x", FieldType::FLOAT32, 1).id != FieldDescriptor("y", FieldType::FLOAT32, 1).id);
        CHECK(FieldDescriptor("x", FieldType::FLOAT32, 1).id == FieldDescriptor("x", FieldType::INT32, 2).id);
        CHECK(fnv1a_64("") == 0xcbf29ce484222325ull);
        CHECK(fnv1a_64("a") == 0xaf63dc4c8601ec8cull);
    });
    run("SOAFieldManager.lifecycle", [] {
        MockBackend b;
        b.initialize();
        SOAFieldManager m(&b);
        auto h = m.allocate(FieldDescriptor("vel", FieldType::FLOAT32, 3), 1000);
        CHECK(h.id != 0 && m.getAllocationSize(h) == 1000);
        char* p0 = (char*)m.getDevicePtr(h, 0);
        char* p1 = (char*)m.getDevicePtr(h, 1);
        CHECK(p1 - p0 == (long)m.getComponentPitch(h) && m.getComponentPitch(h) >= 4000);
        CHECK_THROWS(m.getDevicePtr(h, 3), std::out_of_range);
        uint64_t v1 = m.getVersion(h);
        m.resize(h, 2000);
        CHECK(m.getAllocationSize(h) == 2000 && m.getVersion(h) > v1 && m.isDirty(h));
        m.markClean(h);
        CHECK(!m.isDirty(h));
        size_t before = m.getTotalMemoryUsage();
        auto h2 = m.allocate(FieldDescriptor("rho", FieldType::FLOAT32, 1), 10);
        CHECK(m.getTotalMemoryUsage() > before);
        m.deallocate(h2);
        CHECK(m.getTotalMemoryUsage() == before);
        CHECK_THROWS(m.allocate(FieldDescriptor("vel", FieldType::FLOAT32, 3), 5), std::runtime_error);
    });
    run("FieldRegistry.basic", [] {
        auto& r = registry::FieldRegistry::instance();
        CHECK(&r == &registry::FieldRegistry::instance());
        CHECK(r.registerField(FieldDescriptor("reg_a", FieldType::FLOAT32, 1)));
        CHECK(!r.registerField(FieldDescriptor("reg_a", FieldType::FLOAT32, 1)));
        CHECK(r.exists("reg_a") && !r.exists("nope"));
        CHECK(r.lookupByName("reg_a")->name == "reg_a");
        CHECK(r.lookupById(FieldHandle(fnv1a_64("reg_a")))->name == "reg_a");
    });
    run("PackBufferLayout.knownOffsets", [] {
        // tests/unit/halo/test_pack_layout.cpp:6-47
        halo::PackBufferLayout L;
        L.capacity_bytes = 1024;
        L.addField("density", 1, 4);
        L.addField("velocity", 3, 4);
        CHECK(L.cell_size_bytes == 16 && L.maxCells() == 64);
        CHECK(L.getOffset(0, 0, 0) == 0 && L.getOffset(1, 0, 0) == 256 && L.getOffset(1, 1, 0) == 512 &&
              L.getOffset(1, 2, 0) == 768);
        CHECK(L.getOffset(0, 0, 1) == 4 && L.getOffset(1, 0, 1) == 260);
        CHECK(L.validate());
    });
    run("MockBackend.errors", [] {
        MockBackend b;
        b.initialize();
        CHECK_THROWS(b.allocateBuffer(0), BackendError);
        CHECK_THROWS(b.allocateBuffer(MockBackend::kMemoryLimit + 1), BackendError);
        b.shutdown();
        CHECK_THROWS(b.shutdown(), BackendError);
    });
}

static void hip_tests() {
    auto backend = BackendFactory::createBackend(BackendChoice::HIP);
    backend->initialize(0);
    IBackend* B = backend.get();
    fl_stream q = static_cast<HIPBackend*>(B)->stream();
    run("HIPBackend.hilbertKernelMatchesHost", [&] {
        // test_hilbert_opencl.cpp:44-87: 1000 points at level 8, mt19937(42), int3 layout
        std::mt19937 rng(42);
        std::uniform_int_distribution<int> d(0, 255);
        std::vector<int32_t> xyz(4000);
        for (int i = 0; i < 1000; ++i) {
            xyz[4 * i] = d(rng); xyz[4 * i + 1] = d(rng); xyz[4 * i + 2] = d(rng); xyz[4 * i + 3] = 0;
        }
        auto dc = B->allocateBuffer(xyz.size() * 4, xyz.data());
        auto dk = B->allocateBuffer(8000);
        auto k = B->compileKernel("kernels/hilbert.cl", "test_hilbert_encode");
        B->launchKernel(k, 1000, 0, {IBackend::KernelArg::fromBuffer(dc->getDevicePointer()),
                                     IBackend::KernelArg::fromBuffer(dk->getDevicePointer()),
                                     IBackend::KernelArg::fromScalar<uint32_t>(1000),
                                     IBackend::KernelArg::fromScalar<uint8_t>(8)});
        std::vector<uint64_t> h(1000);
        B->copyDeviceToHost(*dk, h.data(), 8000);
        for (int i = 0; i < 1000; ++i) CHECK(h[i] == or_hilbert_encode(xyz[4 * i], xyz[4 * i + 1], xyz[4 * i + 2], 8));
        std::vector<int32_t> back(4000, -1);
        auto dd = B->allocateBuffer(16000);
        auto kd = B->compileKernel("kernels/hilbert.cl", "test_hilbert_decode");
        B->launchKernel(kd, 1000, 0, {IBackend::KernelArg::fromBuffer(dk->getDevicePointer()),
                                      IBackend::KernelArg::fromBuffer(dd->getDevicePointer()),
                                      IBackend::KernelArg::fromScalar<uint32_t>(1000),
                                      IBackend::KernelArg::fromScalar<uint8_t>(8)});
        B->copyDeviceToHost(*dd, back.data(), 16000);
        for (int i = 0; i < 1000; ++i)
            CHECK(back[4 * i] == xyz[4 * i] && back[4 * i + 1] == xyz[4 * i + 1] && back[4 * i + 2] == xyz[4 * i + 2]);
        CHECK_THROWS(B->compileKernel("kernels/other.cl", "nothing"), BackendError);
    });
    run("HIPBackend.packUnpackRoundTrip", [&] {
        // test_halo_exchanger.cpp:82-127: 1024 floats i*1.5f
        std::vector<float> f(1024);
        for (int i = 0; i < 1024; ++i) f[i] = i * 1.5f;
        auto df = B->allocateBuffer(4096, f.data());
        auto dp = B->allocateBuffer(8192);
        auto dout = B->allocateBuffer(4096);
        auto kp = B->compileKernel("kernels/halo_pack.cl", "pack_fields");
        auto ku = B->compileKernel("kernels/halo_pack.cl", "unpack_fields");
        B->launchKernel(kp, 1024, 0, {IBackend::KernelArg::fromBuffer(df->getDevicePointer()),
                                      IBackend::KernelArg::fromBuffer(dp->getDevicePointer()),
                                      IBackend::KernelArg::fromScalar<uint64_t>(1024),
                                      IBackend::KernelArg::fromScalar<uint64_t>(4096)});
        B->launchKernel(ku, 1024, 0, {IBackend::KernelArg::fromBuffer(dp->getDevicePointer()),
                                      IBackend::KernelArg::fromBuffer(dout->getDevicePointer()),
                                      IBackend::KernelArg::fromScalar<uint64_t>(1024),
                                      IBackend::KernelArg::fromScalar<uint64_t>(4096)});
        std::vector<float> g(1024);
        B->copyDeviceToHost(*dout, g.data(), 4096);
        CHECK(std::memcmp(f.data(), g.data(), 4096) == 0);
    });
    run("HashTableManager.rebuildQuery", [&] {
        hashmap::HashTableManager m(B);
        std::vector<uint64_t> k;
        std::vector<uint32_t> v;
        for (uint32_t i = 0; i < 10000; ++i) { k.push_back(i * 1000ull + 7); v.push_back(i); }
        m.rebuild(k, v);
        CHECK((m.getCapacity() & (m.getCapacity() - 1)) == 0 && m.getCapacity() >= 10000 / 0.6);
        auto r = m.query({7, 1007, 9999007, 3});
        CHECK(r[0] == 0 && r[1] == 1 && r[2] == 9999 && r[3] == hashmap::HASH_INVALID_VALUE);
        CHECK(m.getMetadata().num_cells == 10000 && m.getMetadata().max_probe_count < 32);
    });
    run("RadixSortAndCompaction", [&] {
        hashmap::CompactionEngine c(B);
        std::vector<uint64_t> keys = {50, 10, hashmap::HASH_EMPTY_KEY, 30, 10, 20};
        std::vector<uint32_t> state = {1, 1, 1, 0, 1, 1};
        auto r = c.compactAndPermute(keys, state);
        CHECK(r.num_valid_cells == 4);
        std::vector<uint32_t> exp = {3, 0, hashmap::HASH_INVALID_VALUE, hashmap::HASH_INVALID_VALUE, 1, 2};
        CHECK(r.permutation == exp);
    });
    auto dev = [&](const void* h, size_t n) {
        void* p = nullptr;
        FL_CHECK_STATUS(fl_malloc(&p, n));
        FL_CHECK_STATUS(fl_memcpy_h2d(p, h, n, q));
        return p;
    };
    adaptation::AdaptationConfig cfg;
    run("SplitEngine.oneCellToEight", [&] {
        int32_t x = 0, y = 0, z = 0, flag = 1;
        uint8_t l = 0, s = 0;
        uint32_t m = 0;
        void* p[7] = {dev(&x, 4), dev(&y, 4), dev(&z, 4), dev(&l, 1), dev(&s, 1), dev(&flag, 4), dev(&m, 4)};
        adaptation::SplitEngine e(nullptr, q, cfg);
        auto r = e.split(p[0], p[1], p[2], p[3], p[4], p[5], p[6], 1);
        CHECK(r.success && r.num_children == 8);
        for (int c = 0; c < 8; ++c) CHECK(r.children[c].level == 1 && r.children[c].x == (c & 1));
        for (void* a : p) fl_free(a);
    });
    run("MergeEngine.eightSiblings", [&] {
        std::vector<int32_t> x(8), y(8), z(8), flag(8, -1);
        std::vector<uint8_t> l(8, 1), s(8, 0);
        std::vector<uint32_t> m(8, 0);
        for (int i = 0; i < 8; ++i) { x[i] = i & 1; y[i] = (i >> 1) & 1; z[i] = i >> 2; }
        void* p[7] = {dev(x.data(), 32), dev(y.data(), 32), dev(z.data(), 32), dev(l.data(), 8), dev(s.data(), 8),
                      dev(flag.data(), 32), dev(m.data(), 32)};
        adaptation::MergeEngine e(nullptr, q, cfg);
        auto r = e.merge(p[0], p[1], p[2], p[3], p[4], p[5], p[6], 8);
        CHECK(r.success && r.num_parents_created == 1 && r.num_children_merged == 8);
        CHECK(r.parents[0].level == 0 && r.parents[0].x == 0);
        for (void* a : p) fl_free(a);
    });
    run("BalanceEnforcer.detectViolation", [&] {
        // BalanceEnforcerTest.cpp:37-93 (anchors {0,128} in 2^8 units)
        std::vector<int32_t> x = {0, 4}, y = {0, 0}, z = {0, 0}, flag = {0, 0};
        std::vector<uint8_t> l = {1, 3}, s = {0, 0};
        void* p[6] = {dev(x.data(), 8), dev(y.data(), 8), dev(z.data(), 8), dev(l.data(), 2), dev(s.data(), 2),
                      dev(flag.data(), 8)};
        adaptation::BalanceEnforcer e(nullptr, q, cfg);
        auto r = e.enforce(p[0], p[1], p[2], p[3], p[4], p[5], 2);
        CHECK(r.converged && r.total_violations_detected > 0 && r.total_cells_marked_for_balance > 0);
        std::vector<int32_t> out(2);
        FL_CHECK_STATUS(fl_memcpy_d2h(out.data(), p[5], 8, q));
        CHECK(out[0] == 1 && out[1] == 0);
        for (void* a : p) fl_free(a);
    });
    run("AdaptationEngine.splitThenMerge", [&] {
        // AdaptationIntegrationTest.cpp:62-128
        int32_t x = 0, y = 0, z = 0, flag = 1;
        uint8_t l = 0, s = 0;
        uint32_t m = 0;
        void* p[7] = {dev(&x, 4), dev(&y, 4), dev(&z, 4), dev(&l, 1), dev(&s, 1), dev(&flag, 4), dev(&m, 4)};
        size_t n = 1, cap = 1;
        adaptation::AdaptationEngine e(nullptr, q, cfg);
        adaptation::releaseEvent(e.adapt(&p[0], &p[1], &p[2], &p[3], &p[4], &p[5], &p[6], &n, &cap));
        CHECK(n == 8 && cap >= 8);
        std::vector<uint8_t> lv(8);
        FL_CHECK_STATUS(fl_memcpy_d2h(lv.data(), p[3], 8, q));
        for (int i = 0; i < 8; ++i) CHECK(lv[i] == 1);
        std::vector<int32_t> f(8, -1);
        FL_CHECK_STATUS(fl_memcpy_h2d(p[5], f.data(), 32, q));
        adaptation::releaseEvent(e.adapt(&p[0], &p[1], &p[2], &p[3], &p[4], &p[5], &p[6], &n, &cap));
        CHECK(n == 1);
        FL_CHECK_STATUS(fl_memcpy_d2h(lv.data(), p[3], 1, q));
        CHECK(lv[0] == 0);
        for (void* a : p) fl_free(a);
    });
    run("HaloExchanger.intraDeviceGhosts", [&] {
        auto sb = std::shared_ptr<IBackend>(B, [](IBackend*) {});
        auto fm = std::make_shared<SOAFieldManager>(B);
        auto h = fm->allocate(FieldDescriptor("halo_vel", FieldType::FLOAT32, 3), 100);
        std::vector<float> comp(100);
        for (int c = 0; c < 3; ++c) {
            for (int i = 0; i < 100; ++i) comp[i] = c * 1000 + i;
            FL_CHECK_STATUS(fl_memcpy_h2d(fm->getDevicePtr(h, c), comp.data(), 400, q));
        }
        halo::HaloExchanger ex(sb, fm);
        ex.registerField(h);
        ex.setGhostMap({1, 2, 3}, {97, 98, 99});
        ex.initialize();
        ex.startExchange();
        ex.finishExchange();
        for (int c = 0; c < 3; ++c) {
            FL_CHECK_STATUS(fl_memcpy_d2h(comp.data(), fm->getDevicePtr(h, c), 400, q));
            CHECK(comp[97] == c * 1000 + 1 && comp[98] == c * 1000 + 2 && comp[99] == c * 1000 + 3);
            CHECK(comp[50] == c * 1000 + 50);
        }
    });
}

int main(int argc, char** argv) {
    std::string mode = argc > 1 ? argv[1] : "--mock";
    mock_tests();
    if (mode == "--hip") hip_tests();
    std::printf("%d passed, %d failed\n", g_pass, g_fail);
    return g_fail ? 1 : 0;
}
