// residency_check.cpp -- drives include/aether/residency.hpp the way a user of aether's residency headers
// would, for tests/test_residency.py (mode "host", no GPU) and tests/test_gpu_residency.py (the other
// modes). Built by the package build with g++ against libaether_hip.so. Each mode prints "ok" on success,
// or the first failed check and exits 1.
//
// Modes:
//   host              : padded_block_samples / partition_view closed forms and refusals, host routes
//   shards <parts>    : the cfg0 expression over a PartitionedArray<double,3> of parts x 2^21 samples on a
//                       device list that repeats device 0, per-rank fused Neumaier sums combined over
//                       ranks, against the one-array evaluation bit for bit; gather against it too
//   short             : a short last block of a rank-2 array (N = 3): scatter, mutate, gather round trip
//   transport         : routes; forced staging copies bit-identical to direct ones (sync, async, ordered);
//                       moveTo / exchange through a staging transport; replicas; refusals
#include <aether/aether.hpp>
#include <aether/residency.hpp>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

using aether::Chunk;
using aether::Device;
using aether::PartitionedArray;
using aether::Route;

namespace {

int failures = 0;

#define CHECK(cond)                                                                                                   \
    do {                                                                                                              \
        if (!(cond)) {                                                                                                \
            std::printf("check failed at line %d: %s\n", __LINE__, #cond);                                           \
            ++failures;                                                                                               \
            return;                                                                                                   \
        }                                                                                                             \
    } while (0)

template<class F>
bool throws(F&& f)
{
    try {
        f();
    } catch (const aether::Error&) {
        return true;
    }
    return false;
}

std::uint64_t bits(double x)
{
    std::uint64_t u;
    std::memcpy(&u, &x, 8);
    return u;
}

void host_mode()
{
    using aether::padded_block_samples;
    CHECK(padded_block_samples(100, 4, 32) == 32);  // ceil(100/4) = 25 -> 32
    CHECK(padded_block_samples(100, 3, 32) == 64);  // ceil(100/3) = 34 -> 64
    CHECK(padded_block_samples(96, 3, 32) == 32);   // already a multiple
    CHECK(padded_block_samples(10, 4) == 32);       // default pad_to
    CHECK(padded_block_samples(10, 4, 4) == 4 && padded_block_samples(10, 4, 1) == 3);
    CHECK(padded_block_samples(0, 4, 32) == 0);
    CHECK(throws([] { (void)padded_block_samples(100, 0, 32); }));
    CHECK(throws([] { (void)padded_block_samples(100, 4, 0); }));

    // n = 100 samples of 3 components (pitch 100), 4 blocks of pitch 32: three full, one of 4
    std::vector<double> mem(3 * 100);
    for (std::size_t k = 0; k < mem.size(); ++k) mem[k] = (double)k;
    const aether::View<double, 3> full(mem.data(), 100, 100, false);
    const aether::PartitionSpec spec{4, 32};
    for (std::size_t r = 0; r < 4; ++r) {
        const auto b = aether::partition_view(full, spec, r);
        CHECK(b.data() == mem.data() + 32 * r && b.pitch() == 100);
        CHECK(b.samples() == (r < 3 ? 32u : 4u));
        CHECK(b(2, 0) == (double)(200 + 32 * r)); // component planes keep the whole view's pitch
    }
    CHECK(throws([&] { (void)aether::partition_view(full, spec, 4); }));
    const auto exact = aether::partition_view(aether::View<double, 3>(mem.data(), 96, 100, false), {3, 32}, 2);
    CHECK(exact.samples() == 32 && exact.data() == mem.data() + 64);
    // a block that starts past the end is empty
    CHECK(aether::partition_view(aether::View<double, 3>(mem.data(), 10, 100, false), {4, 8}, 3).samples() == 0);

    CHECK(aether::peerRoute(Device::cpu(), Device::cpu()) == Route::HOST);
    CHECK(aether::peerRoute(Device::cpu(), Device::gpu(0)) == Route::HOST);
    CHECK(aether::peerRoute(Device::pinned(), Device::gpu(0), true) == Route::HOST); // forcing needs two GPUs
    CHECK(throws([] { PartitionedArray<double, 3>(std::vector<Device>{}, 100); }));

    // the rank combine in the device's halving order: 4 ranks give (s0 + s2) + (s1 + s3)
    const std::vector<std::array<double, 2>> four = {{1.0, 0.0}, {1e16, 0.0}, {-1e16, 0.0}, {1.0, 0.0}};
    const auto c4 = aether::combine_ranks(four, aether::SumMode::Neumaier);
    CHECK(c4[0] + c4[1] == 2.0);
    const auto p4 = aether::combine_ranks(four, aether::SumMode::Plain);
    CHECK(p4[0] == ((1.0 + -1e16) + (1e16 + 1.0)) && p4[1] == 0.0);
}

template<class V>
auto cfg0(const V& a, const V& b, const V& c)
{
    using T = typename V::element_type;
    return a + b * c + aether::cwiseSin(a) * aether::cwiseExp(T(-1) * (b * b));
}

void shards_mode(std::size_t parts)
{
    const std::size_t block = std::size_t(1) << 21; // 1024 tiles of 2048 samples: one group of the sum tree
    const std::size_t n = parts * block;
    const aether::random::Generator g(42);
    // the whole arrays, evaluated in one piece
    aether::Array<double, 3> a(n), b(n), c(n), d(n);
    a.deviceView() = g.uniform<double, 3>(-1.0, 1.0);
    b.deviceView() = g.substream(1).uniform<double, 3>(-1.0, 1.0);
    c.deviceView() = g.substream(2).uniform<double, 3>(-1.0, 1.0);
    const auto whole = aether::evaluate_pairs(d.deviceView(), cfg0(a.deviceView(), b.deviceView(), c.deviceView()));
    a.download();
    b.download();
    c.download();
    d.download();

    // the same arrays partitioned over `parts` ranks that all live on device 0
    const std::vector<Device> devs(parts, Device::gpu(0));
    PartitionedArray<double, 3> pa(devs, n), pb(devs, n), pc(devs, n), pd(devs, n);
    CHECK(pa.padded() == block && pa.realCount(parts - 1) == block);
    for (auto [src, dst] : {std::pair{&a, &pa}, std::pair{&b, &pb}, std::pair{&c, &pc}}) {
        auto hv = dst->hostView();
        auto sv = src->hostView();
        for (std::size_t k = 0; k < 3; ++k) std::memcpy(&hv(k, 0), &sv(k, 0), n * sizeof(double));
    }
    pa.scatter();
    pb.scatter();
    pc.scatter();
    std::vector<std::array<double, 2>> per_rank[3];
    for (std::size_t r = 0; r < parts; ++r) {
        const auto s = aether::evaluate_pairs(pd.deviceView(r), cfg0(pa.deviceView(r), pb.deviceView(r),
                                                                   pc.deviceView(r)));
        for (std::size_t k = 0; k < 3; ++k) per_rank[k].push_back(s[k]);
    }
    for (std::size_t k = 0; k < 3; ++k) {
        const auto s = aether::combine_ranks(per_rank[k], aether::SumMode::Neumaier);
        CHECK(bits(s[0]) == bits(whole[k][0]) && bits(s[1]) == bits(whole[k][1]));
    }
    pd.gather();
    auto got = pd.hostView();
    auto want = d.hostView();
    for (std::size_t k = 0; k < 3; ++k) CHECK(std::memcmp(&got(k, 0), &want(k, 0), n * sizeof(double)) == 0);
}

void short_mode()
{
    // n = 100 over 4 ranks of pitch 32: the last rank holds 4 real samples of each of 3 components
    const std::size_t n = 100;
    PartitionedArray<double, 3> pa(std::vector<Device>(4, Device::gpu(0)), n);
    CHECK(pa.parts() == 4 && pa.padded() == 32 && pa.realCount(3) == 4 && pa.realCount(0) == 32);
    auto hv = pa.hostView();
    for (std::size_t k = 0; k < 3; ++k)
        for (std::size_t i = 0; i < n; ++i) hv(k, i) = 1000.0 * (double)k + (double)i;
    pa.scatter();
    // change every real sample on the devices (2 * x + 1, one kernel per rank), then gather
    for (std::size_t r = 0; r < 4; ++r) {
        auto v = pa.deviceView(r);
        CHECK(v.pitch() == 32 && v.samples() == pa.realCount(r));
        v = 2.0 * v + 1.0;
    }
    for (std::size_t k = 0; k < 3; ++k)
        for (std::size_t i = 0; i < n; ++i) hv(k, i) = -1.0;
    std::vector<aether::Stream> streams(4, nullptr);
    pa.gather(streams);
    CHECK(aether::detail::seam().stream_sync(0, nullptr) == 0);
    for (std::size_t k = 0; k < 3; ++k)
        for (std::size_t i = 0; i < n; ++i) CHECK(hv(k, i) == 2.0 * (1000.0 * (double)k + (double)i) + 1.0);
    CHECK(throws([&] { pa.scatter(std::vector<aether::Stream>(3, nullptr)); }));
}

void fill_bytes(Chunk& c, unsigned seed)
{
    std::vector<unsigned char> h(c.size());
    for (std::size_t i = 0; i < h.size(); ++i) h[i] = (unsigned char)((i * 2654435761u + seed) >> 7);
    Chunk src = Chunk::borrow(h.data(), h.size(), Device::cpu());
    aether::copy(c, src);
}

std::vector<unsigned char> bytes_of(const Chunk& c)
{
    std::vector<unsigned char> h(c.size());
    Chunk dst = Chunk::borrow(h.data(), h.size(), Device::cpu());
    aether::copy(dst, c);
    return h;
}

void transport_mode()
{
    CHECK(aether::peerRoute(Device::gpu(0), Device::gpu(0)) == Route::SAME);
    CHECK(aether::peerRoute(Device::gpu(0), Device::gpu(0), true) == Route::STAGED);
    CHECK(aether::peerRoute(Device::gpu(0), Device::pinned()) == Route::HOST);

    const std::size_t bytes = (std::size_t(3) << 20) + 77;
    Chunk src = Chunk::allocate(Device::gpu(0), bytes);
    fill_bytes(src, 5);
    const auto want = bytes_of(src);
    aether::StreamTransport direct, staged(true);
    CHECK(staged.forceStaging() && staged.route(Device::gpu(0), Device::gpu(0)) == Route::STAGED);
    Chunk d1 = Chunk::allocate(Device::gpu(0), bytes), d2 = Chunk::allocate(Device::gpu(0), bytes);
    direct.copy(d1, src);
    staged.copy(d2, src);
    CHECK(bytes_of(d1) == want && bytes_of(d2) == want);
    // asynchronous staged copies in a row share the bounce buffer: each waits for the previous one
    Chunk d3 = Chunk::allocate(Device::gpu(0), bytes), d4 = Chunk::allocate(Device::gpu(0), bytes);
    staged.copyAsync(d3, src, nullptr);
    staged.copyAsync(d4, d3, nullptr);
    CHECK(staged.lastEvent() != nullptr);
    staged.fence();
    CHECK(bytes_of(d4) == want);
    // ordered copies from and to pageable host memory
    std::vector<unsigned char> page(want);
    Chunk hsrc = Chunk::borrow(page.data(), bytes, Device::cpu());
    Chunk d5 = Chunk::allocate(Device::gpu(0), bytes);
    staged.copyOrdered(d5, hsrc, nullptr);
    std::vector<unsigned char> back(bytes, 0);
    Chunk hdst = Chunk::borrow(back.data(), bytes, Device::cpu());
    staged.copyOrdered(hdst, d5, nullptr);
    CHECK(back == want);

    // moveTo and exchange through the staging transport keep every byte, padding included
    const std::size_t n = 1000;
    PartitionedArray<double, 2> pa(std::vector<Device>(3, Device::gpu(0)), n);
    auto hv = pa.hostView();
    for (std::size_t k = 0; k < 2; ++k)
        for (std::size_t i = 0; i < n; ++i) hv(k, i) = (double)(k * n + i);
    pa.scatter();
    pa.moveTo(0, Device::gpu(0), staged);
    CHECK(pa.device(0) == Device::gpu(0));
    pa.exchange(0, 2, staged);
    pa.exchange(1, 1, staged); // a no-op
    for (std::size_t k = 0; k < 2; ++k)
        for (std::size_t i = 0; i < n; ++i) hv(k, i) = -1.0;
    pa.exchange(0, 2, staged, nullptr); // back again, asynchronously
    staged.fence();
    pa.gather();
    for (std::size_t k = 0; k < 2; ++k)
        for (std::size_t i = 0; i < n; ++i) CHECK(hv(k, i) == (double)(k * n + i));
    CHECK(throws([&] { pa.moveTo(3, Device::gpu(0), staged); }));
    CHECK(throws([&] { pa.exchange(0, 5, staged); }));
    CHECK(throws([&] { (void)pa.deviceView(3); }));

    // replicas: every copy equals the source bit for bit
    aether::ReplicaSet set({Device::gpu(0), Device::gpu(0), Device::pinned()}, bytes);
    set.broadcast(src);
    for (std::size_t i = 0; i < set.size(); ++i) CHECK(bytes_of(set.chunk(i)) == want);
    aether::Replica<unsigned char> rep({Device::gpu(0), Device::gpu(0)}, bytes);
    rep.broadcast(src, nullptr);
    rep.transport().fence();
    for (std::size_t i = 0; i < rep.size(); ++i) CHECK(bytes_of(rep.chunk(i)) == want);
    CHECK(throws([&] {
        Chunk small = Chunk::allocate(Device::gpu(0), 16);
        aether::copy(small, src);
    }));
}

} // namespace

int main(int argc, char** argv)
{
    if (argc < 2) return 2;
    const std::string mode = argv[1];
    try {
        if (mode == "host") host_mode();
        else if (mode == "shards" && argc == 3) shards_mode(std::strtoull(argv[2], nullptr, 10));
        else if (mode == "short") short_mode();
        else if (mode == "transport") transport_mode();
        else return 2;
    } catch (const std::exception& e) {
        std::printf("exception: %s\n", e.what());
        return 1;
    }
    if (failures) return 1;
    std::printf("ok\n");
    return 0;
}
