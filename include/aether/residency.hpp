// aether/residency.hpp -- where arrays live, how their blocks are laid out across devices and how bytes
// move between them, in aether's spelling (residency/Partition.h, PartitionedArray.h, PeerAccess.h,
// Transport.h, Replica.h; chunk/Chunk.h, chunk/Copy.h).
//
// Host-only (g++ or clang, C++20, no HIP headers), over the C seam of libaether_hip.so:
//
//   aether::PartitionedArray<double, 3> pa({aether::Device::gpu(0), aether::Device::gpu(1)}, n);
//   fill(pa.hostView());  pa.scatter();
//   for (r ...) { set device r; auto sums = aether::evaluate(pa.deviceView(r), expr); }
//   pa.gather();
//
// A rank's block is `padded_block_samples(n, parts, pad_to)` samples long on its device (the last one
// may hold fewer real samples); the host copy is compact (component c of sample i at c*n + i). Copies
// between devices take the route peerRoute() names: the same device, peer to peer, or staged through a
// pinned bounce buffer (forced with StreamTransport(true), the path of a GPU pair without peer access,
// which one GPU can exercise by repeating its own id).
#pragma once

#include <concepts>
#include <cstddef>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "aether.hpp"

namespace aether {

// ------------------------------------------------------------------ places and chunks
/** Where memory lives (device/Device.h): host memory, page-locked host memory, or a GPU's memory. */
enum class MemoryKind : int { CPU = AETHER_MEM_CPU, GPU = AETHER_MEM_DEVICE, Pinned = AETHER_MEM_PINNED };

class Device {
public:
    constexpr Device() = default;
    constexpr explicit Device(MemoryKind kind, int id = 0) : kind_(kind), id_(id) {}
    static constexpr Device gpu(int id) { return Device(MemoryKind::GPU, id); }
    static constexpr Device cpu() { return Device(MemoryKind::CPU, 0); }
    static constexpr Device pinned() { return Device(MemoryKind::Pinned, 0); }
    constexpr MemoryKind kind() const { return kind_; }
    constexpr int id() const { return id_; }
    constexpr bool is_gpu() const { return kind_ == MemoryKind::GPU; }
    constexpr bool operator==(const Device&) const = default;

private:
    MemoryKind kind_ = MemoryKind::CPU;
    int id_ = 0;
};

/**
 * A byte range in one place (chunk/Chunk.h:53-160): an owner (allocate: freed on release, stream-ordered
 * for GPU memory and deferred while a capture is running) or a borrower (borrow: never frees). Move-only.
 */
class Chunk {
public:
    Chunk() = default;
    static Chunk allocate(Device d, std::size_t bytes, Stream s = nullptr)
    {
        Chunk c;
        c.device_ = d;
        c.bytes_ = bytes;
        if (bytes == 0) return c;
        detail::check(detail::seam().chunk_alloc((int)d.kind(), d.id(), bytes, s, &c.owner_, &c.ptr_),
                      "Chunk::allocate");
        if (d.is_gpu()) detail::check(detail::seam().stream_sync(d.id(), s), "Chunk::allocate");
        return c;
    }
    static Chunk borrow(void* p, std::size_t bytes, Device d)
    {
        Chunk c;
        c.ptr_ = p;
        c.bytes_ = bytes;
        c.device_ = d;
        return c;
    }
    Chunk(Chunk&& o) noexcept { swap(o); }
    Chunk& operator=(Chunk&& o) noexcept
    {
        Chunk gone(std::move(*this));
        swap(o);
        return *this;
    }
    Chunk(const Chunk&) = delete;
    Chunk& operator=(const Chunk&) = delete;
    ~Chunk()
    {
        if (owner_) detail::seam().chunk_release(owner_);
    }
    std::byte* data() const { return static_cast<std::byte*>(ptr_); }
    std::size_t size() const { return bytes_; }
    Device device() const { return device_; }
    bool owning() const { return owner_ != nullptr; }

private:
    void swap(Chunk& o) noexcept
    {
        std::swap(owner_, o.owner_);
        std::swap(ptr_, o.ptr_);
        std::swap(bytes_, o.bytes_);
        std::swap(device_, o.device_);
    }
    void* owner_ = nullptr;
    void* ptr_ = nullptr;
    std::size_t bytes_ = 0;
    Device device_{};
};

namespace detail {
inline void same_size(const Chunk& dst, const Chunk& src, const char* op)
{
    if (dst.size() != src.size())
        throw Error(std::string(op) + ": size mismatch (dst " + std::to_string(dst.size()) + " bytes, src " +
                    std::to_string(src.size()) + " bytes)");
}
inline int sync_device(const Chunk& dst, const Chunk& src)
{
    return dst.device().is_gpu() ? dst.device().id() : src.device().is_gpu() ? src.device().id() : current_device();
}
} // namespace detail

/** Stream-ordered copy (chunk/Copy.h:147): returns once queued; pageable host memory is the exception, the
 *  seam completes such a copy before it returns. */
inline void copyAsync(Chunk& dst, const Chunk& src, Stream s)
{
    detail::same_size(dst, src, "copyAsync");
    if (src.size() == 0) return;
    detail::check(detail::seam().copy((int)dst.device().kind(), dst.device().id(), dst.data(),
                                      (int)src.device().kind(), src.device().id(), src.data(), src.size(), s),
                  "copyAsync");
}

/** Blocking copy between any two places (chunk/Copy.h:87). */
inline void copy(Chunk& dst, const Chunk& src)
{
    copyAsync(dst, src, nullptr);
    if (src.size()) detail::check(detail::seam().stream_sync(detail::sync_device(dst, src), nullptr), "copy");
}

// ------------------------------------------------------------------ partitions
/** Split the sample axis into `parts` blocks whose pitch is rounded up to `pad_to` samples
 *  (residency/Partition.h:52-56). */
struct PartitionSpec {
    std::size_t parts;
    std::size_t pad_to = 32;
};

/** The block pitch in samples: ceil(n / parts) rounded up to a multiple of pad_to (Partition.h:65-75). */
inline std::size_t padded_block_samples(std::size_t n, std::size_t parts, std::size_t pad_to = 32)
{
    if (parts == 0) throw Error("padded_block_samples failed, size=" + std::to_string(n) + " (parts must be > 0)");
    if (pad_to == 0) throw Error("padded_block_samples failed, size=" + std::to_string(n) + " (pad_to must be > 0)");
    const std::size_t per = n / parts + (n % parts != 0);
    return (per / pad_to + (per % pad_to != 0)) * pad_to;
}

/** Block r of a view, sharing its memory: samples [r * pitch, min(n, (r + 1) * pitch)), component
 *  planes as far apart as in the whole view (Partition.h:105-160). */
template<class T, std::size_t N>
View<T, N> partition_view(const View<T, N>& v, const PartitionSpec& spec, std::size_t r)
{
    if (r >= spec.parts)
        throw Error("partition_view: block index " + std::to_string(r) + " >= spec.parts=" +
                    std::to_string(spec.parts));
    const std::size_t n = v.samples();
    const std::size_t pitch = padded_block_samples(n, spec.parts, spec.pad_to);
    const std::size_t first = r * pitch;
    const std::size_t real = first < n ? std::min(pitch, n - first) : 0;
    return View<T, N>(v.data() + first, real, v.pitch(), v.on_device());
}

// ------------------------------------------------------------------ routes and transports
enum class Route : int { SAME = AETH_ROUTE_SAME, P2P = AETH_ROUTE_P2P, STAGED = AETH_ROUTE_STAGED,
                         HOST = AETH_ROUTE_HOST };

/** How bytes travel from a to b (PeerAccess.h:97-129); P2P enables peer access both ways. */
inline Route peerRoute(const Device& a, const Device& b, bool forceStaging = false)
{
    int r = AETH_ROUTE_HOST;
    detail::check(aeth_peer_route((int)a.kind(), a.id(), (int)b.kind(), b.id(), forceStaging ? 1 : 0, &r),
                  "peerRoute");
    return static_cast<Route>(r);
}

/** A pinned host staging buffer that only grows (PeerAccess.h:143-190). */
class BounceBuffer {
public:
    void ensure(std::size_t bytes)
    {
        if (bytes <= chunk_.size()) return;
        chunk_ = Chunk::allocate(Device::pinned(), bytes);
    }
    Chunk& chunk() { return chunk_; }
    std::size_t capacity() const { return chunk_.size(); }

private:
    Chunk chunk_;
};

/** What a transport offers (Transport.h:40-52). */
template<class X>
concept transport = requires(X t, Chunk& dst, const Chunk& src, Stream s) {
    { t.copy(dst, src) } -> std::same_as<void>;
    { t.copyAsync(dst, src, s) } -> std::same_as<void>;
};

/**
 * Copies along peerRoute()'s route (Transport.h:78-287). A STAGED copy goes source -> pinned bounce ->
 * destination. The bounce buffer is shared by this transport's copies, so a staged copy first waits for
 * the previous one that used it. copyAsync records an event per call (lastEvent / wait / fence).
 */
class StreamTransport {
public:
    StreamTransport() = default;
    explicit StreamTransport(bool forceStaging) : forceStaging_(forceStaging) {}
    StreamTransport(const StreamTransport&) = delete;
    StreamTransport& operator=(const StreamTransport&) = delete;
    StreamTransport(StreamTransport&& o) noexcept { take(o); }
    StreamTransport& operator=(StreamTransport&& o) noexcept
    {
        if (this != &o) {
            drop();
            take(o);
        }
        return *this;
    }
    ~StreamTransport() { drop(); }

    Route route(const Device& a, const Device& b) const { return peerRoute(a, b, forceStaging_); }
    bool forceStaging() const { return forceStaging_; }

    void copy(Chunk& dst, const Chunk& src)
    {
        if (route(src.device(), dst.device()) != Route::STAGED) {
            aether::copy(dst, src);
            return;
        }
        Chunk stage = bounce(src.size());
        aether::copy(stage, src);
        aether::copy(dst, stage);
    }

    void copyAsync(Chunk& dst, const Chunk& src, Stream s)
    {
        if (route(src.device(), dst.device()) == Route::STAGED) {
            Chunk stage = bounce(src.size());
            aether::copyAsync(stage, src, s);
            aether::copyAsync(dst, stage, s);
            bounce_event_ = record(s);
        } else {
            aether::copyAsync(dst, src, s);
            record(s);
        }
    }

    /** Ordered on `s` for every pair of places; pageable host memory goes through the bounce buffer so the
     *  GPU side of the copy stays stream-ordered (Transport.h:120-170). */
    void copyOrdered(Chunk& dst, const Chunk& src, Stream s)
    {
        detail::same_size(dst, src, "copyOrdered");
        const bool src_gpu = src.device().is_gpu(), dst_gpu = dst.device().is_gpu();
        const bool src_pageable = src.device().kind() == MemoryKind::CPU;
        const bool dst_pageable = dst.device().kind() == MemoryKind::CPU;
        if (!src_gpu && !dst_gpu) {
            detail::check(detail::seam().stream_sync(detail::current_device(), s), "copyOrdered");
            if (src.size()) std::memcpy(dst.data(), src.data(), src.size());
            return;
        }
        if (!src_pageable && !dst_pageable) {
            aether::copyAsync(dst, src, s);
            return;
        }
        if (src.size() == 0) return;
        Chunk stage = bounce(src.size());
        if (src_pageable) {
            std::memcpy(stage.data(), src.data(), src.size());
            aether::copyAsync(dst, stage, s);
            bounce_event_ = record(s);
        } else {
            aether::copyAsync(stage, src, s);
            detail::check(detail::seam().stream_sync(src.device().id(), s), "copyOrdered");
            std::memcpy(dst.data(), stage.data(), src.size());
        }
    }

    void* lastEvent() const { return events_.empty() ? nullptr : events_.back(); }
    void wait(void* event) const { detail::check(aeth_event_synchronize(event), "StreamTransport::wait"); }
    void fence() const
    {
        for (void* e : events_) wait(e);
    }

private:
    // A view of the bounce buffer for `bytes`, once the last staged copy that used it has completed.
    Chunk bounce(std::size_t bytes)
    {
        if (bounce_event_) {
            wait(bounce_event_);
            bounce_event_ = nullptr;
        }
        bounce_.ensure(bytes);
        return Chunk::borrow(bounce_.chunk().data(), bytes, Device::pinned());
    }
    void* record(Stream s)
    {
        void* e = nullptr;
        detail::check(aeth_event_record(s, &e), "StreamTransport event");
        events_.push_back(e);
        return e;
    }
    void drop()
    {
        for (void* e : events_) (void)aeth_event_destroy(e);
        events_.clear();
        bounce_event_ = nullptr;
    }
    void take(StreamTransport& o)
    {
        forceStaging_ = o.forceStaging_;
        bounce_ = std::move(o.bounce_);
        events_ = std::move(o.events_);
        bounce_event_ = o.bounce_event_;
        o.events_.clear();
        o.bounce_event_ = nullptr;
    }

    bool forceStaging_ = false;
    BounceBuffer bounce_;
    std::vector<void*> events_;
    void* bounce_event_ = nullptr; // the last staged copy's completion, while the bounce buffer is in use
};
static_assert(transport<StreamTransport>, "StreamTransport must satisfy aether::transport");

// ------------------------------------------------------------------ replicas
/** One chunk per device, filled together by broadcast() (Replica.h:56-100). */
class ReplicaSet {
public:
    ReplicaSet() = default;
    ReplicaSet(const std::vector<Device>& devices, std::size_t bytes)
    {
        chunks_.reserve(devices.size());
        for (const Device& d : devices) chunks_.push_back(Chunk::allocate(d, bytes));
    }
    std::size_t size() const { return chunks_.size(); }
    Chunk& chunk(std::size_t i) { return chunks_[i]; }
    const Chunk& chunk(std::size_t i) const { return chunks_[i]; }
    void broadcast(const Chunk& src)
    {
        for (Chunk& c : chunks_) aether::copy(c, src);
    }
    void broadcast(const Chunk& src, Stream s)
    {
        for (Chunk& c : chunks_) aether::copyAsync(c, src, s);
    }

private:
    std::vector<Chunk> chunks_;
};

/** `count` elements of T on each device, broadcast through a transport (Replica.h:120-181). */
template<class T, transport Transport = StreamTransport>
class Replica {
public:
    Replica() = default;
    Replica(std::vector<Device> devices, std::size_t count) : devices_(std::move(devices))
    {
        chunks_.reserve(devices_.size());
        for (const Device& d : devices_) chunks_.push_back(Chunk::allocate(d, count * sizeof(T)));
    }
    std::size_t size() const { return chunks_.size(); }
    Chunk& chunk(std::size_t i) { return chunks_[i]; }
    const Chunk& chunk(std::size_t i) const { return chunks_[i]; }
    Transport& transport() { return transport_; }
    void broadcast(const Chunk& src)
    {
        for (Chunk& c : chunks_) transport_.copy(c, src);
    }
    void broadcast(const Chunk& src, Stream s)
    {
        for (Chunk& c : chunks_) transport_.copyAsync(c, src, s);
    }

private:
    std::vector<Device> devices_;
    std::vector<Chunk> chunks_;
    Transport transport_{};
};

// ------------------------------------------------------------------ partitioned arrays
/**
 * Array<T, N> split along the sample axis over a device list (PartitionedArray.h:96-416): a compact pinned
 * host copy of n samples, and on each rank's device a block of padded() samples per component. scatter()
 * and gather() move each rank's real samples, never the padding; moveTo() relocates a whole block to
 * another device and exchange() swaps two ranks' blocks, both through a transport.
 */
template<class T, std::size_t N = 1>
class PartitionedArray {
public:
    PartitionedArray(std::vector<Device> devices, std::size_t n, std::size_t pad_to = 32)
        : devices_(std::move(devices)), n_(n), padded_(padded_block_samples(n, devices_.size(), pad_to)),
          host_(Chunk::allocate(Device::pinned(), N * n * sizeof(T)))
    {
        blocks_.reserve(devices_.size());
        for (const Device& d : devices_) blocks_.push_back(Chunk::allocate(d, N * padded_ * sizeof(T)));
    }

    std::size_t parts() const { return devices_.size(); }
    std::size_t samples() const { return n_; }
    std::size_t padded() const { return padded_; }
    std::size_t realCount(std::size_t r) const
    {
        const std::size_t first = r * padded_;
        return first < n_ ? std::min(padded_, n_ - first) : 0;
    }
    Device device(std::size_t r) const { return devices_[r]; }

    /** The compact host copy: fill it before scatter(), read it after gather(). */
    View<T, N> hostView() { return View<T, N>(reinterpret_cast<T*>(host_.data()), n_, n_, false); }
    /** Rank r's block: realCount(r) samples, planes padded() elements apart. */
    View<T, N> deviceView(std::size_t r)
    {
        check_rank(r, "deviceView");
        return View<T, N>(reinterpret_cast<T*>(blocks_[r].data()), realCount(r), padded_,
                          devices_[r].is_gpu());
    }

    void scatter()
    {
        for (std::size_t r = 0; r < parts(); ++r) rows(r, true, nullptr, false);
    }
    void gather()
    {
        for (std::size_t r = 0; r < parts(); ++r) rows(r, false, nullptr, false);
    }
    /** Asynchronous forms, one stream per rank; the caller synchronises. */
    void scatter(const std::vector<Stream>& streams)
    {
        check_streams(streams, "scatter");
        for (std::size_t r = 0; r < parts(); ++r) rows(r, true, streams[r], true);
    }
    void gather(const std::vector<Stream>& streams)
    {
        check_streams(streams, "gather");
        for (std::size_t r = 0; r < parts(); ++r) rows(r, false, streams[r], true);
    }

    template<transport Transport>
    void moveTo(std::size_t r, Device d, Transport& t)
    {
        check_rank(r, "moveTo");
        Chunk moved = Chunk::allocate(d, blocks_[r].size());
        t.copy(moved, blocks_[r]);
        blocks_[r] = std::move(moved);
        devices_[r] = d;
    }
    template<transport Transport>
    void moveTo(std::size_t r, Device d, Transport& t, Stream s)
    {
        check_rank(r, "moveTo");
        Chunk moved = Chunk::allocate(d, blocks_[r].size(), s);
        t.copyAsync(moved, blocks_[r], s);
        blocks_[r] = std::move(moved);
        devices_[r] = d;
    }

    template<transport Transport>
    void exchange(std::size_t ra, std::size_t rb, Transport& t)
    {
        check_rank(ra, "exchange");
        check_rank(rb, "exchange");
        if (ra == rb) return;
        Chunk a_on_b = Chunk::allocate(devices_[rb], blocks_[ra].size());
        Chunk b_on_a = Chunk::allocate(devices_[ra], blocks_[rb].size());
        t.copy(a_on_b, blocks_[ra]);
        t.copy(b_on_a, blocks_[rb]);
        blocks_[ra] = std::move(b_on_a);
        blocks_[rb] = std::move(a_on_b);
    }
    template<transport Transport>
    void exchange(std::size_t ra, std::size_t rb, Transport& t, Stream s)
    {
        check_rank(ra, "exchange");
        check_rank(rb, "exchange");
        if (ra == rb) return;
        Chunk a_on_b = Chunk::allocate(devices_[rb], blocks_[ra].size(), s);
        Chunk b_on_a = Chunk::allocate(devices_[ra], blocks_[rb].size(), s);
        t.copyAsync(a_on_b, blocks_[ra], s);
        t.copyAsync(b_on_a, blocks_[rb], s);
        blocks_[ra] = std::move(b_on_a);
        blocks_[rb] = std::move(a_on_b);
    }

private:
    // Component planes of rank r's real samples between the host copy (pitch n) and the block (pitch
    // padded), one copy per component.
    void rows(std::size_t r, bool to_device, Stream s, bool async)
    {
        const std::size_t real = realCount(r);
        if (real == 0) return;
        for (std::size_t c = 0; c < N; ++c) {
            Chunk h = Chunk::borrow(host_.data() + (c * n_ + r * padded_) * sizeof(T), real * sizeof(T),
                                    host_.device());
            Chunk b = Chunk::borrow(blocks_[r].data() + c * padded_ * sizeof(T), real * sizeof(T), devices_[r]);
            Chunk& dst = to_device ? b : h;
            const Chunk& src = to_device ? h : b;
            if (async) aether::copyAsync(dst, src, s);
            else aether::copy(dst, src);
        }
    }
    void check_rank(std::size_t r, const char* op) const
    {
        if (r >= parts())
            throw Error(std::string("PartitionedArray::") + op + ": rank " + std::to_string(r) + " >= parts()=" +
                        std::to_string(parts()));
    }
    void check_streams(const std::vector<Stream>& streams, const char* op) const
    {
        if (streams.size() != parts())
            throw Error(std::string("PartitionedArray::") + op + ": one stream per rank (" +
                        std::to_string(parts()) + "), got " + std::to_string(streams.size()));
    }

    std::vector<Device> devices_;
    std::size_t n_, padded_;
    Chunk host_;
    std::vector<Chunk> blocks_;
};

// ------------------------------------------------------------------ sums over ranks
/**
 * Combine per-rank (hi, comp) sums in the order of the device reduction's group tree (a halving tree:
 * pairs (t, t + stride) for stride = 512 .. 1), with the combine of `mode` (pairCombine_ for Neumaier,
 * accum/atomic.h:157-164; the float-float add for FloatFloat; a plain add otherwise). When every rank holds
 * a whole number of 1024^k-tile groups of the sample axis (2048-sample tiles) and the ranks are in sample
 * order, the result is the one-device sum bit for bit. Up to 1024 ranks.
 */
template<class T>
std::array<T, 2> combine_ranks(std::vector<std::array<T, 2>> s, SumMode mode)
{
    if (s.empty()) return {T(0), T(0)};
    if (s.size() > 1024) throw Error("combine_ranks: at most 1024 ranks");
    auto two_sum_err = [](T a, T b, T sum) {
        const T bb = sum - a;
        const T e = (a - (sum - bb)) + (b - bb);
        return e == T(0) ? T(0) : e;
    };
    auto combine = [&](std::array<T, 2> a, std::array<T, 2> b) -> std::array<T, 2> {
        if (mode == SumMode::Plain) return {a[0] + b[0], T(0)};
        if (mode == SumMode::FloatFloat) {
            auto ts = [](T x, T y, T& sum, T& err) {
                sum = x + y;
                const T bp = sum - x;
                const T ap = sum - bp;
                err = (x - ap) + (y - bp);
            };
            T sh, sl, th, tl, vh, vl;
            ts(a[0], b[0], sh, sl);
            ts(a[1], b[1], th, tl);
            const T c = sl + th;
            ts(sh, c, vh, vl);
            const T w = vl + tl;
            const T hi = vh + w;
            const T bp = hi - vh;
            return {hi, w - bp};
        }
        const T sum = a[0] + b[0];
        const T e = two_sum_err(a[0], b[0], sum);
        return {sum, std::isfinite(sum) ? (a[1] + b[1]) + e : T(0)};
    };
    const std::size_t m = s.size();
    for (std::size_t stride = 512; stride >= 1; stride >>= 1)
        for (std::size_t t = 0; t < stride && t + stride < m; ++t) s[t] = combine(s[t], s[t + stride]);
    return s[0];
}

} // namespace aether
