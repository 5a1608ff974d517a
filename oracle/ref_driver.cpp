// ref_driver.cpp -- drives the REFERENCE's own header-only CPU path.
// TEST INFRASTRUCTURE ONLY. Compiled by oracle/Makefile (target `ref`) against
// the reference headers where they lie (-I <aether source tree>,
// -DAETHER_CPP_MODE=1); the binary goes to oracle/_ref/ and is used to mint
// small golden vectors (tests/golden/mint_ref_vectors.py) and, on the GPU
// host, as the timed CPU baseline of the fused FP64 expression.
//
// Modes (all output is raw little-endian binary on stdout unless noted):
//   cfg0   <n> <seedA> <seedB> <seedC>  : a,b,c,d (3 planes x n each, pitch = n
//                                         compacted) then 3 tree sums of d
//   cfg0g  <n> <seed>                   : same, inputs from Generator(seed).uniform
//   gnormal <n> <seed>                  : Generator(seed).normal<double,3>(0.5, 2) planes, then
//                                         Generator(seed).substream(3).normal<float,3>(-1, 0.25) planes
//   cfg0f  <n> <seedA> <seedB> <seedC>  : the float variant
//   vec3   <n> <seedA> <seedB> <seedC>  : e (n doubles)
//   keys   <seed> <streamId> <k>        : generator key, substream key (text)
//   rows   <seed> <n> <counter>         : per global i<n: bits64, u01d bits,
//                                         u01f bits, normal d, normal f
//   fft    <n> <seed> <fwd>             : pocketfft c2c of a length-n signal
//                                         (entries u01-derived), in then out
//   bench  <n> <reps>                   : times the OpenMP fused loop (text)
//   members[f] <n> <sQ> <sR> <sV>       : q, r (4 planes), v (3 planes) of raw-seed uniforms, then the
//                                         per-sample members and elementwise functions listed in
//                                         members() below, each as its planes, in that order
//                                         (membersf: in float)
#include <aether/aether.h>
#include <aether/fft/Fft.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <omp.h>

using aether::SampleIndex;
namespace rnd = aether::random::detail;
namespace at = aether::accum::atomic::detail;

namespace {

// The raw-seed uniform leaf: key = seed ^ salt so the leaf draws u01(seed, i, c).
template<class T>
auto rawUniform(std::uint64_t seed)
{
    using Leaf = rnd::UniformLeaf<T, 3, T, T>;
    return Leaf(seed ^ Leaf::salt, T(-1), T(1));
}

template<class T, std::size_t N>
auto rawUniformN(std::uint64_t seed)
{
    using Leaf = rnd::UniformLeaf<T, N, T, T>;
    return Leaf(seed ^ Leaf::salt, T(-1), T(1));
}

template<class T, std::size_t N = 3>
void writePlanes(const aether::Array<T, N>& a, std::size_t n)
{
    auto v = a.hostView();
    for (std::size_t c = 0; c < N; ++c)
        for (std::size_t i = 0; i < n; ++i) {
            const T x = v(c, i);
            std::fwrite(&x, sizeof(T), 1, stdout);
        }
}

template<class T>
T treeSum(const aether::Array<T, 3>& d, std::size_t n, std::size_t c)
{
    auto v = d.hostView();
    T h[32] = {}, cc[32] = {};
    for (std::size_t i = 0; i < n; ++i)
        at::neumaierAdd_(h[i % 32], cc[i % 32], static_cast<T>(v(c, i)));
    T oh, oc;
    at::reduce32Host_(h, cc, oh, oc);
    return oh + oc;
}

template<class T, class LA, class LB, class LC>
void cfg0(std::size_t n, const LA& la, const LB& lb, const LC& lc)
{
    aether::Array<T, 3> a(n), b(n), c(n), d(n);
    auto av = a.hostView();
    auto bv = b.hostView();
    auto cv = c.hostView();
    auto dv = d.hostView();
#pragma omp parallel for schedule(static)
    for (std::ptrdiff_t k = 0; k < static_cast<std::ptrdiff_t>(n); ++k) {
        const auto i = SampleIndex::make(static_cast<std::size_t>(k));
        av[i] = la;
        bv[i] = lb;
        cv[i] = lc;
    }
#pragma omp parallel for schedule(static)
    for (std::ptrdiff_t k = 0; k < static_cast<std::ptrdiff_t>(n); ++k) {
        const auto i = SampleIndex::make(static_cast<std::size_t>(k));
        dv[i] = av + bv * cv + aether::cwiseSin(av) * aether::cwiseExp(T(-1) * (bv * bv));
    }
    writePlanes(a, n);
    writePlanes(b, n);
    writePlanes(c, n);
    writePlanes(d, n);
    for (std::size_t k = 0; k < 3; ++k) {
        const T s = treeSum(d, n, k);
        std::fwrite(&s, sizeof(T), 1, stdout);
    }
}

void vec3(std::size_t n, std::uint64_t sa, std::uint64_t sb, std::uint64_t sc)
{
    aether::Array<double, 3> a(n), b(n), c(n);
    aether::Array<double> e(n);
    auto av = a.hostView();
    auto bv = b.hostView();
    auto cv = c.hostView();
    auto ev = e.hostView();
    for (std::size_t k = 0; k < n; ++k) {
        const auto i = SampleIndex::make(k);
        av[i] = rawUniform<double>(sa);
        bv[i] = rawUniform<double>(sb);
        cv[i] = rawUniform<double>(sc);
    }
    for (std::size_t k = 0; k < n; ++k) {
        const auto i = SampleIndex::make(k);
        const auto ai = av[i].get();
        const auto bi = bv[i].get();
        const auto ci = cv[i].get();
        const double r = ai.cross(bi).dot(ci) + aether::math::sqrt(ai.dot(ai) + 1.0);
        ev(k) = r;
        std::fwrite(&r, sizeof(double), 1, stdout);
    }
}

// The reference's own members and elementwise functions, per sample through its host spelling:
// 4-planes: quatMul(q, r), quatConj(q), quatReciprocal(q), asPureQuaternion(v);
// 3-planes: quatRotate(q, v), unitVector(v), asBack3DVector(q), head<3>(q), cwiseMin(v, 0.25),
//           cwiseMax(v, -0.25), clamp(v, -0.5, 0.5), cwiseSign(v), cwiseDiv(v, back3(q)), cwiseMul(v, head3(q)),
//           cwiseMin(v, back3(q)), cwiseMax(v, back3(q)), cwiseTan(v), cwiseAtan(v), cwiseAsin(v),
//           cwiseAcos(v), cwiseTanh(v), cwiseRsqrt(|v|), cwisePow(|v|, 1.5), cwisePow(|v|, back3(q));
// 1-planes: maxNorm(q), sum(v), norm(v), squaredNorm(v), rNorm(v), rSquaredNorm(v), cubedNorm(v),
//           rCubedNorm(v).
template<class T>
void members(std::size_t n, std::uint64_t sq, std::uint64_t sr, std::uint64_t sv)
{
    aether::Array<T, 4> q(n), r(n);
    aether::Array<T, 3> v(n);
    auto qv = q.hostView();
    auto rv = r.hostView();
    auto vv = v.hostView();
    for (std::size_t k = 0; k < n; ++k) {
        const auto i = SampleIndex::make(k);
        qv[i] = rawUniformN<T, 4>(sq);
        rv[i] = rawUniformN<T, 4>(sr);
        vv[i] = rawUniformN<T, 3>(sv);
    }
    writePlanes(q, n);
    writePlanes(r, n);
    writePlanes(v, n);
    // members are sample-free: they are called on each sample's materialised Item (Expression.h:155-160;
    // called on a view, a reduction such as rNorm() would read sample 0 for every sample)
    auto out4 = [&](auto make) {
        aether::Array<T, 4> o(n);
        auto ov = o.hostView();
        for (std::size_t k = 0; k < n; ++k) {
            const auto i = SampleIndex::make(k);
            ov[i] = make(qv[i].get(), rv[i].get(), vv[i].get());
        }
        writePlanes(o, n);
    };
    auto out3 = [&](auto make) {
        aether::Array<T, 3> o(n);
        auto ov = o.hostView();
        for (std::size_t k = 0; k < n; ++k) {
            const auto i = SampleIndex::make(k);
            ov[i] = make(qv[i].get(), rv[i].get(), vv[i].get());
        }
        writePlanes(o, n);
    };
    out4([](const auto& qi, const auto& ri, const auto&) { return qi.quatMul(ri); });
    out4([](const auto& qi, const auto&, const auto&) { return qi.quatConj(); });
    out4([](const auto& qi, const auto&, const auto&) { return qi.quatReciprocal(); });
    out4([](const auto&, const auto&, const auto& vi) { return vi.asPureQuaternion(); });
    out3([](const auto& qi, const auto&, const auto& vi) { return qi.quatRotate(vi); });
    out3([](const auto&, const auto&, const auto& vi) { return vi.unitVector(); });
    out3([](const auto& qi, const auto&, const auto&) { return qi.asBack3DVector(); });
    out3([](const auto& qi, const auto&, const auto&) { return qi.template head<3>(); });
    out3([](const auto&, const auto&, const auto& vi) { return aether::cwiseMin(vi, T(0.25)); });
    out3([](const auto&, const auto&, const auto& vi) { return aether::cwiseMax(vi, T(-0.25)); });
    out3([](const auto&, const auto&, const auto& vi) { return aether::clamp(vi, T(-0.5), T(0.5)); });
    out3([](const auto&, const auto&, const auto& vi) { return aether::cwiseSign(vi); });
    out3([](const auto& qi, const auto&, const auto& vi) { return aether::cwiseDiv(vi, qi.asBack3DVector()); });
    out3([](const auto& qi, const auto&, const auto& vi) { return vi.cwiseMul(qi.template head<3>()); });
    out3([](const auto& qi, const auto&, const auto& vi) { return aether::cwiseMin(vi, qi.asBack3DVector()); });
    out3([](const auto& qi, const auto&, const auto& vi) { return aether::cwiseMax(vi, qi.asBack3DVector()); });
    out3([](const auto&, const auto&, const auto& vi) { return aether::cwiseTan(vi); });
    out3([](const auto&, const auto&, const auto& vi) { return aether::cwiseAtan(vi); });
    out3([](const auto&, const auto&, const auto& vi) { return aether::cwiseAsin(vi); });
    out3([](const auto&, const auto&, const auto& vi) { return aether::cwiseAcos(vi); });
    out3([](const auto&, const auto&, const auto& vi) { return aether::cwiseTanh(vi); });
    out3([](const auto&, const auto&, const auto& vi) { return aether::cwiseRsqrt(aether::cwiseAbs(vi)); });
    out3([](const auto&, const auto&, const auto& vi) { return aether::cwisePow(aether::cwiseAbs(vi), T(1.5)); });
    out3([](const auto& qi, const auto&, const auto& vi) {
        return aether::cwisePow(aether::cwiseAbs(vi), qi.asBack3DVector());
    });
    std::vector<T> s(8 * n);
    for (std::size_t k = 0; k < n; ++k) {
        const auto i = SampleIndex::make(k);
        const auto qi = qv[i].get();
        const auto vi = vv[i].get();
        s[0 * n + k] = qi.maxNorm();
        s[1 * n + k] = vi.sum();
        s[2 * n + k] = vi.norm();
        s[3 * n + k] = vi.squaredNorm();
        s[4 * n + k] = vi.rNorm();
        s[5 * n + k] = vi.rSquaredNorm();
        s[6 * n + k] = vi.cubedNorm();
        s[7 * n + k] = vi.rCubedNorm();
    }
    std::fwrite(s.data(), sizeof(T), s.size(), stdout);
}

void rows(std::uint64_t seed, std::size_t n, std::uint64_t counter)
{
    for (std::size_t g = 0; g < n; ++g) {
        const auto gi = static_cast<aether::offset_t>(g);
        const std::uint64_t b = rnd::randomBits64(seed, gi, counter);
        const double ud = rnd::uniform01<double>(seed, gi, counter);
        const float uf = rnd::uniform01<float>(seed, gi, counter);
        const double nd = rnd::standardNormal<double>(seed, gi, counter);
        const float nf = rnd::standardNormal<float>(seed, gi, counter);
        std::fwrite(&b, 8, 1, stdout);
        std::fwrite(&ud, 8, 1, stdout);
        std::fwrite(&uf, 4, 1, stdout);
        std::fwrite(&nf, 4, 1, stdout);
        std::fwrite(&nd, 8, 1, stdout);
    }
}

void fft(std::size_t n, std::uint64_t seed, bool forward)
{
    std::vector<double> in(2 * n), out(2 * n);
    for (std::size_t i = 0; i < n; ++i) {
        in[2 * i] = rnd::standardNormal<double>(seed, static_cast<aether::offset_t>(i), 0);
        in[2 * i + 1] = rnd::standardNormal<double>(seed, static_cast<aether::offset_t>(i), 1);
    }
    const std::size_t shape[1] = { n };
    const std::size_t axes[1] = { 0 };
    aether::fft::c2c<double>(in.data(), out.data(), shape, axes, forward, aether::fft::Norm::Backward);
    std::fwrite(in.data(), 8, 2 * n, stdout);
    std::fwrite(out.data(), 8, 2 * n, stdout);
}

void bench(std::size_t n, int reps)
{
    aether::Array<double, 3> a(n), b(n), c(n), d(n);
    auto av = a.hostView();
    auto bv = b.hostView();
    auto cv = c.hostView();
    auto dv = d.hostView();
    const auto la = rawUniform<double>(42), lb = rawUniform<double>(43), lc = rawUniform<double>(44);
#pragma omp parallel for schedule(static)
    for (std::ptrdiff_t k = 0; k < static_cast<std::ptrdiff_t>(n); ++k) {
        const auto i = SampleIndex::make(static_cast<std::size_t>(k));
        av[i] = la;
        bv[i] = lb;
        cv[i] = lc;
        dv[i] = la; // first touch of d by the same thread layout
    }
    std::vector<double> ms, evalMs;
    double sums[3] = {0, 0, 0};
    for (int r = 0; r <= reps; ++r) {
        const auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(static)
        for (std::ptrdiff_t k = 0; k < static_cast<std::ptrdiff_t>(n); ++k) {
            const auto i = SampleIndex::make(static_cast<std::size_t>(k));
            dv[i] = av + bv * cv + aether::cwiseSin(av) * aether::cwiseExp(-1.0 * (bv * bv));
        }
        const auto t1 = std::chrono::steady_clock::now();
        for (std::size_t k = 0; k < 3; ++k) sums[k] = treeSum(d, n, k);
        const auto t2 = std::chrono::steady_clock::now();
        if (r > 0) {
            ms.push_back(std::chrono::duration<double, std::milli>(t2 - t0).count());
            evalMs.push_back(std::chrono::duration<double, std::milli>(t1 - t0).count());
        }
    }
    std::sort(ms.begin(), ms.end());
    std::sort(evalMs.begin(), evalMs.end());
    std::printf("{\"n\": %zu, \"threads\": %d, \"median_ms\": %.6f, \"min_ms\": %.6f, \"eval_median_ms\": %.6f, "
                "\"sum0\": %.17g}\n",
        n, omp_get_max_threads(), ms[ms.size() / 2], ms[0], evalMs[evalMs.size() / 2], sums[0]);
}

} // namespace

int main(int argc, char** argv)
{
    if (argc < 2) return 2;
    const std::string mode = argv[1];
    auto u = [&](int k) { return static_cast<std::uint64_t>(std::strtoull(argv[k], nullptr, 0)); };
    if (mode == "cfg0" && argc == 6) {
        cfg0<double>(u(2), rawUniform<double>(u(3)), rawUniform<double>(u(4)), rawUniform<double>(u(5)));
    } else if (mode == "cfg0f" && argc == 6) {
        cfg0<float>(u(2), rawUniform<float>(u(3)), rawUniform<float>(u(4)), rawUniform<float>(u(5)));
    } else if (mode == "cfg0g" && argc == 4) {
        const aether::random::Generator g(u(3));
        const auto leaf = g.uniform<double, 3>(-1.0, 1.0);
        cfg0<double>(u(2), leaf, g.substream(1).uniform<double, 3>(-1.0, 1.0),
            g.substream(2).uniform<double, 3>(-1.0, 1.0));
    } else if (mode == "gnormal" && argc == 4) {
        // Generator(seed).normal<double,3>(0.5, 2) and Generator(seed).substream(3).normal<float,3>(-1, 0.25)
        const std::size_t n = u(2);
        const aether::random::Generator g(u(3));
        aether::Array<double, 3> x(n);
        aether::Array<float, 3> y(n);
        auto xv = x.hostView();
        auto yv = y.hostView();
        const auto lx = g.normal<double, 3>(0.5, 2.0);
        const auto ly = g.substream(3).normal<float, 3>(-1.0f, 0.25f);
        for (std::size_t k = 0; k < n; ++k) {
            const auto i = SampleIndex::make(k);
            xv[i] = lx;
            yv[i] = ly;
        }
        writePlanes(x, n);
        writePlanes(y, n);
    } else if (mode == "vec3" && argc == 6) {
        vec3(u(2), u(3), u(4), u(5));
    } else if (mode == "keys" && argc == 5) {
        const aether::random::Generator g(u(2), u(3));
        std::printf("%llu %llu\n", static_cast<unsigned long long>(g.key()),
            static_cast<unsigned long long>(g.substream(u(4)).key()));
    } else if (mode == "rows" && argc == 5) {
        rows(u(2), u(3), u(4));
    } else if (mode == "fft" && argc == 5) {
        fft(u(2), u(3), u(4) != 0);
    } else if (mode == "members" && argc == 6) {
        members<double>(u(2), u(3), u(4), u(5));
    } else if (mode == "membersf" && argc == 6) {
        members<float>(u(2), u(3), u(4), u(5));
    } else if (mode == "bench" && argc == 4) {
        bench(u(2), static_cast<int>(u(3)));
    } else {
        return 2;
    }
    return 0;
}
