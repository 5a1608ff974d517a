// api_check.cpp -- drives the C++ front end (include/aether/aether.hpp) the way
// a user of aether's headers would, for the tests in tests/test_cpp_api.py.
// Built by __graft_entry__.build() with g++ against libaether_hip.so.
//
// Modes (binary little-endian on stdout unless noted):
//   tape                       : lowered programs of fixed expressions (text; no GPU)
//   cfg0g  <n> <seed>          : a, b, c from Generator(seed) (substreams 1, 2) on the
//                                device, d = cfg0 with fused Neumaier sums; writes
//                                a, b, c, d (3 x n each) then 3 x (hi, comp)
//   host   <n> <seed>          : the same inputs and d through the per-sample host
//                                spelling (std:: math); writes a, b, c, d
//   vec3   <n> <sA> <sB> <sC>  : raw-seed uniform inputs, e = (a x b).c + sqrt(a.a + 1)
//                                on the device; writes e (n)
//   fft    <n> <fwd>           : reads n complex (re, im) doubles on stdin, writes c2c
//   rfft   <rows> <n>          : reads rows x n doubles on stdin; r2c over the rows' axis,
//                                writes the half spectrum, then c2r back, writes the reals
//   append <n>                 : push_back / commit / clear on the host (text)
//   float  <n> <seed>          : cfg0 in float with float-float sums; writes d, sums
//   resize <n>                 : upload/reserve/resize/download round trip (text)
//   plane  <n>                 : three producers into an AccumPlane<3>; max error of the delivered
//                                sum in ulps of the exact one (text)
//   where  <n> <threshold>     : appendWhere over Generator(5) uniforms into an Array<uint32_t>'s
//                                spare capacity; checks the committed set on the host (text)
//   keyed  <n> <run>           : sorted keyed sum of runs of `run` lanes, compensated; max error over
//                                eps * sum|t| per slot (text)
//   runtime <rows> <cols>      : runtimeEval AddScaled of a transposed and a dense view, and an
//                                Assign cast to float, checked on the host (text)
//   tune   <n> <seed>          : aether::tune with Neumaier sums and without, then the assignments:
//                                values and sums equal the untuned evaluation's (text)
//   ext    <n> <seed>          : q, r (quaternions) and a (3-vector) from Generator(seed) on the
//                                device; quaternion, fold, norm, min/max/clamp and slice members
//                                on the device, then the same per sample on the host; writes q, r,
//                                a, the four device results, the four host results, then
//                                tan(a) + |a|^1.5 from the device
#include <aether/aether.hpp>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

using aether::Array;
using aether::SampleIndex;

namespace {

template<class T, std::size_t N>
void write_planes(const Array<T, N>& a)
{
    auto v = a.hostView();
    for (std::size_t c = 0; c < N; ++c) std::fwrite(&v(c, 0), sizeof(T), a.samples(), stdout);
}

template<class V>
auto cfg0(const V& a, const V& b, const V& c)
{
    using T = typename V::element_type;
    return a + b * c + aether::cwiseSin(a) * aether::cwiseExp(T(-1) * (b * b));
}

void print_tape(const char* name, const aether::detail::TapeCtx& x, const void* const* ids, int nids)
{
    std::printf("expr %s programs %zu leaves %zu consts %zu rngs %zu\n", name, x.programs.size(), x.leaves.size(),
                x.consts.size(), x.rngs.size());
    std::printf("leaves");
    for (auto* p : x.leaves) {
        int k = 0;
        while (k < nids && ids[k] != p) ++k;
        std::printf(" %d", k);
    }
    std::printf("\n");
    for (const auto& p : x.programs) {
        std::printf("program %d %d %d %d :", p.n_loads, p.n_insns, p.out_leaf, p.out_comp);
        for (int k = 0; k < p.n_loads; ++k) std::printf(" %d.%d", p.load_leaf[k], p.load_comp[k]);
        std::printf(" :");
        for (int k = 0; k < p.n_insns; ++k) std::printf(" %d.%d", p.insns[k].op, p.insns[k].arg);
        std::printf("\n");
    }
    for (double v : x.consts) std::printf("const %.17g\n", v);
    for (const auto& r : x.rngs)
        std::printf("rng %llu %.17g %.17g %u %u\n", (unsigned long long)r.seed, r.p0, r.p1, r.comp, r.dist);
}

int tape()
{
    // fake device planes: lowering reads no memory
    static double buf[8][160];
    const void* ids[8];
    for (int k = 0; k < 8; ++k) ids[k] = buf[k];
    aether::View<double, 3> a(buf[0], 10, 32, true), b(buf[1], 10, 32, true), c(buf[2], 10, 32, true),
        d(buf[3], 10, 32, true);
    aether::View<double, 1> e(buf[4], 10, 32, true), s(buf[5], 10, 32, true);
    aether::View<double, 4> q(buf[6], 10, 32, true), r(buf[7], 10, 32, true);
    namespace ad = aether::detail;
    print_tape("cfg0", ad::lower<double, 3>(buf[3], 32, cfg0(a, b, c)), ids, 8);
    print_tape("vec3", ad::lower<double, 1>(buf[4], 32, a.cross(b).dot(c) + aether::sqrt(a.dot(a) + 1.0)), ids, 8);
    print_tape("negdiv", ad::lower<double, 1>(buf[5], 32, -e / 4.0), ids, 8);
    print_tape("ops",
               ad::lower<double, 1>(buf[5], 32,
                                    aether::cwiseLog(aether::cwiseAbs(e) + 0.5) -
                                        aether::cwiseCos(e) * aether::cwiseSqrt(e * e + 1.0) / (e + 2.0)),
               ids, 8);
    const aether::random::Generator g(7);
    print_tape("random",
               ad::lower<double, 3>(buf[3], 32,
                                    g.uniform<double, 3>(-2.0, 5.0) + g.substream(3).normal<double, 3>(1.0, 0.5) * a),
               ids, 8);
    print_tape("norm", ad::lower<double, 1>(buf[4], 32, a.norm() - b.squaredNorm()), ids, 8);
    print_tape("inplace", ad::lower<double, 3>(buf[0], 32, a * 2.0 + 1.0), ids, 8);
    using namespace aether;
    print_tape("cwise",
               ad::lower<double, 3>(buf[3], 32,
                                    cwiseTan(a) + cwiseAtan(b) - cwiseAsin(c) * cwiseAcos(a) + cwiseTanh(b) +
                                        cwiseSign(c) + cwiseRsqrt(a)),
               ids, 8);
    print_tape("minmax",
               ad::lower<double, 3>(buf[3], 32,
                                    clamp(a, -0.5, 0.5) + cwiseMin(a, b) * cwiseMax(b, 0.25) + cwiseDiv(a, c) +
                                        cwisePow(cwiseAbs(a), 1.5) + cwisePow(cwiseAbs(b), c)),
               ids, 8);
    print_tape("norms1", ad::lower<double, 1>(buf[4], 32, a.rNorm() + a.rSquaredNorm()), ids, 8);
    print_tape("norms2", ad::lower<double, 1>(buf[4], 32, a.cubedNorm() - a.rCubedNorm()), ids, 8);
    print_tape("norms3", ad::lower<double, 1>(buf[4], 32, q.maxNorm() + a.sum()), ids, 8);
    print_tape("unit", ad::lower<double, 3>(buf[3], 32, a.unitVector() + a.cwiseMul(b)), ids, 8);
    print_tape("quat", ad::lower<double, 4>(buf[6], 32, q.quatMul(r) + q.quatConj()), ids, 8);
    print_tape("qrecip", ad::lower<double, 4>(buf[7], 32, q.quatReciprocal()), ids, 8);
    print_tape("qrot", ad::lower<double, 3>(buf[3], 32, q.quatRotate(a)), ids, 8);
    print_tape("slices",
               ad::lower<double, 3>(buf[3], 32, q.asBack3DVector() + q.head<3>() - a.asPureQuaternion().tail<3>()),
               ids, 8);
    return 0;
}

template<class T>
void fill_generator(Array<T, 3>& a, Array<T, 3>& b, Array<T, 3>& c, std::uint64_t seed)
{
    const aether::random::Generator g(seed);
    a.deviceView() = g.uniform<T, 3>(T(-1), T(1));
    b.deviceView() = g.substream(1).uniform<T, 3>(T(-1), T(1));
    c.deviceView() = g.substream(2).uniform<T, 3>(T(-1), T(1));
}

int cfg0g(std::size_t n, std::uint64_t seed)
{
    Array<double, 3> a(n), b(n), c(n), d(n);
    fill_generator(a, b, c, seed);
    const auto sums = aether::evaluate(d.deviceView(), cfg0(a.deviceView(), b.deviceView(), c.deviceView()),
                                       aether::SumMode::Neumaier);
    a.download();
    b.download();
    c.download();
    d.download();
    write_planes(a);
    write_planes(b);
    write_planes(c);
    write_planes(d);
    std::fwrite(sums.data(), sizeof(double), 3, stdout);
    return 0;
}

// aether::tune with and without sums, then the same assignments: values and sums bit-identical to the
// untuned evaluation (text: "tune ok <tile> <bytes> <waves>" per form, or the first mismatch)
int tune(std::size_t n, std::uint64_t seed)
{
    Array<double, 3> a(n), b(n), c(n), d1(n), d2(n), d3(n);
    fill_generator(a, b, c, seed);
    const auto e = cfg0(a.deviceView(), b.deviceView(), c.deviceView());
    aether::prepare(d1.deviceView(), e, aether::SumMode::Neumaier); // compiles and loads, runs nothing
    const auto s1 = aether::evaluate(d1.deviceView(), e, aether::SumMode::Neumaier);
    const aeth_tape_shape ts = aether::tune(d2.deviceView(), e, nullptr, aether::SumMode::Neumaier);
    const auto s2 = aether::evaluate(d2.deviceView(), e, aether::SumMode::Neumaier);
    const aeth_tape_shape tn = aether::tune(d3.deviceView(), e);
    d3.deviceView() = e;
    d1.download();
    d2.download();
    d3.download();
    auto v1 = d1.hostView(), v2 = d2.hostView(), v3 = d3.hostView();
    for (std::size_t k = 0; k < 3; ++k)
        if (s1[k] != s2[k]) {
            std::printf("sum %zu differs\n", k);
            return 1;
        }
    for (std::size_t comp = 0; comp < 3; ++comp)
        for (std::size_t i = 0; i < n; ++i)
            if (v1(comp, i) != v2(comp, i) || v1(comp, i) != v3(comp, i)) {
                std::printf("value (%zu, %zu) differs\n", comp, i);
                return 1;
            }
    if (ts.tile != 2048 || ts.vector_bytes != 16) {
        std::printf("tuned with sums at %u x %u\n", ts.tile, ts.vector_bytes);
        return 1;
    }
    std::printf("tune ok %u %u %u\ntune ok %u %u %u\n", ts.tile, ts.vector_bytes, ts.waves_per_simd, tn.tile,
                tn.vector_bytes, tn.waves_per_simd);
    return 0;
}

int host(std::size_t n, std::uint64_t seed)
{
    Array<double, 3> a(n), b(n), c(n), d(n);
    const aether::random::Generator g(seed);
    const auto la = g.uniform<double, 3>(-1.0, 1.0);
    const auto lb = g.substream(1).uniform<double, 3>(-1.0, 1.0);
    const auto lc = g.substream(2).uniform<double, 3>(-1.0, 1.0);
    auto av = a.hostView(), bv = b.hostView(), cv = c.hostView(), dv = d.hostView();
    for (std::size_t k = 0; k < n; ++k) {
        const auto i = SampleIndex::make(k);
        av[i] = la;
        bv[i] = lb;
        cv[i] = lc;
    }
    for (std::size_t k = 0; k < n; ++k) dv[SampleIndex::make(k)] = cfg0(av, bv, cv);
    write_planes(a);
    write_planes(b);
    write_planes(c);
    write_planes(d);
    return 0;
}

int vec3(std::size_t n, std::uint64_t sa, std::uint64_t sb, std::uint64_t sc)
{
    Array<double, 3> a(n), b(n), c(n);
    Array<double> e(n);
    const std::uint64_t seeds[3] = {sa, sb, sc};
    Array<double, 3>* arrs[3] = {&a, &b, &c};
    for (int k = 0; k < 3; ++k)
        aether::detail::check(aeth_array_fill(AETH_DIST_UNIFORM, AETH_F64, seeds[k], -1.0, 1.0, 3, n,
                                              arrs[k]->capacity(), arrs[k]->deviceView().data(), nullptr),
                              "aeth_array_fill");
    auto av = a.deviceView(), bv = b.deviceView(), cv = c.deviceView();
    e.deviceView() = av.cross(bv).dot(cv) + aether::sqrt(av.dot(av) + 1.0);
    e.download();
    write_planes(e);
    return 0;
}

int fft(std::size_t n, bool forward)
{
    std::vector<double> in(2 * n), out(2 * n);
    if (std::fread(in.data(), sizeof(double), 2 * n, stdin) != 2 * n) return 3;
    const std::size_t shape[1] = {n}, axes[1] = {0};
    aether::fft::c2c<double>(in.data(), out.data(), shape, axes, forward);
    std::fwrite(out.data(), sizeof(double), 2 * n, stdout);
    return 0;
}

int rfft(std::size_t rows, std::size_t n)
{
    std::vector<double> in(rows * n), spec(rows * (n / 2 + 1) * 2), back(rows * n);
    if (std::fread(in.data(), sizeof(double), in.size(), stdin) != in.size()) return 3;
    const std::size_t shape[2] = {rows, n}, axes[1] = {1};
    aether::fft::r2c<double>(in.data(), spec.data(), shape, axes);
    aether::fft::c2r<double>(spec.data(), back.data(), shape, axes);
    std::fwrite(spec.data(), sizeof(double), spec.size(), stdout);
    std::fwrite(back.data(), sizeof(double), back.size(), stdout);
    return 0;
}

int append(std::size_t n)
{
    Array<double, 3> x(0);
    bool ok = x.samples() == 0 && x.capacity() == 0;
    for (std::size_t i = 0; i < n; ++i) x.push_back({double(i), 2.0 * double(i), -double(i)});
    ok = ok && x.samples() == n && x.capacity() >= n && x.capacity() % 32 == 0;
    auto v = x.hostView();
    for (std::size_t i = 0; i < n; ++i) ok = ok && v(0, i) == double(i) && v(1, i) == 2.0 * double(i) && v(2, i) == -double(i);
    // the host spelling on a SampleRef: += and -=
    auto r = v[SampleIndex::make(n / 2)];
    r += v;         // x -> 2x
    r -= 0.25 * v;  // reads the updated sample: 2x - x/2
    const auto g = r.get();
    ok = ok && g(0) == 1.5 * double(n / 2) && g(1) == 3.0 * double(n / 2);
    const std::size_t spare = x.capacity() - x.samples();
    auto sp = x.hostSpare();
    if (spare > 0) sp(0, 0) = 42.0;
    x.commit(spare > 0 ? 1 : 0);
    if (spare > 0) ok = ok && x.hostView()(0, n) == 42.0;
    bool threw = false;
    try {
        x.commit(x.capacity() - x.samples() + 1);
    } catch (const aether::Error&) {
        threw = true;
    }
    x.clear();
    ok = ok && threw && x.samples() == 0 && x.capacity() >= n;
    std::printf("append %s samples %zu capacity %zu\n", ok ? "ok" : "BAD", x.samples(), x.capacity());
    return ok ? 0 : 1;
}

int plane(std::size_t n)
{
    aether::AccumPlane<3> p(n);
    std::vector<Array<double, 3>> terms;
    const aether::random::Generator g(11);
    for (int k = 0; k < 3; ++k) {
        terms.emplace_back(n);
        terms.back().deviceView() = g.substream(k).normal<double, 3>(0.0, std::pow(10.0, k));
        p.add(terms.back());
    }
    Array<double, 3> out(n);
    p.delivered(out);
    out.download();
    for (auto& t : terms) t.download();
    double worst = 0.0;
    for (std::size_t c = 0; c < 3; ++c)
        for (std::size_t i = 0; i < n; ++i) {
            long double e = 0.0L;
            for (auto& t : terms) e += (long double)t.hostView()(c, i);
            const double ex = (double)e;
            const double ulp = std::nextafter(std::fabs(ex), INFINITY) - std::fabs(ex);
            worst = std::max(worst, std::fabs(out.hostView()(c, i) - ex) / ulp);
        }
    std::printf("plane %s worst_ulp %.3f\n", worst <= 1.0 ? "ok" : "BAD", worst);
    return worst <= 1.0 ? 0 : 1;
}

int where(std::size_t n, double threshold)
{
    Array<double> x(n);
    x.deviceView() = aether::random::Generator(5).uniform<double, 1>(0.0, 1.0);
    Array<std::uint32_t> out(0);
    out.reserve(n);
    aether::accum::AppendCounter ctr(out.capacity());
    aether::accum::appendWhere(x.deviceView(), threshold, ctr, out);
    const std::size_t raw = ctr.rawCount();
    const std::size_t k = ctr.commit(out);
    out.download();
    x.download();
    std::vector<char> seen(n, 0);
    bool ok = k == raw && out.samples() == k && !ctr.overflowed();
    for (std::size_t j = 0; j < k && ok; ++j) {
        const std::uint32_t i = out.hostView()(0, j);
        ok = i < n && !seen[i] && x.hostView()(0, i) > threshold;
        if (ok) seen[i] = 1;
    }
    std::size_t expect = 0;
    for (std::size_t i = 0; i < n; ++i) expect += x.hostView()(0, i) > threshold;
    ok = ok && expect == k;
    std::printf("where %s count %zu\n", ok ? "ok" : "BAD", k);
    return ok ? 0 : 1;
}

int keyed(std::size_t n, std::size_t run)
{
    Array<double> terms(n), target(n / run + 1), comp(n / run + 1);
    Array<std::int64_t> keys(n);
    terms.deviceView() = aether::random::Generator(9).uniform<double, 1>(-1.0, 1.0);
    for (std::size_t i = 0; i < n; ++i) keys.hostView()(0, i) = (std::int64_t)(i / run);
    keys.upload();
    const std::size_t slots = n / run + 1;
    aether::accum::detail::zero_device(target.deviceView().data(), target.capacity() * 8, nullptr);
    aether::accum::detail::zero_device(comp.deviceView().data(), comp.capacity() * 8, nullptr);
    aether::accum::keyedSum(aether::accum::Run::Sorted, keys.deviceView().data(), terms.deviceView().data(), n,
                            target.deviceView().data(), slots, comp.deviceView().data());
    target.download();
    comp.download();
    terms.download();
    double worst = 0.0;
    for (std::size_t k = 0; k < slots; ++k) {
        long double e = 0.0L, mass = 0.0L;
        for (std::size_t i = k * run; i < std::min(n, (k + 1) * run); ++i) {
            e += (long double)terms.hostView()(0, i);
            mass += std::fabs((long double)terms.hostView()(0, i));
        }
        if (mass == 0.0L) continue;
        const long double got = (long double)target.hostView()(0, k) + (long double)comp.hostView()(0, k);
        worst = std::max(worst, (double)(std::fabs(got - e) / (mass * 0x1p-52L)));
    }
    std::printf("keyed %s worst %.3f\n", worst <= 1.0 ? "ok" : "BAD", worst);
    return worst <= 1.0 ? 0 : 1;
}

int runtime(std::size_t rows, std::size_t cols)
{
    const std::size_t n = rows * cols;
    Array<double> a(n), b(n), out(n);
    Array<float> f(n);
    a.deviceView() = aether::random::Generator(3).uniform<double, 1>(-1.0, 1.0);
    b.deviceView() = aether::random::Generator(4).uniform<double, 1>(-1.0, 1.0);
    // a read as the transpose of a (cols x rows) matrix: shape (rows, cols), strides (1, rows)
    const std::size_t shape[2] = {rows, cols};
    const std::int64_t tstr[2] = {1, (std::int64_t)rows};
    const auto va = aether::makeRuntimeView<double>(a.deviceView().data(), shape, tstr);
    const auto vb = aether::makeRuntimeView<double>(b.deviceView().data(), shape);
    const auto vo = aether::makeRuntimeView<double>(out.deviceView().data(), shape);
    aether::eval::runtimeEval(aether::eval::RuntimeOp::AddScaled, va, vb, vo, 0.25);
    aether::eval::runtimeEval(aether::eval::RuntimeOp::Assign, vo, aether::RuntimeView{},
                              aether::makeRuntimeView<float>(f.deviceView().data(), shape));
    a.download();
    b.download();
    out.download();
    f.download();
    bool ok = true;
    for (std::size_t i = 0; i < rows && ok; ++i)
        for (std::size_t j = 0; j < cols && ok; ++j) {
            const double want = a.hostView()(0, j * rows + i) + 0.25 * b.hostView()(0, i * cols + j);
            ok = out.hostView()(0, i * cols + j) == want && f.hostView()(0, i * cols + j) == (float)want;
        }
    bool threw = false;
    try {
        aether::eval::runtimeEval(aether::eval::RuntimeOp::Add, va, vb,
                                  aether::makeRuntimeView<float>(f.deviceView().data(), shape));
    } catch (const aether::Error&) {
        threw = true;
    }
    std::printf("runtime %s\n", ok && threw ? "ok" : "BAD");
    return ok && threw ? 0 : 1;
}

int flt(std::size_t n, std::uint64_t seed)
{
    Array<float, 3> a(n), b(n), c(n), d(n);
    fill_generator(a, b, c, seed);
    const auto sums = aether::evaluate(d.deviceView(), cfg0(a.deviceView(), b.deviceView(), c.deviceView()),
                                       aether::SumMode::FloatFloat);
    d.download();
    write_planes(d);
    std::fwrite(sums.data(), sizeof(double), 3, stdout);
    return 0;
}

int resize(std::size_t n)
{
    Array<double, 2> x(n);
    auto hv = x.hostView();
    for (std::size_t i = 0; i < n; ++i) {
        hv(0, i) = double(i);
        hv(1, i) = -double(i);
    }
    x.upload();
    x.deviceView() = x.deviceView() * 3.0;
    const std::size_t cap0 = x.capacity();
    x.resize(n + 100);
    x.download();
    auto v = x.hostView();
    bool ok = x.samples() == n + 100 && x.capacity() >= n + 100 && x.capacity() % 32 == 0 && cap0 % 32 == 0;
    for (std::size_t i = 0; i < n; ++i) ok = ok && v(0, i) == 3.0 * double(i) && v(1, i) == -3.0 * double(i);
    // the packed views: pitch samples() when the capacity is the sample count, refused otherwise
    Array<double, 2> full(64);
    const auto pk = full.devicePacked();
    ok = ok && pk.pitch() == 64 && pk.samples() == 64 && full.hostPacked().data() == full.hostView().data();
    bool refused = false;
    try {
        (void)x.hostPacked(); // capacity() > samples() after the resize
    } catch (const aether::Error&) {
        refused = x.capacity() != x.samples();
    }
    ok = ok && refused;
    std::printf("resize %s samples %zu capacity %zu\n", ok ? "ok" : "BAD", x.samples(), x.capacity());
    return ok ? 0 : 1;
}

} // namespace

template<class Q, class V>
auto ext1(const Q& q, const Q& r)
{
    return q.quatMul(r) + q.quatConj() - aether::cwiseMax(q.quatReciprocal(), -2.0);
}

template<class Q, class V>
auto ext2(const Q& q, const V& a)
{
    return q.quatRotate(a) + q.template head<3>() * a.sum() - a.asPureQuaternion().template tail<3>();
}

template<class Q, class V>
auto ext3(const Q& q, const V& a)
{
    return q.maxNorm() + a.sum() - a.rCubedNorm();
}

template<class Q, class V>
auto ext4(const Q& q, const V& a)
{
    return a.unitVector() + aether::cwiseMin(a, 0.25) - aether::clamp(a, -0.5, 0.5) +
           aether::cwiseDiv(a, q.asBack3DVector()) + aether::cwiseSign(a);
}

int ext(std::size_t n, std::uint64_t seed)
{
    Array<double, 4> q(n), r(n), o1(n), h1(n);
    Array<double, 3> a(n), o2(n), o4(n), o5(n), h2(n), h4(n);
    Array<double> o3(n), h3(n);
    const aether::random::Generator g(seed);
    q.deviceView() = g.uniform<double, 4>(-1.0, 1.0);
    r.deviceView() = g.substream(1).uniform<double, 4>(-1.0, 1.0);
    a.deviceView() = g.substream(2).uniform<double, 3>(-1.0, 1.0);
    const auto qd = q.deviceView(), rd = r.deviceView();
    const auto ad = a.deviceView();
    o1.deviceView() = ext1<decltype(qd), decltype(ad)>(qd, rd);
    o2.deviceView() = ext2(qd, ad);
    o3.deviceView() = ext3(qd, ad);
    o4.deviceView() = ext4(qd, ad);
    o5.deviceView() = aether::cwiseTan(ad) + aether::cwisePow(aether::cwiseAbs(ad), 1.5);
    for (auto* x : {&q, &r, &o1}) x->download();
    for (auto* x : {&a, &o2, &o4, &o5}) x->download();
    o3.download();
    const auto qh = q.hostView(), rh = r.hostView();
    const auto ah = a.hostView();
    auto v1 = h1.hostView();
    auto v2 = h2.hostView();
    auto v3 = h3.hostView();
    auto v4 = h4.hostView();
    for (std::size_t k = 0; k < n; ++k) {
        const auto i = SampleIndex::make(k);
        v1[i] = ext1<decltype(qh), decltype(ah)>(qh, rh);
        v2[i] = ext2(qh, ah);
        v3[i] = ext3(qh, ah);
        v4[i] = ext4(qh, ah);
    }
    write_planes(q);
    write_planes(r);
    write_planes(a);
    write_planes(o1);
    write_planes(o2);
    write_planes(o3);
    write_planes(o4);
    write_planes(h1);
    write_planes(h2);
    write_planes(h3);
    write_planes(h4);
    write_planes(o5);
    return 0;
}

// Bundles against per-sample assignment on host views over plain buffers (no device involved): n = 37 is odd,
// so every width has a masked tail bundle; W = 2 and 4 for doubles, 4 for floats; an axpy, a transcendental
// tree and a quaternion chain; and views whose planes start off a 16-byte boundary (the lane-by-lane store).
template<class T, std::size_t N, std::size_t W, class F>
int bundle_case(const char* what, std::size_t offset, F make_expr)
{
    constexpr std::size_t n = 37, pitch = 40;
    std::vector<T> a(N * pitch + offset), b(N * pitch + offset), x(N * pitch + offset, T(0)), y(N * pitch + offset, T(0));
    for (std::size_t c = 0; c < N; ++c)
        for (std::size_t i = 0; i < n; ++i) {
            a[offset + c * pitch + i] = T(0.75) * T(i) + T(0.1) * T(c) - T(3);
            b[offset + c * pitch + i] = T(-0.5) * T(i) + T(c) + T(1.25);
        }
    const aether::View<T, N> av(a.data() + offset, n, pitch, false), bv(b.data() + offset, n, pitch, false);
    const aether::View<T, N> xv(x.data() + offset, n, pitch, false), yv(y.data() + offset, n, pitch, false);
    const auto e = make_expr(av, bv);
    for (std::size_t k = 0; k * W < n; ++k) {
        const auto bi = aether::BundleIndex<W>::make(k);
        if (bi.base() + W <= n) xv[bi] = e;
        else aether::bundleAssign(xv, e, bi, bi.mask((aether::offset_t)n));
    }
    for (std::size_t i = 0; i < n; ++i) yv[SampleIndex::make(i)] = e;
    if (std::memcmp(x.data(), y.data(), x.size() * sizeof(T)) != 0) {
        std::printf("mismatch: bundle %s W=%zu offset %zu\n", what, W, offset);
        return 1;
    }
    return 0;
}

int bundle()
{
    int bad = 0;
    auto axpy = [](auto a, auto b) { return a + 2.5 * b; };
    auto trig = [](auto a, auto b) { return a * aether::cwiseSin(b) + aether::cwiseExp(-1.0 * (a * a)); };
    auto quat = [](auto a, auto b) { return a.quatMul(b).quatMul(a.quatConj()) + b.quatConj(); };
    for (std::size_t off : {std::size_t(0), std::size_t(1)}) {
        bad += bundle_case<double, 3, 2>("axpy", off, axpy);
        bad += bundle_case<double, 3, 4>("axpy", off, axpy);
        bad += bundle_case<double, 3, 2>("trig", off, trig);
        bad += bundle_case<double, 4, 2>("quat", off, quat);
        bad += bundle_case<float, 3, 4>("axpy f32", off, [](auto a, auto b) { return a + 2.5f * b; });
    }
    // the masks and the packet index
    const auto tail = aether::BundleIndex<4>::make(9); // samples 36..39 of a 37-sample array: one lane
    const auto m = tail.mask(37);
    if (!(m.lane(0) && !m.lane(1) && !m.full() && aether::BundleIndex<4>::make(8).mask(37).full() &&
          tail.lane(0).global() == 36 && tail.lane(0).work() == 0 && aether::BundleIndex<4>::make(10).mask(37).active_ == 0 &&
          aether::BundleIndex<2>::make(3, 2, 64).base() == 2 * (2 * 64 + 3))) {
        std::printf("mismatch: bundle masks\n");
        ++bad;
    }
    const auto pk = aether::PacketIndex<8>::makeTail(32, 5);
    if (!(aether::PacketIndex<8>::make(8).full() && !pk.full() && pk.scalar(4).global() == 36)) {
        std::printf("mismatch: packet index\n");
        ++bad;
    }
    if (!bad) std::printf("bundles ok\n");
    return bad ? 1 : 0;
}

// Item (view/Item.h) and scalar broadcast through SampleRef (tests/test_ScalarBroadcast.cpp's rows) on host
// views over plain buffers
int item()
{
    using aether::Vec3d;
    int bad = 0;
    auto fail = [&](const char* what) {
        std::printf("mismatch: item %s\n", what);
        ++bad;
    };
    constexpr std::size_t n = 4;
    std::vector<double> buf(3 * n, -1.0), wbuf(3 * n, -9.0);
    const aether::View<double, 3> v(buf.data(), n, n, false), w(wbuf.data(), n, n, false);
    for (std::size_t i = 0; i < n; ++i) v[SampleIndex::make(i)] = 2.0;
    for (std::size_t i = 0; i < n; ++i)
        for (std::size_t c = 0; c < 3; ++c)
            if (v(c, i) != 2.0) fail("scalar assignment");
    const Vec3d a{1.0, 2.0, 3.0};
    v[SampleIndex::make(1)] = a;
    v[SampleIndex::make(1)] += 0.5;
    if (!(v(0, 1) == 1.5 && v(1, 1) == 2.5 && v(2, 1) == 3.5)) fail("+= scalar");
    v[SampleIndex::make(1)] -= 0.25;
    if (!(v(0, 1) == 1.25 && v(1, 1) == 2.25 && v(2, 1) == 3.25)) fail("-= scalar");
    // a scalar store and the same store of a filled Item give the same bits
    w[SampleIndex::make(0)] = Vec3d::filled(2.0);
    if (std::memcmp(&v(0, 0), &w(0, 0), sizeof(double)) || std::memcmp(&v(2, 0), &w(2, 0), sizeof(double)))
        fail("scalar store vs filled item");
    // an Item from a sample-free tree, accessors, get()
    const Vec3d out = a + Vec3d::filled(10.0);
    if (!(out(0) == 11.0 && out.get<1>() == 12.0 && out.data()[2] == 13.0 && Vec3d::size() == 3)) fail("tree value");
    const Vec3d b{0.5, -1.0, 4.0};
    const Vec3d cr = a.cross(b);
    if (!(cr(0) == 2.0 * 4.0 - 3.0 * -1.0 && cr(1) == 3.0 * 0.5 - 1.0 * 4.0 && cr(2) == 1.0 * -1.0 - 2.0 * 0.5))
        fail("cross");
    const aether::Vec<double, 1> d1 = a.dot(b);
    if (d1(0) != 3.0 * 4.0 + (2.0 * -1.0 + 1.0 * 0.5)) fail("dot fold order");
    if (!(Vec3d::Zeros()(1) == 0.0 && Vec3d::Ones()(2) == 1.0 && Vec3d{}(0) == 0.0)) fail("factories");
    const Vec3d g = v[SampleIndex::make(1)].get();
    if (!(g(0) == 1.25 && g(2) == 3.25)) fail("get");
    // an Item as a leaf of a per-sample tree: the same value at every sample
    for (std::size_t i = 0; i < n; ++i) w[SampleIndex::make(i)] = v * 2.0 + a;
    for (std::size_t i = 0; i < n; ++i)
        for (std::size_t c = 0; c < 3; ++c)
            if (w(c, i) != v(c, i) * 2.0 + a(c)) fail("leaf");
    // the free distributions are Generator(DEFAULT_SEED)'s (random/free.h:46-79)
    const aether::random::Generator dg(aether::random::DEFAULT_SEED);
    for (std::size_t i = 0; i < 3; ++i)
        if (aether::random::uniform<double, 3>(-1.0, 1.0).eval(1, i) != dg.uniform<double, 3>(-1.0, 1.0).eval(1, i) ||
            aether::random::bernoulli<double, 3>(0.5).eval(2, i) != dg.bernoulli<double, 3>(0.5).eval(2, i) ||
            aether::random::exponential<double, 3>(2.0).eval(0, i) != dg.exponential<double, 3>(2.0).eval(0, i))
            fail("free distributions");
    if (!bad) std::printf("items ok\n");
    return bad ? 1 : 0;
}

int main(int argc, char** argv)
{
    if (argc < 2) return 2;
    const std::string mode = argv[1];
    auto u = [&](int k) { return (std::uint64_t)std::strtoull(argv[k], nullptr, 0); };
    try {
        if (mode == "tape") return tape();
        if (mode == "bundle") return bundle();
        if (mode == "item") return item();
        if (mode == "cfg0g" && argc == 4) return cfg0g(u(2), u(3));
        if (mode == "host" && argc == 4) return host(u(2), u(3));
        if (mode == "vec3" && argc == 6) return vec3(u(2), u(3), u(4), u(5));
        if (mode == "fft" && argc == 4) return fft(u(2), u(3) != 0);
        if (mode == "float" && argc == 4) return flt(u(2), u(3));
        if (mode == "rfft" && argc == 4) return rfft(u(2), u(3));
        if (mode == "append" && argc == 3) return append(u(2));
        if (mode == "plane" && argc == 3) return plane(u(2));
        if (mode == "where" && argc == 4) return where(u(2), std::strtod(argv[3], nullptr));
        if (mode == "keyed" && argc == 4) return keyed(u(2), u(3));
        if (mode == "runtime" && argc == 4) return runtime(u(2), u(3));
        if (mode == "resize" && argc == 3) return resize(u(2));
        if (mode == "ext" && argc == 4) return ext(u(2), u(3));
        if (mode == "tune" && argc == 4) return tune(u(2), u(3));
    } catch (const aether::Error& e) {
        std::fprintf(stderr, "aether error: %s\n", e.what());
        return 1;
    }
    return 2;
}
