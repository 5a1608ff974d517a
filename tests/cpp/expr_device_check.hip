// expr_device_check.hip -- the C++ front end's expressions evaluated per sample inside the caller's own HIP
// kernels (`dv[SampleIndex::make(threadIdx.x, blockIdx.x, blockDim.x)] = expr`, expr/Assign.h:251-346 on the
// device), compared bit for bit with the same expression assigned to the whole view (`dv = expr`: one
// runtime-compiled kernel). Covers the fused benchmark tree and the Vec3 tree in double and float, the
// elementwise functions, folds and norms, quaternions, Philox distribution leaves, `+=` / `-=`, `get()`, and a
// node type of the caller's own (per sample, as a whole-array assignment and with fused sums).
// Built by the package build with hipcc against libaether_hip.so; prints "ok" or the first mismatch, exit 1.
#include <aether/aether.hpp>

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

using aether::Array;
using aether::SampleIndex;
using aether::View;

int failures = 0;

void launch_check(const char* what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw aether::Error(std::string(what) + ": " + hipGetErrorString(e));
    if (hipDeviceSynchronize() != hipSuccess) throw aether::Error(std::string(what) + ": synchronize");
}

dim3 grid_of(std::size_t n) { return dim3((unsigned)((n + 255) / 256)); }

// the n samples of every component, as bytes
template<class T, std::size_t N>
std::vector<unsigned char> bytes_of(Array<T, N>& a)
{
    a.download();
    std::vector<unsigned char> out;
    const auto hv = a.hostView();
    for (std::size_t c = 0; c < N; ++c) {
        const unsigned char* p = reinterpret_cast<const unsigned char*>(&hv(c, 0));
        out.insert(out.end(), p, p + a.samples() * sizeof(T));
    }
    return out;
}

template<class T, std::size_t N>
void expect_same(const char* what, Array<T, N>& got, Array<T, N>& want)
{
    if (bytes_of(got) != bytes_of(want)) {
        std::printf("mismatch: %s\n", what);
        ++failures;
    }
}

// ---- a node type of the caller's own: smoothstep x * x * (3 - 2 x), per component
template<class E>
struct Smoothstep : aether::Expr<Smoothstep<E>> {
    using element_type = aether::elem_t<E>;
    static constexpr std::size_t ncomp = E::ncomp;
    E e;
    AETHER_DEVICEHOST() explicit Smoothstep(const E& x) : e(x) {}
    AETHER_DEVICEHOST() element_type eval(std::size_t c, std::size_t i) const
    {
        const element_type x = e.eval(c, i);
        return x * x * (element_type(3) - element_type(2) * x);
    }
};

template<class E>
AETHER_DEVICEHOST() Smoothstep<E> smoothstep(const E& e)
{
    return Smoothstep<E>(e);
}

// ---- the caller's own kernels
template<class T>
__global__ void cfg0_kernel(View<T, 3> d, View<T, 3> a, View<T, 3> b, View<T, 3> c, std::uint32_t n)
{
    const SampleIndex i = SampleIndex::make(threadIdx.x, blockIdx.x, blockDim.x);
    if (i.global() < n) d[i] = a + b * c + aether::cwiseSin(a) * aether::cwiseExp(T(-1) * (b * b));
}

// an Item leaf: the same three values at every sample
template<class T>
__global__ void item_kernel(View<T, 3> d, View<T, 3> a, View<T, 3> b, aether::Item<T, 3> k, std::uint32_t n)
{
    const SampleIndex i = SampleIndex::make(threadIdx.x, blockIdx.x, blockDim.x);
    if (i.global() < n) d[i] = a + k * b - aether::cwiseSin(k);
}

// the same tree a bundle of W samples a thread (16 bytes of each plane: W = 2 doubles, 4 floats); the tail
// bundle takes the masked form
template<class T>
__global__ void bundle_kernel(View<T, 3> d, View<T, 3> a, View<T, 3> b, View<T, 3> c, std::uint32_t n)
{
    constexpr std::size_t W = 16 / sizeof(T);
    const auto bi = aether::BundleIndex<W>::make(threadIdx.x, blockIdx.x, blockDim.x);
    const auto m = bi.mask(n);
    const auto e = a + b * c + aether::cwiseSin(a) * aether::cwiseExp(T(-1) * (b * b));
    if (m.full()) d[bi] = e;
    else if (m.active_ > 0) aether::bundleAssign(d, e, bi, m);
}

template<class T>
__global__ void vec3_kernel(View<T, 1> e, View<T, 3> a, View<T, 3> b, View<T, 3> c, std::uint32_t n)
{
    const SampleIndex i = SampleIndex::make(threadIdx.x, blockIdx.x, blockDim.x);
    if (i.global() < n) e[i] = a.cross(b).dot(c) + aether::sqrt(a.dot(a) + T(1));
}

template<class T>
__global__ void elementwise_kernel(View<T, 3> d, View<T, 3> a, View<T, 3> b, std::uint32_t n)
{
    const SampleIndex i = SampleIndex::make(threadIdx.x, blockIdx.x, blockDim.x);
    if (i.global() < n)
        d[i] = aether::cwiseTan(a) + aether::cwiseAtan(b) * aether::cwiseAsin(T(0.5) * a) -
               aether::cwiseAcos(b) / T(3) + aether::cwiseTanh(T(4) * b) +
               aether::cwisePow(aether::cwiseAbs(a) + T(0.25), b) + aether::clamp(a - b, T(-0.5), T(0.75)) +
               aether::cwiseSign(b) * aether::cwiseLog(aether::cwiseAbs(a) + T(1e-3)) +
               aether::cwiseRsqrt(aether::cwiseAbs(b) + T(0.1)) + aether::cwiseMax(a, b) -
               aether::cwiseCos(aether::cwiseMin(a, b)) + aether::cwiseDiv(a, b + T(3));
}

template<class T>
__global__ void folds_kernel(View<T, 1> s, View<T, 3> u, View<T, 3> a, View<T, 3> b, std::uint32_t n)
{
    const SampleIndex i = SampleIndex::make(threadIdx.x, blockIdx.x, blockDim.x);
    if (i.global() < n) {
        s[i] = a.maxNorm() + b.sum() * a.rNorm() - a.cubedNorm() + b.rSquaredNorm() * T(0.125) + a.norm();
        u[i] = a.unitVector() + b.cwiseMul(a) * b.rCubedNorm();
    }
}

__global__ void quat_kernel(View<double, 3> r, View<double, 4> m, View<double, 4> q, View<double, 4> p,
                            View<double, 3> v, std::uint32_t n)
{
    const SampleIndex i = SampleIndex::make(threadIdx.x, blockIdx.x, blockDim.x);
    if (i.global() < n) {
        r[i] = q.quatRotate(v) + p.quatMul(q).tail<3>();
        m[i] = q.quatMul(p) + p.quatReciprocal() - v.asPureQuaternion() + q.quatConj();
    }
}

template<class T>
__global__ void random_kernel(View<T, 3> d, View<T, 3> a, aether::random::Generator g, std::uint32_t n)
{
    const SampleIndex i = SampleIndex::make(threadIdx.x, blockIdx.x, blockDim.x);
    if (i.global() < n)
        d[i] = g.uniform<T, 3>(T(-1), T(1)) * a + g.substream(3).normal<T, 3>(T(0.5), T(2));
}

// the other distributions (lognormal, exponential, bernoulli), scalar and per-component parameters
template<class T>
__global__ void dist_kernel(View<T, 3> d, aether::random::Generator g, std::uint32_t n)
{
    const SampleIndex i = SampleIndex::make(threadIdx.x, blockIdx.x, blockDim.x);
    const aether::Item<T, 3> mu{T(0.25), T(-0.5), T(1)}, lam{T(1.5), T(0.5), T(3)};
    if (i.global() < n)
        d[i] = g.lognormal<T, 3>(T(0.25), T(0.5)) + g.exponential<T, 3>(T(1.5)) * g.bernoulli<T, 3>(T(0.3)) +
               g.lognormal(mu, lam) - g.exponential(lam) + g.uniform(mu, lam) * g.normal(mu, lam);
}

template<class T>
__global__ void update_kernel(View<T, 3> d, View<T, 3> a, View<T, 3> b, std::uint32_t n)
{
    const SampleIndex i = SampleIndex::make(threadIdx.x, blockIdx.x, blockDim.x);
    if (i.global() < n) {
        d[i] += a * b;
        d[i] -= aether::cwiseExp(T(0.5) * b);
    }
}

template<class T>
__global__ void user_node_kernel(View<T, 3> d, View<T, 3> a, View<T, 3> b, std::uint32_t n)
{
    const SampleIndex i = SampleIndex::make(threadIdx.x, blockIdx.x, blockDim.x);
    if (i.global() < n) d[i] = smoothstep(a) + b * aether::cwiseSin(smoothstep(b));
}

template<class T>
__global__ void get_kernel(View<T, 1> s, View<T, 3> a, std::uint32_t n)
{
    const SampleIndex i = SampleIndex::make(threadIdx.x, blockIdx.x, blockDim.x);
    if (i.global() < n) {
        const auto item = a[i].get();
        s(i.global()) = item.template get<2>() + (item(1) + item(0));
    }
}

// ---- cases
template<class T>
void run_cases(std::size_t n)
{
    const aether::random::Generator g(42);
    Array<T, 3> a(n), b(n), c(n);
    a.deviceView() = g.uniform<T, 3>(T(-1), T(1));
    b.deviceView() = g.substream(1).uniform<T, 3>(T(-1), T(1));
    c.deviceView() = g.substream(2).uniform<T, 3>(T(-1), T(1));
    const auto av = a.deviceView(), bv = b.deviceView(), cv = c.deviceView();
    const std::uint32_t n32 = (std::uint32_t)n;
    const char* tn = sizeof(T) == 8 ? "double" : "float";
    std::string w;

    {   // the fused benchmark tree
        Array<T, 3> got(n), want(n);
        hipLaunchKernelGGL(cfg0_kernel<T>, grid_of(n), dim3(256), 0, 0, got.deviceView(), av, bv, cv, n32);
        launch_check("cfg0_kernel");
        want.deviceView() = av + bv * cv + aether::cwiseSin(av) * aether::cwiseExp(T(-1) * (bv * bv));
        w = std::string("cfg0 ") + tn;
        expect_same(w.c_str(), got, want);
    }
    {   // an Item leaf: per sample in a kernel against the whole-array assignment, where it lowers to constants
        Array<T, 3> got(n), want(n);
        const aether::Item<T, 3> k{T(0.75), T(-1.5), T(2.25)};
        hipLaunchKernelGGL(item_kernel<T>, grid_of(n), dim3(256), 0, 0, got.deviceView(), av, bv, k, n32);
        launch_check("item_kernel");
        want.deviceView() = av + k * bv - aether::cwiseSin(k);
        w = std::string("item leaf ") + tn;
        expect_same(w.c_str(), got, want);
    }
    {   // the same tree in bundles of 16 bytes a plane, with a masked tail bundle
        Array<T, 3> got(n), want(n);
        const std::size_t bundles = (n + 16 / sizeof(T) - 1) / (16 / sizeof(T));
        hipLaunchKernelGGL(bundle_kernel<T>, grid_of(bundles), dim3(256), 0, 0, got.deviceView(), av, bv, cv, n32);
        launch_check("bundle_kernel");
        want.deviceView() = av + bv * cv + aether::cwiseSin(av) * aether::cwiseExp(T(-1) * (bv * bv));
        w = std::string("bundles ") + tn;
        expect_same(w.c_str(), got, want);
    }
    {   // Vec3 geometry
        Array<T, 1> got(n), want(n);
        hipLaunchKernelGGL(vec3_kernel<T>, grid_of(n), dim3(256), 0, 0, got.deviceView(), av, bv, cv, n32);
        launch_check("vec3_kernel");
        want.deviceView() = av.cross(bv).dot(cv) + aether::sqrt(av.dot(av) + T(1));
        w = std::string("vec3 ") + tn;
        expect_same(w.c_str(), got, want);
    }
    {   // elementwise functions
        Array<T, 3> got(n), want(n);
        hipLaunchKernelGGL(elementwise_kernel<T>, grid_of(n), dim3(256), 0, 0, got.deviceView(), av, bv, n32);
        launch_check("elementwise_kernel");
        want.deviceView() = aether::cwiseTan(av) + aether::cwiseAtan(bv) * aether::cwiseAsin(T(0.5) * av) -
                            aether::cwiseAcos(bv) / T(3) + aether::cwiseTanh(T(4) * bv) +
                            aether::cwisePow(aether::cwiseAbs(av) + T(0.25), bv) +
                            aether::clamp(av - bv, T(-0.5), T(0.75)) +
                            aether::cwiseSign(bv) * aether::cwiseLog(aether::cwiseAbs(av) + T(1e-3)) +
                            aether::cwiseRsqrt(aether::cwiseAbs(bv) + T(0.1)) + aether::cwiseMax(av, bv) -
                            aether::cwiseCos(aether::cwiseMin(av, bv)) + aether::cwiseDiv(av, bv + T(3));
        w = std::string("elementwise ") + tn;
        expect_same(w.c_str(), got, want);
    }
    {   // folds and norms
        Array<T, 1> s(n), s_want(n);
        Array<T, 3> u(n), u_want(n);
        hipLaunchKernelGGL(folds_kernel<T>, grid_of(n), dim3(256), 0, 0, s.deviceView(), u.deviceView(), av, bv, n32);
        launch_check("folds_kernel");
        s_want.deviceView() = av.maxNorm() + bv.sum() * av.rNorm() - av.cubedNorm() + bv.rSquaredNorm() * T(0.125) +
                              av.norm();
        u_want.deviceView() = av.unitVector() + bv.cwiseMul(av) * bv.rCubedNorm();
        w = std::string("folds ") + tn;
        expect_same(w.c_str(), s, s_want);
        w = std::string("unit vector ") + tn;
        expect_same(w.c_str(), u, u_want);
    }
    {   // Philox leaves inside the kernel
        Array<T, 3> got(n), want(n);
        hipLaunchKernelGGL(random_kernel<T>, grid_of(n), dim3(256), 0, 0, got.deviceView(), av, g, n32);
        launch_check("random_kernel");
        want.deviceView() = g.uniform<T, 3>(T(-1), T(1)) * av + g.substream(3).normal<T, 3>(T(0.5), T(2));
        w = std::string("random leaves ") + tn;
        expect_same(w.c_str(), got, want);
    }
    {   // lognormal / exponential / bernoulli and the per-component forms inside the kernel
        Array<T, 3> got(n), want(n);
        hipLaunchKernelGGL(dist_kernel<T>, grid_of(n), dim3(256), 0, 0, got.deviceView(), g, n32);
        launch_check("dist_kernel");
        const aether::Item<T, 3> mu{T(0.25), T(-0.5), T(1)}, lam{T(1.5), T(0.5), T(3)};
        want.deviceView() = g.lognormal<T, 3>(T(0.25), T(0.5)) + g.exponential<T, 3>(T(1.5)) * g.bernoulli<T, 3>(T(0.3)) +
                            g.lognormal(mu, lam) - g.exponential(lam) + g.uniform(mu, lam) * g.normal(mu, lam);
        w = std::string("distributions ") + tn;
        expect_same(w.c_str(), got, want);
    }
    {   // += and -= on the device against the same updates as whole-array assignments
        Array<T, 3> got(n), want(n);
        got.deviceView() = cv * T(1);
        want.deviceView() = cv * T(1);
        hipLaunchKernelGGL(update_kernel<T>, grid_of(n), dim3(256), 0, 0, got.deviceView(), av, bv, n32);
        launch_check("update_kernel");
        const auto wv = want.deviceView();
        want.deviceView() = wv + av * bv;
        want.deviceView() = wv - aether::cwiseExp(T(0.5) * bv);
        w = std::string("+= and -= ") + tn;
        expect_same(w.c_str(), got, want);
    }
    {   // a node type of the caller's own: per sample in a kernel, as a whole-array assignment (a kernel of this
        // translation unit) and with fused sums, against the same arithmetic spelled with built-in nodes
        Array<T, 3> got(n), whole(n), want(n);
        hipLaunchKernelGGL(user_node_kernel<T>, grid_of(n), dim3(256), 0, 0, got.deviceView(), av, bv, n32);
        launch_check("user_node_kernel");
        const auto ss = [](auto x) { return x * x * (T(3) - T(2) * x); };
        const auto ps = aether::evaluate_pairs(whole.deviceView(), smoothstep(av) + bv * aether::cwiseSin(smoothstep(bv)));
        const auto pw = aether::evaluate_pairs(want.deviceView(), ss(av) + bv * aether::cwiseSin(ss(bv)));
        w = std::string("user node, per sample ") + tn;
        expect_same(w.c_str(), got, want);
        w = std::string("user node, whole array ") + tn;
        expect_same(w.c_str(), whole, want);
        if (std::memcmp(&ps, &pw, sizeof ps) != 0) {
            std::printf("mismatch: user node sums %s\n", tn);
            ++failures;
        }
    }
    {   // a tree holding a user node that reads its own destination: component 0 of a sample is stored before
        // component 1 is evaluated, as the built-in lowering does (one launch a component in that case)
        Array<T, 3> got(n), want(n);
        got.deviceView() = cv * T(1);
        want.deviceView() = cv * T(1);
        const auto gv = got.deviceView(), wv = want.deviceView();
        const auto ss = [](auto x) { return x * x * (T(3) - T(2) * x); };
        got.deviceView() = smoothstep(gv).cross(bv);
        want.deviceView() = ss(wv).cross(bv);
        w = std::string("user node reading its destination ") + tn;
        expect_same(w.c_str(), got, want);
    }
    {   // get(): the sample's components, read in the kernel
        Array<T, 1> got(n), want(n);
        hipLaunchKernelGGL(get_kernel<T>, grid_of(n), dim3(256), 0, 0, got.deviceView(), av, n32);
        launch_check("get_kernel");
        want.deviceView() = av.sum();
        w = std::string("get ") + tn;
        expect_same(w.c_str(), got, want);
    }
}

void run_quaternions(std::size_t n)
{
    const aether::random::Generator g(7);
    Array<double, 4> q(n), p(n);
    Array<double, 3> v(n);
    q.deviceView() = g.uniform<double, 4>(-1.0, 1.0);
    p.deviceView() = g.substream(1).uniform<double, 4>(-1.0, 1.0);
    v.deviceView() = g.substream(2).uniform<double, 3>(-1.0, 1.0);
    const auto qv = q.deviceView(), pv = p.deviceView();
    const auto vv = v.deviceView();
    Array<double, 3> r(n), r_want(n);
    Array<double, 4> m(n), m_want(n);
    hipLaunchKernelGGL(quat_kernel, grid_of(n), dim3(256), 0, 0, r.deviceView(), m.deviceView(), qv, pv, vv,
                       (std::uint32_t)n);
    launch_check("quat_kernel");
    r_want.deviceView() = qv.quatRotate(vv) + pv.quatMul(qv).tail<3>();
    m_want.deviceView() = qv.quatMul(pv) + pv.quatReciprocal() - vv.asPureQuaternion() + qv.quatConj();
    expect_same("quaternion rotate", r, r_want);
    expect_same("quaternion products", m, m_want);
}

} // namespace

int main()
{
    try {
        for (std::size_t n : {std::size_t(1), std::size_t(257), std::size_t(100003)}) {
            run_cases<double>(n);
            run_cases<float>(n);
            run_quaternions(n);
        }
    } catch (const std::exception& e) {
        std::printf("error: %s\n", e.what());
        return 1;
    }
    if (failures) return 1;
    std::printf("ok\n");
    return 0;
}
