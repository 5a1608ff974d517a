"""The C++ front end (include/aether/aether.hpp) through its test driver
tests/cpp/api_check.cpp, built by the package build with g++.

CPU: the C++ lowering of `dest = expr` gives the same programs as the Python
lowering for the same trees (postfix order, load and leaf numbering,
constants, salted random leaves).
GPU: device assignment through the headers is bit-exact with the oracle, the
per-sample host spelling reproduces the reference's own output for the same
Generator seed (ref_cfg0_generator_f64.npz, libm route), raw-seed vec3
geometry matches the reference (ref_vec3_f64.npz), fft::c2c matches the
reference's pocketfft output within its error bound, and upload / resize /
download keep the samples.
"""
import hashlib
import os
import subprocess

import numpy as np
import pytest
import torch

import aether_amd as ae
from aether_amd import build as B
from aether_amd import expr as E
from conftest import GOLDEN

NAMES = {v: k for k, v in E.OP.items()}


def run(*args, stdin=None):
    assert os.path.exists(B.API_CHECK), "api_check not built (run the package build)"
    r = subprocess.run([B.API_CHECK, *map(str, args)], input=stdin, capture_output=True, timeout=120)
    assert r.returncode == 0, r.stderr.decode()[-2000:]
    return r.stdout


def render(name, lowered, ids):
    progs, leaves, consts, rngs = lowered
    out = [f"expr {name} programs {len(progs)} leaves {len(leaves)} consts {len(consts)} rngs {len(rngs)}",
           "leaves " + " ".join(str(next(k for k, x in enumerate(ids) if x is lf)) for lf in leaves)]
    for p in progs:
        loads = " ".join(f"{p.load_leaf[k]}.{p.load_comp[k]}" for k in range(p.n_loads))
        insns = " ".join(f"{p.insns[k].op}.{p.insns[k].arg}" for k in range(p.n_insns))
        out.append(f"program {p.n_loads} {p.n_insns} {p.out_leaf} {p.out_comp} : {loads} : {insns}")
    out += [f"const {v:.17g}" for v in consts]
    out += [f"rng {s} {p0:.17g} {p1:.17g} {c} {d}" for (s, p0, p1, c, d) in rngs]
    return out


def test_bundles_equal_per_sample_assignment():
    """BundleIndex / bundleAssign / View[BundleIndex] (index/BundleIndex.h:56-112, backend/cuda/bundle/Assign.h)
    on host views: every bundle width and a masked tail give the per-sample assignment's bits, aligned and
    not (the reference's own gate, tests/test_Bundle.cpp, with n = 37); the masks and PacketIndex."""
    assert run("bundle").decode().strip() == "bundles ok"


def test_items_and_scalar_broadcast():
    """Item / Vec3d (view/Item.h, aether.h:234-260) and scalar = += -= through SampleRef (the rows of the
    reference's tests/test_ScalarBroadcast.cpp), Items in sample-free trees (cross, the dot fold order),
    get() as an Item, and an Item as a leaf of a per-sample tree, on host views."""
    assert run("item").decode().strip() == "items ok"


def test_cpp_lowering_equals_python_lowering():
    got = run("tape").decode().splitlines()
    arr = lambda k: ae.Array(10, k, torch.float64, device="cpu")  # noqa: E731
    a, b, c, d, e, s, q, r = arr(3), arr(3), arr(3), arr(3), arr(1), arr(1), arr(4), arr(4)
    ids = [a, b, c, d, e, s, q, r]
    g = ae.Generator.from_seed(7)
    want = []
    want += render("cfg0", E.lower(d, a + b * c + E.sin(a) * E.exp(-1.0 * (b * b))), ids)
    want += render("vec3", E.lower(e, a.cross(b).dot(c) + E.sqrt(a.dot(a) + 1.0)), ids)
    want += render("negdiv", E.lower(s, -E.Leaf(e) / 4.0), ids)
    want += render("ops", E.lower(s, E.log(E.abs_(e) + 0.5) - E.cos(e) * E.sqrt(e * e + 1.0) / (e + 2.0)), ids)
    want += render("random", E.lower(d, E.uniform(g, 3, -2.0, 5.0) + E.normal(g.substream(3), 3, 1.0, 0.5) * a),
                   ids)
    want += render("norm", E.lower(e, E.Leaf(a).norm() - E.Leaf(b).squared_norm()), ids)
    want += render("inplace", E.lower(a, a * 2.0 + 1.0), ids)
    L = E.Leaf
    want += render("cwise", E.lower(d, E.tan(a) + E.atan(b) - E.asin(c) * E.acos(a) + E.tanh(b) + E.sign(c) +
                                    E.rsqrt(a)), ids)
    want += render("minmax", E.lower(d, E.clamp(a, -0.5, 0.5) + E.minimum(a, b) * E.maximum(b, 0.25) + E.div(a, c) +
                                     E.pow_(E.abs_(a), 1.5) + E.pow_(E.abs_(b), c)), ids)
    want += render("norms1", E.lower(e, L(a).rNorm() + L(a).rSquaredNorm()), ids)
    want += render("norms2", E.lower(e, L(a).cubedNorm() - L(a).rCubedNorm()), ids)
    want += render("norms3", E.lower(e, L(q).maxNorm() + L(a).sum()), ids)
    want += render("unit", E.lower(d, L(a).unitVector() + L(a).cwiseMul(b)), ids)
    want += render("quat", E.lower(q, L(q).quatMul(r) + L(q).quatConj()), ids)
    want += render("qrecip", E.lower(r, L(q).quatReciprocal()), ids)
    want += render("qrot", E.lower(d, L(q).quatRotate(a)), ids)
    want += render("slices", E.lower(d, L(q).asBack3DVector() + L(q).head(3) - L(a).asPureQuaternion().tail(3)), ids)
    assert got == want


def test_header_compiles_standalone(tmp_path):
    # the header needs nothing beyond the C ABI header and the standard library
    src = tmp_path / "t.cpp"
    src.write_text('#include <aether/aether.hpp>\nint main() { return 0; }\n')
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    r = subprocess.run(["g++", "-std=c++20", "-Wall", "-Wextra", "-Werror", "-fsyntax-only", "-I", inc, str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


N = 2000


def planes(buf, dtype, count, n=N):
    x = np.frombuffer(buf, dtype=dtype, count=count * 3 * n)
    return x.reshape(count, 3, n)


@pytest.mark.gpu
def test_device_assignment_bit_exact(dev, orc):
    z = np.load(os.path.join(GOLDEN, "ref_cfg0_generator_f64.npz"))
    out = run("cfg0g", N, 42)
    abcd = planes(out, np.float64, 4)
    sums = np.frombuffer(out[abcd.nbytes:], dtype=np.float64)
    # inputs: the reference's Generator(42) draws (sha256 of its compact planes)
    for k, key in enumerate(("sha_a", "sha_b", "sha_c")):
        assert hashlib.sha256(np.ascontiguousarray(abcd[k]).tobytes()).hexdigest() == str(z[key])
    pitch = -(-N // 32) * 32
    pad = lambda x: np.pad(x, ((0, 0), (0, pitch - N))).reshape(-1)  # noqa: E731
    ref = orc.eval_cfg0(pad(abcd[0]), pad(abcd[1]), pad(abcd[2]), 3, N, pitch, mathmode=1).reshape(3, pitch)[:, :N]
    assert np.array_equal(abcd[3].view(np.uint64), np.ascontiguousarray(ref).view(np.uint64))
    for k in range(3):
        h, c = orc.device_sum(np.ascontiguousarray(ref[k]))
        assert sums[k] == h + c


@pytest.mark.gpu
def test_host_spelling_reproduces_reference(dev):
    z = np.load(os.path.join(GOLDEN, "ref_cfg0_generator_f64.npz"))
    abcd = planes(run("host", N, 42), np.float64, 4)
    assert np.array_equal(abcd[3].view(np.uint64), z["d"].view(np.uint64))


@pytest.mark.gpu
def test_vec3_geometry_matches_reference(dev):
    z = np.load(os.path.join(GOLDEN, "ref_vec3_f64.npz"))
    sa, sb, sc = (int(v) for v in z["seeds"])
    e = np.frombuffer(run("vec3", N, sa, sb, sc), dtype=np.float64)
    assert np.array_equal(e.view(np.uint64), z["e"].view(np.uint64))


@pytest.mark.gpu
@pytest.mark.parametrize("n", [64, 1024])
def test_fft_c2c_matches_reference(dev, n):
    z = np.load(os.path.join(GOLDEN, f"ref_fft_n{n}.npz"))
    x = np.ascontiguousarray(z["x"])
    y = np.frombuffer(run("fft", n, 1, stdin=x.tobytes()), dtype=np.complex128)
    # pocketfft and a radix-4 Stockham differ by rounding: bound ~ log2(n) ulp of sum |x|
    bound = 8 * np.log2(n) * 2.0**-52 * np.abs(x).sum()
    assert np.max(np.abs(y - z["y"])) <= bound
    back = np.frombuffer(run("fft", n, 0, stdin=y.tobytes()), dtype=np.complex128)
    assert np.max(np.abs(back - x)) <= bound  # backward norm scales by 1/n


@pytest.mark.gpu
def test_float_assignment_with_float_float_sums(dev, orc):
    out = run("float", N, 42)
    d = np.frombuffer(out, dtype=np.float32, count=3 * N).reshape(3, N)
    sums = np.frombuffer(out[d.nbytes:], dtype=np.float64)
    # inputs: the float conversion of the same Generator(42) streams (random/Generator.h:64-90)
    key = orc.generator_key(42, 0)
    seeds = [key ^ 0x243F6A8885A308D3, orc.substream_key(key, 1) ^ 0x243F6A8885A308D3,
             orc.substream_key(key, 2) ^ 0x243F6A8885A308D3]
    pitch = -(-N // 32) * 32
    A, B_, C_ = (orc.fill_array_uniform(s, 3, N, pitch, dtype=np.float32) for s in seeds)
    ref = orc.eval_cfg0(A, B_, C_, 3, N, pitch, mathmode=1).reshape(3, pitch)[:, :N]
    assert np.array_equal(d.view(np.uint32), np.ascontiguousarray(ref).view(np.uint32))
    for k in range(3):
        h, c = orc.device_sum(np.ascontiguousarray(ref[k]), mode="ff")
        assert sums[k] == float(h) + float(c)


@pytest.mark.gpu
def test_upload_resize_download(dev):
    assert run("resize", 5000).decode().startswith("resize ok")


@pytest.mark.gpu
@pytest.mark.parametrize("rows,n", [(3, 64), (1, 1024), (5, 2)])
def test_rfft_round_trip(dev, rows, n):
    x = np.random.default_rng(n).standard_normal((rows, n))
    out = run("rfft", rows, n, stdin=x.tobytes())
    h = n // 2 + 1
    spec = np.frombuffer(out, dtype=np.complex128, count=rows * h).reshape(rows, h)
    back = np.frombuffer(out[spec.nbytes:], dtype=np.float64).reshape(rows, n)
    c = (8 + 4 * np.log2(n + 1)) * 2.0**-52 * np.linalg.norm(x)
    assert np.max(np.abs(spec - np.fft.rfft(x, axis=1))) <= c * np.sqrt(n)
    assert np.max(np.abs(back - x)) <= 2 * c


@pytest.mark.gpu
def test_append_commit_clear(dev):
    assert run("append", 100).decode().startswith("append ok")


@pytest.mark.gpu
def test_accum_plane_delivers_the_exact_sum(dev):
    assert run("plane", 100_003).decode().startswith("plane ok")


@pytest.mark.gpu
def test_append_where_commits_the_selected_samples(dev):
    assert run("where", 300_007, 0.75).decode().startswith("where ok")


@pytest.mark.gpu
def test_keyed_sum_compensated_within_one_eps(dev):
    assert run("keyed", 1 << 18, 777).decode().startswith("keyed ok")


@pytest.mark.gpu
def test_runtime_eval_strided_views(dev):
    assert run("runtime", 123, 77).decode().startswith("runtime ok")


@pytest.mark.gpu
def test_tune_with_and_without_sums_keeps_the_bits(dev):
    out = run("tune", (1 << 21) + 333, 9).decode().split("\n")
    assert out[0].startswith("tune ok 2048 16 ") and out[1].startswith("tune ok"), out
