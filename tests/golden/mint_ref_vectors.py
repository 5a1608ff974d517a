"""Mint golden vectors from the REFERENCE's own CPU path.

Runs oracle/_ref/aether_ref (the reference's header-only CPU path compiled
from its sources by `make -C oracle ref`) and stores small input/output
vectors next to this script. Run in the build container, where the reference
tree exists: python tests/golden/mint_ref_vectors.py [members | gnormal]
"""
import hashlib
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "aether_ref")


def run(*args):
    res = subprocess.run([REF, *map(str, args)], capture_output=True, check=True)
    return res.stdout


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def mint_cfg0(mode, n, seeds, dtype, name):
    raw = run(mode, n, *seeds)
    x = np.frombuffer(raw, dtype=dtype)
    planes = x[: 4 * 3 * n].reshape(4, 3, n)
    sums = x[4 * 3 * n:]
    np.savez_compressed(os.path.join(HERE, name), n=n, seeds=np.array(seeds, np.uint64), d=planes[3], sums=sums,
                        sha_a=sha(planes[0]), sha_b=sha(planes[1]), sha_c=sha(planes[2]),
                        head_a=planes[0][:, :8], head_b=planes[1][:, :8], head_c=planes[2][:, :8],
                        mode=mode)
    print(name, planes.shape, sums)


def mint_vec3(n, seeds):
    e = np.frombuffer(run("vec3", n, *seeds), np.float64)
    np.savez_compressed(os.path.join(HERE, "ref_vec3_f64.npz"), n=n, seeds=np.array(seeds, np.uint64), e=e)
    print("vec3", e.shape)


def mint_rows(seed, n, counter):
    rec = np.dtype([("bits", "<u8"), ("ud", "<f8"), ("uf", "<f4"), ("nf", "<f4"), ("nd", "<f8")])
    r = np.frombuffer(run("rows", seed, n, counter), rec)
    np.savez_compressed(os.path.join(HERE, f"ref_rows_s{seed}_c{counter}.npz"), seed=seed, counter=counter,
                        bits=r["bits"], ud=r["ud"], uf=r["uf"], nd=r["nd"], nf=r["nf"])
    print("rows", seed, counter, len(r))


def mint_keys():
    rows = []
    for seed in (42, 7, 0, 2**64 - 1):
        for stream in (0, 1, 5):
            for k in (0, 1, 2, 1000):
                key, sub = map(int, run("keys", seed, stream, k).split())
                rows.append((seed, stream, k, key, sub))
    np.savez_compressed(os.path.join(HERE, "ref_keys.npz"), rows=np.array(rows, np.uint64))
    print("keys", len(rows))


def mint_fft(n, seed):
    x = np.frombuffer(run("fft", n, seed, 1), np.float64)
    np.savez_compressed(os.path.join(HERE, f"ref_fft_n{n}.npz"), n=n, seed=seed,
                        x=x[: 2 * n].view(np.complex128), y=x[2 * n:].view(np.complex128))
    print("fft", n)


MEMBER_PLANES = [  # (name, components), in the order the driver writes them after q, r, v
    ("quat_mul", 4), ("quat_conj", 4), ("quat_reciprocal", 4), ("as_pure_quaternion", 4),
    ("quat_rotate", 3), ("unit_vector", 3), ("as_back3", 3), ("head3", 3), ("min_025", 3), ("max_m025", 3),
    ("clamp_05", 3), ("sign", 3), ("div_back3", 3), ("mul_head3", 3), ("min_back3", 3), ("max_back3", 3),
    ("tan", 3), ("atan", 3), ("asin", 3), ("acos", 3), ("tanh", 3), ("rsqrt_abs", 3), ("pow_abs_15", 3),
    ("pow_abs_back3", 3),
    ("max_norm", 1), ("sum", 1), ("norm", 1), ("squared_norm", 1), ("r_norm", 1), ("r_squared_norm", 1),
    ("cubed_norm", 1), ("r_cubed_norm", 1)]


def mint_members(n, seeds, dtype=np.float64):
    """Inputs are raw-seed uniforms on (-1, 1) (regenerated by the tests), stored as hashes and heads."""
    mode, fname = ("members", "ref_members_f64.npz") if dtype == np.float64 else ("membersf", "ref_members_f32.npz")
    x = np.frombuffer(run(mode, n, *seeds), dtype)
    out, off = {"n": n, "seeds": np.array(seeds, np.uint64)}, 0
    for name, k in [("q", 4), ("r", 4), ("v", 3)] + MEMBER_PLANES:
        plane = x[off: off + k * n].reshape(k, n)
        if name in ("q", "r", "v"):
            out["sha_" + name], out["head_" + name] = sha(plane), plane[:, :4].copy()
        else:
            out[name] = plane
        off += k * n
    assert off == x.size
    np.savez_compressed(os.path.join(HERE, fname), **out)
    print(mode, n, len(MEMBER_PLANES))


def mint_generator_normal(n, seed):
    """Generator(seed).normal<double,3>(0.5, 2) and .substream(3).normal<float,3>(-1, 0.25) planes."""
    raw = run("gnormal", n, seed)
    d = np.frombuffer(raw[: 3 * n * 8], np.float64).reshape(3, n)
    f = np.frombuffer(raw[3 * n * 8:], np.float32).reshape(3, n)
    np.savez_compressed(os.path.join(HERE, "ref_generator_normal.npz"), n=n, seed=seed, d=d, f=f)
    print("gnormal", d.shape, f.shape)


def main(parts):
    if not os.path.exists(REF):
        sys.exit("build the reference driver first: make -C oracle ref")
    if parts == ["gnormal"]:
        mint_generator_normal(1000, 42)
        return
    if parts == ["members"]:
        mint_members(512, (11, 12, 13))
        mint_members(512, (11, 12, 13), np.float32)
        return
    mint_cfg0("cfg0", 2000, (42, 43, 44), np.float64, "ref_cfg0_f64.npz")
    mint_cfg0("cfg0f", 2000, (42, 43, 44), np.float32, "ref_cfg0_f32.npz")
    raw = run("cfg0g", 2000, 42)
    x = np.frombuffer(raw, np.float64)
    planes = x[: 4 * 3 * 2000].reshape(4, 3, 2000)
    np.savez_compressed(os.path.join(HERE, "ref_cfg0_generator_f64.npz"), n=2000, seed=42, d=planes[3],
                        sums=x[4 * 3 * 2000:], sha_a=sha(planes[0]), sha_b=sha(planes[1]), sha_c=sha(planes[2]))
    print("cfg0g", x[4 * 3 * 2000:])
    mint_vec3(2000, (42, 43, 44))
    for seed in (42, 7):
        for counter in (0, 3):
            mint_rows(seed, 512, counter)
    mint_keys()
    for n in (64, 1024):
        mint_fft(n, 7)
    mint_members(512, (11, 12, 13))
    mint_members(512, (11, 12, 13), np.float32)
    mint_generator_normal(1000, 42)


if __name__ == "__main__":
    main(sys.argv[1:])
