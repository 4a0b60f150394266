"""The oracle's site random-number streams and lattice fills against the
reference's golden vectors (test/site_rng.jl), Random123's and PCG's published
known answers, closed-form identities and the reference's statistical and
fill tests (test/random_fill.jl)."""
import itertools
import json
import os

import numpy as np
import pytest

import oracle as orc

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "rng_known_answers.json")))
KEY = tuple(GOLDEN["site_rng"]["key"])  # test/site_rng.jl:14
ALGS = ("pcg32", "xoshiro256++", "philox4x32")


def test_philox4x32_10_known_answers():
    # Random123 kat_vectors for philox4x32_10; the first is also test/site_rng.jl:136-145
    for v in GOLDEN["philox4x32_10_kat"]["vectors"]:
        assert orc.philox4x32_10(tuple(v["ctr"]), tuple(v["key"])) == tuple(v["out"])


def test_pcg32_step_reference_sequence():
    # pcg32_srandom_r(rng, 42, 54) then six draws: the PCG reference demo
    demo = GOLDEN["pcg32_demo"]
    inc = (demo["stream"] << 1) | 1
    st, _ = orc.pcg32_step(0, inc)
    st, _ = orc.pcg32_step((st + demo["seed"]) % (1 << 64), inc)
    out = []
    for _ in range(6):
        st, v = orc.pcg32_step(st, inc)
        out.append(v)
    assert out == demo["out"]


def test_site_stream_golden_words():
    gs = GOLDEN["site_rng"]
    site = gs["site"]
    pcg = orc.site_rng_draw(KEY, "pcg32", site, 1, words=8)["words"][0]  # test/site_rng.jl:23-37
    assert list(pcg) == gs["pcg32_u32"]
    philox = orc.site_rng_draw(KEY, "philox4x32", site, 1, words=8)["words"][0]  # :146-160
    assert list(philox) == gs["philox4x32_u32"]
    # xoshiro256++ u64 words (:120-134): u32 draws are their high halves and
    # Float64 uniforms their top 53 bits
    x64 = gs["xoshiro256pp_u64"]
    xw = orc.site_rng_draw(KEY, "xoshiro256++", site, 1, words=4)["words"][0]
    assert list(xw) == [v >> 32 for v in x64]
    xu = orc.site_rng_draw(KEY, "xoshiro256++", site, 1, uniform64=4)["uniform64"][0]
    assert [int(u * 2.0 ** 53) for u in xu] == [v >> 11 for v in x64]


def test_philox_site_stream_is_counter_blocks():
    # the stream at site s hands out block k = philox(ctr = (k, 0, s lo, s hi),
    # key = splitmix64(stream seed ^ salt)) word by word
    h = orc.rng_stream_seed(*KEY)
    k = orc.splitmix64(h ^ 0x9216D5D98979FB1B)
    site = (7 << 32) | 41
    words = orc.site_rng_draw(KEY, "philox4x32", site, 1, words=12)["words"][0]
    for b in range(3):
        blk = orc.philox4x32_10((b, 0, site & 0xFFFFFFFF, site >> 32), (k & 0xFFFFFFFF, k >> 32))
        assert tuple(words[4 * b:4 * b + 4]) == blk


def test_key_tags_and_sites_change_the_stream():
    base = list(orc.site_rng_draw(KEY, "pcg32", 41, 1, words=4)["words"][0])
    for i in range(5):  # test/site_rng.jl:96-121
        k = list(KEY)
        k[i] += 1
        assert list(orc.site_rng_draw(tuple(k), "pcg32", 41, 1, words=4)["words"][0]) != base
    assert list(orc.site_rng_draw(KEY, "pcg32", 42, 1, words=4)["words"][0]) != base


@pytest.mark.parametrize("alg", ALGS)
def test_uniform_identities_and_ranges(alg):
    n = 64
    w = orc.site_rng_draw(KEY, alg, 5, 3, words=2 * n)["words"].astype(np.uint64)
    u32 = orc.site_rng_draw(KEY, alg, 5, 3, uniform32=n)["uniform32"]
    assert np.array_equal(u32, ((w[:, :n] >> 8).astype(np.float32) * np.float32(2.0 ** -24)))
    u64 = orc.site_rng_draw(KEY, alg, 5, 3, uniform64=n)["uniform64"]
    if alg == "xoshiro256++":  # 64-bit outputs: checked against the golden words above
        assert np.all((u64 >= 0) & (u64 < 1))
    else:
        full = (w[:, 0::2] << np.uint64(32)) | w[:, 1::2]
        assert np.array_equal(u64, (full >> np.uint64(11)).astype(np.float64) * 2.0 ** -53)
    big = orc.site_rng_draw(KEY, alg, 0, 64, uniform32=500, uniform64=500)
    for k in ("uniform32", "uniform64"):
        assert big[k].min() >= 0 and big[k].max() < 1


@pytest.mark.parametrize("alg", ALGS)
def test_box_muller_pairs(alg):
    # a normal pair is Box-Muller of (open uniform, uniform) drawn in that
    # order: 8 pairs use the same 16 uniforms a fresh stream hands out
    n64 = orc.site_rng_draw(KEY, alg, 9, 16, normal64=16)["normal64"]
    u64 = orc.site_rng_draw(KEY, alg, 9, 16, uniform64=16)["uniform64"]
    r, a = np.sqrt(-2.0 * np.log(u64[:, 0::2])), 6.283185307179586 * u64[:, 1::2]
    assert np.allclose(n64[:, 0::2], r * np.cos(a), rtol=4e-15, atol=4e-15)
    assert np.allclose(n64[:, 1::2], r * np.sin(a), rtol=4e-15, atol=4e-15)
    n32 = orc.site_rng_draw(KEY, alg, 9, 16, normal32=16)["normal32"]
    u32 = orc.site_rng_draw(KEY, alg, 9, 16, uniform32=16)["uniform32"]
    r32 = np.sqrt(-2.0 * np.log(u32[:, 0::2].astype(np.float64)))
    a32 = (np.float32(6.283185307179586) * u32[:, 1::2]).astype(np.float64)  # the angle is rounded to Float32
    assert np.allclose(n32[:, 0::2], r32 * np.cos(a32), rtol=2e-6, atol=2e-6)
    assert np.allclose(n32[:, 1::2], r32 * np.sin(a32), rtol=2e-6, atol=2e-6)


def test_normal_moments():
    # test/site_rng.jl:70-83: 100000 Philox normals at site 73
    z = orc.site_rng_draw(KEY, "philox4x32", 73, 1, normal64=100_000)["normal64"][0]
    assert abs(z.mean()) < 0.015 and abs(z.std() - 1) < 0.015


@pytest.mark.parametrize("alg", ALGS)
def test_bounded_counts(alg):
    b = orc.site_rng_draw(KEY, alg, 41, 1, bounded=7000, upper=7)["bounded"][0]  # test/site_rng.jl:84-94
    counts = np.bincount(b, minlength=7)
    assert b.max() < 7 and np.all(counts > 800) and np.all(counts < 1200)


def _site_loop_fill(gsize, nc1, nc2, key, alg, dtype, gaussian, sigma=1.0):
    """test/random_fill.jl:1-43 restated on top of the oracle's per-site draws:
    one stream per global site (x fastest), entries (ic fastest, jc)."""
    dt = np.dtype(dtype)
    real = np.float32 if dt in (np.float32, np.complex64) else np.float64
    cplx = dt.kind == "c"
    ne = nc1 * nc2
    V = int(np.prod(gsize))
    out = np.zeros((V, nc2, nc1), dtype=dt)
    for s in range(V):
        if not gaussian:
            cnt = 2 * ne if cplx else ne
            k = "uniform32" if real is np.float32 else "uniform64"
            u = orc.site_rng_draw(key, alg, s, 1, **{k: cnt})[k][0]
            vals = (u[0::2] - real(0.5)) + 1j * (u[1::2] - real(0.5)) if cplx else u - real(0.5)
        else:
            cnt = 2 * ne if cplx else ne + (ne % 2)
            k = "normal32" if real is np.float32 else "normal64"
            z = orc.site_rng_draw(key, alg, s, 1, **{k: cnt})[k][0]
            sg = real(sigma)
            vals = (sg * z[0::2]) + 1j * (sg * z[1::2]) if cplx else (sg * z)[:ne]
        out[s] = np.asarray(vals, dtype=dt).reshape(nc2, nc1)
    return out


@pytest.mark.parametrize("alg", ALGS)
def test_fills_match_site_loop(alg):
    # the shapes and keys of test/random_fill.jl:47-114 on the (8, 3) lattice
    ukey = (0x8F3C2A19, 7, 2, 1, 0x55)
    gkey = (0x51F15E, 11, 4, 0, 0x66)
    for dtype in (np.complex128, np.complex64, np.float64, np.float32):
        u = orc.randomize_fill(dtype, 3, 2, (8, 3), ukey, algorithm=alg)
        assert np.array_equal(u, _site_loop_fill((8, 3), 3, 2, ukey, alg, dtype, False))
        assert u.real.min() >= -0.5 and u.real.max() < 0.5
        g = orc.randomize_fill(dtype, 5, 1, (8, 3), gkey, gaussian=True, sigma=1.75, algorithm=alg)
        assert np.array_equal(g, _site_loop_fill((8, 3), 5, 1, gkey, alg, dtype, True, 1.75))
    changed = orc.randomize_fill(np.complex128, 3, 2, (8, 3), (0x8F3C2A19, 7, 3, 1, 0x55), algorithm=alg)
    assert not np.array_equal(changed, orc.randomize_fill(np.complex128, 3, 2, (8, 3), ukey, algorithm=alg))


def test_fill_is_decomposition_independent():
    # any box of the global lattice, placed by its offset, equals that slice of
    # the whole fill (test/site_rng.jl:162-172 and the MPI fills of random_fill.jl)
    G = (12, 10, 8, 6)
    key = (123, 9, 2, 0, 0)
    whole = orc.randomize_fill(np.complex128, 3, 4, G, key, gaussian=True, sigma=0.5)
    W = whole.reshape(G[3], G[2], G[1], G[0], 4, 3)
    for split in ((6, 10, 8, 6), (12, 5, 8, 6), (12, 10, 4, 3)):
        for off in itertools.product(*[range(0, g, s) for g, s in zip(G, split)]):
            box = orc.randomize_fill(np.complex128, 3, 4, split, key, gaussian=True, sigma=0.5, offset=off, gdims=G)
            ref = W[off[3]:off[3] + split[3], off[2]:off[2] + split[2], off[1]:off[1] + split[1],
                    off[0]:off[0] + split[0]]
            assert np.array_equal(box.reshape(ref.shape), ref)


def test_gaussian_fill_statistics():
    # test/random_fill.jl:117-138: 64 x 32 real 8x1 field, sigma 1.6, seed 0x31415926
    v = orc.randomize_fill(np.float64, 8, 1, (64, 32), (0x31415926, 0, 0, 0, 0), gaussian=True, sigma=1.6)
    m = v.mean()
    sd = np.sqrt((v ** 2).mean() - m * m)
    assert abs(m) < 0.04 * 1.6 and abs(sd - 1.6) < 0.04 * 1.6


def test_real_gaussian_uses_both_normals_of_a_pair():
    # 5 real entries = pairs (z0, z1), (z0, z1), (z0, -): the odd tail drops z1
    key = (5, 0, 0, 0, 0)
    z = orc.site_rng_draw(key, "philox4x32", 0, 1, normal64=6)["normal64"][0]
    g = orc.randomize_fill(np.float64, 5, 1, (1,), key, gaussian=True, sigma=2.0)
    assert np.array_equal(g.reshape(-1), 2.0 * z[:5])
