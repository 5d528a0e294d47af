"""SHARC radiance cache: key hashing, claims, deposits, EMA resolve, eviction, queries.

The reference's own GPU test (crates/arcadia-render/tests/sharc_gpu.rs:268-349: racing deposits,
resolve, jittered / unjittered queries, EMA blend, aging to eviction) is restated here against
the oracle (CPU) and against the device kernels; the layout KATs of
crates/arcadia-render/src/sharc.rs:441-475 pin the buffer sizes. Device and oracle are compared
per surfel key: which way of a bucket a new key takes depends on the order of racing claims,
the resolved contents per key and every query result do not (integer deposit sums are
order-independent).
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import arcadia_amd as A
from oracle import oracle as O

POS = (40.13, 12.07, 55.31)
CAM = (41.0, 13.0, 56.0)


def samples(rows):
    s = np.zeros(len(rows), O.SHARC_SAMPLE_DTYPE)
    for i, (p, n, r) in enumerate(rows):
        s[i]["pos"], s[i]["normal"], s[i]["radiance"] = p, n, r
    return s


def kat_probes():
    rows = [(POS, (1, 0, 0), (0, 0)), (POS, (1, 0, 0), (0.9, -0.8)), (POS, (-1, 0, 0), (0, 0)),
            ((POS[0] + 50.0, POS[1], POS[2]), (1, 0, 0), (0, 0)), (POS, (0, 1, 0), (0, 0))]
    q = np.zeros(len(rows), O.SHARC_PROBE_DTYPE)
    for i, (p, n, j) in enumerate(rows):
        q[i]["pos"], q[i]["normal"], q[i]["jitter"] = p, n, j
    return q


def first_frame():
    # 64 threads deposit the +X surfel, every second one also the -X face (sharc_gpu.rs:20-26)
    rows = []
    for i in range(64):
        rows.append((POS, (1, 0, 0), (1.0, 2.0, 3.0)))
        if i % 2 == 0:
            rows.append((POS, (-1, 0, 0), (0.25, 0.5, 0.75)))
    return samples(rows)


def second_frame():
    return samples([(POS, (1, 0, 0), (3.0, 2.0, 1.0))] * 16)


def check_kat(deposit, resolve, query, keys, accum):
    """sharc_deposit_resolve_query_roundtrip_and_eviction (sharc_gpu.rs:268-349)."""
    deposit(first_frame())
    assert np.count_nonzero(keys()) == 2
    counts = sorted(c for c in accum().reshape(-1, 4)[:, 3] if c > 0)
    assert counts == [32, 64]
    resolve(1)
    assert not accum().any()
    q = query()
    close = lambda a, e: np.allclose(a, e, atol=1e-3)
    assert close(q[0], [1, 2, 3, 1]) and close(q[1], [1, 2, 3, 1]) and close(q[2], [0.25, 0.5, 0.75, 1])
    assert close(q[3], [0, 0, 0, 0]) and close(q[4], [0, 0, 0, 0])
    deposit(second_frame())
    resolve(1)
    q = query()
    assert close(q[0], [2, 2, 2, 1]) and close(q[2], [0.25, 0.5, 0.75, 1])
    resolve(63)
    q = query()
    assert close(q[2], [0.25, 0.5, 0.75, 1]) and close(q[0], [2, 2, 2, 1])
    resolve(1)
    q = query()
    assert close(q[2], [0, 0, 0, 0]) and close(q[0], [2, 2, 2, 1])
    assert np.count_nonzero(keys()) == 1


def test_layout_kats():  # sharc.rs:441-450
    slots = 4 << 15
    assert slots == 131_072 and slots * 4 + slots * 16 + slots * 16 == 4_718_592
    big = 4 << 20
    assert big == 1 << 22 and (big * 36) / 2**20 == 144.0
    assert A.SHARC_SAMPLE_DTYPE.itemsize == 40 and A.SHARC_PROBE_DTYPE.itemsize == 32


def test_cascade_level_is_floor_log2():
    """The level is floor(log2(max(d, 4) / 4)) clamped to 0..3 (sharc.wgsl:47-49), exactly."""
    sh = O.Sharc(15)
    rng = np.random.default_rng(0)
    for d in np.concatenate([rng.uniform(0.0, 80.0, 300), [4.0, 8.0, 16.0, 32.0, 7.999, 8.001, 64.0, 1e6]]):
        d = float(np.float32(d))
        want = int(min(max(math.floor(math.log2(max(d, 4.0) * 0.25)), 0), 3))
        # the level sets the cell size: the key changes exactly when the level does
        cam = (0.0, 0.0, 0.0)
        pos = (d, 0.0, 0.0)
        s = 0.25 * (1 << want)
        ax = float(np.float32(pos[0]) - np.float32(1.0) * np.float32(0.02))  # binary32 anchor
        cell = [math.floor(ax / s) + 16384, 16384, 16384]
        h = lambda x: int(x) & 0xFFFFFFFF
        pcg = O_pcg
        b = pcg(h(cell[0] * 73856093) ^ h(cell[1] * 19349663) ^ h(cell[2] * 83492791) ^ 0 ^ h(want * 38069)) & 32767
        assert sh.make_key(pos, (1, 0, 0), cam)[0] == b, d


def O_pcg(v):
    state = (v * 747796405 + 2891336453) & 0xFFFFFFFF
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & 0xFFFFFFFF
    return (word >> 22) ^ word


def test_key_hash_against_direct_formula():
    """Bucket / key of sharc_make_key recomputed in Python integers for axis normals."""
    sh = O.Sharc(15)
    rng = np.random.default_rng(1)
    for _ in range(200):
        p = tuple(float(v) for v in rng.uniform(0, 60, 3).astype(np.float32))
        axis, sign = int(rng.integers(0, 3)), float(rng.choice([-1.0, 1.0]))
        n = [0.0, 0.0, 0.0]
        n[axis] = sign
        cam = (30.0, 10.0, 30.0)
        d = float(np.float32(math.sqrt(sum((np.float32(p[k]) - np.float32(cam[k])) ** 2 for k in range(3)))))
        lod = int(min(max(math.floor(math.log2(max(d, 4.0) * 0.25)), 0), 3))
        s = 0.25 * (1 << lod)
        a = [np.float32(p[k]) - np.float32(n[k]) * np.float32(0.02) for k in range(3)]
        cell = [(math.floor(float(a[k]) / s) + 16384) & 0xFFFFFFFF for k in range(3)]
        nid = 2 * axis + (0 if sign > 0 else 1)
        m = lambda x: x & 0xFFFFFFFF
        bucket = O_pcg(m(cell[0] * 73856093) ^ m(cell[1] * 19349663) ^ m(cell[2] * 83492791) ^ m(nid * 2654435761)
                       ^ m(lod * 38069)) & 32767
        key = O_pcg(bucket ^ m(cell[0] * 1013904223) ^ m(cell[1] * 1664525) ^ m(cell[2] * 2891336453) ^ (nid << 20)
                    ^ (lod << 24)) | 0x80000000
        assert sh.make_key(p, n, cam) == (bucket, key)


def test_oracle_reference_kat():
    sh = O.Sharc(15)
    check_kat(lambda s: sh.deposit(s, CAM), lambda k: [sh.resolve() for _ in range(k)],
              lambda: sh.query(kat_probes(), CAM), lambda: sh.keys, lambda: sh.accum)


def test_oracle_ema_and_fixed_point():
    """Three frames of one surfel: fixed point truncation, alpha = 1/hist, f16 rounding."""
    sh = O.Sharc(10)
    vals = [(0.7, 1.3, 40.0), (0.1, 0.2, 0.3), (5.0, 6.0, 7.0)]
    ema = np.zeros(3)
    for f, v in enumerate(vals):
        sh.deposit(samples([(POS, (0, 1, 0), v), (POS, (0, 1, 0), (0.0, 0.0, 0.0))]), CAM)
        sh.resolve()
        fixed = np.floor(np.clip(np.array(v, np.float32), 0, 32) * np.float32(1024.0))
        smp = (fixed / (2 * 1024.0)).astype(np.float32)
        alpha = 1.0 / min(f + 1, 16)
        ema = (ema * (1 - alpha) + smp * alpha).astype(np.float16).astype(np.float32)
        q = sh.query(kat_probes()[4:5], CAM)[0]
        assert np.allclose(q[:3], ema, rtol=2e-3, atol=1e-4) and q[3] == 1.0


def random_deposits(rng, n, cam):
    s = np.zeros(n, O.SHARC_SAMPLE_DTYPE)
    s["pos"] = (np.asarray(cam, np.float32) + rng.uniform(-20, 20, (n, 3))).astype(np.float32)
    axis = rng.integers(0, 3, n)
    sign = rng.choice([-1.0, 1.0], n)
    nrm = np.zeros((n, 3), np.float32)
    nrm[np.arange(n), axis] = sign
    s["normal"] = nrm
    s["radiance"] = rng.uniform(0, 40, (n, 3)).astype(np.float32)
    return s


@pytest.mark.gpu
def test_gpu_reference_kat(ctx):
    import torch

    sh = A.Sharc(15)

    def deposit(s):
        d = torch.from_numpy(s.view(np.uint8).copy()).cuda()
        sh.deposit(d.data_ptr(), len(s), CAM)
        torch.cuda.synchronize()

    probes = torch.from_numpy(kat_probes().view(np.uint8).copy()).cuda()

    def query():
        out = torch.zeros(5 * 4, dtype=torch.float32, device="cuda")
        sh.query(probes.data_ptr(), 5, CAM, out.data_ptr())
        torch.cuda.synchronize()
        return out.cpu().numpy().reshape(5, 4)

    def resolve(k):
        for _ in range(k):
            sh.resolve()
        torch.cuda.synchronize()

    check_kat(deposit, resolve, query, lambda: sh.read()[0], lambda: sh.read()[1])
    sh.close()


@pytest.mark.gpu
def test_gpu_sharc_matches_oracle_per_key(ctx):
    import torch

    log2 = 13  # 32768 slots for 6000 surfels: some buckets fill, most do not
    sh, ref = A.Sharc(log2), O.Sharc(log2)
    rng = np.random.default_rng(7)
    cam = (30.0, 12.0, 30.0)
    pool = random_deposits(rng, 6000, cam)

    for frame in range(5):
        s = pool[rng.integers(0, len(pool), 30000)].copy()  # surfels revisited every frame
        s["radiance"] = rng.uniform(0, 40, (len(s), 3)).astype(np.float32)
        d = torch.from_numpy(s.view(np.uint8).copy()).cuda()
        sh.deposit(d.data_ptr(), len(s), cam)
        ref.deposit(s, cam)
        sh.resolve()
        ref.resolve()
        torch.cuda.synchronize()
        k, _, r = sh.read()
        dkeys, dres = k.reshape(-1, 4), r.reshape(-1, 4, 4)
        rkeys, rres = ref.keys.reshape(-1, 4), ref.resolved.reshape(-1, 4, 4)
        # buckets whose claims could race differently: more distinct keys deposited than ways
        assert np.count_nonzero(dkeys) > 1000
        full = (np.count_nonzero(rkeys, axis=1) == 4) | (np.count_nonzero(dkeys, axis=1) == 4)
        print(f"[sharc] frame {frame}: {np.count_nonzero(rkeys)} surfels, {full.sum()} full buckets", flush=True)
        for bk in np.nonzero(~full)[0]:
            a = {int(r[0]): tuple(r) for r in dres[bk] if r[0]}
            b = {int(r[0]): tuple(r) for r in rres[bk] if r[0]}
            assert a == b, (frame, bk, a, b)
        probes = np.zeros(len(s), O.SHARC_PROBE_DTYPE)
        probes["pos"], probes["normal"] = s["pos"], s["normal"]
        probes["jitter"][::3] = rng.uniform(-1, 1, (len(probes[::3]), 2))
        dp = torch.from_numpy(probes.view(np.uint8).copy()).cuda()
        out = torch.zeros(len(probes) * 4, dtype=torch.float32, device="cuda")
        sh.query(dp.data_ptr(), len(probes), cam, out.data_ptr())
        got = out.cpu().numpy().reshape(-1, 4)
        want = ref.query(probes, cam)
        # compare probes whose (jittered and unjittered) buckets are not full in either cache
        ok = np.ones(len(probes), bool)
        for i in range(len(probes)):
            for j in ((0.0, 0.0), tuple(probes["jitter"][i])):
                bk, _ = ref.make_key(tuple(probes["pos"][i]), tuple(probes["normal"][i]), cam, j)
                if full[bk]:
                    ok[i] = False
        assert ok.sum() > 0.5 * len(ok)
        assert np.array_equal(got[ok].view(np.uint32), want[ok].view(np.uint32))
    sh.close()


@pytest.mark.gpu
def test_gpu_invalidate_evicts_surfels_inside_spheres(ctx):
    import torch

    sh = A.Sharc(14)
    rng = np.random.default_rng(3)
    cam = (30.0, 12.0, 30.0)
    s = random_deposits(rng, 4000, cam)
    d = torch.from_numpy(s.view(np.uint8).copy()).cuda()
    sh.deposit(d.data_ptr(), len(s), cam)
    sh.resolve()
    spheres = np.array([[30.0, 12.0, 30.0, 6.0], [40.0, 20.0, 22.0, 3.0]], np.float32)
    probes = np.zeros(len(s), O.SHARC_PROBE_DTYPE)
    probes["pos"], probes["normal"] = s["pos"], s["normal"]
    dp = torch.from_numpy(probes.view(np.uint8).copy()).cuda()
    out = torch.zeros(len(s) * 4, dtype=torch.float32, device="cuda")
    sh.query(dp.data_ptr(), len(s), cam, out.data_ptr())
    before = out.cpu().numpy().reshape(-1, 4)[:, 3]
    sh.invalidate(spheres)
    sh.query(dp.data_ptr(), len(s), cam, out.data_ptr())
    after = out.cpu().numpy().reshape(-1, 4)[:, 3]
    p = s["pos"].astype(np.float64)
    dist = np.min(np.linalg.norm(p[:, None, :] - spheres[None, :, :3], axis=2) - spheres[None, :, 3], axis=1)
    cheb = np.min(np.abs(p[:, None, :] - spheres[None, :, :3]).max(axis=2) - spheres[None, :, 3], axis=1)
    inside, far = dist < -0.05, cheb > 2.0 + 0.1  # far: outside each sphere's box widened by the largest cell
    assert before[inside].all() and inside.sum() > 30
    assert not after[inside].any()
    assert np.array_equal(after[far], before[far]) and far.sum() > 1000
    sh.close()
