"""Request server: every request's result is the same, bit for bit, whether it
runs alone (the single-batch entry points) or in the served mix (batched
rounds, overlapping rounds, per-request policy, capped rounds); admission
never holds more than the memory budget; views kept on the device hold their
memory until released; trace replay completes every request."""
import time

import numpy as np
import pytest

from rkifield_amd import layouts as L
from rkifield_amd import march as M
from rkifield_amd import scenes

pytestmark = pytest.mark.gpu

PLANES = ("position", "normal", "material", "motion")


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def three_scenes(ctx):
    out = []
    for seed in (21, 22, 23):
        prims = scenes.primitive_field(seed=seed, n=40, half_cube=3.0, size_range=(0.3, 1.5), group=4)
        lo, hi = scenes.prim_bounds(prims, 0.5)
        out.append(M.SparseScene.build(ctx, prims, 0.05, lo, hi, smin_k=0.1, mat_seed=scenes.material_seed(seed)))
    yield out
    for s in out:
        s.destroy()


def _ext(shadows):
    light = np.array([0.5, 1.0, 0.3])
    return L.scene_ext(hit_eps_mode=L.EPS_RELATIVE, shadow_enable=1 if shadows else 0,
                       light_dir=light / np.linalg.norm(light))


class Mix:
    """A mixed request set with pinned host buffers and its alone results."""

    def __init__(self, ctx, sc, seed=7):
        rng = np.random.default_rng(seed)
        self.items = []  # (kind, scene index, params, host buffers)
        for w, h, s, sh, keep in ((72, 40, 0, True, True), (33, 57, 1, False, False), (64, 64, 2, True, False),
                                  (128, 96, 1, True, True), (1, 1, 0, True, False)):
            cam = M.camera_look_at(rng.uniform(4, 7, 3) * rng.choice([-1, 1], 3), (0, 0, 0), 60.0, w, h)
            host = None
            if not keep:
                n = w * h
                host = dict(position=M.PinnedArray((n, 4), np.float32), normal=M.PinnedArray((n, 4), np.uint16),
                            material=M.PinnedArray((n,), np.uint32), motion=M.PinnedArray((n, 4), np.float32))
                if sh:
                    host["shadow"] = M.PinnedArray((n,), np.float32)
            self.items.append(("view", s, dict(cam=cam, su=L.scene_uniforms(num_objects=1), ext=_ext(sh), w=w, h=h,
                                                keep=keep), host))
        for s, n in ((0, 3000), (1, 65), (2, 20000), (1, 1)):
            x = M.PinnedArray((n, 3), np.float32)
            x.a[:] = rng.uniform(-5, 5, (n, 3))
            self.items.append(("points", s, dict(n=n), dict(xyz=x, out=M.PinnedArray((n,), L.POINT_RESULT))))
        for s, n in ((1, 1500), (2, 40)):
            c = M.PinnedArray((n,), L.SPHERE_CAST)
            c.a["origin"] = rng.uniform(-5, 5, (n, 3))
            d = rng.normal(size=(n, 3))
            c.a["dir"] = d / np.linalg.norm(d, axis=1, keepdims=True)
            c.a["radius"] = rng.uniform(0.02, 0.6, n)
            c.a["max_t"] = 10.0
            self.items.append(("casts", s, dict(n=n), dict(casts=c, out=M.PinnedArray((n,), L.CAST_RESULT))))
        for s, n in ((0, 1000), (2, 77)):
            o = M.PinnedArray((n, 3), np.float32)
            o.a[:] = rng.uniform(-4, 4, (n, 3))
            self.items.append(("shadow", s, dict(n=n, ext=_ext(True)), dict(origins=o, vis=M.PinnedArray((n,), np.float32))))
        self.alone = [self._alone(ctx, sc, it) for it in self.items]

    @staticmethod
    def _alone(ctx, sc, it):
        kind, s, p, host = it
        if kind == "view":
            gb = M.GBuffer(ctx, p["w"], p["h"], shadow=bool(p["ext"]["shadow_enable"]))
            sc[s].march(p["cam"], p["su"], p["ext"], gb, variant=0)
            r = gb.read()
            return r
        if kind == "points":
            xd = ctx.to_device(host["xyz"].a)
            out = ctx.array((p["n"],), L.POINT_RESULT)
            sc[s].query(xd, out, p["n"])
            return out.download()
        if kind == "casts":
            cd = ctx.to_device(host["casts"].a)
            out = ctx.array((p["n"],), L.CAST_RESULT)
            sc[s].sphere_cast(cd, out, p["n"])
            return out.download()
        od = ctx.to_device(host["origins"].a)
        vd = ctx.array((p["n"],), np.float32)
        b = M.Batch(ctx)
        b.add_shadow(sc[s], od, vd, p["n"], p["ext"])
        b.run()
        return vd.download()

    def request(self, k):
        kind, s, p, host = self.items[k]
        if kind == "view":
            return M.Server.view_request(s, p["cam"], p["su"], p["ext"], p["w"], p["h"], keep_on_device=p["keep"],
                                         host=host, tag=k)
        if kind == "points":
            return M.Server.points_request(s, host["xyz"], host["out"], tag=k)
        if kind == "casts":
            return M.Server.cast_request(s, host["casts"], host["out"], tag=k)
        return M.Server.shadow_request(s, host["origins"], host["vis"], p["ext"], tag=k)

    def clear(self):
        """Poisons every output buffer (inputs are left alone)."""
        for _, _, _, host in self.items:
            for name, arr in (host or {}).items():
                if name not in ("xyz", "origins", "casts"):
                    arr.a.view(np.uint8)[...] = 0xAB

    def check(self, srv, k, rid):
        kind, s, p, host = self.items[k]
        ref = self.alone[k]
        if kind == "view":
            got = srv.read_view(rid) if p["keep"] else {name: arr.a for name, arr in host.items()}
            names = PLANES + (("shadow",) if p["ext"]["shadow_enable"] else ())
            for name in names:
                assert np.array_equal(got[name].view(np.uint8), ref[name].view(np.uint8)), (k, name)
        elif kind in ("points", "casts"):
            assert np.array_equal(host["out"].a.view(np.uint8), ref.view(np.uint8)), k
        else:
            assert np.array_equal(host["vis"].a.view(np.uint32), ref.view(np.uint32)), k


def _serve_all(srv, mix, order, split=None):
    """Submits mix requests in `order` (the part after `split` once the first
    part is launched) and checks every result."""
    mix.clear()
    ids = {}
    first = order if split is None else order[:split]
    for k in first:
        ids[k] = srv.submit(mix.request(k))
    if split is not None:
        t0 = time.time()
        while all(srv.status(i)["state"] < L.REQ_RUNNING for i in ids.values()) and time.time() - t0 < 10:
            time.sleep(0.0005)
        for k in order[split:]:
            ids[k] = srv.submit(mix.request(k))
    rounds = []
    for k, rid in ids.items():
        st = srv.wait(rid, timeout=60.0)
        assert st["state"] == L.REQ_DONE, (k, st)
        assert st["tag"] == k
        assert st["t_submit"] <= st["t_admit"] <= st["t_launch"] <= st["t_done"]
        rounds.append((int(st["round"]), int(st["round_requests"])))
        mix.check(srv, k, rid)
        srv.release(rid)
    return rounds


def test_served_equals_alone(ctx, three_scenes):
    mix = Mix(ctx, three_scenes)
    assert sum(int((r["position"][:, 3] != L.MAX_FLOAT).sum()) for r in mix.alone if isinstance(r, dict)) > 500
    order = list(range(len(mix.items)))
    for name, kw, split in (("batched, 2 rounds in flight", dict(rounds_in_flight=2), 4),
                            ("batched, 1 round", dict(rounds_in_flight=1), None),
                            ("batched, capped rounds", dict(rounds_in_flight=3, max_round_work=5000), 3),
                            ("persistent schedule", dict(tuning=(L.SCHEDULE_PERSISTENT, 16, 4, 8)), 5),
                            ("per request", dict(policy=L.SERVE_PER_REQUEST), None)):
        _say(f"case: server {name}")
        srv = M.Server(ctx, three_scenes, memory_budget=64 << 20, **kw)
        try:
            rounds = _serve_all(srv, mix, order[::-1] if split == 3 else order, split)
            st = srv.stats()
            assert st["completed"] == len(order) and st["failed"] == 0
            assert st["arena_in_use"] == 0 and st["arena_peak"] <= 64 << 20
            if kw.get("policy") == L.SERVE_PER_REQUEST:
                assert all(n == 1 for _, n in rounds) and st["rounds"] == len(order)
            if "max_round_work" in kw:
                assert st["rounds"] >= 3
        finally:
            srv.close()


def test_queued_requests_join_one_round(ctx, three_scenes):
    """Requests queued behind a held view are admitted together when its
    memory is released and run in one round."""
    mix = Mix(ctx, three_scenes, seed=8)
    gate = M.Server.view_request(0, mix.items[0][2]["cam"], L.scene_uniforms(), _ext(False), 256, 256)
    budget = 256 * 256 * (16 + 8 + 4 + 16)
    _say("case: server gated round")
    srv = M.Server(ctx, three_scenes, memory_budget=budget)
    try:
        g = srv.submit(gate)
        assert srv.wait(g, timeout=60.0)["state"] == L.REQ_DONE
        mix.clear()
        ids = [srv.submit(mix.request(k)) for k in range(len(mix.items))]
        assert all(srv.status(i)["state"] == L.REQ_QUEUED for i in ids)
        srv.release(g)
        for k, rid in enumerate(ids):
            st = srv.wait(rid, timeout=60.0)
            assert st["state"] == L.REQ_DONE and st["round_requests"] == len(ids)
            mix.check(srv, k, rid)
            srv.release(rid)
        assert srv.stats()["rounds"] == 2
    finally:
        srv.close()


def test_memory_budget_admission(ctx, three_scenes):
    mix = Mix(ctx, three_scenes, seed=9)
    k20 = [k for k, it in enumerate(mix.items) if it[0] == "points" and it[2]["n"] == 20000][0]
    need = 20000 * 12 + 20000 * 24
    budget = 2 * need + need // 2  # two such batches at a time
    _say("case: server budget admission")
    srv = M.Server(ctx, three_scenes, memory_budget=budget, rounds_in_flight=2)
    try:
        ids = [srv.submit(mix.request(k20)) for _ in range(8)]
        sts = [srv.wait(i, timeout=60.0) for i in ids]
        assert all(s["state"] == L.REQ_DONE for s in sts)
        for i in ids:
            mix.check(srv, k20, i)  # identical inputs -> identical outputs, last writer included
            srv.release(i)
        st = srv.stats()
        assert st["arena_peak"] <= budget and st["arena_peak"] >= need
        # admitted in arrival order; never more than two held at once
        admits = sorted((s["t_admit"], s["t_done"]) for s in sts)
        for j in range(2, len(admits)):
            assert admits[j][0] >= min(a[1] for a in admits[:j]) - 1e-9
        big = M.Server.view_request(0, mix.items[0][2]["cam"], L.scene_uniforms(), _ext(True), 1920, 1080)
        with pytest.raises(M.RkfMarchError) as e:
            srv.submit(big)
        assert e.value.code == M.ERR["OOM"]
    finally:
        srv.close()


def test_keep_on_device_holds_memory(ctx, three_scenes):
    mix = Mix(ctx, three_scenes, seed=11)
    kv = 3  # 128x96 view kept on the device
    assert mix.items[kv][0] == "view" and mix.items[kv][2]["keep"]
    need = 128 * 96 * (16 + 8 + 4 + 16 + 4) + 5 * 256
    _say("case: server keep-on-device release")
    srv = M.Server(ctx, three_scenes, memory_budget=need + need // 2)
    try:
        a = srv.submit(mix.request(kv))
        assert srv.wait(a, timeout=60.0)["state"] == L.REQ_DONE
        b = srv.submit(mix.request(kv))
        time.sleep(0.05)
        assert srv.status(b)["state"] == L.REQ_QUEUED  # a still holds the memory
        mix.check(srv, kv, a)
        srv.release(a)
        assert srv.wait(b, timeout=60.0)["state"] == L.REQ_DONE
        mix.check(srv, kv, b)
        srv.release(b)
        assert srv.stats()["arena_in_use"] == 0
    finally:
        srv.close()


def test_replay_and_errors(ctx, three_scenes):
    mix = Mix(ctx, three_scenes, seed=13)
    rng = np.random.default_rng(3)
    n = 120
    picks = rng.integers(0, len(mix.items), n)
    reqs = np.array([mix.request(int(k)) for k in picks], L.REQUEST)
    arrivals = np.cumsum(rng.exponential(1.0 / 2000.0, n))
    _say("case: server replay")
    srv = M.Server(ctx, three_scenes, memory_budget=64 << 20)
    try:
        lat, wall = srv.replay(reqs, arrivals)
        assert np.all(np.isfinite(lat)) and np.all(lat > 0), lat
        assert wall >= arrivals[-1]
        st = srv.stats()
        assert st["completed"] == n and st["arena_in_use"] == 0
        bad = mix.request(0)
        bad["scene"] = 3
        with pytest.raises(M.RkfMarchError):
            srv.submit(bad)
        bad = mix.request(0)
        bad["kind"] = 9
        with pytest.raises(M.RkfMarchError):
            srv.submit(bad)
        bad = mix.request(len(mix.items) - 1)
        bad["in"] = 0
        with pytest.raises(M.RkfMarchError):
            srv.submit(bad)
        with pytest.raises(M.RkfMarchError):
            srv.release(123456789)
        # a view asking for per-tile ray intervals is refused before admission
        # and does not fail the requests submitted around it
        mix.clear()
        ids = [srv.submit(mix.request(0))]
        bad = mix.request(2)
        bad["ext"]["march_flags"] = L.MARCH_TILE_CULL
        with pytest.raises(M.RkfMarchError) as e:
            srv.submit(bad)
        assert e.value.code == M.ERR["INVALID_ARG"]
        ids += [srv.submit(mix.request(2)), srv.submit(mix.request(5))]
        for k, rid in zip((0, 2, 5), ids):
            st = srv.wait(rid, timeout=60.0)
            assert st["state"] == L.REQ_DONE, (k, st)
            mix.check(srv, k, rid)
            srv.release(rid)
    finally:
        srv.close()
