"""GPU tests of the geometric consistency filter (hp3d_pm_geom_filter).

Parity: bit-exact with the oracle's FP32 restatement of
geometric_consistency_filter_kernel (patchmatch_mvs.cu:668-781) applied view
by view in order, in place, as the reference's driver does (:1267-1384).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def view_set(pkg, scenes):
    """Five views on an arc with GT maps; depth perturbed so that some pixels
    disagree, the last view at a different size (skipped as a source)."""
    rng = np.random.default_rng(5)
    W, H, f = 240, 180, 225.0
    K = scenes.pinhole(f, W, H)
    scene = scenes.sphere_scene(8, f)
    maps = []
    for k, a in enumerate((-6.0, -3.0, 0.0, 3.0, 6.0)):
        R, t = scenes.arc_camera(a)
        w, h = (W, H) if k < 4 else (200, 150)
        KK = K if k < 4 else scenes.pinhole(f, w, h)
        _, gd, gn = pkg.render(scene, scenes._camera(KK, R, t, w, h))
        d = (gd * rng.choice([1.0, 1.0, 1.0, 1.3], size=gd.shape)).astype(np.float32)
        d[rng.random(gd.shape) < 0.02] = 0.0
        maps.append({"depth": d, "normal": gn, "confidence": rng.random(gd.shape).astype(np.float32),
                     "K": KK, "R": R, "t": t})
    srcs = [[1, 2], [0, 2, 3], [1, 3, 4], [2, 1], [3]]
    for m, s in zip(maps, srcs):
        m["src"] = s
    return maps


def oracle_filter(orc, maps, tol=0.12, min_consistent=2):
    out = [{k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in m.items()} for m in maps]
    for m in out:
        srcs = [out[s] for s in m["src"] if out[s]["depth"].shape == m["depth"].shape]
        if len(srcs) < min_consistent:
            continue
        orc.geom_filter_fp32(m["depth"], m["normal"], m["confidence"], np.linalg.inv(m["K"]),
                             m["R"], m["t"],
                             [(s["K"], s["R"], s["t"], s["depth"], s["normal"]) for s in srcs],
                             depth_tol=tol, normal_tol_cos=np.cos(np.float32(np.pi)),
                             min_consistent=min_consistent)
    return out


def test_filter_bit_exact_with_oracle_in_order(pkg, orc, ctx, view_set):
    gpu = [{k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in m.items()} for m in view_set]
    retained = ctx.geom_filter(gpu)
    ref = oracle_filter(orc, view_set)
    for i, (g, r) in enumerate(zip(gpu, ref)):
        assert np.array_equal(g["depth"], r["depth"]), i
        assert np.array_equal(g["confidence"], r["confidence"]), i
    assert retained[4] == -1                       # one usable source only: left as is
    assert all(0 < retained[i] < view_set[i]["depth"].size for i in range(4))
    # perturbed pixels are dropped, and zeroed pixels stay zero with zero confidence
    for g in gpu[:4]:
        assert np.all(g["confidence"][g["depth"] == 0] == 0)


def test_filter_order_dependence(pkg, orc, ctx, view_set):
    """Reversing the order changes results exactly as the oracle predicts."""
    order = [3, 2, 1, 0]
    sub = [view_set[i] for i in order]
    remap = {old: new for new, old in enumerate(order)}
    rev = []
    for m in sub:
        mm = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in m.items()}
        mm["src"] = [remap[s] for s in m["src"] if s in remap]
        rev.append(mm)
    expect = oracle_filter(orc, rev)
    ctx.geom_filter(rev)
    for g, r in zip(rev, expect):
        assert np.array_equal(g["depth"], r["depth"])


def test_filter_refuses_self_and_repeated_sources(pkg, ctx, view_set):
    """A view filtered in place cannot also be read as its own source, and a
    source listed twice would be counted twice: both are refused before any
    map is touched."""
    for bad in ([0, 1], [1, 1]):
        maps = [{k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in m.items()}
                for m in view_set[:4]]
        maps[0]["src"] = bad
        before = maps[0]["depth"].copy()
        with pytest.raises(pkg.HP3DError) as e:
            ctx.geom_filter(maps)
        assert e.value.code == pkg.ERR_INVALID_ARG, bad
        assert np.array_equal(maps[0]["depth"], before)
