"""The fast-path window test of the tap loop is sound (CPU property test).

The half-pass drops per-tap masking for a wave when `window_inside`
(hemiphoto3d-private_amd/csrc/pm_pixel.h) holds on every lane: it bounds the
window's projective coordinates in closed form, with margins for FP32
rounding.  Here the same test is evaluated in FP32 (FMAs emulated exactly
enough: the product of two floats is exact in float64) on random windows --
ordinary ones, windows straddling the image edges, grazing ones whose terms
nearly cancel -- and every window it accepts is checked tap by tap the way
the tap loop computes it: two FMAs per coordinate, a reciprocal off by up to
one ulp either way (v_rcp_f32), a rounded product, floor.  Every tap of an
accepted window must be valid (|w_z| >= 1e-8, 0 <= floor(u) <= W-2,
0 <= floor(v) <= H-2); rejected windows are merely sent to the masked loop.
The same holds with the local tap coordinates of windows of up to 5x5 taps
(an integer base taken out of the centre and columns in FP64).  The same
holds for `z_clear`, which lets the FP16-storage masked loop drop
its per-tap |w_z| >= 1e-8 test: every window it accepts, of w_z terms that
nearly cancel at the guard, has |w_z| >= 1e-8 at every tap.
"""
import numpy as np

f32 = np.float32
EPS = 2.0 ** -21
SHRINK = f32(1.0 - 2.0 ** -20)


def fma(a, b, c):
    d = np.float64
    return (np.asarray(a, d) * np.asarray(b, d) + np.asarray(c, d)).astype(f32)


def window_inside(c, h, R, W, H):
    """FP32 restatement of window_inside; c (n,3), h (n,3,2) float32."""
    cx, cy, cz = c[:, 0], c[:, 1], c[:, 2]
    h00, h01, h10, h11, h20, h21 = (h[:, 0, 0], h[:, 0, 1], h[:, 1, 0], h[:, 1, 1],
                                    h[:, 2, 0], h[:, 2, 1])
    R = f32(R)
    Wq, Hq = f32(f32(W - 1) * SHRINK), f32(f32(H - 1) * SHRINK)
    ax = (np.abs(h00) + np.abs(h01)).astype(f32)
    ay = (np.abs(h10) + np.abs(h11)).astype(f32)
    az = (np.abs(h20) + np.abs(h21)).astype(f32)
    Tx, Ty, Tz = fma(R, ax, np.abs(cx)), fma(R, ay, np.abs(cy)), fma(R, az, np.abs(cz))
    lox, loy, loz = fma(-R, ax, cx), fma(-R, ay, cy), fma(-R, az, cz)
    gx = fma(R, (np.abs(fma(-Wq, h20, h00)) + np.abs(fma(-Wq, h21, h01))).astype(f32),
             fma(-Wq, cz, cx))
    gy = fma(R, (np.abs(fma(-Hq, h20, h10)) + np.abs(fma(-Hq, h21, h11))).astype(f32),
             fma(-Hq, cz, cy))
    e = f32(EPS)
    return ((fma(-e, Tx, lox) >= 0) & (fma(-e, Ty, loy) >= 0) & (fma(-e, Tz, loz) >= f32(1e-8))
            & (fma(e, fma(Wq, Tz, Tx), gx) < 0) & (fma(e, fma(Hq, Tz, Ty), gy) < 0))


def taps_valid(c, h, NT, ST, W, H):
    """Every tap of every window, computed as the tap loop does, with the
    reciprocal taken 1 ulp low and 1 ulp high (the worst cases of v_rcp_f32)."""
    R = (NT - 1) * ST // 2
    ok = np.ones(len(c), bool)
    for j in range(NT):
        oy = f32(j * ST - R)
        b = [fma(h[:, k, 1], oy, c[:, k]) for k in range(3)]
        for i in range(NT):
            ox = f32(i * ST - R)
            wx, wy, wz = (fma(h[:, k, 0], ox, b[k]) for k in range(3))
            with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
                r = (f32(1) / wz).astype(f32)
                for rr in (np.nextafter(r, f32(-np.inf)), np.nextafter(r, f32(np.inf))):
                    u = (wx * rr).astype(f32)
                    v = (wy * rr).astype(f32)
                    ix, iy = np.floor(u), np.floor(v)
                    ok &= (np.abs(wz) >= f32(1e-8)) & (ix >= 0) & (ix <= W - 2) & \
                          (iy >= 0) & (iy <= H - 2)
    return ok


def _windows(rng, n, W, H, kind):
    """Centre w(x,y) and tap-step columns of H for n windows."""
    z = rng.uniform(0.5, 2.0, n)
    if kind == "edge":      # centres within a few px of the image border
        u = np.where(rng.random(n) < 0.5, rng.uniform(-3, 8, n), rng.uniform(W - 9, W + 2, n))
        v = rng.uniform(-3, H + 2, n)
    elif kind == "tight":   # the outermost tap within 1e-6 .. 1e-2 px of u = 0 or W-1
        d = 10.0 ** rng.uniform(-6, -2, n) * rng.choice([-1.0, 1.0], n)
        Rw = 5.0
        u = np.where(rng.random(n) < 0.5, Rw + d, (W - 1) - Rw - d)
        v = rng.uniform(Rw + 1, H - Rw - 2, n)
    else:
        u, v = rng.uniform(0, W - 1, n), rng.uniform(0, H - 1, n)
    c = np.stack([u * z, v * z, z], 1)
    h = np.empty((n, 3, 2))
    scale = rng.uniform(0.5, 2.0, (n, 1))
    h[:, 0, 0] = z * scale[:, 0] + rng.normal(0, 0.05, n)
    h[:, 1, 1] = z * scale[:, 0] + rng.normal(0, 0.05, n)
    h[:, 0, 1] = rng.normal(0, 0.1, n) * z
    h[:, 1, 0] = rng.normal(0, 0.1, n) * z
    h[:, 2, :] = rng.normal(0, 1e-3, (n, 2)) * z[:, None]
    if kind == "tight":     # a pure scaling by z: the outermost tap is at u +- 5
        h[:, 0, 0], h[:, 1, 1] = z, z
        h[:, 0, 1] = h[:, 1, 0] = h[:, 2, 0] = h[:, 2, 1] = 0.0
    if kind == "grazing":
        # w_z nearly vanishes at one side of the window (a plane seen almost
        # edge-on), with w_x, w_y = (u, v) w_z + a small affine part: large
        # terms that cancel, mapping close to (u, v) or far off
        R = 5.0
        f = 1.0 - 10.0 ** rng.uniform(-6, -1, n)
        cz = z * 10.0 ** rng.uniform(0, 3, n)
        hz = -rng.choice([-1.0, 1.0], n) * cz / R * f
        c[:, 2], h[:, 2, 0], h[:, 2, 1] = cz, hz, 0.0
        for k, (p, q) in enumerate(((u, (0, 0)), (v, (1, 1)))):
            c[:, k] = p * cz + rng.normal(0, 1, n)
            h[:, k, 0] = p * hz + (z if k == 0 else 0) * rng.uniform(0.5, 2, n)
            h[:, k, 1] = (z if k == 1 else 0) * rng.uniform(0.5, 2, n)
    return c.astype(f32), h.astype(f32)


def _accepted_invalid(rng, eps=None, shrink=None):
    global EPS, SHRINK
    saved = EPS, SHRINK
    if eps is not None:
        EPS, SHRINK = eps, shrink
    try:
        bad = 0
        for W, H in ((1600, 1200), (4000, 3000)):
            for kind in ("edge", "tight"):
                c, h = _windows(rng, 20000, W, H, kind)
                ins = window_inside(c, h, 5, W, H)
                bad += int((~taps_valid(c[ins], h[ins], 11, 1, W, H)).sum())
        return bad
    finally:
        EPS, SHRINK = saved


def test_property_check_detects_missing_margins():
    """The check has teeth: without the rounding margins (2^-21 and the
    1 - 2^-20 shrink) the same test accepts windows with invalid taps."""
    assert _accepted_invalid(np.random.default_rng(5), 0.0, f32(1.0)) > 100
    assert _accepted_invalid(np.random.default_rng(5)) == 0


def test_accepted_windows_have_only_valid_taps():
    rng = np.random.default_rng(11)
    accepted = 0
    for W, H in ((1600, 1200), (97, 61), (4000, 3000)):
        for NT, ST in ((6, 2), (11, 1), (3, 1)):
            R = (NT - 1) * ST // 2
            for kind in ("interior", "edge", "grazing", "tight"):
                if kind == "tight" and R != 5:
                    continue
                c, h = _windows(rng, 20000, W, H, kind)
                ins = window_inside(c, h, R, W, H)
                if ins.any():
                    ok = taps_valid(c[ins], h[ins], NT, ST, W, H)
                    assert ok.all(), (W, H, NT, ST, kind, int((~ok).sum()))
                accepted += int(ins.sum())
    assert accepted > 100000    # the test is not vacuous


def test_margin_is_small_for_ordinary_windows():
    """Windows well inside pass, and the test gives up at most a few
    thousandths of a pixel at the edge (an identity warp at the boundary)."""
    W, H, R = 1600, 1200, 5
    # taps at x-5 .. x+5 are valid for x in [5, 1594); the test keeps about
    # 1e-5 px at the low edge and 3e-3 px at the high edge
    for x, y, want in ((800.0, 600.0, True), (5.01, 600.0, True), (5.0, 600.0, False),
                       (1593.99, 600.0, True), (1593.998, 600.0, False), (1594.0, 600.0, False)):
        c = np.array([[x, y, 1.0]], f32)
        h = np.array([[[1, 0], [0, 1], [0, 0]]], f32)
        assert bool(window_inside(c, h, R, W, H)[0]) == want, (x, want)


def test_non_finite_windows_are_rejected():
    c = np.array([[np.nan, 10, 1], [10, 10, np.inf], [10, 10, -1]], f32)
    h = np.tile(np.array([[1, 0], [0, 1], [0, 0]], f32), (3, 1, 1))
    with np.errstate(invalid="ignore"):
        assert not window_inside(c, h, 5, 100, 100).any()


# ---------------------------------------------------------------------------
# Windows of up to 5x5 taps compute their taps relative to an integer base
# (bu, bv) near the centre's source position: the base comes out of the
# centre and the columns in FP64, the taps run on u - bu in FP32, and the
# base is added back to floor(u - bu).  window_inside is still evaluated on
# the absolute centre and columns; it must stay sound for that arithmetic.
# ---------------------------------------------------------------------------
def taps_valid_local(c, h, NT, ST, W, H, dbase=0):
    """Every tap as the tap loop computes it with local coordinates (c, h
    taken as the exact FP64 values the kernel rounds); dbase moves the base
    by one either way, as a reciprocal 1 ulp off could."""
    R = (NT - 1) * ST // 2
    C, Hm = c.astype(np.float64), h.astype(np.float64)
    with np.errstate(all="ignore"):
        rc = (f32(1) / c[:, 2]).astype(f32)
        pu, pv = (c[:, 0] * rc).astype(f32), (c[:, 1] * rc).astype(f32)
        bu = np.where(np.abs(pu) < f32(4194304.0), np.rint(pu), 0).astype(np.int64) + dbase
        bv = np.where(np.abs(pv) < f32(4194304.0), np.rint(pv), 0).astype(np.int64) + dbase
    tx, ty = (C[:, 0] - bu * C[:, 2]).astype(f32), (C[:, 1] - bv * C[:, 2]).astype(f32)
    t00, t01 = (Hm[:, 0, 0] - bu * Hm[:, 2, 0]).astype(f32), (Hm[:, 0, 1] - bu * Hm[:, 2, 1]).astype(f32)
    t10, t11 = (Hm[:, 1, 0] - bv * Hm[:, 2, 0]).astype(f32), (Hm[:, 1, 1] - bv * Hm[:, 2, 1]).astype(f32)
    ok = np.ones(len(c), bool)
    for j in range(NT):
        oy = f32(j * ST - R)
        bx, by, bz = fma(t01, oy, tx), fma(t11, oy, ty), fma(h[:, 2, 1], oy, c[:, 2])
        for i in range(NT):
            ox = f32(i * ST - R)
            wx, wy, wz = fma(t00, ox, bx), fma(t10, ox, by), fma(h[:, 2, 0], ox, bz)
            with np.errstate(all="ignore"):
                r = (f32(1) / wz).astype(f32)
                for rr in (np.nextafter(r, f32(-np.inf)), np.nextafter(r, f32(np.inf))):
                    ix = bu + np.floor((wx * rr).astype(f32))
                    iy = bv + np.floor((wy * rr).astype(f32))
                    ok &= (np.abs(wz) >= f32(1e-8)) & (ix >= 0) & (ix <= W - 2) & \
                          (iy >= 0) & (iy <= H - 2)
    return ok


def _tight_windows(rng, n, W, H, R):
    """Identity-scaled windows whose outermost tap is 1e-7 .. 1e-2 px from an
    image border, in u or in v."""
    d = 10.0 ** rng.uniform(-7, -2, n) * rng.choice([-1.0, 1.0], n)
    low = rng.random(n) < 0.5
    z = rng.uniform(0.5, 2.0, n)
    t_u = np.where(low, R + d, (W - 1) - R - d)
    t_v = np.where(low, R + d, (H - 1) - R - d)
    in_v = rng.random(n) < 0.5
    u = np.where(in_v, rng.uniform(R + 1, W - R - 2, n), t_u)
    v = np.where(in_v, t_v, rng.uniform(R + 1, H - R - 2, n))
    c = np.stack([u * z, v * z, z], 1)
    h = np.zeros((n, 3, 2))
    h[:, 0, 0] = h[:, 1, 1] = z
    return c.astype(f32), h.astype(f32)


def test_accepted_windows_have_only_valid_taps_with_local_coordinates():
    rng = np.random.default_rng(3)
    accepted = 0
    for W, H in ((1600, 1200), (97, 61), (4000, 3000), (23, 17)):
        for NT, ST in ((3, 1), (5, 1), (2, 2), (3, 2), (4, 2), (5, 2)):
            R = (NT - 1) * ST // 2
            if 2 * R + 3 > min(W, H):
                continue
            for kind in ("interior", "edge", "grazing", "tight"):
                c, h = (_tight_windows(rng, 10000, W, H, R) if kind == "tight"
                        else _windows(rng, 10000, W, H, kind))
                ins = window_inside(c, h, R, W, H)
                accepted += int(ins.sum())
                for db in (0, -1, 1):
                    ok = taps_valid_local(c[ins], h[ins], NT, ST, W, H, db)
                    assert ok.all(), (W, H, NT, ST, kind, db, int((~ok).sum()))
    assert accepted > 300000    # the test is not vacuous


def test_local_coordinate_check_detects_missing_margins():
    """Without window_inside's rounding margins, windows a sub-pixel fraction
    from the border are accepted with taps the local arithmetic puts outside."""
    global EPS, SHRINK
    saved = EPS, SHRINK
    EPS, SHRINK = 0.0, f32(1.0)
    try:
        rng = np.random.default_rng(3)
        bad = 0
        for W, H in ((1600, 1200), (4000, 3000)):
            for NT, ST in ((3, 1), (5, 1), (4, 2)):
                R = (NT - 1) * ST // 2
                c, h = _tight_windows(rng, 20000, W, H, R)
                ins = window_inside(c, h, R, W, H)
                bad += int((~taps_valid_local(c[ins], h[ins], NT, ST, W, H)).sum())
        assert bad > 100
    finally:
        EPS, SHRINK = saved


# ---------------------------------------------------------------------------
# z_clear: the FP16-storage masked loop drops its per-tap |w_z| >= 1e-8 test
# when this closed-form bound holds on every lane.
# ---------------------------------------------------------------------------
def z_clear(c, h, R, eps=EPS):
    """FP32 restatement of z_clear (pm_pixel.h)."""
    cz, h20, h21 = c[:, 2], h[:, 2, 0], h[:, 2, 1]
    R = f32(R)
    e = f32(eps)
    az = (np.abs(h20) + np.abs(h21)).astype(f32)
    Tz = fma(R, az, np.abs(cz))
    return (fma(-e, Tz, fma(-R, az, cz)) >= f32(1e-8)) | (fma(e, Tz, fma(R, az, cz)) <= f32(-1e-8))


def taps_wz_clear(c, h, NT, ST):
    """|w_z| >= 1e-8 for every tap, evaluated as the tap loop does."""
    R = (NT - 1) * ST // 2
    ok = np.ones(len(c), bool)
    for j in range(NT):
        bz = fma(h[:, 2, 1], f32(j * ST - R), c[:, 2])
        for i in range(NT):
            ok &= np.abs(fma(h[:, 2, 0], f32(i * ST - R), bz)) >= f32(1e-8)
    return ok


def _z_windows(rng, n):
    """w_z near the 1e-8 guard somewhere in the window, of either sign, with
    terms from 1e-8 to 1e3 that nearly cancel."""
    R = 5.0
    mag = 10.0 ** rng.uniform(-8, 3, n)
    cz = mag * rng.choice([-1.0, 1.0], n)
    # the extreme tap lands within a factor (1 +- 1e-6 .. 1e-1) of the guard
    f = 1.0 + rng.choice([-1.0, 1.0], n) * 10.0 ** rng.uniform(-7, -1, n)
    target = np.sign(cz) * 1e-8 * f
    a = (cz - target) / R                      # |h20| + |h21| that puts the extreme at target
    split = rng.uniform(0, 1, n)
    h = np.zeros((n, 3, 2))
    h[:, 2, 0] = a * split * rng.choice([-1.0, 1.0], n)
    h[:, 2, 1] = a * (1 - split) * rng.choice([-1.0, 1.0], n)
    c = np.zeros((n, 3))
    c[:, 2] = cz
    return c.astype(f32), h.astype(f32)


def test_z_clear_windows_have_no_tap_near_the_guard():
    rng = np.random.default_rng(21)
    accepted = rejected_fine = 0
    for NT, ST in ((6, 2), (11, 1), (3, 1)):
        R = (NT - 1) * ST // 2
        c, h = _z_windows(rng, 200000)
        clear = z_clear(c, h, R)
        fine = taps_wz_clear(c, h, NT, ST)
        assert fine[clear].all(), (NT, ST, int((~fine[clear]).sum()))
        accepted += int(clear.sum())
        rejected_fine += int((fine & ~clear).sum())
    assert accepted > 100000                     # not vacuous
    print("z_clear: accepted", accepted, "rejected although clear", rejected_fine)


def test_z_clear_check_detects_a_missing_margin():
    """Without the 2^-21 margin the same bound accepts windows with a tap
    below the guard."""
    rng = np.random.default_rng(21)
    bad = 0
    for NT, ST in ((6, 2), (11, 1)):
        R = (NT - 1) * ST // 2
        c, h = _z_windows(rng, 200000)
        clear = z_clear(c, h, R, eps=0.0)
        bad += int((~taps_wz_clear(c, h, NT, ST)[clear]).sum())
    assert bad > 0
