"""Closed-form pins of the oracle functions that the GPU-vs-oracle comparisons cannot pin on their
own, because both sides were written from the same reading of the reference:

* the bounce loop with Russian roulette (trace_path, src/pathtracer.cu:295-331);
* the occluded branch of the shadow test (traverse_occluded, src/pathtracer.cu:162-191, as
  shade_direct calls it at :243-250);
* the G-buffer encodings (gbuffer_kernel, src/pathtracer.cu:366-411).

Each check derives its expectation from geometry in FP64, not from the oracle's formulas, and
each is shown to reject a build of the oracle with one planted mistake (ORC_MUTANT in
oracle/oracle.cpp), so a dropped 1/q weight, an inverted shadow test, a bounce origin pushed
through the surface or an unflipped normal would fail here."""
import numpy as np
import pytest

from oracle import oracle as orc


# ----------------------------------------------------------------------------- scene helpers
def icosphere(level, radius):
    """Vertices (V, 3) and faces (F, 3) of a subdivided icosahedron on a sphere of `radius`."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.asarray(p, np.float64) / np.linalg.norm(p) for p in v]
    for _ in range(level):
        mid = {}

        def m(a, b):
            k = (min(a, b), max(a, b))
            if k not in mid:
                p = verts[a] + verts[b]
                verts.append(p / np.linalg.norm(p))
                mid[k] = len(verts) - 1
            return mid[k]

        nf = []
        for a, b, c in f:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    return np.asarray(verts) * radius, np.asarray(f)


def camera(W, H, pos, tgt, fov=60.0):
    import lgi
    cam = lgi.Camera()
    cam.width, cam.height, cam.fov_y_deg, cam.aspect = W, H, fov, W / H
    lgi.camera_look_at(cam, pos, tgt)
    return cam


def sun_light(direction, intensity, color=(1.0, 0.9, 0.8)):
    import lgi
    sun = lgi.DirLight()
    sun.dir[:] = list(direction)
    sun.color[:] = list(color)
    sun.intensity = intensity
    return sun


def quad(x0, x1, y0, y1, z, normal_z, mat):
    t = np.zeros(2, orc.TRI_DTYPE)
    P = np.array([[x0, y0, z], [x1, y0, z], [x1, y1, z], [x0, y1, z]], np.float32)
    t["v"][0] = P[[0, 1, 2]]
    t["v"][1] = P[[0, 2, 3]]
    t["n"][:, :, 2] = normal_z
    t["material"] = mat
    return t


def pixel_rays(cam, jitter=None):
    """FP64 world directions of the renderer's rays (src/pathtracer.cu:284-293): pixel centres, or
    the first two draws of each pixel's sample-0 stream when jitter is True."""
    W, H = cam.width, cam.height
    th = np.float64(orc.tan_half_fov(cam.fov_y_deg))
    fwd, left, up = (np.asarray(cam.camera_to_world[4 * k:4 * k + 3], np.float64) for k in range(3))
    py, px = np.mgrid[0:H, 0:W]
    if jitter:
        u = np.stack([orc.pixel_randf(0, p, 0, 2) for p in range(W * H)]).astype(np.float64).reshape(H, W, 2)
        jx, jy = u[..., 0], u[..., 1]
    else:
        jx = jy = np.full((H, W), 0.5)
    ndcx = 2 * (px + jx) / W - 1
    ndcy = 1 - 2 * (py + jy) / H
    d = (fwd[None, None] - left[None, None] * (ndcx * cam.aspect * th)[..., None]
         + up[None, None] * (ndcy * th)[..., None])
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


# ----------------------------------------------------------------- (i) bounces and roulette
SPHERE_RHO = np.array([0.3, 0.5, 0.8])
SPHERE_I, SPHERE_R_LIGHT = 10.0, 2.0


def sphere_interior_scene():
    """Closed sphere of radius 1 (level-5 icosphere, radial vertex normals) with a point light of
    radius 2 at its centre and a sun outside it. The interpolated normal at a hit point P is P/|P|
    (normals are the vertices over the radius), so toward the light n.l = 1 at every hit, and the
    direct term is D(|P|) = rho I (1 - |P|/2)^2, with |P| between the faces' inradius and 1. The
    sun's shadow rays all end on the far side of the sphere."""
    V, F = icosphere(5, 1.0)
    t = np.zeros(len(F), orc.TRI_DTYPE)
    t["v"] = V[F].astype(np.float32)
    t["n"] = (V[F] / np.linalg.norm(V[F], axis=-1, keepdims=True)).astype(np.float32)
    pts = np.zeros(1, orc.POINT_DTYPE)
    pts[0] = ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), SPHERE_I, SPHERE_R_LIGHT)
    cam = camera(48, 32, (0.2, -0.1, 0.05), (1.0, 0.3, 0.2))
    sc = orc.Scene(t, np.array([SPHERE_RHO], np.float32), pts, sun_light((0.3, -0.2, -1.0), 3.0), cam)
    # the smallest distance from the centre to a face: to its plane, whose foot lies inside
    # these near-equilateral triangles
    a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    n = np.cross(b - a, c - a)
    inradius = np.min(np.abs(np.einsum("ij,ij->i", n, a)) / np.linalg.norm(n, axis=1))
    return sc, inradius


def check_sphere_bounce_mean(sc, inradius, spp=24):
    out = sc.render(spp, 2)
    m = out["sum64"].reshape(-1, 3) / spp            # per-pixel means, independent across pixels
    mean = m.mean(0)
    se = m.std(0, ddof=1) / np.sqrt(len(m))
    D = lambda d: SPHERE_RHO * SPHERE_I * (1 - d / SPHERE_R_LIGHT) ** 2
    series = 1 + SPHERE_RHO + SPHERE_RHO ** 2        # direct + one bounce + two bounces
    lo, hi = series * D(1.0), series * D(inradius)
    assert np.all(mean >= lo - 3 * se) and np.all(mean <= hi + 3 * se), (mean, lo, hi, se)
    return mean, se


def test_bounce_estimator_closed_form_inside_lit_sphere():
    sc, r_in = sphere_interior_scene()
    assert 0.9995 < r_in < 1.0
    mean, se = check_sphere_bounce_mean(sc, r_in)
    assert np.all(se / mean < 2e-3)                  # the test can tell a few per mille apart
    # without bounces the mean is the direct term alone (the sun never reaches the inside)
    out = sc.render(4, 0)
    m0 = out["sum64"].reshape(-1, 3).mean(0) / 4
    D = SPHERE_RHO * SPHERE_I * (1 - np.array([1.0, r_in])[:, None] / SPHERE_R_LIGHT) ** 2
    assert np.all(m0 >= D[0] * (1 - 1e-5)) and np.all(m0 <= D[1] * (1 + 1e-5))


# ------------------------------------------------------------------------- (ii) occlusion
def plate_scene(W=96, H=72, ground=50.0, plate_normal_z=1.0):
    tris = np.concatenate([quad(-ground, ground, -ground, ground, 0.0, 1.0, 0),
                           quad(-1.0, 1.0, -1.0, 1.0, 1.0, plate_normal_z, 1)])
    mats = np.array([[0.3, 0.6, 0.45], [0.7, 0.2, 0.55]], np.float32)
    cam = camera(W, H, (0.0, -4.0, 6.0), (0.0, 0.0, 0.0))
    sun = sun_light((0.3, -0.2, -1.0), 3.0)
    return orc.Scene(tris, mats, np.zeros(0, orc.POINT_DTYPE), sun, cam), cam, sun, mats


def check_plate_shadow(sc, cam, sun, mats, margin=2e-3):
    """Direct light, 1 spp: a ground point whose ray toward the sun crosses the plate gets exactly 0;
    every other visible point gets albedo * colour * I * cos(theta)."""
    L = sc.render(1, 0)["sum"]
    d = pixel_rays(cam, jitter=True)
    o = np.asarray(cam.pos[:], np.float64)
    to_sun = -np.asarray(sun.dir[:], np.float64)
    to_sun /= np.linalg.norm(to_sun)
    lit = np.asarray(sun.color[:], np.float64) * sun.intensity * to_sun[2]
    t_plate = (1.0 - o[2]) / d[..., 2]
    Pp = o + t_plate[..., None] * d
    on_plate = (np.abs(Pp[..., 0]) < 1 - margin) & (np.abs(Pp[..., 1]) < 1 - margin)
    off_plate = (np.abs(Pp[..., 0]) > 1 + margin) | (np.abs(Pp[..., 1]) > 1 + margin)
    Pg = o + (-o[2] / d[..., 2])[..., None] * d
    Q = Pg + ((1.0 - Pg[..., 2]) / to_sun[2])[..., None] * to_sun   # where the sun ray meets z = 1
    shadowed = off_plate & (np.abs(Q[..., 0]) < 1 - margin) & (np.abs(Q[..., 1]) < 1 - margin)
    sunny = off_plate & ((np.abs(Q[..., 0]) > 1 + margin) | (np.abs(Q[..., 1]) > 1 + margin))
    assert shadowed.sum() > 50 and sunny.sum() > 1000 and on_plate.sum() > 50
    assert np.all(L[shadowed] == 0.0), "ground inside the plate's shadow must receive no sun"
    assert np.allclose(L[sunny], mats[0] * lit, rtol=3e-6)
    assert np.allclose(L[on_plate], mats[1] * lit, rtol=3e-6)


def test_sun_shadow_of_a_plate_closed_form():
    check_plate_shadow(*plate_scene())


# ------------------------------------------------------------------------- (iii) G-buffer
def check_gbuffer_plate(sc, cam, mats, margin=2e-3):
    """Pixel-centre G-buffer of the plate over a 4 x 4 m ground patch (src/pathtracer.cu:366-411,
    host extents :728-734): misses are (0,0,0) / (128,128,255) / (0,0,0); hits encode albedo,
    the normal turned toward the camera as N/2 + 1/2, and (P - box min) / box extent, each as
    (u8)(clamp(v, 0, 1) * 255)."""
    alb, nrm, wp = sc.gbuffer()
    d = pixel_rays(cam)
    o = np.asarray(cam.pos[:], np.float64)
    Pp = o + ((1.0 - o[2]) / d[..., 2])[..., None] * d
    Pg = o + (-o[2] / d[..., 2])[..., None] * d
    in_sq = lambda P, h: (np.abs(P[..., 0]) < h - margin) & (np.abs(P[..., 1]) < h - margin)
    out_sq = lambda P, h: (np.abs(P[..., 0]) > h + margin) | (np.abs(P[..., 1]) > h + margin)
    plate = in_sq(Pp, 1.0)
    ground = out_sq(Pp, 1.0) & in_sq(Pg, 2.0)
    miss = out_sq(Pp, 1.0) & out_sq(Pg, 2.0)
    assert plate.sum() > 50 and ground.sum() > 200 and miss.sum() > 200
    enc = lambda v: np.floor(np.clip(v, 0, 1) * 255).astype(np.int64)
    # misses
    assert np.all(alb[miss] == 0) and np.all(wp[miss] == 0)
    assert np.all(nrm[miss] == [128, 128, 255])
    # both surfaces face +z toward a camera above them; 0.5 * 255 = 127.5 truncates to 127
    for region, mat in ((plate, 1), (ground, 0)):
        assert np.all(alb[region] == enc(mats[mat].astype(np.float32).astype(np.float64)))
        assert np.all(nrm[region] == [127, 127, 255])
    # world position over the scene box (-2, -2, 0) .. (2, 2, 1)
    for region, P in ((plate, Pp), (ground, Pg)):
        e = enc((P[region] - [-2.0, -2.0, 0.0]) / [4.0, 4.0, 1.0])
        diff = np.abs(wp[region].astype(np.int64) - e)
        assert diff.max() <= 1                       # FP32 vs FP64 at an integer boundary
        assert np.mean(diff == 0) > 0.97


def test_gbuffer_encodings_closed_form():
    # the plate's stored normals point down (-z): the encoding must turn them toward the camera
    sc, cam, sun, mats = plate_scene(W=80, H=60, ground=2.0, plate_normal_z=-1.0)
    check_gbuffer_plate(sc, cam, mats)


# --------------------------------------------------------------- the pins reject mistakes
@pytest.fixture(scope="module")
def mutants(tmp_path_factory):
    d = tmp_path_factory.mktemp("oracle_mutants")
    return {k: orc.build_mutant(k, d) for k in (1, 2, 3, 4)}


def test_pins_reject_planted_mistakes(mutants):
    sphere, r_in = sphere_interior_scene()
    plate = plate_scene()
    gplate = plate_scene(W=80, H=60, ground=2.0, plate_normal_z=-1.0)
    checks = {
        1: lambda: check_sphere_bounce_mean(sphere, r_in),          # roulette without 1/q
        2: lambda: check_plate_shadow(*plate),                      # shadow test inverted
        3: lambda: check_sphere_bounce_mean(sphere, r_in),          # bounce origin through the wall
        4: lambda: check_gbuffer_plate(gplate[0], gplate[1], gplate[3]),  # normal not flipped
    }
    for k, check in checks.items():
        with orc.library(mutants[k]):
            with pytest.raises(AssertionError):
                check()
    # and the unmutated oracle passes all of them
    for check in checks.values():
        check()
