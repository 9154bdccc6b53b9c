"""Pins of the oracle's shading, lighting, texture, sky and ray-generation functions against what
the mathematics or the reference's own definitions fix, independently of how the oracle computes
them: closed forms (null and camera-slice emission times, aberration of the rest-frame cosines,
inverse-square-like attenuation), special cases (hands and notches of the clock dial at chosen
proper times, the light-clock photon at chosen phases, checker parity of negative cells, sky grid
lines), library routines (colorsys for the HSV map) and a second formulation (the bloom as
products of 1D filter matrices). CPU only.

Every value below is computed here from the reference's definitions in renderer.rs, bloom.rs and
minkowski.rs (cited per test), in float64, and compared within float32 rounding (1e-6 relative
unless stated); the inputs are chosen away from every branch boundary."""
import colorsys
import math

import numpy as np
import pytest

from specrel_amd import scene as S
from oracle import pyoracle as O

TOL = 1e-6

# common_materials (scene.rs:181-280) entries used here: base, checker, emission, grid, wireframe
CHECKER = 0  # (0.22, 0.25, 0.30), (0.13, 0.15, 0.19), 0.18, 1.0, 0.012
CLOCK = 1    # (0.20, 0.72, 0.95), (0.10, 0.25, 0.45), ClockFace, grid 1
LIGHT_CLOCK = 5  # (0.20, 0.95, 0.65), LightClock, cavity 1.0
CHECKER2 = 6  # (0.85, 0.82, 0.72), (0.22, 0.52, 0.62), grid 2
BASE0, CHK0 = np.array([0.22, 0.25, 0.30]), np.array([0.13, 0.15, 0.19])
BASE1, CHK1 = np.array([0.20, 0.72, 0.95]), np.array([0.10, 0.25, 0.45])
BASE5 = np.array([0.20, 0.95, 0.65])

LIGHT = np.array([3.0, 4.0, 0.0])
EVENT = np.array([10.0, 0.0, 0.0, 0.0])  # the hit: light at spatial distance 5
TO_LIGHT_NULL = np.array([-1.0, 0.6, 0.8, 0.0])  # (emission - hit) / 5 on the past light cone
VIEW_DIR = np.array([-1.0, 0.0, -0.6, -0.8])  # a past-directed null ray arriving at the hit


def close(a, b, tol=TOL, what=""):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    err = np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-3))
    assert err <= tol, f"{what}: {a} vs {b} (rel {err:.3g})"


def occluder_at(b, p, r=0.5):
    """A tetrahedron of the hyperplane through p spanned by t, z and the spatial direction (0.8,
    -0.6, 0): a ray through p along a direction with spatial part not in that span crosses it at p,
    its barycentric centre."""
    d1, d2, d3 = np.array([1.0, 0, 0, 0]), np.array([0, 0, 0, 1.0]), np.array([0, 0.8, -0.6, 0])
    verts = np.array([p + r * d1, p + r * d2, p + r * d3, p - r * (d1 + d2 + d3)], np.float32)
    assert b.add_simplex(verts, material=CHECKER)


def scene_with(points=()):
    b = S.Builder()
    b.set_light(*LIGHT)
    for p in points:
        occluder_at(b, np.asarray(p, np.float64))
    if not points:  # a scene needs a tree: a tetrahedron far from everything
        occluder_at(b, np.array([0.0, 500.0, 500.0, 500.0]))
    return b.finish()


def frame(mode=S.RelativisticShaded, shadow=S.ShadowLorentzInvariant, view=S.PastLightCone, e0=None, dyn=None,
          w=64, h=40):
    cam = S.Camera4D.new((0.0, 0.0, 0.0, 0.0))
    fr = S.build_frame(cam, S.RenderConfig(w, h, mode, shadow, view), dyn)
    if e0 is not None:
        for i in range(4):
            fr.e[0][i] = e0[i]
    return fr


def shade(scene, fr, **kw):
    direction = kw.pop("direction", VIEW_DIR)
    return O.shade_record(scene, fr, direction, O.record(**kw))


# ---------------------------------------------------------------- lighting (renderer.rs:585-653)

def test_null_shadow_emission_time_cosines_and_attenuation():
    sc = scene_with()
    col, p = shade(sc, frame(), event=EVENT, lam=0.0, material=CHECKER)
    assert p["t_emit"] == 5.0 and p["dist"] == 5.0  # t_emit = t_hit - |x_L - x_hit|
    close(p["cos_light"], 0.8, what="cos light")  # spatial (0.6, 0.8, 0) against normal +y
    close(p["cos_view"], 0.6, what="cos view")  # -dir points back along (0, 0.6, 0.8)
    atten = 1.0 / (1.0 + 0.006 * 25.0)
    lighting = 0.18 + 0.30 * 0.6 + 0.80 * 1.0 * 0.8 * atten
    close(p["lighting"], lighting, what="lighting")
    close(p["albedo"], BASE0, what="albedo")
    close(col, BASE0 * lighting, what="colour (no fog at lambda 0)")
    assert p["light_shadowed"] == 0


def test_null_shadow_ray_is_blocked_on_the_light_cone_and_factor_is_0_32():
    # an occluder half way along the null segment to the emission event
    sc = scene_with([EVENT + 2.5 * TO_LIGHT_NULL])
    _, p = shade(sc, frame(), event=EVENT, lam=0.0, material=CHECKER)
    assert p["light_shadowed"] == 1
    close(p["lighting"], 0.18 + 0.18 + 0.80 * 0.32 * 0.8 / 1.15, what="shadowed lighting")
    # shadows off: no shadow ray at all; back-facing: no shadow ray and no light term
    _, q = shade(sc, frame(shadow=S.ShadowOff), event=EVENT, lam=0.0, material=CHECKER)
    assert q["light_shadowed"] == 0
    close(q["lighting"], 0.18 + 0.18 + 0.80 * 0.8 / 1.15, what="shadows off")
    _, r = shade(sc, frame(), event=EVENT, lam=0.0, material=CHECKER, normal=(0, 0, -1, 0))
    assert r["light_shadowed"] == 0 and r["cos_light"] == 0.0 and r["cos_view"] == 0.0
    close(r["lighting"], 0.18, what="back-facing")


def test_camera_slice_emission_time_uses_the_camera_velocity():
    # e0 = (1.25, 0.75, 0, 0): beta = (0.6, 0, 0); t_emit = t_hit + beta . (x_L - x_hit) = 10 + 1.8
    e0 = (1.25, 0.75, 0.0, 0.0)
    slice_dir = np.array([1.8, 3.0, 4.0, 0.0]) / 5.0
    on_slice = scene_with([EVENT + 2.5 * slice_dir])
    on_cone = scene_with([EVENT + 2.5 * TO_LIGHT_NULL])
    fr = frame(shadow=S.ShadowCameraSlice, view=S.CameraSimultaneity, e0=e0)
    _, p = shade(on_slice, fr, event=EVENT, lam=0.0, material=CHECKER)
    close(p["t_emit"], 11.8, what="slice t_emit")
    close(p["cos_light"], 0.8, what="slice cos light")  # the same spatial direction at rest
    assert p["light_shadowed"] == 1
    _, q = shade(on_cone, fr, event=EVENT, lam=0.0, material=CHECKER)
    assert q["light_shadowed"] == 0  # the null segment's occluder is off the slice segment
    _, r = shade(on_slice, frame(), event=EVENT, lam=0.0, material=CHECKER)
    assert r["light_shadowed"] == 0  # and the slice segment's occluder off the light cone


def test_rest_frame_cosines_follow_aberration_for_a_moving_surface():
    # surface moving at beta = 0.5 along x, normal +y (orthogonal to u): the rest-frame cosine of a
    # null direction k = (-1, k_hat) against a transverse normal is k_y / (gamma (1 + beta k_x))
    beta = 0.5
    g = 1.0 / math.sqrt(1.0 - beta * beta)
    u = (g, g * beta, 0.0, 0.0)
    sc = scene_with()
    _, p = shade(sc, frame(), event=EVENT, lam=0.0, material=CHECKER, four_velocity=u)
    close(p["cos_light"], 0.8 / (g * (1.0 + beta * 0.6)), 2e-6, "aberrated cos light")
    # the view ray: -dir has spatial part (0, 0.6, 0.8) and dir.t = -1: k = -dir is future-pointing
    close(p["cos_view"], 0.6 / (g * (1.0 - beta * 0.0)), 2e-6, "aberrated cos view")


def test_clock_materials_are_self_lit():
    sc = scene_with()
    _, p = shade(sc, frame(), event=EVENT, lam=0.0, material=CLOCK, uv=(0.05, 0.3))
    close(p["lighting"], 0.72 + 0.38 * 0.6 + 0.22 * 0.8, what="clock lighting")
    assert p["light_shadowed"] == 0


# ---------------------------------------------------------------- flashes (renderer.rs:658-688)

def flash_frame(t_at, width=2.0, intensity=2.0, color=(1.0, 0.5, 0.25), shadow=S.ShadowLorentzInvariant):
    dyn = S.FrameDynamics(flashes=[S.FlashEvent((t_at, *LIGHT), width, intensity, color)])
    return frame(shadow=shadow, dyn=dyn)


def test_flash_on_the_past_cone_pulse_attenuation_and_shadow():
    color = np.array([1.0, 0.5, 0.25])
    atten = 1.0 / (1.0 + 0.015 * 25.0)
    sc = scene_with()
    # the flash 5 s before the hit at spatial distance 5: on the hit's past light cone, pulse 1
    _, p = shade(sc, flash_frame(5.0), event=EVENT, lam=0.0, material=CHECKER)
    close(p["flash"], color * (2.0 * 1.0 * atten * (0.25 + 0.75 * 0.8)), what="flash on the cone")
    # half a half-width off the shell: k = 1/2, pulse 1/4
    _, q = shade(sc, flash_frame(5.5), event=EVENT, lam=0.0, material=CHECKER)
    close(q["flash"], color * (2.0 * 0.25 * atten * (0.25 + 0.75 * 0.8)), what="flash off the shell")
    # outside the shell, and a flash after the hit: nothing
    _, r = shade(sc, flash_frame(7.0), event=EVENT, lam=0.0, material=CHECKER)
    _, s = shade(sc, flash_frame(12.0), event=EVENT, lam=0.0, material=CHECKER)
    assert not r["flash"].any() and not s["flash"].any()
    # an occluder on the null segment to the flash: shadow factor 0.25
    blocked = scene_with([EVENT + 2.5 * TO_LIGHT_NULL])
    col, t = shade(blocked, flash_frame(5.0), event=EVENT, lam=0.0, material=CHECKER)
    close(t["flash"], color * (2.0 * atten * (0.25 + 0.75 * 0.25 * 0.8)), what="shadowed flash")
    # composition: albedo * lighting + (0.65 albedo + 0.35) * flash (renderer.rs:427-434)
    close(col, BASE0 * t["lighting"] + (0.65 * BASE0 + 0.35) * t["flash"], what="lit with flash")


# ---------------------------------------------------------------- surface textures (renderer.rs:544-580)

@pytest.mark.parametrize("uv,mat,expect", [
    ((0.5, 0.5), CHECKER, BASE0), ((1.5, 0.5), CHECKER, CHK0), ((-0.3, 0.2), CHECKER, CHK0),
    ((-0.3, -0.2), CHECKER, BASE0), ((-1.7, 0.9), CHECKER, BASE0), ((2.2, -0.1), CHECKER, CHK0),
    ((0.3, 0.2), CHECKER2, np.array([0.85, 0.82, 0.72])), ((0.6, 0.2), CHECKER2, np.array([0.22, 0.52, 0.62])),
])
def test_checker_parity_uses_floor_and_euclidean_remainder(uv, mat, expect):
    _, p = shade(scene_with(), frame(), event=EVENT, lam=0.0, material=mat, uv=uv)
    close(p["albedo"], expect, what=f"checker at {uv}")


def test_wireframe_blend():
    sc = scene_with()
    for mtb, t in ((0.006, 0.5), (0.0, 0.0), (0.003, 0.25)):
        _, p = shade(sc, frame(), event=EVENT, lam=0.0, material=CHECKER, tri_bary=(0.5, mtb, 0.5 - mtb))
        close(p["albedo"], 0.07 * (1.0 - t) + BASE0 * t, what=f"wire at {mtb}")
    _, p = shade(sc, frame(), event=EVENT, lam=0.0, material=CHECKER, tri_bary=(0.4, 0.013, 0.587))
    close(p["albedo"], BASE0, what="outside the band")


# ---------------------------------------------------------------- clock face (renderer.rs:737-943)

CYAN, YELLOW, RED = np.array([0.25, 0.98, 1.0]), np.array([1.0, 0.95, 0.30]), np.array([1.0, 0.35, 0.25])
DIAL = np.array([0.04, 0.06, 0.10])


def clock_albedo(uv, tau, grad=(0.0, 0.0), uvws=(0.0, 0.0)):
    _, p = shade(scene_with(), frame(), event=EVENT, lam=0.0, material=CLOCK, uv=uv, tau=tau, tau_uv_grad=grad,
                 uv_world_scale=uvws)
    return p["albedo"]


def test_dial_hands_at_known_proper_times():
    right = (0.6, 0.57)  # 0.1 to the right of the pivot (0.5, 0.57): clockwise angle pi/2 from the top
    close(clock_albedo(right, 0.25), CYAN, what="fast hand (1 s per turn) at a quarter second")
    close(clock_albedo(right, 2.5), YELLOW, what="main hand (10 s per turn) at 2.5 s")
    close(clock_albedo(right, 25.0), RED, what="slow hand (100 s per turn) at 25 s")
    close(clock_albedo(right, 0.75), DIAL, what="no hand, outside the swept sector")
    close(clock_albedo(right, 5.0), DIAL + BASE1 * 0.22, what="inside the main hand's swept sector")


def test_dial_notches_and_rings():
    top, bottom = (0.5, 0.82), (0.5, 0.32)
    a = 3 * 2 * math.pi / 10  # notch 3, clockwise from the top
    third = (0.5 + 0.25 * math.sin(a), 0.57 + 0.25 * math.cos(a))
    close(clock_albedo(top, 0.0), np.array([1.0, 0.88, 0.15]), what="gold notch 0")
    close(clock_albedo(bottom, 0.0), np.array([1.0, 1.0, 1.0]), what="white notch 5")
    close(clock_albedo(third, 0.0), np.array([0.75, 0.80, 0.88]), what="notch 3")
    close(clock_albedo((0.81, 0.57), 0.0), BASE1 * 1.15, what="bezel ring")
    close(clock_albedo((0.05, 0.3), 0.0), CHK1 * 0.78, what="casing, odd quarter cell")
    close(clock_albedo((0.3, 0.3), 0.0), BASE1 * 0.78, what="casing, even quarter cell")
    close(clock_albedo((0.01, 0.5), 0.0), BASE1 * 0.78 * 0.35, what="tile border")


def test_dial_phase_is_retarded_by_the_rest_frame_distance_to_the_pivot():
    # uv_world_scale 2.5: the point 0.1 from the pivot is 0.25 light-seconds away, so a local proper
    # time of 0.5 shows the dial phase of 0.25 (fast hand at a quarter turn)
    close(clock_albedo((0.6, 0.57), 0.5, uvws=(2.5, 2.5)), CYAN, what="retarded fast hand")
    assert not np.allclose(clock_albedo((0.6, 0.57), 0.5), CYAN)


def test_corner_clock_shows_its_own_proper_time():
    # 0.04 right of the (0.865, 0.865) corner clock's centre, proper-time gradient 2.5 per tile along
    # u: the corner oscillator runs 0.04 * 2.5 = 0.1 s behind the local proper time 0.35, i.e. at
    # 0.25: its fast hand points at the sampled point
    close(clock_albedo((0.905, 0.865), 0.35, grad=(2.5, 0.0)), CYAN, what="corner fast hand")
    assert not np.allclose(clock_albedo((0.905, 0.865), 0.35), CYAN)


# ---------------------------------------------------------------- light clock (renderer.rs:983-1052)

def light_clock_albedo(uv, tau):
    _, p = shade(scene_with(), frame(), event=EVENT, lam=0.0, material=LIGHT_CLOCK, uv=uv, tau=tau)
    return p["albedo"]


def test_light_clock_photon_bounces_with_period_2L():
    bg, guide = np.array([0.05, 0.09, 0.13]), np.array([0.06, 0.16, 0.16])
    core = bg + guide + BASE5 * 0.9 + np.array([1.0, 1.0, 0.85]) * 1.4  # photon at the sampled point
    for tau in (0.5, 1.5, 2.5, -1.5):  # a quarter period (L = 1, period 2): the photon mid-cavity
        close(light_clock_albedo((0.5, 0.5), tau), core, what=f"photon at v = 0.5, tau {tau}")
    close(light_clock_albedo((0.5, 0.3), 0.25), core, what="photon at v = 0.3")
    close(light_clock_albedo((0.5, 0.5), 1.0), bg + guide, what="photon at the top mirror")
    close(light_clock_albedo((0.5, 0.07), 0.0), np.array([0.75, 0.90, 0.98]), what="mirror plate")
    close(light_clock_albedo((0.5, 0.02), 0.0), np.array([0.08, 0.12, 0.16]) * 0.35, what="outside, tile border")
    close(light_clock_albedo((0.05, 0.5), 0.0), BASE5 * 0.65, what="side frame")
    close(light_clock_albedo((0.15, 0.2), 1.0), bg + np.array([0.20, 0.35, 0.35]), what="ruler tick")


# ---------------------------------------------------------------- sky and fog (renderer.rs:691-695, 1056-1079)

ZENITH, HORIZON, GRID = np.array([0.04, 0.06, 0.13]), np.array([0.12, 0.16, 0.25]), np.array([0.06, 0.09, 0.14])


def test_sky_gradient_end_points():
    close(O.sky_gradient((-1.0, 0.0, 3.0, 0.0)), ZENITH, what="straight up (unnormalised)")
    close(O.sky_gradient((-1.0, 0.0, -1.0, 0.0)), HORIZON, what="straight down")
    close(O.sky_gradient((-1.0, 1.0, 0.0, 0.0)), (ZENITH + HORIZON) / 2, what="level")


def sky_dir(theta, phi):
    return np.array([-1.0, math.sin(theta) * math.cos(phi), math.cos(theta), math.sin(theta) * math.sin(phi)])


@pytest.mark.parametrize("theta,phi,line", [
    (math.pi / 12, math.pi / 24, False), (math.pi / 12, 5 * math.pi / 12 + 0.002, True),
    (math.pi / 12, 5 * math.pi / 12 - 0.002, False), (math.pi / 6 + 0.003, math.pi / 24, True),
    (3 * math.pi / 4, -7 * math.pi / 24, False), (3 * math.pi / 4, -math.pi / 4 + 0.001, True),
])
def test_sky_grid_lines_every_15_and_30_degrees(theta, phi, line):
    d = sky_dir(theta, phi)
    out = np.zeros(3, np.float32)
    dd = np.ascontiguousarray(d, np.float32)
    O.lib().oracle_shade_sky(dd.ctypes.data_as(O._FP), out.ctypes.data_as(O._FP))
    base = O.sky_gradient(d)
    close(out, base + GRID if line else base, what=f"sky at theta {theta:.4f} phi {phi:.4f}")


def test_distance_fog_is_exponential_in_lambda():
    c = np.array([0.9, 0.5, 0.1])
    d = (-1.0, 0.0, 1.0, 0.0)  # straight up: the sky gradient is the zenith colour
    close(O.distance_fog(c, d, 0.0), c, what="no fog at lambda 0")
    f = math.exp(-1.0)  # lambda 125: exp(-0.008 * 125)
    close(O.distance_fog(c, d, 125.0), c * f + ZENITH * (1 - f), what="one e-fold")
    f = math.exp(-1.44)
    close(O.distance_fog(c, d, 180.0), c * f + ZENITH * (1 - f), what="at the ray's end")


# ---------------------------------------------------------------- other modes (renderer.rs:437-472)

@pytest.mark.parametrize("lam", [0.0, 2.5, 5.0, 7.5, 1.2])
def test_isochrone_colour_is_the_hsv_map_of_lambda(lam):
    col, _ = shade(scene_with(), frame(mode=S.LookbackIsochrones), event=(0.5, 0, 0, 0), lam=lam, tau=0.5)
    phase = (lam * 0.12) % 1.0
    close(col, colorsys.hsv_to_rgb(phase * 300.0 / 360.0, 0.8, 0.85), 2e-6, f"isochrone base at lambda {lam}")


def test_isochrone_contours():
    fr = frame(mode=S.LookbackIsochrones)
    for t in (3.02, -0.98, 6.97):  # t mod 1 within 0.05 of an integer: white ring
        col, _ = shade(scene_with(), fr, event=(t, 0, 0, 0), lam=1.0, tau=0.5)
        close(col, np.ones(3), what=f"t ring at {t}")
    for tau in (7.995, -2.98):  # tau mod 1 within 0.04 of an integer: dark ring
        col, _ = shade(scene_with(), fr, event=(0.5, 0, 0, 0), lam=1.0, tau=tau)
        close(col, np.full(3, 0.1), what=f"tau ring at {tau}")


def test_inspector_tint_and_edges():
    fr = frame(mode=S.Simplex4DInspector)
    b = (0.1, 0.2, 0.3, 0.4)
    col, _ = shade(scene_with(), fr, event=EVENT, lam=1.0, tet_bary=b)
    close(col, (0.25 + 0.65 * 0.2, 0.25 + 0.65 * 0.3, 0.35 + 0.55 * 0.4), what="tint at rest")
    col, _ = shade(scene_with(), fr, event=EVENT, lam=1.0, tet_bary=(0.02, 0.2, 0.3, 0.48))
    close(col, np.array([0.25 + 0.13, 0.25 + 0.195, 0.35 + 0.264]) * 0.12, what="near a facet")
    g = 1.5  # speed boost (gamma - 1) * 0.8 = 0.4
    col, _ = shade(scene_with(), fr, event=EVENT, lam=1.0, tet_bary=b, four_velocity=(g, math.sqrt(g * g - 1), 0, 0))
    close(col, (0.25 + 0.13 + 0.12, 0.25 + 0.195, 0.35 + 0.55 * 0.4 * 0.8), what="tint when moving")


# ---------------------------------------------------------------- ray generation (renderer.rs:334-384)

def boosted_frame(view, w=321, h=201):
    cam = S.Camera4D.new((1.0, 2.0, -3.0, 0.5), 65.0)
    cam.tetrad.boost_local((0.3, -0.2, 0.6))
    cam.tetrad.rotate_local(0.4, -0.1, 0.2)
    return S.build_frame(cam, S.RenderConfig(w, h, S.RelativisticShaded, S.ShadowOff, view))


def mdot(a, b):
    return -a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3]


@pytest.mark.parametrize("view", [S.PastLightCone, S.CameraSimultaneity, S.Frontier, S.ProphecySlice])
def test_pixel_rays_against_the_camera_geometry(view):
    fr = boosted_frame(view)
    e = np.array([[fr.e[a][i] for i in range(4)] for a in range(4)], np.float64)
    w, h = 321, 201
    half_h = math.tan(math.radians(65.0) / 2)
    half_w = half_h * w / h
    close(fr.half_wh[1], half_h, what="half height = tan(fov / 2)")
    close(fr.half_wh[0], half_w, what="half width = aspect * half height")
    cam = np.array([fr.cam_pos[i] for i in range(4)], np.float64)
    for px, py in ((160, 100), (0, 0), (320, 200), (0, 200), (37, 151)):
        sx = (2 * (px + 0.5) / w - 1) * half_w
        sy = (1 - 2 * (py + 0.5) / h) * half_h
        n = np.array([sx, sy, 1.0]) / math.sqrt(sx * sx + sy * sy + 1.0)
        o, d = O.pixel_ray(fr, px, py)
        if view == S.PastLightCone:
            want = -e[0] + n @ e[1:]
        elif view == S.Frontier:
            want = e[0] + n @ e[1:]
        else:
            want = n @ e[1:]
        close(d, want, 3e-6, f"ray of pixel ({px}, {py})")
        if view in (S.PastLightCone, S.Frontier):
            assert abs(mdot(d.astype(np.float64), d.astype(np.float64))) < 1e-5  # null
            assert (d[0] < 0) == (view == S.PastLightCone)  # past or future pointing
        if view == S.Frontier:
            close(o, cam - e[0] * fr.frontier_view_offset, what="frontier origin")
        elif view == S.ProphecySlice:
            close(o, cam + e[0] * fr.prophecy_delta, what="prophecy origin")
        else:
            close(o, cam, what="origin")
    # the centre pixel of an odd-sized view looks straight along e3
    _, d = O.pixel_ray(fr, 160, 100)
    if view == S.CameraSimultaneity:
        close(d, e[3], 1e-6, "centre ray")


# ---------------------------------------------------------------- bloom (bloom.rs:22-125)

def down_matrix(n):
    """1D [1 3 3 1] / 8 downsample of n samples into ceil(n / 2), taps 2x - 1 .. 2x + 2 clamped."""
    m = (n + 1) // 2
    D = np.zeros((m, n))
    for x in range(m):
        for i, wgt in enumerate((0.125, 0.375, 0.375, 0.125)):
            D[x, min(max(2 * x - 1 + i, 0), n - 1)] += wgt
    return D


def up_matrix(n_fine, n_coarse):
    """1D bilinear 2x upsample: fine sample i takes coarse i // 2 at 3/4 and its neighbour on i's
    side (clamped) at 1/4."""
    U = np.zeros((n_fine, n_coarse))
    for i in range(n_fine):
        m = i // 2
        other = min(max(m - 1 if i % 2 == 0 else m + 1, 0), n_coarse - 1)
        U[i, m] += 0.75
        U[i, other] += 0.25
    return U


def bloom_by_matrices(img):
    """out = 0.96 img + (0.04 / 6) sum_k Up_0<-1 ... Up_k-1<-k Down_k<-k-1 ... Down_1<-0 img, each 2D
    step the product of its row and column matrices."""
    h, w = img.shape
    ws, hs = [w], [h]
    for _ in range(6):
        ws.append((ws[-1] + 1) // 2)
        hs.append((hs[-1] + 1) // 2)
    out = 0.96 * img
    for k in range(1, 7):
        Ly, Lx = np.eye(h), np.eye(w)
        for j in range(1, k + 1):  # down to level k
            Ly, Lx = down_matrix(hs[j - 1]) @ Ly, down_matrix(ws[j - 1]) @ Lx
        for j in range(k, 0, -1):  # back up to level 0
            Ly, Lx = up_matrix(hs[j - 1], hs[j]) @ Ly, up_matrix(ws[j - 1], ws[j]) @ Lx
        out = out + (0.04 / 6) * (Ly @ img @ Lx.T)
    return out


@pytest.mark.parametrize("w,h,points", [(37, 23, [(11, 7)]), (64, 40, [(0, 0), (63, 39)]), (97, 64, [(40, 30)]),
                                        (5, 3, [(2, 1)])])
def test_bloom_delta_response_equals_the_filter_products(w, h, points):
    img = np.zeros((h, w), np.float64)
    for x, y in points:
        img[y, x] = 1000.0
    got = O.bloom(np.repeat(img[:, :, None], 3, axis=2).astype(np.float32))
    want = bloom_by_matrices(img)
    for c in range(3):
        err = np.abs(got[:, :, c] - want).max()
        assert err <= 1e-6 * 1000.0, f"{w}x{h}: max error {err}"
    if w * h >= 512:  # energy within 1% (bloom.rs:131-143); clamped borders lose more on tiny images
        assert np.abs(got[:, :, 0].sum() - img.sum()) <= 1e-2 * img.sum()
